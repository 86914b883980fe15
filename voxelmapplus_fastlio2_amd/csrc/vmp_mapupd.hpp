// Map update: VoxelMap::update / build (voxel_map.cpp:200-256) with
// VoxelGrid::pushPoint / updatePlane / merge (voxel_map.cpp:42-186), executed
// as a pipeline of data-parallel stages plus two short ordered replays.
//
//   K1 ins_keys      per point: fp32 insertion transform (PCL order), key,
//                    hash find-or-insert -> (hash slot, point index)
//   K2 sort          radix sort by hash slot (stable: point order kept)
//   K3 seg_heads     segment = one key's points in scan order
//   K4 seg_info      existing voxel of each segment; new keys flagged at their
//                    first point
//   K5 create        creation events in point order: parallel when the LRU
//                    cannot overflow, else one wave replays creations and
//                    evictions against the LRU queue (oldest stamp first,
//                    skipping voxels already touched earlier in the scan)
//   K6 replay        one lane per voxel walks its points: the pushPoint state
//                    machine; refits from centre-local moments (no point pool)
//   K7 candidates    merge() calls that can change state: the first call after
//                    each event that could make a previously idle call act
//   K8 merge replay  one wave evaluates those calls in global point order,
//                    adding the follow-up calls of every voxel a merge touched
//   K9 finalize      dirty bits, LRU queue append, hash tombstones, free list
// Every ordering decision (LRU order, victims, group ids, merge order) is
// therefore the reference's point order.
#pragma once
#include "vmp_common.hpp"
#include "vmp_observe.hpp"

namespace vmp {

constexpr int kMapThreads = 256;
constexpr int kHeapCap = 4096;
constexpr int32_t kNever = 0x7fffffff;

struct MapCounters {
  int64_t size;           // live voxels
  uint64_t group_counter; // VoxelGrid::count
  int64_t point_base;     // global point index of this scan's point 0
  int64_t ring_head, ring_tail;
  int32_t free_top, bump;
  int32_t used_entries;   // non-empty hash slots, tombstones included
  int32_t tombs;
  int32_t tag;            // scan tag
  int32_t error;          // bit 1 hash probe overflow, 2 range, 4 capacity, 8 pool, 16 heap
  // per-scan outputs
  int32_t nseg, ncreate_ev, ncre, nevict, ncand;
  int32_t refits, merge_calls, merge_evals, merges, dropped;
  int32_t fast;  // 1: this scan's creations took the parallel (no-eviction) path
  int32_t nlru;  // voxels appended to the LRU queue by this scan
  int32_t pad[2];
};

struct MapDev {
  HashEntry* table;
  uint32_t hmask;
  int32_t pool;
  VoxHot* hot;
  VoxState* vs;
  double* plane;      // [pool][kPlaneD]
  double* mom;        // [pool][kMomD]
  int32_t* freelist;  // [pool]
  int32_t* ring_vid;  // [qcap]
  uint64_t* ring_stamp;
  int64_t qcap;
  MapCounters* ctr;
  int32_t capacity;
  int32_t max_pt, upd_pt;
  double plane_thresh, merge_a, merge_d, vs64;
};

struct ScanBufs {
  const float* x;
  const float* y;
  const float* z;
  int32_t n;
  int32_t is_build;
  const UpdParams* prm;
  uint32_t* skey_in;   // hash slot per point
  int32_t* sval_in;    // point index
  uint32_t* skey_out;
  int32_t* sval_out;
  int32_t* flags;      // scratch flag / value array [n]
  int32_t* seg_start;  // [n]
  int32_t* seg_vid;    // [n]
  int32_t* seg_entry;  // [n]
  int32_t* cre_t;      // creation events: point index
  int32_t* cre_seg;    //                   segment
  int32_t* co_vid;     // resolved creations: vid, seg, t, prev, group lo
  int32_t* co_seg;
  int32_t* co_t;
  int32_t* co_prev;
  uint64_t* co_group;
  int32_t* ev_vid;     // evictions in order
  int32_t* ev_t;
  uint64_t* cand_in;   // merge candidates (t << 32 | vid)
  uint64_t* cand_out;
  int32_t* lru_sel;    // compaction output for the LRU append
};

// ------------------------------------------------------------ helpers -----
__device__ inline int32_t hash_find_or_insert(HashEntry* table, uint32_t mask, uint64_t key, MapCounters* ctr) {
  uint32_t h = hash_slot(key, mask);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    unsigned long long* kp = reinterpret_cast<unsigned long long*>(&table[h].key);
    unsigned long long k = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == key) return int32_t(h);
    if (k == kKeyEmpty) {
      unsigned long long old = atomicCAS(kp, (unsigned long long)kKeyEmpty, (unsigned long long)key);
      if (old == kKeyEmpty) {
        atomicAdd(&ctr->used_entries, 1);
        return int32_t(h);
      }
      if (old == key) return int32_t(h);
    }
    h = (h + 1) & mask;
  }
  atomicOr(&ctr->error, 1);
  return -1;
}

__device__ inline int32_t hash_find_slot(const HashEntry* table, uint32_t mask, uint64_t key) {
  uint32_t h = hash_slot(key, mask);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const uint64_t k = table[h].key;
    if (k == key) return int32_t(h);
    if (k == kKeyEmpty) return -1;
    h = (h + 1) & mask;
  }
  return -1;
}

// symmetric 3x3 index (xx xy xz yy yz zz)
__host__ __device__ constexpr int s6(int a, int b) {
  return a <= b ? (a == 0 ? b : (a == 1 ? 2 + b : 5)) : (b == 0 ? a : (b == 1 ? 2 + a : 5));
}

// Cyclic Jacobi on a symmetric 3x3 (same method and sweep order as the
// oracle's; eigenvalues ascending, eigenvectors in columns of V row-major).
__device__ inline void eigh3_dev(const double Ain[9], double ev[3], double V[9]) {
  double A[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i * 3 + j] = 0.5 * (Ain[i * 3 + j] + Ain[j * 3 + i]);
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  const int Pi[3] = {0, 0, 1}, Qi[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
    double dia = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
    if (off <= 1e-36 * dia || off == 0.0) break;
    for (int r = 0; r < 3; ++r) {
      const int p = Pi[r], q = Qi[r];
      const double apq = A[p * 3 + q];
      if (apq == 0.0) continue;
      const double theta = (A[q * 3 + q] - A[p * 3 + p]) / (2.0 * apq);
      const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      for (int k = 0; k < 3; ++k) {
        const double akp = A[k * 3 + p], akq = A[k * 3 + q];
        A[k * 3 + p] = c * akp - s * akq;
        A[k * 3 + q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; ++k) {
        const double apk = A[p * 3 + k], aqk = A[q * 3 + k];
        A[p * 3 + k] = c * apk - s * aqk;
        A[q * 3 + k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; ++k) {
        const double vkp = V[k * 3 + p], vkq = V[k * 3 + q];
        V[k * 3 + p] = c * vkp - s * vkq;
        V[k * 3 + q] = s * vkp + c * vkq;
      }
    }
  }
  int i0 = 0, i1 = 1, i2 = 2;
  const double d[3] = {A[0], A[4], A[8]};
  // stable 3-element sort by eigenvalue (ties keep index order)
  if (d[i1] < d[i0]) { int t = i0; i0 = i1; i1 = t; }
  if (d[i2] < d[i1]) { int t = i1; i1 = i2; i2 = t; }
  if (d[i1] < d[i0]) { int t = i0; i0 = i1; i1 = t; }
  const int id[3] = {i0, i1, i2};
  double Vs[9];
  for (int c = 0; c < 3; ++c) {
    ev[c] = d[id[c]];
    for (int r = 0; r < 3; ++r) Vs[r * 3 + c] = V[r * 3 + id[c]];
  }
  for (int i = 0; i < 9; ++i) V[i] = Vs[i];
}

__device__ inline void write_hot(const MapDev& m, int32_t vid, uint32_t flags) {
  const double* pl = m.plane + size_t(vid) * kPlaneD;
  int64_t kx, ky, kz;
  unpack_key(m.vs[vid].key, kx, ky, kz);
  VoxHot h;
  h.mx = float(pl[0] - double(kx) * m.vs64);
  h.my = float(pl[1] - double(ky) * m.vs64);
  h.mz = float(pl[2] - double(kz) * m.vs64);
  h.flags = flags;
  h.nx = float(pl[3]);
  h.ny = float(pl[4]);
  h.nz = float(pl[5]);
  h.pad = 0.f;
  m.hot[vid] = h;
}

__global__ void k_begin_scan(MapDev m) {
  MapCounters* c = m.ctr;
  c->tag += 1;
  c->fast = 0;
  c->nseg = c->ncreate_ev = c->ncre = c->nevict = c->ncand = 0;
  c->refits = c->merge_calls = c->merge_evals = c->merges = c->dropped = 0;
}
__global__ void k_fill_i32(int32_t* p, int32_t v, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------ K1 ----
__global__ void k_ins_keys(MapDev m, ScanBufs b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  const UpdParams& P = *b.prm;
  float pw[3];
  ins_transform(P.m, P.t, b.x[i], b.y[i], b.z[i], pw);
  int64_t k[3];
  bool ok = true;
  for (int r = 0; r < 3; ++r) {
    k[r] = int64_t(floor(double(pw[r]) / m.vs64));
    ok &= key_ok(k[r]);
  }
  int32_t e = -1;
  if (ok) e = hash_find_or_insert(m.table, m.hmask, pack_key(k[0], k[1], k[2]), m.ctr);
  else atomicOr(&m.ctr->error, 2);
  b.skey_in[i] = uint32_t(e < 0 ? 0 : e);
  b.sval_in[i] = i;
}

// ------------------------------------------------------------------ K3 ----
__global__ void k_seg_heads(ScanBufs b) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= b.n) return;
  b.flags[j] = (j == 0 || b.skey_out[j] != b.skey_out[j - 1]) ? 1 : 0;
}

// ------------------------------------------------------------------ K4 ----
__global__ void k_seg_info(MapDev m, ScanBufs b) {
  const int nseg = m.ctr->nseg;
  const int tag = m.ctr->tag;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
    const int start = b.seg_start[s];
    const uint32_t e = b.skey_out[start];
    const int first = b.sval_out[start];
    b.seg_entry[s] = int32_t(e);
    const int32_t vid = m.table[e].vid;
    b.seg_vid[s] = vid;
    if (vid >= 0) {
      VoxState& st = m.vs[vid];
      st.tag = tag;
      st.create_t = -1;
      st.evict_t = kNever;
      st.freeze_t = -1;
      st.prev_vid = -1;
      st.first_t = first;
      st.seg = s;
      st.merge_pos = kNever;
      st.dirty_t = -2;
    }
  }
}
// marks the first point of each new key with its segment (+1); the flag array
// was cleared beforehand
__global__ void k_mark_new(MapDev m, ScanBufs b) {
  const int nseg = m.ctr->nseg;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
    if (b.seg_vid[s] < 0) b.flags[b.sval_out[b.seg_start[s]]] = s + 1;
  }
}
__global__ void k_cre_split(MapDev m, ScanBufs b) {
  // lru_sel holds the compacted point indices of creation events
  const int nc = m.ctr->ncreate_ev;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nc; r += gridDim.x * blockDim.x) {
    const int t = b.lru_sel[r];
    b.cre_t[r] = t;
    b.cre_seg[r] = b.flags[t] - 1;
  }
}

// ------------------------------------------------------------------ K5 ----
// Parallel creation when no eviction can happen in this scan.
__global__ void k_create_fast(MapDev m, ScanBufs b) {
  const int nc = m.ctr->ncreate_ev;
  if (m.ctr->size + nc > m.capacity) return;
  const int ftop = m.ctr->free_top, bump = m.ctr->bump;
  const uint64_t gc = m.ctr->group_counter;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nc; r += gridDim.x * blockDim.x) {
    const int32_t vid = r < ftop ? m.freelist[ftop - 1 - r] : bump + (r - ftop);
    b.co_vid[r] = vid;
    b.co_seg[r] = b.cre_seg[r];
    b.co_t[r] = b.cre_t[r];
    b.co_prev[r] = -1;
    b.co_group[r] = gc + uint64_t(r);
  }
}
__global__ void k_create_fast_ctr(MapDev m) {
  MapCounters* c = m.ctr;
  const int nc = c->ncreate_ev;
  if (c->size + nc > m.capacity) return;
  const int ftop = c->free_top;
  if (nc <= ftop) c->free_top = ftop - nc;
  else {
    c->bump += nc - ftop;
    c->free_top = 0;
  }
  if (c->bump > m.pool) atomicOr(&c->error, 8);
  c->size += nc;
  c->group_counter += uint64_t(nc);
  c->ncre = nc;
  c->nevict = 0;
  c->fast = 1;
}

// Ordered creation + eviction (one wave).  LRU order is the descending order
// of last-touch stamps; the ring holds (vid, stamp) in ascending stamp order
// with stale entries skipped lazily.
__global__ void __launch_bounds__(64) k_create_evict(MapDev m, ScanBufs b) {
  MapCounters* c = m.ctr;
  const int ncre_ev = c->ncreate_ev;
  if (c->fast) return;
  __shared__ uint64_t heap[kHeapCap];
  const int lane = threadIdx.x;
  const int tag = c->tag;
  int hn = 0;
  int64_t size = c->size;
  uint64_t gc = c->group_counter;
  int ftop = c->free_top, bump = c->bump;
  int64_t ptr = c->ring_head;
  const int64_t tail = c->ring_tail;
  int64_t base = -1;       // first ring index held by the lanes
  int32_t l_vid = -1;
  int32_t l_first = kNever;  // first touch this scan (kNever if untouched)
  bool l_live = false;
  int ci = 0, nco = 0, nev = 0;
  int err = 0;
  while (true) {
    const int t_s = ci < ncre_ev ? b.cre_t[ci] : kNever;
    const int t_h = hn > 0 ? int(heap[0] >> 32) : kNever;
    if (t_s == kNever && t_h == kNever) break;
    int t, seg, prev;
    if (t_s < t_h) {
      t = t_s;
      seg = b.cre_seg[ci];
      prev = -1;
      ++ci;
    } else {
      const uint64_t top = heap[0];
      t = int(top >> 32);
      prev = int32_t(top & 0xffffffffu);  // evicted incarnation of this key
      seg = m.vs[prev].seg;
      __syncthreads();
      if (lane == 0) {  // pop
        const uint64_t last = heap[hn - 1];
        int i = 0;
        const int nn = hn - 1;
        while (true) {
          int l = 2 * i + 1, r = l + 1, s = i;
          uint64_t sv = last;
          if (l < nn && heap[l] < sv) { s = l; sv = heap[l]; }
          if (r < nn && heap[r] < sv) { s = r; sv = heap[r]; }
          if (s == i) break;
          heap[i] = heap[s];
          i = s;
        }
        if (nn > 0) heap[i] = last;
      }
      __syncthreads();
      --hn;
    }
    int32_t vid;
    if (ftop > 0) vid = m.freelist[--ftop];
    else vid = bump++;
    if (vid >= m.pool) { err |= 8; break; }
    if (lane == 0) {
      b.co_vid[nco] = vid;
      b.co_seg[nco] = seg;
      b.co_t[nco] = t;
      b.co_prev[nco] = prev;
      b.co_group[nco] = gc;
    }
    ++gc;
    ++nco;
    ++size;
    if (size > m.capacity) {
      int32_t victim = -1;
      while (victim < 0) {
        if (ptr >= tail) { err |= 4; break; }
        if (base < 0 || ptr >= base + 64) {
          base = ptr;
          const int64_t idx = base + lane;
          l_vid = -1;
          l_live = false;
          l_first = kNever;
          if (idx < tail) {
            const int64_t ph = idx % m.qcap;
            const int32_t v = m.ring_vid[ph];
            const uint64_t stp = m.ring_stamp[ph];
            const VoxState& st = m.vs[v];
            l_vid = v;
            l_live = (st.flags & kAlive) && st.stamp == stp && !(st.tag == tag && st.evict_t != kNever);
            l_first = (st.tag == tag && st.create_t < 0) ? st.first_t : kNever;
          }
        }
        const bool elig = (base + lane >= ptr) && (base + lane < tail) && l_live && (l_first > t);
        const unsigned long long bal = __ballot(elig);
        if (bal) {
          const int w = __ffsll(bal) - 1;
          victim = __shfl(l_vid, w, 64);
          ptr = base + w + 1;
          if (lane == w) l_live = false;
        } else {
          ptr = base + 64;
        }
      }
      if (victim < 0) break;
      VoxState& vst = m.vs[victim];
      const int32_t vfirst = (vst.tag == tag && vst.seg >= 0) ? vst.first_t : kNever;
      __syncthreads();
      if (lane == 0) {
        if (vst.tag != tag) {
          vst.tag = tag;
          vst.create_t = -1;
          vst.freeze_t = -1;
          vst.prev_vid = -1;
          vst.first_t = kNever;
          vst.seg = -1;
          vst.merge_pos = kNever;
          vst.dirty_t = -2;
        }
        vst.evict_t = t;
        b.ev_vid[nev] = victim;
        b.ev_t[nev] = t;
      }
      ++nev;
      --size;
      if (vfirst != kNever) {  // the key comes back later in this scan
        if (hn >= kHeapCap) { err |= 16; break; }
        __syncthreads();
        if (lane == 0) {
          int i = hn;
          const uint64_t v = (uint64_t(uint32_t(vfirst)) << 32) | uint32_t(victim);
          while (i > 0) {
            int p = (i - 1) >> 1;
            if (heap[p] <= v) break;
            heap[i] = heap[p];
            i = p;
          }
          heap[i] = v;
        }
        __syncthreads();
        ++hn;
      }
    }
  }
  __syncthreads();
  if (lane == 0) {
    c->size = size;
    c->group_counter = gc;
    c->free_top = ftop;
    c->bump = bump;
    c->ring_head = ptr;
    c->ncre = nco;
    c->nevict = nev;
    if (err) atomicOr(&c->error, err);
  }
}

// Initialise created voxels and point their hash slots at them.
__global__ void k_create_init(MapDev m, ScanBufs b) {
  const int nco = m.ctr->ncre;
  const int tag = m.ctr->tag;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nco; r += gridDim.x * blockDim.x) {
    const int32_t vid = b.co_vid[r];
    const int s = b.co_seg[r];
    if (vid < 0 || vid >= m.pool || s < 0) continue;  // error already flagged
    const int32_t e = b.seg_entry[s];
    VoxState st;
    st.key = m.table[e].key;
    st.group_id = b.co_group[r];
    st.stamp = 0;
    st.n = 0;
    st.newly_add = 0;
    st.flags = kEnable | kAlive;
    st.tag = tag;
    st.create_t = b.co_t[r];
    st.evict_t = kNever;
    st.freeze_t = -1;
    st.prev_vid = b.co_prev[r];
    st.first_t = b.co_t[r];
    st.seg = s;
    st.merge_pos = kNever;
    st.dirty_t = -2;
    m.vs[vid] = st;
    double* pl = m.plane + size_t(vid) * kPlaneD;
    for (int k = 0; k < kPlaneD; ++k) pl[k] = 0.0;
    double* mo = m.mom + size_t(vid) * kMomD;
    for (int k = 0; k < kMomD; ++k) mo[k] = 0.0;
    VoxHot h{};
    h.flags = kEnable | kAlive;
    m.hot[vid] = h;
    m.table[e].vid = vid;
    b.seg_vid[s] = vid;
  }
}

// ------------------------------------------------------------------ K6 ----
struct Moments {
  double s1[3], s2[6], c0[6], c1[3][6], c2[6][6];
};

__device__ inline void mom_load(const double* src, Moments& M) {
  int o = 0;
  for (int a = 0; a < 3; ++a) M.s1[a] = src[o++];
  for (int a = 0; a < 6; ++a) M.s2[a] = src[o++];
  for (int a = 0; a < 6; ++a) M.c0[a] = src[o++];
  for (int a = 0; a < 3; ++a)
    for (int l = 0; l < 6; ++l) M.c1[a][l] = src[o++];
  for (int a = 0; a < 6; ++a)
    for (int l = 0; l < 6; ++l) M.c2[a][l] = src[o++];
}
__device__ inline void mom_store(double* dst, const Moments& M) {
  int o = 0;
  for (int a = 0; a < 3; ++a) dst[o++] = M.s1[a];
  for (int a = 0; a < 6; ++a) dst[o++] = M.s2[a];
  for (int a = 0; a < 6; ++a) dst[o++] = M.c0[a];
  for (int a = 0; a < 3; ++a)
    for (int l = 0; l < 6; ++l) dst[o++] = M.c1[a][l];
  for (int a = 0; a < 6; ++a)
    for (int l = 0; l < 6; ++l) dst[o++] = M.c2[a][l];
}
__device__ inline void mom_add(Moments& M, const float q[3], const float C[6]) {
  const double qd[3] = {double(q[0]), double(q[1]), double(q[2])};
  double qq[6];
  for (int a = 0; a < 3; ++a)
    for (int bb = a; bb < 3; ++bb) qq[s6(a, bb)] = qd[a] * qd[bb];
  double Cd[6];
  for (int l = 0; l < 6; ++l) Cd[l] = double(C[l]);
  for (int a = 0; a < 3; ++a) M.s1[a] += qd[a];
  for (int a = 0; a < 6; ++a) M.s2[a] += qq[a];
  for (int l = 0; l < 6; ++l) M.c0[l] += Cd[l];
  for (int a = 0; a < 3; ++a)
    for (int l = 0; l < 6; ++l) M.c1[a][l] += qd[a] * Cd[l];
  for (int a = 0; a < 6; ++a)
    for (int l = 0; l < 6; ++l) M.c2[a][l] += qq[a] * Cd[l];
}

// updatePlane (voxel_map.cpp:97-136) from moments.  Plane::cov is kept as the
// traces of its two diagonal blocks: the normal block's trace
// sum_i tr(F_i C_i F_i^T) is a quadratic form of d_i = p_i - mean, rebuilt
// from sum q q C, sum q C, sum C; the centre block adds tr(C_i)/n^2.
__device__ inline void refit(const MapDev& m, Moments& M, int n, uint32_t& flags, double* pl,
                             const double center[3], int& refits) {
  if (n < m.upd_pt) return;
  flags |= kInit;
  ++refits;
  const double invn = 1.0 / double(n);
  double mu[3];
  for (int a = 0; a < 3; ++a) mu[a] = M.s1[a] * invn;
  double cov[9];
  for (int a = 0; a < 3; ++a)
    for (int bb = 0; bb < 3; ++bb) cov[a * 3 + bb] = M.s2[s6(a, bb)] * invn - mu[a] * mu[bb];
  double ev[3], V[9];
  eigh3_dev(cov, ev, V);
  if (ev[0] > m.plane_thresh) {
    flags &= ~kPlane;
    return;
  }
  flags |= kPlane;
  // T[ab][lm] = sum_i d_a d_b C_lm
  double T[6][6];
  for (int a = 0; a < 3; ++a)
    for (int bb = a; bb < 3; ++bb)
      for (int l = 0; l < 6; ++l)
        T[s6(a, bb)][l] = M.c2[s6(a, bb)][l] - mu[a] * M.c1[bb][l] - mu[bb] * M.c1[a][l] + mu[a] * mu[bb] * M.c0[l];
  const double nn[3] = {V[0], V[3], V[6]};
  // G(w,x)[ab] = sum_lm T[ab][lm] w_l x_m ; Q = u^T G v
  auto G = [&](const double* w, const double* x, double g[6]) {
    for (int ab = 0; ab < 6; ++ab) {
      double s = 0.0;
      for (int l = 0; l < 3; ++l)
        for (int mm = 0; mm < 3; ++mm) s += T[ab][s6(l, mm)] * w[l] * x[mm];
      g[ab] = s;
    }
  };
  auto Q = [&](const double* g, const double* u, const double* v) {
    double s = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int bb = 0; bb < 3; ++bb) s += g[s6(a, bb)] * u[a] * v[bb];
    return s;
  };
  double gnn[6];
  G(nn, nn, gnn);
  double tn_inc = 0.0;
  for (int mm = 1; mm < 3; ++mm) {
    const double e[3] = {V[mm], V[3 + mm], V[6 + mm]};
    double gen[6], gee[6];
    G(e, nn, gen);
    G(e, e, gee);
    const double num = Q(gnn, e, e) + 2.0 * Q(gen, e, nn) + Q(gee, nn, nn);
    const double den = double(n) * (ev[0] - ev[mm]);
    tn_inc += num / (den * den);
  }
  const double tm_inc = (M.c0[0] + M.c0[3] + M.c0[5]) * invn * invn;
  pl[6] += tn_inc;
  pl[7] += tm_inc;
  double mw[3];
  for (int a = 0; a < 3; ++a) mw[a] = center[a] + mu[a];
  double axis = -(mw[0] * nn[0] + mw[1] * nn[1] + mw[2] * nn[2]);
  const double sg = axis < 0.0 ? -1.0 : 1.0;
  for (int a = 0; a < 3; ++a) pl[3 + a] = sg * nn[a];
}

__global__ void __launch_bounds__(kMapThreads) k_replay(MapDev m, ScanBufs b) {
  MapCounters* c = m.ctr;
  const int nseg = c->nseg;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  int refits = 0, calls = 0, dropped = 0;
  const int32_t vid = s < nseg ? b.seg_vid[s] : -1;
  if (vid >= 0) {
    const int start = b.seg_start[s];
    const int end = s + 1 < nseg ? b.seg_start[s + 1] : b.n;
    VoxState st = m.vs[vid];
    double* pl = m.plane + size_t(vid) * kPlaneD;
    double plv[kPlaneD];
    for (int k = 0; k < kPlaneD; ++k) plv[k] = pl[k];
    Moments M;
    mom_load(m.mom + size_t(vid) * kMomD, M);
    int64_t kx, ky, kz;
    unpack_key(st.key, kx, ky, kz);
    const double center[3] = {(double(kx) + 0.5) * m.vs64, (double(ky) + 0.5) * m.vs64, (double(kz) + 0.5) * m.vs64};
    const UpdParams& P = *b.prm;
    uint32_t flags = st.flags;
    int n = st.n, newly = st.newly_add;
    int merge_pos = kNever, freeze_t = -1;
    bool moved = false;
    for (int pos = start; pos < end; ++pos) {
      if ((flags & kInit) && !(flags & kEnable) && !b.is_build) {
        if (flags & kPlane) {
          merge_pos = pos - start;
          calls += end - pos;
        } else {
          dropped += end - pos;
        }
        break;
      }
      const int i = b.sval_out[pos];
      const float px = b.x[i], py = b.y[i], pz0 = b.z[i];
      float pw[3];
      ins_transform(P.m, P.t, px, py, pz0, pw);
      const float q[3] = {float(double(pw[0]) - center[0]), float(double(pw[1]) - center[1]),
                          float(double(pw[2]) - center[2])};
      const float pz = pz0 == 0.0f ? 0.001f : pz0;
      float C[6];
      point_world_cov(P.rwl32, P.Prr, P.Ppp, P.s2, P.rv, px, py, pz, C);
      mom_add(M, q, C);
      ++n;
      moved = true;
      if (b.is_build) continue;
      if (!(flags & kInit)) {
        refit(m, M, n, flags, plv, center, refits);
        continue;
      }
      ++newly;
      if (newly >= m.upd_pt) {
        refit(m, M, n, flags, plv, center, refits);
        newly = 0;
      }
      if (n >= m.max_pt) {
        flags &= ~kEnable;
        freeze_t = i;
      }
    }
    if (b.is_build) refit(m, M, n, flags, plv, center, refits);
    if (moved) {
      const double invn = 1.0 / double(n);
      for (int a = 0; a < 3; ++a) plv[a] = center[a] + M.s1[a] * invn;
      mom_store(m.mom + size_t(vid) * kMomD, M);
    }
    for (int k = 0; k < kPlaneD; ++k) pl[k] = plv[k];
    st.flags = flags;
    st.n = n;
    st.newly_add = newly;
    st.freeze_t = freeze_t;
    st.merge_pos = merge_pos;
    m.vs[vid] = st;
    write_hot(m, vid, flags);
  }
  // block-level counter reduction
  refits = wave_sum_i(refits);
  calls = wave_sum_i(calls);
  dropped = wave_sum_i(dropped);
  if ((threadIdx.x & 63) == 0) {
    if (refits) atomicAdd(&c->refits, refits);
    if (calls) atomicAdd(&c->merge_calls, calls);
    if (dropped) atomicAdd(&c->dropped, dropped);
  }
}

// ------------------------------------------------------------------ K7 ----
// incarnation of a key at point time t (-1: none)
__device__ inline int32_t vid_at(const MapDev& m, int32_t vid, int tag, int t) {
  if (vid < 0) return -1;
  const VoxState& st = m.vs[vid];
  if (st.tag == tag && st.create_t > t) vid = st.prev_vid;
  if (vid < 0) return -1;
  const VoxState& s2 = m.vs[vid];
  if (s2.tag == tag && s2.evict_t <= t) return -1;
  return vid;
}

// first merge() call of segment s strictly after point t (kNever if none)
__device__ inline int first_call_after(const ScanBufs& b, int nseg, int s, int merge_pos, int t) {
  const int start = b.seg_start[s];
  const int end = s + 1 < nseg ? b.seg_start[s + 1] : b.n;
  int lo = start + merge_pos, hi = end;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (b.sval_out[mid] > t) hi = mid;
    else lo = mid + 1;
  }
  return lo < end ? b.sval_out[lo] : kNever;
}

__device__ inline int32_t lookup_vid(const MapDev& m, uint64_t key) {
  const int32_t e = hash_find_slot(m.table, m.hmask, key);
  return e < 0 ? -1 : m.table[e].vid;
}

__global__ void k_candidates(MapDev m, ScanBufs b) {
  MapCounters* c = m.ctr;
  const int nseg = c->nseg, tag = c->tag;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
    const int32_t vid = b.seg_vid[s];
    if (vid < 0) continue;
    VoxState& st = m.vs[vid];
    int dirty = (st.flags & kDirty) ? -1 : -2;
    const bool own_freeze = st.freeze_t >= 0 && (st.flags & kPlane);
    if (own_freeze && st.freeze_t > dirty) dirty = st.freeze_t;
    const bool has_calls = st.merge_pos != kNever;
    int emitted[7];
    int ne = 0;
    auto emit = [&](int after) {
      const int t = first_call_after(b, nseg, s, st.merge_pos, after);
      if (t == kNever) return;
      for (int q = 0; q < ne; ++q)
        if (emitted[q] == t) return;  // one candidate per distinct call
      emitted[ne++] = t;
      const int slot = atomicAdd(&c->ncand, 1);
      if (slot < b.n) b.cand_in[slot] = (uint64_t(uint32_t(t)) << 32) | uint32_t(vid);
      else atomicOr(&c->error, 16);
    };
    if (has_calls && ((st.flags & kDirty) || own_freeze)) emit(-1);
    for (int j = 0; j < 6; ++j) {
      if (!key_neighbour_ok(st.key, j)) continue;
      const int32_t nv = lookup_vid(m, key_neighbour(st.key, j));
      if (nv < 0) continue;
      const VoxState& ns = m.vs[nv];
      if (ns.tag != tag || ns.freeze_t < 0 || !(ns.flags & kPlane)) continue;
      if (ns.freeze_t > dirty) dirty = ns.freeze_t;
      if (has_calls) emit(ns.freeze_t);
    }
    st.dirty_t = dirty;
  }
}

// voxels not touched in this scan whose neighbour froze as a plane
__global__ void k_dirty_untouched(MapDev m, ScanBufs b) {
  MapCounters* c = m.ctr;
  const int nseg = c->nseg, tag = c->tag;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
    if (b.seg_vid[s] < 0) continue;
    const VoxState& st = m.vs[b.seg_vid[s]];
    if (st.freeze_t < 0 || !(st.flags & kPlane)) continue;
    for (int j = 0; j < 6; ++j) {
      if (!key_neighbour_ok(st.key, j)) continue;
      const int32_t nv = lookup_vid(m, key_neighbour(st.key, j));
      if (nv < 0) continue;
      VoxState& ns = m.vs[nv];
      if (ns.tag != tag && (ns.flags & kAlive)) atomicOr(&ns.flags, kDirty);
    }
  }
}

// ------------------------------------------------------------------ K8 ----
struct NbrRec {
  int32_t vid;
  int32_t ok;  // eligible: exists, frozen and plane at time t
  uint64_t group;
};

__global__ void __launch_bounds__(64) k_merge_replay(MapDev m, ScanBufs b) {
  MapCounters* c = m.ctr;
  const int ncand = c->ncand, nseg = c->nseg, tag = c->tag;
  __shared__ uint64_t heap[kHeapCap];
  __shared__ NbrRec nb[6];
  __shared__ int32_t clos[64];
  __shared__ int nclos;
  __shared__ int merged_any;
  const int lane = threadIdx.x;
  int hn = 0, ci = 0, evals = 0, merges = 0;
  uint64_t last = ~0ull;
  int err = 0;
  while (true) {
    const uint64_t vs_ = ci < ncand ? b.cand_out[ci] : ~0ull;
    const uint64_t vh = hn > 0 ? heap[0] : ~0ull;
    if (vs_ == ~0ull && vh == ~0ull) break;
    uint64_t cur;
    if (vs_ <= vh) {
      cur = vs_;
      ++ci;
    } else {
      cur = vh;
      __syncthreads();
      if (lane == 0) {
        const uint64_t lastv = heap[hn - 1];
        const int nn = hn - 1;
        int i = 0;
        while (true) {
          int l = 2 * i + 1, r = l + 1, s = i;
          uint64_t sv = lastv;
          if (l < nn && heap[l] < sv) { s = l; sv = heap[l]; }
          if (r < nn && heap[r] < sv) { s = r; sv = heap[r]; }
          if (s == i) break;
          heap[i] = heap[s];
          i = s;
        }
        if (nn > 0) heap[i] = lastv;
      }
      __syncthreads();
      --hn;
    }
    if (cur == last) continue;
    last = cur;
    const int t = int(cur >> 32);
    const int32_t V = int32_t(cur & 0xffffffffu);
    ++evals;
    const uint64_t vkey = m.vs[V].key;
    // ---- neighbours at time t (lanes 0..5)
    if (lane < 6) {
      NbrRec r{-1, 0, 0};
      if (key_neighbour_ok(vkey, lane)) {
        const int32_t nv = vid_at(m, lookup_vid(m, key_neighbour(vkey, lane)), tag, t);
        if (nv >= 0) {
          const VoxState& ns = m.vs[nv];
          const bool frozen = !(ns.flags & kEnable) && !(ns.tag == tag && ns.freeze_t > t);
          r.vid = nv;
          r.ok = frozen && (ns.flags & kPlane) && (ns.flags & kInit);
          r.group = ns.group_id;
        }
      }
      nb[lane] = r;
    }
    if (lane == 0) {
      merged_any = 0;
      nclos = 0;
    }
    __syncthreads();
    // ---- merge(): sequential over the six neighbours (lane 0)
    if (lane == 0) {
      double* pv = m.plane + size_t(V) * kPlaneD;
      const uint64_t gV = m.vs[V].group_id;
      for (int j = 0; j < 6; ++j) {
        const NbrRec r = nb[j];
        if (r.vid < 0 || !r.ok) continue;
        if (m.vs[r.vid].group_id == gV) continue;
        double* pn = m.plane + size_t(r.vid) * kPlaneD;
        double dn = pn[3] * pv[3];
        dn += pn[4] * pv[4];
        dn += pn[5] * pv[5];
        const double norm_distance = 1.0 - dn;
        double a1 = pn[3] * pn[0];
        a1 += pn[4] * pn[1];
        a1 += pn[5] * pn[2];
        double a0 = pv[3] * pv[0];
        a0 += pv[4] * pv[1];
        a0 += pv[5] * pv[2];
        const double axis_distance = fabs(a1 - a0);
        if (norm_distance > m.merge_a || axis_distance > m.merge_d) continue;
        const double tn0 = pv[6], tm0 = pv[7], tn1 = pn[6], tm1 = pn[7];
        const double tc0 = tn0 + tm0, tc1 = tn1 + tm1;
        double nm[3], nrm[3];
        for (int k = 0; k < 3; ++k) {
          nm[k] = tm0 * pn[k] + (tm1 * pv[k]) / (tm0 + tm1);        // precedence as in the reference
          nrm[k] = tn0 * pn[3 + k] + (tn1 * pv[3 + k]) / (tn0 + tn1);
        }
        const double den = (tc0 + tc1) * (tc0 + tc1);
        const double ntn = (tc0 * tc0 * tn1 + tc1 * tc1 * tn0) / den;
        const double ntm = (tc0 * tc0 * tm1 + tc1 * tc1 * tm0) / den;
        m.vs[r.vid].group_id = gV;
        m.vs[r.vid].flags |= kMerged;
        m.vs[V].flags |= kMerged;
        double dd = nm[0] * nrm[0];
        dd += nm[1] * nrm[1];
        dd += nm[2] * nrm[2];
        if (-dd < 0.0)
          for (int k = 0; k < 3; ++k) nrm[k] = -nrm[k];
        for (int k = 0; k < 3; ++k) {
          pv[k] = nm[k];
          pv[3 + k] = nrm[k];
          pn[k] = nm[k];
          pn[3 + k] = nrm[k];
        }
        pv[6] = ntn;
        pv[7] = ntm;
        pn[6] = ntn;
        pn[7] = ntm;
        write_hot(m, r.vid, m.vs[r.vid].flags);
        ++merges;
        clos[nclos++] = r.vid;
        merged_any = 1;
      }
      if (merged_any) {
        write_hot(m, V, m.vs[V].flags);
        clos[nclos++] = V;
      }
    }
    __syncthreads();
    if (!merged_any) continue;
    // ---- follow-up calls: V, merged neighbours and their six neighbours
    const int nc0 = nclos;
    __syncthreads();
    for (int q0 = 0; q0 < nc0 * 7; q0 += 64) {
      const int q = q0 + lane;
      int32_t X = -1;
      if (q < nc0 * 7) {
        const int32_t base_v = clos[q / 7];
        const int j = q % 7;
        X = base_v;
        if (j < 6) {
          const uint64_t bk = m.vs[base_v].key;
          X = key_neighbour_ok(bk, j) ? vid_at(m, lookup_vid(m, key_neighbour(bk, j)), tag, t) : -1;
        }
      }
      int tn = kNever;
      if (X >= 0) {
        VoxState& xs = m.vs[X];
        if (xs.tag == tag && xs.seg >= 0 && xs.evict_t == kNever) {
          if (xs.dirty_t < t) xs.dirty_t = t;  // benign race: all writers store t
          if (xs.merge_pos != kNever) tn = first_call_after(b, nseg, xs.seg, xs.merge_pos, t);
        } else if (xs.flags & kAlive) {
          atomicOr(&xs.flags, kDirty);
        }
      }
      // serialise heap pushes through lane 0
      for (int w = 0; w < 64; ++w) {
        const int tw = __shfl(tn, w, 64);
        const int32_t xw = __shfl(X, w, 64);
        if (tw == kNever) continue;
        if (lane == 0) {
          if (hn >= kHeapCap) {
            err |= 16;
          } else {
            const uint64_t v = (uint64_t(uint32_t(tw)) << 32) | uint32_t(xw);
            int i = hn;
            while (i > 0) {
              int p = (i - 1) >> 1;
              if (heap[p] <= v) break;
              heap[i] = heap[p];
              i = p;
            }
            heap[i] = v;
          }
        }
        hn = __shfl(hn + ((lane == 0 && hn < kHeapCap) ? 1 : 0), 0, 64);
      }
    }
    __syncthreads();
    if (err) break;
  }
  if (lane == 0) {
    c->merge_evals += evals;
    c->merges += merges;
    if (err) atomicOr(&c->error, err);
  }
}

// ------------------------------------------------------------------ K9 ----
__global__ void k_finalize_segs(MapDev m, ScanBufs b) {
  MapCounters* c = m.ctr;
  const int nseg = c->nseg;
  const int64_t base = c->point_base;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
    const int start = b.seg_start[s];
    const int end = s + 1 < nseg ? b.seg_start[s + 1] : b.n;
    const int32_t vid = b.seg_vid[s];
    if (vid < 0) continue;
    VoxState& st = m.vs[vid];
    const int last_t = b.sval_out[end - 1];
    const bool calls = st.merge_pos != kNever;
    const bool frozen_plane = (st.flags & kPlane) && !(st.flags & kEnable) && (st.flags & kInit);
    // a dirtying event is resolved by a later (evaluated) merge() call
    const bool dirty = frozen_plane && st.dirty_t > -2 && !(calls && last_t > st.dirty_t);
    if (dirty) st.flags |= kDirty;
    else st.flags &= ~kDirty;
    st.stamp = uint64_t(base + last_t);
    b.flags[last_t] = vid + 1;  // LRU append marker (array cleared before)
  }
}

__global__ void k_evict_finalize(MapDev m, ScanBufs b) {
  MapCounters* c = m.ctr;
  const int nev = c->nevict;
  const int ftop = c->free_top;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nev; r += gridDim.x * blockDim.x) {
    const int32_t v = b.ev_vid[r];
    VoxState& st = m.vs[v];
    const int32_t e = hash_find_slot(m.table, m.hmask, st.key);
    if (e >= 0 && m.table[e].vid == v) {
      m.table[e].vid = -1;
      m.table[e].key = kKeyTomb;
    }
    st.flags = 0;
    m.hot[v].flags = 0;
    m.freelist[ftop + r] = v;
  }
}

__global__ void k_evict_finalize_ctr(MapDev m) {
  MapCounters* c = m.ctr;
  c->free_top += c->nevict;
  c->tombs += c->nevict;
}

__global__ void k_lru_append(MapDev m, ScanBufs b) {
  MapCounters* c = m.ctr;
  const int nsel = c->nlru;  // every touched voxel appears exactly once
  const int64_t tail = c->ring_tail;
  const int64_t base = c->point_base;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nsel; r += gridDim.x * blockDim.x) {
    const int t = b.lru_sel[r];
    const int32_t vid = b.flags[t] - 1;
    const int64_t ph = (tail + r) % m.qcap;
    m.ring_vid[ph] = vid;
    m.ring_stamp[ph] = uint64_t(base + t);
  }
}

__global__ void k_end_scan(MapDev m, ScanBufs b) {
  MapCounters* c = m.ctr;
  c->ring_tail += c->nlru;
  c->point_base += b.n;
}

// hash rebuild: every live voxel is re-inserted into a cleared table
__global__ void k_rehash_collect(MapDev m, int32_t* live, int32_t* nlive) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.pool; v += gridDim.x * blockDim.x) {
    if (m.vs[v].flags & kAlive) live[atomicAdd(nlive, 1)] = v;
  }
}
__global__ void k_rehash_clear(MapDev m) {
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h <= m.hmask; h += gridDim.x * blockDim.x) {
    m.table[h].key = kKeyEmpty;
    m.table[h].vid = -1;
  }
}
__global__ void k_rehash_insert(MapDev m, const int32_t* live, const int32_t* nlive) {
  const int n = *nlive;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const int32_t v = live[r];
    const int32_t e = hash_find_or_insert(m.table, m.hmask, m.vs[v].key, m.ctr);
    if (e >= 0) m.table[e].vid = v;
  }
}
// ring compaction: keep live entries in order (src -> dst), one pass
__global__ void k_ring_flags(MapDev m, int32_t* keep, int64_t head, int64_t count) {
  for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r < count; r += int64_t(gridDim.x) * blockDim.x) {
    const int64_t ph = (head + r) % m.qcap;
    const int32_t v = m.ring_vid[ph];
    const VoxState& st = m.vs[v];
    keep[r] = ((st.flags & kAlive) && st.stamp == m.ring_stamp[ph]) ? 1 : 0;
  }
}

}  // namespace vmp
