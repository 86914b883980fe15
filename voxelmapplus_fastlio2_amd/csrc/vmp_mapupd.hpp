// Map update: VoxelMap::update / build (voxel_map.cpp:200-256) with
// VoxelGrid::pushPoint / updatePlane / merge (voxel_map.cpp:42-186), executed
// for a batch of independent sessions as a pipeline of data-parallel stages
// plus two short ordered replays (one wave per session).
//
// Points of all sessions in a batch share one global index space, session
// after session (g = off[s] + i); times inside a session are its point
// indices i, which is the order the reference loops in.
//
//   K1 ins_keys      per point: fp32 insertion transform (PCL order), key,
//                    hash find-or-insert in the session's table
//   K2 sort          per session, stable radix sort by hash slot: each key's
//                    points become a contiguous segment in point order
//   K3/K4 segments   segment heads, the existing voxel of each segment, new
//                    keys flagged at their first point
//   K5 create        creation events in point order: parallel when the LRU
//                    cannot overflow, else one wave per session replays
//                    creations and evictions against the LRU queue
//   K6 replay        one lane per voxel walks its points through the pushPoint
//                    state machine; the voxels whose walk reaches a refit
//                    (known from counts) take the FP64 moment path, the others
//                    a light one that queues their points (no point pool)
//   K7/K8 merge      see vmp_merge.hpp
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
  int64_t size;            // live voxels
  uint64_t group_counter;  // VoxelGrid::count
  int64_t point_base;      // global point index of this scan's point 0 (LRU stamps)
  int64_t ring_head, ring_tail;
  int32_t free_top, bump;
  int32_t used_entries;    // non-empty hash slots, tombstones included
  int32_t tombs;
  int32_t tag;             // scan tag
  int32_t error;           // bit 1 hash probe overflow, 2 range, 4 capacity, 8 pool, 16 replay queue,
                           // 32 replay class (cannot happen)
  // per-scan outputs
  int32_t nseg, ncreate_ev, ncre, nevict, ncand;
  int32_t refits, merge_calls, merge_evals, merges, dropped;
  int32_t fast;            // 1: this scan's creations took the parallel (no-eviction) path
  int32_t nlru;            // voxels appended to the LRU queue by this scan
  int32_t merge_windows;   // windows of the merge replay
  int32_t merge_requeued;  // window entries sent back for re-evaluation
  int32_t merge_max_acting;  // most acting entries in one window of the merge replay
  int32_t pad_ctr;
};

// one session's map (device pointers)
struct MapDev {
  HashEntry* table;
  uint32_t hmask;
  int32_t pool;
  VoxHot* hot;
  VoxState* vs;
  double* plane;      // [pool][kPlaneD]
  double* mom;        // [pool][kMomD]
  float* pend;        // [pool][pend_cap][kPendF] points added since the moments' last update
  int32_t pend_cap;   // update_size_thresh - 1
  int32_t* freelist;  // [pool]
  int32_t* ring_vid;  // [qcap]
  uint64_t* ring_stamp;
  int64_t qcap;
  MapCounters* ctr;
  int32_t capacity;
  int32_t max_pt, upd_pt;
  int32_t hbits;      // log2 of the table size
  double plane_thresh, merge_a, merge_d, vs64;
};

// one session of a batch
struct SessInfo {
  const float* x;
  const float* y;
  const float* z;
  const UpdParams* prm;
  int32_t n, off;     // points, global offset of point 0
  int32_t is_build;
  int32_t seg_lo, seg_hi;  // its segments in the batch segment list
  int32_t cre_lo, cre_hi;  // its creation events in the batch creation list
  int32_t lru_lo, lru_hi;  // its LRU appends in the batch append list
  int32_t spl_lo, spl_hi;  // its acting merge candidates in the compacted list
  int32_t ftop0, bump0;    // free-list top and bump pointer before its fast creations
};

// batch workspace, all arrays sized for the batch's total points
struct BatchBufs {
  int32_t ntot;
  int32_t nsess;
  int32_t hbits;       // bits of the largest hash-slot index in the batch
  int32_t cand_sshift; // position of the session bits in a batch candidate key (25 + time bits)
  uint64_t* skey_in;   // (session << hbits) | hash slot, per point
  int32_t* sval_in;    // global point index
  uint64_t* skey_out;
  int32_t* sval_out;
  int32_t* flags;      // scratch flag / value array
  int32_t* sel;        // compaction output (ascending global indices)
  int32_t* cnt;        // counts: [0] segments, [1] creation events, [2] LRU appends
  int32_t* seg_start;  // segment -> first sorted position
  int32_t* seg_end;
  int32_t* seg_sess;
  int32_t* seg_vid;
  int32_t* seg_entry;
  int32_t* cre_t;      // creation events: local point index, segment
  int32_t* cre_seg;
  int4* co;            // resolved creations in order, per session at its offset:
                       // (vid, segment, point time, evicted incarnation or -1);
                       // the r-th takes group id group_counter - ncre + r
  int32_t* ev_vid;     // evicted voxels in order, per session at its offset
  uint64_t* cand_in;   // merge candidates, see cand_key(); per session at its offset
  uint64_t* cand_out;  // sorted: session-major, then point time, then voxel
  int32_t* cand_nbr;   // per sorted candidate: 6 eligible neighbour ids at its time
  // per point (session offset + point time): the 6 eligible neighbour ids of
  // the merge() call at that time, + 2 (0: not looked up yet in this scan);
  // they cannot change during the merge replay, so a call that returns to
  // the pool reads them back instead of repeating the lookups
  int32_t* tnbr;
  int32_t* cand_spec;  // per sorted candidate: 1 if it merges on the pre-replay state
  int32_t* cand_start; // per session: first sorted candidate
  int32_t* spl;        // positions of candidates with cand_spec = 1, ascending
  int32_t* seg_ct;     // per segment: its candidate times (up to 7, ascending)
  int32_t* seg_nct;
  int32_t* seg_nbr;    // per segment: the six neighbour voxels (current incarnation, -1 none)
  int4* hlist;         // refitting voxels as (vid, start, length, session), see k_seg_lenkeys
  uint32_t* len_in;    // replay order: segments by descending length
  uint32_t* len_out;
  int32_t* ord_in;
  int32_t* ord;
  // per acting candidate (its index in spl, the first pre_cap of them): for
  // each record r (its six eligible neighbours, then itself) and each
  // neighbour j of r (j = 6: r itself), the voxel X there at the candidate's
  // time, X's segment in this scan (-1: none or evicted) and X's first
  // merge() call after that time; none of these depends on merges
  int32_t* pre_x;
  int32_t* pre_seg;
  int32_t* pre_tn;
  int32_t pre_cap;
  int32_t pad1;
  // live LRU ring entries from the head of each evicting session: (vid,
  // first touch in this scan or kNever, segment, ring position - head).
  // lv holds them per 64-entry ring chunk (chunk k of a session at
  // 2 * off + 64 k, two entries per point; lv_cnt: its count, at
  // lv_cnt_base()), lv2 all of them in ring order at 2 * off (lv_n: count)
  int4* lv;
  int4* lv2;
  int32_t* lv_cnt;
  int32_t* lv_n;
  int32_t* prim;       // work entries of the compaction / radix sort (prim_ws_entries)
  // per point in key-sorted order: its insertion point relative to its
  // voxel's centre (fp32, as the replays fold it in), its index in the scan
  // and its fp32 world covariance, (q.x, q.y, q.z, index bits, C0..C5, -, -),
  // formed once by k_seg_heads
  float4* qc;
};
__host__ __device__ inline int64_t lv_cnt_base(int off, int s) { return (int64_t(2) * off >> 6) + s; }

// sessions of one batch: the merge-candidate key holds a session in 15 bits
constexpr int kMaxBatchSessions = 1 << 15;

// merge-candidate key: session (15 bits) | point time (24 bits) | voxel (25 bits)
__host__ __device__ inline uint64_t cand_key(int s, int t, int32_t vid) {
  return (uint64_t(uint32_t(s)) << 49) | (uint64_t(uint32_t(t)) << 25) | uint64_t(uint32_t(vid));
}
__host__ __device__ inline int cand_t(uint64_t k) { return int((k >> 25) & 0xFFFFFFull); }
__host__ __device__ inline int32_t cand_vid(uint64_t k) { return int32_t(k & 0x1FFFFFFull); }
// In the batch's candidate list the session sits right above the point time
// (at bit cand_sshift, 25 + the bits of the largest scan), so the sort reads
// no constant bits; the local part (time, voxel) is cand_key(0, t, vid).
__host__ __device__ inline uint64_t cand_key_batch(int s, int t, int32_t vid, int sshift) {
  return (uint64_t(uint32_t(s)) << sshift) | (uint64_t(uint32_t(t)) << 25) | uint64_t(uint32_t(vid));
}
__host__ __device__ inline uint64_t cand_local(uint64_t k, int sshift) { return k & ((uint64_t(1) << sshift) - 1); }

__device__ inline int sess_of(const SessInfo* si, int nsess, int g) {
  int lo = 0, hi = nsess - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (si[mid].off <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------ helpers -----
// Grid-stride loop whose trip count is the same for every lane of a block,
// so wave-level operations inside see the whole wave; `valid` marks the
// lanes with an item.
#define VMP_UNIFORM_LOOP(i, valid, n)                                                            \
  for (int i##_0 = blockIdx.x * blockDim.x; i##_0 < (n); i##_0 += gridDim.x * blockDim.x)         \
    for (int i = i##_0 + int(threadIdx.x), valid = (i < (n)), i##_once = 1; i##_once; i##_once = 0)

// Fetch-add on counters, aggregated over the wave: the lanes that add to the
// same counter do so with one atomic, and each lane gets the counter's value
// before its own contribution (lanes in lane order).  Every lane of the wave
// calls it; a lane with v = 0 contributes nothing.
__device__ inline int wave_fetch_add(int* p, int v) {
  const int lane = threadIdx.x & 63;
  uint64_t pending = __ballot(v != 0);
  const uint64_t pv = reinterpret_cast<uint64_t>(p);
  int out = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t lo = uint32_t(__shfl(int(uint32_t(pv)), leader, 64));
    const uint32_t hi = uint32_t(__shfl(int(uint32_t(pv >> 32)), leader, 64));
    const uint64_t pl = uint64_t(lo) | (uint64_t(hi) << 32);
    const bool in = ((pending >> lane) & 1ull) && pv == pl;
    const uint64_t grp = __ballot(in);
    const int x = in ? v : 0;
    int incl = x;
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    const int total = __shfl(incl, 63, 64);
    int base = 0;
    if (lane == leader) base = atomicAdd(reinterpret_cast<int*>(pl), total);
    base = __shfl(base, leader, 64);
    if (in) out = base + incl - x;
    pending &= ~grp;
  }
  return out;
}

// LDS written by some lanes of a wave and read by others of the same wave
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Least value over the 64 lanes of a wave (every lane active): row shifts
// within each 16-lane row, then the row broadcasts, leave it in lane 63.
__device__ inline uint32_t wave_min_u32(uint32_t v) {
  constexpr int kId = -1;  // 0xffffffff, the identity of min
  v = min(v, uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x111, 0xf, 0xf, false)));  // row_shr:1
  v = min(v, uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x112, 0xf, 0xf, false)));  // row_shr:2
  v = min(v, uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x114, 0xf, 0xf, false)));  // row_shr:4
  v = min(v, uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x118, 0xf, 0xf, false)));  // row_shr:8
  v = min(v, uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x142, 0xa, 0xf, false)));  // row_bcast:15
  v = min(v, uint32_t(__builtin_amdgcn_update_dpp(kId, int(v), 0x143, 0xc, 0xf, false)));  // row_bcast:31
  return uint32_t(__builtin_amdgcn_readlane(int(v), 63));
}
__device__ inline uint64_t wave_min_u64(uint64_t v) {
  const uint32_t hi = wave_min_u32(uint32_t(v >> 32));
  const uint32_t lo = wave_min_u32(uint32_t(v >> 32) == hi ? uint32_t(v) : 0xffffffffu);
  return (uint64_t(hi) << 32) | lo;
}

// *inserted = 1 when this call claimed the empty slot (the caller counts
// used_entries); with inserted == nullptr the count is added here
__device__ inline int32_t hash_find_or_insert(HashEntry* table, uint32_t mask, uint64_t key, MapCounters* ctr,
                                              int* inserted = nullptr) {
  uint32_t h = hash_slot(key, mask);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    unsigned long long* kp = reinterpret_cast<unsigned long long*>(&table[h].key);
    unsigned long long k = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == key) return int32_t(h);
    if (k == kKeyEmpty) {
      unsigned long long old = atomicCAS(kp, (unsigned long long)kKeyEmpty, (unsigned long long)key);
      if (old == kKeyEmpty) {
        if (inserted) *inserted = 1;
        else atomicAdd(&ctr->used_entries, 1);
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

#define VMP_GRID_LOOP(i, n) for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

__global__ void k_begin_scan(const MapDev* maps, int nsess) {
  if (threadIdx.x != 0 || int(blockIdx.x) >= nsess) return;
  MapCounters* c = maps[blockIdx.x].ctr;
  c->tag += 1;
  c->fast = 0;
  c->nseg = c->ncreate_ev = c->ncre = c->nevict = c->ncand = c->nlru = 0;
  c->refits = c->merge_calls = c->merge_evals = c->merges = c->dropped = 0;
  c->merge_windows = c->merge_requeued = c->merge_max_acting = 0;
}
__global__ void k_fill_i32(int32_t* p, int32_t v, int n) {
  VMP_GRID_LOOP(i, n) p[i] = v;
}

// A session whose creation events all fit under its capacity (no LRU
// eviction in this scan) takes its new voxels from the free-list top down,
// then from the bump pointer.  Its counters advance as soon as its creation
// count is known; the free-list top and bump pointer they started from stay
// in S for k_create_fast.  Other sessions keep fast = 0 and go through
// k_create_evict.
__device__ inline void create_fast_counters(const MapDev& m, SessInfo& S, int nc) {
  MapCounters* c = m.ctr;
  if (c->size + nc > m.capacity) return;
  const int ftop = c->free_top, bump = c->bump;
  S.ftop0 = ftop;
  S.bump0 = bump;
  if (nc <= ftop) {
    c->free_top = ftop - nc;
  } else {
    c->bump = bump + (nc - ftop);
    c->free_top = 0;
  }
  if (c->bump > m.pool) atomicOr(&c->error, 8);
  c->size += nc;
  c->group_counter += uint64_t(nc);
  c->ncre = nc;
  c->nevict = 0;
  c->fast = 1;
}

// per-session [lo, hi) of an ascending list of n global indices
// which: 0 segments, 1 creation events, 2 LRU appends
__device__ inline void session_range(const MapDev* maps, SessInfo* si, int s, const int32_t* list, int n,
                                     int which) {
  auto lb = [&](int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (list[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const int lo = lb(si[s].off), hi = lb(si[s].off + si[s].n);
  MapCounters* c = maps[s].ctr;
  if (which == 0) {
    si[s].seg_lo = lo;
    si[s].seg_hi = hi;
    c->nseg = hi - lo;
  } else if (which == 1) {
    si[s].cre_lo = lo;
    si[s].cre_hi = hi;
    c->ncreate_ev = hi - lo;
    create_fast_counters(maps[s], si[s], hi - lo);
  } else {
    si[s].lru_lo = lo;
    si[s].lru_hi = hi;
    c->nlru = hi - lo;
  }
}
__global__ void k_ranges(const MapDev* maps, SessInfo* si, int nsess, const int32_t* list, const int32_t* count,
                         int which) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsess) session_range(maps, si, s, list, *count, which);
}
// The same, inside the compaction that produces the list: a session's first
// list position is the number of items selected before its first point,
// which the thread holding that point knows once its tile's prefix is known
// (sfirst[t] .. sfirst[t + 1] - 1 are the sessions starting in tile t); the
// last tile to finish sets each session's end (the next one's start, or the
// count) and its counter.  Sessions starting at or past nitems (empty
// trailing sessions) start at the count.
__device__ inline int32_t* range_lo(SessInfo& S, int which) {
  return which == 0 ? &S.seg_lo : (which == 1 ? &S.cre_lo : &S.lru_lo);
}
struct TailRanges {
  const MapDev* maps;
  SessInfo* si;
  int nsess;
  const int32_t* list;
  int which;
  const int32_t* sfirst;
  int nitems;
  static constexpr bool kActive = true;
  __device__ void at_tile(int tile, int base, int pos, const int32_t (&v)[kPrimIpt]) const {
    const int s1 = sfirst[tile + 1];
    for (int s = sfirst[tile]; s < s1; ++s) {
      const int o = si[s].off;
      if (o < base || o >= base + kPrimIpt) continue;
      int p = pos;
#pragma unroll
      for (int k = 0; k < kPrimIpt; ++k)
        if (k < o - base && v[k] != 0) ++p;
      *range_lo(si[s], which) = p;
    }
  }
  __device__ void operator()(int n) const {
    for (int s = threadIdx.x; s < nsess; s += blockDim.x) {
      const int lo = si[s].off >= nitems ? n : *range_lo(si[s], which);
      const int hi = s + 1 < nsess ? (si[s + 1].off >= nitems ? n : *range_lo(si[s + 1], which)) : n;
      if (si[s].off >= nitems) *range_lo(si[s], which) = lo;
      MapCounters* c = maps[s].ctr;
      if (which == 0) {
        si[s].seg_hi = hi;
        c->nseg = hi - lo;
      } else if (which == 1) {
        si[s].cre_hi = hi;
        c->ncreate_ev = hi - lo;
        create_fast_counters(maps[s], si[s], hi - lo);
      } else {
        si[s].lru_hi = hi;
        c->nlru = hi - lo;
      }
    }
  }
};

// ------------------------------------------------------------------ K1 ----
__global__ void k_ins_keys(const MapDev* maps, const SessInfo* si, int nsess, BatchBufs b) {
  VMP_UNIFORM_LOOP(g, valid, b.ntot) {
    int ins = 0;
    const int s = valid ? sess_of(si, nsess, g) : 0;
    if (valid) {
    const SessInfo& S = si[s];
    const MapDev m = maps[s];
    const int i = g - S.off;
    const UpdParams& P = *S.prm;
    float pw[3];
    ins_transform(P.m, P.t, S.x[i], S.y[i], S.z[i], pw);
    int64_t k[3];
    bool ok = true;
    for (int r = 0; r < 3; ++r) {
      k[r] = int64_t(floor(double(pw[r]) / m.vs64));
      ok &= key_ok(k[r]);
    }
    int32_t e = -1;
    if (ok) e = hash_find_or_insert(m.table, m.hmask, pack_key(k[0], k[1], k[2]), m.ctr, &ins);
    else atomicOr(&m.ctr->error, 2);
    b.skey_in[g] = (uint64_t(s) << b.hbits) | uint64_t(uint32_t(e < 0 ? 0 : e));
    b.sval_in[g] = g;
    }
    wave_fetch_add(&maps[s].ctr->used_entries, ins);
  }
}

// ------------------------------------------------------------------ K3 ----
// Segment heads of the key-sorted points, and the replays' input of every
// point in that order (so a voxel's points are contiguous): the point
// relative to its voxel's centre, its world covariance (calcBodyCov's z = 0
// -> 0.001 applies to the covariance only, commons.cpp:20-23) and its index
// in the scan, (q.x, q.y, q.z, index bits, C0..C5, -, -).  The transform and
// the key are k_ins_keys' expressions, so the voxel is the one the point was
// keyed to.
__global__ void k_seg_heads(const MapDev* maps, const SessInfo* si, int nsess, BatchBufs b) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {  // refitting voxels, counted by k_seg_lenkeys
    b.cnt[5] = 0;
    b.cnt[6] = 0;
  }
  VMP_GRID_LOOP(j, b.ntot) {
    const int s = sess_of(si, nsess, j);
    const SessInfo& S = si[s];
    b.flags[j] = (j == S.off || b.skey_out[j] != b.skey_out[j - 1]) ? 1 : 0;
    const double vs64 = maps[s].vs64;
    const UpdParams& P = *S.prm;
    const int i = b.sval_out[j] - S.off;
    const float x = S.x[i], y = S.y[i], z = S.z[i];
    float pw[3];
    ins_transform(P.m, P.t, x, y, z, pw);
    float qv[3];
    for (int r = 0; r < 3; ++r) {
      const int64_t k = int64_t(floor(double(pw[r]) / vs64));
      qv[r] = float(double(pw[r]) - (double(k) + 0.5) * vs64);
    }
    const float pz = z == 0.0f ? 0.001f : z;
    float C[6];
    point_world_cov(P.rwl32, P.Prr, P.Ppp, P.s2, P.rv, x, y, pz, C);
    b.qc[3 * size_t(j)] = make_float4(qv[0], qv[1], qv[2], __int_as_float(i));
    b.qc[3 * size_t(j) + 1] = make_float4(C[0], C[1], C[2], C[3]);
    b.qc[3 * size_t(j) + 2] = make_float4(C[4], C[5], 0.0f, 0.0f);
  }
}

// ------------------------------------------------------------------ K4 ----
__global__ void k_seg_info(const MapDev* maps, const SessInfo* si, int nsess, BatchBufs b) {
  const int nseg = b.cnt[0];
  VMP_GRID_LOOP(q, nseg) {
    const int start = b.sel[q];
    const int s = sess_of(si, nsess, start);
    const SessInfo& S = si[s];
    const int lim = S.off + S.n;
    const int end = (q + 1 < nseg && b.sel[q + 1] < lim) ? b.sel[q + 1] : lim;
    const MapDev m = maps[s];
    const uint32_t e = uint32_t(b.skey_out[start] & ((uint64_t(1) << b.hbits) - 1));
    const int first = b.sval_out[start] - S.off;
    b.seg_start[q] = start;
    b.seg_end[q] = end;
    b.seg_sess[q] = s;
    b.seg_entry[q] = int32_t(e);
    const int32_t vid = m.table[e].vid;
    b.seg_vid[q] = vid;
    if (vid >= 0) {  // this scan's fields (three vector stores)
      VoxState* st = m.vs + vid;
      st_global2(&st->merge_pos, kNever, -2);
      st_global4(&st->tag, m.ctr->tag, -1, kNever, -1);
      st_global4(&st->prev_vid, -1, first, q, kNever);
    } else {
      b.flags[S.off + first] = q + 1;  // creation event at its first point (array cleared before)
    }
  }
}

__global__ void k_cre_split(const SessInfo* si, int nsess, BatchBufs b) {
  const int nc = b.cnt[1];
  VMP_GRID_LOOP(r, nc) {
    const int g = b.sel[r];
    const int s = sess_of(si, nsess, g);
    b.cre_t[r] = g - si[s].off;
    b.cre_seg[r] = b.flags[g] - 1;
  }
}
// the same as the output of the compaction of the creation events (v: the
// event's flag, its segment + 1)
struct OutCre {
  int32_t* sel;
  int32_t* cre_t;
  int32_t* cre_seg;
  const SessInfo* si;
  int nsess;
  __device__ void operator()(int r, int g, int32_t v) const {
    sel[r] = g;
    cre_t[r] = g - si[sess_of(si, nsess, g)].off;
    cre_seg[r] = v - 1;
  }
};

// ------------------------------------------------------------------ K5 ----
// Parallel creation for sessions whose LRU cannot overflow in this scan
// (fast, set with their counters where the creation events were counted).
__global__ void k_create_fast(const MapDev* maps, const SessInfo* si, int nsess, BatchBufs b) {
  const int nc = b.cnt[1];
  VMP_GRID_LOOP(r, nc) {
    const int s = sess_of(si, nsess, b.sel[r]);
    const MapDev m = maps[s];
    if (!m.ctr->fast) continue;
    const int rl = r - si[s].cre_lo;
    const int ftop = si[s].ftop0, bump = si[s].bump0;
    const int32_t vid = rl < ftop ? m.freelist[ftop - 1 - rl] : bump + (rl - ftop);
    b.co[si[s].off + rl] = make_int4(vid, b.cre_seg[r], b.cre_t[r], -1);
  }
}

// Chunks of the ring that k_ring_live compacts for a session: enough for the
// creations of the scan at a few stale entries per victim, within the
// session's two entries per point; the ordered pass reads the ring directly
// beyond them.
__device__ inline int64_t lv_chunks(const MapCounters* c, const SessInfo& S) {
  const int64_t used = c->ring_tail - c->ring_head;
  const int64_t want = int64_t(6) * (S.cre_hi - S.cre_lo) + 4096;
  return min(min((used + 63) >> 6, (want + 63) >> 6), (int64_t(2) * S.n) >> 6);
}

constexpr int kLvWaves = 4;   // waves per block of k_ring_live
constexpr int kLvBlocks = 32; // blocks per session

// Liveness of the ring entries ahead of the head (grid: kLvBlocks x
// sessions), one wave per 64-entry chunk, compacted within the chunk.  An entry is live while its voxel
// is alive, still carries the entry's stamp and has not been evicted in this
// scan; nothing the ordered pass changes makes another entry live.
__global__ void __launch_bounds__(64 * kLvWaves) k_ring_live(const MapDev* maps, const SessInfo* si, int nsess,
                                                              BatchBufs b) {
  const int s = blockIdx.y;
  if (s >= nsess) return;
  const MapDev m = maps[s];
  const MapCounters* c = m.ctr;
  if (c->fast) return;
  const SessInfo S = si[s];
  const int64_t nch = lv_chunks(c, S);
  const int64_t head = c->ring_head, tail = c->ring_tail;
  const int tag = c->tag;
  const int lane = threadIdx.x & 63;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int4* out0 = b.lv + int64_t(2) * S.off;
  for (int64_t k = int64_t(blockIdx.x) * kLvWaves + (threadIdx.x >> 6); k < nch;
       k += int64_t(gridDim.x) * kLvWaves) {
    const int64_t idx = head + k * 64 + lane;
    int32_t v = -1, first = kNever, seg = -1;
    bool live = false;
    if (idx < tail) {
      const int64_t ph = idx % m.qcap;
      v = m.ring_vid[ph];
      const uint64_t stp = m.ring_stamp[ph];
      const VoxState& st = m.vs[v];
      const bool cur = st.tag == tag;
      live = (st.flags & kAlive) && st.stamp == stp && !(cur && st.evict_t != kNever);
      first = (cur && st.create_t < 0) ? st.first_t : kNever;
      seg = st.seg;
    }
    const uint64_t bal = __ballot(live);
    if (live) out0[k * 64 + __popcll(bal & below)] = make_int4(v, first, seg, int32_t(k * 64 + lane));
    if (lane == 0) b.lv_cnt[lv_cnt_base(S.off, s) + k] = __popcll(bal);
  }
}

constexpr int kLpThreads = 1024;  // k_ring_pack block

// The chunks' live entries packed in ring order (one block per session):
// counts scanned per group of 1024 chunks, then one wave copies a chunk.
__global__ void __launch_bounds__(kLpThreads) k_ring_pack(const MapDev* maps, const SessInfo* si, int nsess,
                                                          BatchBufs b) {
  const int s = blockIdx.x;
  if (s >= nsess) return;
  const MapCounters* c = maps[s].ctr;
  if (c->fast) return;
  const SessInfo S = si[s];
  const int64_t nch = lv_chunks(c, S);
  const int4* src = b.lv + int64_t(2) * S.off;
  int4* dst = b.lv2 + int64_t(2) * S.off;
  const int32_t* cnt = b.lv_cnt + lv_cnt_base(S.off, s);
  __shared__ int32_t sh[kLpThreads];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int32_t base = 0;
  for (int64_t g0 = 0; g0 < nch; g0 += kLpThreads) {
    const int32_t cj = g0 + tid < nch ? cnt[g0 + tid] : 0;
    sh[tid] = cj;
    __syncthreads();
    for (int d = 1; d < kLpThreads; d <<= 1) {  // inclusive scan
      const int32_t v = tid >= d ? sh[tid - d] : 0;
      __syncthreads();
      sh[tid] += v;
      __syncthreads();
    }
    for (int q = wave; q < kLpThreads && g0 + q < nch; q += kLpThreads / 64) {
      const int32_t hi = sh[q], lo = q > 0 ? sh[q - 1] : 0;
      if (lane < hi - lo) dst[base + lo + lane] = src[(g0 + q) * 64 + lane];
    }
    base += sh[kLpThreads - 1];
    __syncthreads();
  }
  if (tid == 0) b.lv_n[s] = base;
}

// Ordered creation + eviction, one workgroup per session that may overflow.
// LRU order is the descending order of last-touch stamps; the ring holds
// (vid, stamp) in ascending stamp order with stale entries skipped lazily.
// The victim of each creation is the oldest live voxel not yet touched by an
// earlier point of this scan; a victim touched later comes back as a new
// voxel at its first touch (a creation event pushed on a min-heap).
//
// Wave 0 decides.  It holds 64-entry windows of the live ring entries and of
// the creation events, and 128 entries of the free list, in its lanes, and
// takes them from LDS stream buffers that wave 1 fills from global memory
// ahead of it, so the deciding wave waits on no global load while its own
// stores drain.  While the next creations precede every pending
// re-creation, up to 64 of them take their victims in one step.
constexpr int kStreamSlots = 8;  // chunks of each stream buffered in LDS
constexpr int kProdBurst = 4;    // chunks of each stream the producer loads per round trip
// Waits for every vector memory operation of the wave.  The decider's rare
// paths (direct ring reads, creations below capacity) end with it: their
// global loads are then complete on every path into the decision loop, so
// the loop's common path (LDS only) carries no wait on the scattered stores
// of earlier steps (stores and loads share one counter on this target).
__device__ __forceinline__ void drain_vm() { __builtin_amdgcn_s_waitcnt(0); }

__global__ void __launch_bounds__(128) k_create_evict(const MapDev* maps, const SessInfo* si, int nsess, BatchBufs b) {
  const int s = blockIdx.x;
  if (s >= nsess) return;
  const MapDev m = maps[s];
  MapCounters* c = m.ctr;
  if (c->fast) return;
  const int clo = si[s].cre_lo, ncre_ev = si[s].cre_hi - clo;
  const int off = si[s].off;
  const int64_t npre = lv_chunks(c, si[s]);
  const int nlive = b.lv_n[s];
  const int ftop0 = c->free_top;
  // streams: 0 packed live ring entries, 1 creation events (time, segment),
  // 2 the free list from its top down; 64 entries per chunk
  const int nwin = (nlive + 63) >> 6, ncch = (ncre_ev + 63) >> 6, nfch = (ftop0 + 63) >> 6;
  __shared__ uint64_t heap[kHeapCap];     // (first touch << 32) | evicted vid
  __shared__ int32_t heap_seg[kHeapCap];  // the key's segment in this scan
  __shared__ int4 s_live[kStreamSlots][64];
  __shared__ int32_t s_cre[kStreamSlots][128];
  __shared__ int32_t s_free[kStreamSlots][64];
  __shared__ int32_t s_seq[3];  // chunks produced per stream
  __shared__ int32_t s_use[3];  // first chunk per stream the decider still reads
  __shared__ int32_t s_done;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x < 3) {
    s_seq[threadIdx.x] = 0;
    s_use[threadIdx.x] = 0;
  }
  if (threadIdx.x == 0) s_done = 0;
  __syncthreads();

  if (threadIdx.x >= 64) {
    // ---- stream producer: keeps each stream up to kStreamSlots chunks
    //      ahead of the decider (skipping chunks it no longer needs), up to
    //      kProdBurst chunks of each stream per round trip, and stops when
    //      the decider is done or every chunk is out
    const int4* lvs = b.lv2 + int64_t(2) * off;
    const __attribute__((address_space(1))) int32_t* fl = as_global(m.freelist);
    int p0 = 0, p1 = 0, p2 = 0;
    while (__hip_atomic_load(&s_done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == 0) {
      const int u0 = __hip_atomic_load(&s_use[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      const int u1 = __hip_atomic_load(&s_use[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      const int u2 = __hip_atomic_load(&s_use[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      p0 = max(p0, u0);
      p1 = max(p1, u1);
      p2 = max(p2, u2);
      const int n0 = max(0, min(kProdBurst, min(nwin, u0 + kStreamSlots) - p0));
      const int n1 = max(0, min(kProdBurst, min(ncch, u1 + kStreamSlots) - p1));
      const int n2 = max(0, min(kProdBurst, min(nfch, u2 + kStreamSlots) - p2));
      int4 e[kProdBurst];
      int32_t ct[kProdBurst], cs[kProdBurst], fv[kProdBurst];
#pragma unroll
      for (int k = 0; k < kProdBurst; ++k) {  // every load of the round first
        const int i0 = 64 * (p0 + k) + lane, i1 = 64 * (p1 + k) + lane, i2 = 64 * (p2 + k) + lane;
        e[k] = make_int4(-1, kNever, -1, 0);
        ct[k] = kNever;
        cs[k] = -1;
        fv[k] = -1;
        if (k < n0 && i0 < nlive) e[k] = lvs[i0];
        if (k < n1 && i1 < ncre_ev) {
          ct[k] = b.cre_t[clo + i1];
          cs[k] = b.cre_seg[clo + i1];
        }
        if (k < n2 && i2 < ftop0) fv[k] = fl[ftop0 - 1 - i2];
      }
#pragma unroll
      for (int k = 0; k < kProdBurst; ++k) {
        if (k < n0) s_live[(p0 + k) % kStreamSlots][lane] = e[k];
        if (k < n1) {
          s_cre[(p1 + k) % kStreamSlots][lane] = ct[k];
          s_cre[(p1 + k) % kStreamSlots][64 + lane] = cs[k];
        }
        if (k < n2) s_free[(p2 + k) % kStreamSlots][lane] = fv[k];
      }
      p0 += n0;
      p1 += n1;
      p2 += n2;
      if (lane == 0) {
        __hip_atomic_store(&s_seq[0], p0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_store(&s_seq[1], p1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_store(&s_seq[2], p2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      if (p0 >= nwin && p1 >= ncch && p2 >= nfch) break;
      if (n0 + n1 + n2 == 0) __builtin_amdgcn_s_sleep(2);
    }
    return;
  }

  // ---- the decider (wave 0)
  const int tag = c->tag;
  const int64_t cap = m.capacity;
  int hn = 0;
  int64_t size = c->size;
  uint64_t gc = c->group_counter;
  int ftop = ftop0, bump = c->bump;
  int64_t ptr = c->ring_head;
  const int64_t tail = c->ring_tail;
  const int64_t head0 = ptr;
  // ring window: per lane an entry's vid, liveness, first touch in this scan
  // (kNever if untouched), segment, ring position - head0, and whether the
  // voxel state carries this tag; wend is the ring position after the
  // window.  The live entries of the first npre ring chunks come packed, 64
  // per window; the ring is read directly after them.
  int64_t wend = head0;
  int lw = 0;  // next packed window
  int32_t l_vid = -1, l_first = kNever, l_seg = -1, l_rel = 0;
  bool l_live = false, l_cur = false;
  // creation chunk cq in lanes; free-list chunks fq, fq + 1 (allocation
  // indices 64 fq ... 64 fq + 127, allocation a taking freelist[ftop0 - 1 - a])
  int cq = -1, fq = -1;
  int32_t c_t = kNever, c_seg = -1, f_lo = -1, f_hi = -1;
  int ci = 0, nco = 0, nev = 0;
  int err = 0;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));

  auto await = [&](int st, int need) {  // until chunk need - 1 of stream st is out
    while (__hip_atomic_load(&s_seq[st], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < need)
      __builtin_amdgcn_s_sleep(1);
  };
  auto release = [&](int st, int first) {
    if (lane == 0) __hip_atomic_store(&s_use[st], first, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  auto load_ring = [&]() {  // the next window, from ptr on
    if (lw < nwin) {
      release(0, lw);
      await(0, lw + 1);
      const int4 e = s_live[lw % kStreamSlots][lane];
      l_vid = e.x;
      l_first = e.y;
      l_seg = e.z;
      l_rel = e.w;
      l_live = e.x >= 0;
      l_cur = e.y != kNever;  // a live voxel carries this tag iff touched in this scan
      wend = head0 + __builtin_amdgcn_readlane(e.w, min(63, nlive - 64 * lw - 1)) + 1;
      ++lw;
      return;
    }
    const int64_t rbase = max(ptr, head0 + 64 * npre);
    ptr = rbase;
    wend = min<int64_t>(rbase + 64, tail);
    const int64_t idx = rbase + lane;
    l_vid = -1;
    l_live = false;
    l_cur = false;
    l_first = kNever;
    l_seg = -1;
    l_rel = int32_t(idx - head0);
    if (idx < tail) {
      const int64_t ph = idx % m.qcap;
      const int32_t v = *as_global(m.ring_vid + ph);
      const uint64_t stp = *as_global(m.ring_stamp + ph);
      const auto* st = as_global(m.vs + v);
      l_vid = v;
      l_cur = st->tag == tag;
      l_live = (st->flags & kAlive) && st->stamp == stp && !(l_cur && st->evict_t != kNever);
      l_first = (l_cur && st->create_t < 0) ? st->first_t : kNever;
      l_seg = st->seg;
    }
    drain_vm();
  };
  auto load_cre = [&]() {  // the chunk of creation ci
    cq = ci >> 6;
    release(1, cq);
    await(1, cq + 1);
    c_t = s_cre[cq % kStreamSlots][lane];
    c_seg = s_cre[cq % kStreamSlots][64 + lane];
  };
  // make the lanes hold the next k free-list entries (k <= 64, k <= ftop)
  auto load_free = [&](int k) {
    const int a = ftop0 - ftop;
    if (fq >= 0 && a >= 64 * fq && a + k <= 64 * fq + 128) return;
    fq = a >> 6;
    release(2, fq);
    await(2, min(fq + 2, nfch));
    f_lo = s_free[fq % kStreamSlots][lane];
    f_hi = fq + 1 < nfch ? s_free[(fq + 1) % kStreamSlots][lane] : -1;
  };
  auto free_at = [&](int a) {  // entry of allocation a (inside the window), per lane
    const int q = a - 64 * fq;
    const int32_t x = __shfl(f_lo, q & 63, 64), y = __shfl(f_hi, q & 63, 64);
    return q < 64 ? x : y;
  };
  auto free_at_u = [&](int a) {  // the same for an allocation uniform over the wave
    const int q = a - 64 * fq;
    return q < 64 ? __builtin_amdgcn_readlane(f_lo, q) : __builtin_amdgcn_readlane(f_hi, q - 64);
  };
  auto evict_write = [&](int32_t victim, bool cur, int t) {  // caller: one lane
    VoxState* st = m.vs + victim;
    if (cur) {
      *as_global(&st->evict_t) = t;
    } else {  // this scan's fields start here (three vector stores)
      st_global2(&st->merge_pos, kNever, -2);
      st_global4(&st->tag, tag, -1, t, -1);
      st_global4(&st->prev_vid, -1, kNever, -1, kNever);
    }
  };

  // pending re-creations, (first touch << 32) | evicted vid, in slots
  // [0, hn) of heap (segment in heap_seg); slot i belongs to lane i & 63,
  // which keeps the least key of its slots (lmin, at lpos), so the earliest
  // one is a wave minimum away
  int th = kNever;  // time of the earliest pending re-creation
  uint64_t lmin = ~0ull;
  int lpos = -1;
  auto lane_rescan = [&](bool me) {
    if (me) {
      lmin = ~0ull;
      lpos = -1;
      for (int i = lane; i < hn; i += 64) {
        const uint64_t k = heap[i];
        if (k < lmin) {
          lmin = k;
          lpos = i;
        }
      }
    }
  };
  // a re-creation at first touch f of the evicted voxel v (segment sg)
  auto heap_push = [&](int32_t f, int32_t v, int32_t sg) {
    const uint64_t key = (uint64_t(uint32_t(f)) << 32) | uint32_t(v);
    if (lane == (hn & 63)) {
      heap[hn] = key;
      heap_seg[hn] = sg;
      if (key < lmin) {
        lmin = key;
        lpos = hn;
      }
    }
    ++hn;
    th = min(th, f);
  };

  while (true) {
    if (ci < ncre_ev && (ci >> 6) != cq) load_cre();
    const int d = ci & 63;  // creation ci sits in lane d of its chunk
    const int t_s = ci < ncre_ev ? __builtin_amdgcn_readlane(c_t, d) : kNever;
    const int t_h = th;
    if (t_s == kNever && t_h == kNever) break;
    if (t_s < t_h) {
      // ---- bulk: creations that need no victim (below capacity, where no
      //      re-created key can be pending)
      if (size < cap && hn == 0) {
        const int mm = int(min<int64_t>(cap - size, int64_t(ncre_ev - ci)));
        for (int j0 = 0; j0 < mm; j0 += 64) {
          const int j = j0 + lane;
          if (j < mm) {
            const int fj = ftop - 1 - j;  // the same allocation order as one by one
            const int32_t vid = fj >= 0 ? *as_global(m.freelist + fj) : bump + (j - ftop);
            if (vid >= m.pool) err |= 8;
            b.co[off + nco + j] = make_int4(vid, b.cre_seg[clo + ci + j], b.cre_t[clo + ci + j], -1);
          }
        }
        drain_vm();
        if (__ballot(err != 0)) {
          err |= 8;
          break;
        }
        if (mm <= ftop) ftop -= mm;
        else {
          bump += mm - ftop;
          ftop = 0;
        }
        ci += mm;
        nco += mm;
        gc += uint64_t(mm);
        size += mm;
        continue;
      }
      // ---- bulk: creations before the next re-creation take their victims
      //      from the live ring entries ahead.  An entry first touched before
      //      t_s is passed over by all of them (creation times only grow);
      //      an untouched one is the victim of whichever creation reaches
      //      it; the first entry touched after t_s ("hard") ends the step,
      //      as the victim of the next creation if that comes before the
      //      entry's first touch (its key then returns at that touch).
      if (size >= cap && ptr < tail) {
        if (ptr >= wend) load_ring();
        const bool ahead = l_live && head0 + l_rel >= ptr;
        const uint64_t hard = __ballot(ahead && l_first != kNever && l_first > t_s);
        const int tpos = hard ? __ffsll((unsigned long long)hard) - 1 : 64;
        const uint64_t usable = __ballot(ahead && l_first == kNever) & (tpos >= 64 ? ~0ull : ((1ull << tpos) - 1));
        const int ncr = __popcll(__ballot(lane >= d && c_t < t_h));  // creations of this chunk before t_h
        const int nu = __popcll(usable);
        int mm = min(nu, ncr);
        int32_t hf = kNever, hv = -1, hs = -1;
        if (mm == nu && nu < ncr && tpos < 64) {
          hf = __builtin_amdgcn_readlane(l_first, tpos);
          if (__builtin_amdgcn_readlane(c_t, d + nu) < hf) {
            hv = __builtin_amdgcn_readlane(l_vid, tpos);
            hs = __builtin_amdgcn_readlane(l_seg, tpos);
            ++mm;
          }
        }
        const uint64_t vmask = usable | (hv >= 0 ? (1ull << tpos) : 0ull);
        if (mm > 0) {
          // the r-th entry of vmask is the victim of creation ci + r
          const bool mine = ((vmask >> lane) & 1ull) != 0;
          const int r = __popcll(vmask & below);
          const bool act = mine && r < mm;
          const int src = d + (act ? r : 0);
          const int t = __shfl(c_t, src, 64);
          const int sg = __shfl(c_seg, src, 64);
          const int nf = min(mm, ftop);  // victims' slots from the free list, then the bump
          if (nf > 0) load_free(nf);
          const int a0 = ftop0 - ftop;  // allocation index of creation ci
          const int32_t fv = nf > 0 ? free_at(a0 + min(act ? r : 0, nf - 1)) : -1;
          if (act) {
            const int32_t vid = r < ftop ? fv : bump + (r - ftop);
            if (vid >= m.pool) err |= 8;
            b.co[off + nco + r] = make_int4(vid, sg, t, -1);
            evict_write(l_vid, l_cur, t);
            b.ev_vid[off + nev + r] = l_vid;
            l_live = false;
          }
          if (__ballot(err != 0)) {
            err |= 8;
            break;
          }
          // the ring position after the mm-th victim (the lane of vmask's mm-th bit)
          const uint64_t lastb = __ballot(mine && r == mm - 1);
          ptr = head0 + __builtin_amdgcn_readlane(l_rel, __ffsll((unsigned long long)lastb) - 1) + 1;
          if (hv >= 0) {
            if (hn >= kHeapCap) {
              err |= 16;
              break;
            }
            heap_push(hf, hv, hs);
          }
          if (mm <= ftop) ftop -= mm;
          else {
            bump += mm - ftop;
            ftop = 0;
          }
          ci += mm;
          nco += mm;
          nev += mm;
          gc += uint64_t(mm);
          continue;
        }
        // no victim left in the window: the rest was touched before t_s
        ptr = wend;
        continue;
      }
    }
    // ---- single step: the earliest event, then its victim if over capacity
    int t, seg, prev;
    if (t_s < t_h) {
      t = t_s;
      seg = __builtin_amdgcn_readlane(c_seg, d);
      prev = -1;
      ++ci;
    } else {  // the earliest pending re-creation leaves its slot to the last one
      const uint64_t top = wave_min_u64(lmin);
      const int w = __ffsll((unsigned long long)__ballot(lmin == top)) - 1;  // keys are unique (vid)
      const int gp = __builtin_amdgcn_readlane(lpos, w);
      t = int(top >> 32);
      prev = int32_t(top & 0xffffffffu);  // evicted incarnation of this key
      seg = heap_seg[gp];
      const int last = hn - 1, lo = last & 63;
      wave_lds_sync();
      if (lane == lo && last != gp) {
        heap[gp] = heap[last];
        heap_seg[gp] = heap_seg[last];
      }
      wave_lds_sync();
      --hn;
      lane_rescan(lane == w || lane == lo);
      th = hn > 0 ? int(wave_min_u64(lmin) >> 32) : kNever;
    }
    int32_t vid;
    if (ftop > 0) {
      load_free(1);
      vid = free_at_u(ftop0 - ftop);
      --ftop;
    } else {
      vid = bump++;
    }
    if (vid >= m.pool) {
      err |= 8;
      break;
    }
    if (lane == 0) b.co[off + nco] = make_int4(vid, seg, t, prev);
    ++gc;
    ++nco;
    ++size;
    if (size > cap) {
      int32_t victim = -1, vfirst = kNever, vseg = -1;
      bool vcur = false;
      while (true) {
        if (ptr >= tail) {
          err |= 4;
          break;
        }
        if (ptr >= wend) load_ring();
        const uint64_t bal = __ballot(l_live && head0 + l_rel >= ptr && l_first > t);
        if (bal) {
          const int w = __ffsll((unsigned long long)bal) - 1;
          victim = __builtin_amdgcn_readlane(l_vid, w);
          vfirst = __builtin_amdgcn_readlane(l_first, w);
          vseg = __builtin_amdgcn_readlane(l_seg, w);
          vcur = __builtin_amdgcn_readlane(int(l_cur), w) != 0;
          ptr = head0 + __builtin_amdgcn_readlane(l_rel, w) + 1;
          if (lane == w) l_live = false;
          break;
        }
        ptr = wend;
      }
      if (victim < 0) break;
      if (lane == 0) {
        evict_write(victim, vcur, t);
        b.ev_vid[off + nev] = victim;
      }
      ++nev;
      --size;
      if (vfirst != kNever) {  // the key comes back later in this scan
        if (hn >= kHeapCap) {
          err |= 16;
          break;
        }
        heap_push(vfirst, victim, vseg);
      }
    }
  }
  if (lane == 0) {
    c->size = size;
    c->group_counter = gc;
    c->free_top = ftop;
    c->bump = bump;
    c->ring_head = ptr;
    c->ncre = nco;
    c->nevict = nev;
    if (err) atomicOr(&c->error, err);
    __hip_atomic_store(&s_done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// Initialise created voxels and point their hash slots at them.
__global__ void k_create_init(const MapDev* maps, const SessInfo* si, int nsess, BatchBufs b) {
  VMP_GRID_LOOP(p, b.ntot) {
    const int s = sess_of(si, nsess, p);
    const MapDev m = maps[s];
    const int r = p - si[s].off;
    const int ncre = m.ctr->ncre;
    if (r >= ncre) continue;
    const int4 co = b.co[p];
    const int32_t vid = co.x;
    const int q = co.y;
    if (vid < 0 || vid >= m.pool || q < 0) continue;  // error already flagged
    const int32_t e = b.seg_entry[q];
    VoxState st;
    st.key = m.table[e].key;
    st.group_id = m.ctr->group_counter - uint64_t(ncre) + uint64_t(r);
    st.stamp = 0;
    st.n = 0;
    st.newly_add = 0;
    st.flags = kEnable | kAlive;
    st.pad = 0;
    st.tag = m.ctr->tag;
    st.create_t = co.z;
    st.evict_t = kNever;
    st.freeze_t = -1;
    st.prev_vid = co.w;
    st.first_t = co.z;
    st.seg = q;
    st.merge_pos = kNever;
    st.dirty_t = -2;
    st.taint_t = kNever;
    m.vs[vid] = st;
    // the plane record and the moments are not cleared here: the voxel has
    // points in this scan, and the replay starts a voxel with n = 0 from zero
    VoxHot h{};
    h.flags = kEnable | kAlive;
    m.hot[vid] = h;
    m.table[e].vid = vid;
    b.seg_vid[q] = vid;
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
__device__ inline void mom_zero(Moments& M) {
  for (int a = 0; a < 3; ++a) M.s1[a] = 0.0;
  for (int a = 0; a < 6; ++a) M.s2[a] = 0.0;
  for (int a = 0; a < 6; ++a) M.c0[a] = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int l = 0; l < 6; ++l) M.c1[a][l] = 0.0;
  for (int a = 0; a < 6; ++a)
    for (int l = 0; l < 6; ++l) M.c2[a][l] = 0.0;
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
// A point's contribution to every moment is one fused multiply-add of two
// of 16 factors, F = (1, q, q q^T (rounded, xx xy xz yy yz zz), C): each
// moment c takes F[mom_i1(c)] * F[mom_i2(c)].  Both the per-lane walk and
// the per-wave walk accumulate exactly these terms in point order, so the
// two give the same bits.  The moments are this library's own intermediate,
// not a quantity the reference rounds in a particular order.
constexpr int kMomC = 69;  // moment components (S1q S2q S0C S1C S2C)
constexpr int kMomF = 16;  // factors per point
__host__ __device__ constexpr int mom_i1(int c) {
  return c < 3 ? 1 + c
               : (c < 9 ? 1 + (c - 3 < 3 ? 0 : (c - 3 < 5 ? 1 : 2))
                        : (c < 15 ? 10 + (c - 9) : (c < 33 ? 1 + (c - 15) / 6 : 4 + (c - 33) / 6)));
}
__host__ __device__ constexpr int mom_i2(int c) {
  return c < 3 ? 0
               : (c < 9 ? 1 + (c - 3 < 3 ? c - 3 : (c - 3 < 5 ? c - 5 : 2))
                        : (c < 15 ? 0 : 10 + (c < 33 ? (c - 15) % 6 : (c - 33) % 6)));
}
__device__ inline void mom_factors(const float q[3], const float C[6], double F[kMomF]) {
  F[0] = 1.0;
  for (int a = 0; a < 3; ++a) F[1 + a] = double(q[a]);
  for (int a = 0; a < 3; ++a)
    for (int bb = a; bb < 3; ++bb) F[4 + s6(a, bb)] = F[1 + a] * F[1 + bb];
  for (int l = 0; l < 6; ++l) F[10 + l] = double(C[l]);
}
__device__ inline void mom_add(Moments& M, const float q[3], const float C[6]) {
  double F[kMomF];
  mom_factors(q, C, F);
  for (int a = 0; a < 3; ++a) M.s1[a] = fma(F[mom_i1(a)], F[mom_i2(a)], M.s1[a]);
  for (int a = 0; a < 6; ++a) M.s2[a] = fma(F[mom_i1(3 + a)], F[mom_i2(3 + a)], M.s2[a]);
  for (int l = 0; l < 6; ++l) M.c0[l] = fma(F[mom_i1(9 + l)], F[mom_i2(9 + l)], M.c0[l]);
  for (int a = 0; a < 3; ++a)
    for (int l = 0; l < 6; ++l) M.c1[a][l] = fma(F[mom_i1(15 + 6 * a + l)], F[mom_i2(15 + 6 * a + l)], M.c1[a][l]);
  for (int a = 0; a < 6; ++a)
    for (int l = 0; l < 6; ++l) M.c2[a][l] = fma(F[mom_i1(33 + 6 * a + l)], F[mom_i2(33 + 6 * a + l)], M.c2[a][l]);
}

// updatePlane (voxel_map.cpp:97-136) from moments.  Plane::cov is kept as the
// traces of its two diagonal blocks: the normal block's trace
// sum_i tr(F_i C_i F_i^T) is a quadratic form of d_i = p_i - mean, rebuilt
// from sum q q C, sum q C, sum C; the centre block adds tr(C_i)/n^2.
struct RefitOut {
  double tn, tm;   // increments of the two traces
  double nrm[3];   // the oriented normal
  bool plane;
};
__device__ inline RefitOut refit_eval(const MapDev& m, const Moments& M, int n, const double center[3]) {
  RefitOut o;
  o.plane = false;
  const double invn = 1.0 / double(n);
  double mu[3];
  for (int a = 0; a < 3; ++a) mu[a] = M.s1[a] * invn;
  double cov[9];
  for (int a = 0; a < 3; ++a)
    for (int bb = 0; bb < 3; ++bb) cov[a * 3 + bb] = M.s2[s6(a, bb)] * invn - mu[a] * mu[bb];
  double ev[3], V[9];
  eigh3_dev(cov, ev, V);
  if (ev[0] > m.plane_thresh) return o;
  o.plane = true;
  // the covariance moments are read after the eigen solve, not hoisted
  // above it (they would hold 120 registers through the Jacobi sweeps)
  asm volatile("" ::: "memory");
  // With d_i = q_i - mu and T[ab][lm] = sum_i d_a d_b C_lm
  //   = S2C[ab][lm] - mu_a S1C[b][lm] - mu_b S1C[a][lm] + mu_a mu_b S0C[lm],
  // and W(w,x) the symmetric weights of w x^T (xx xy xz yy yz zz; the
  // off-diagonal ones w_a x_b + w_b x_a), the numerator for eigenvector e is
  //   sum_{ab,lm} T[ab][lm] (W(n,n)[lm] W(e,e)[ab] + 2 W(e,n)[lm] W(e,n)[ab]
  //                          + W(e,e)[lm] W(n,n)[ab]),
  // formed one row ab of T at a time (T is not stored).
  const double nn[3] = {V[0], V[3], V[6]};
  auto W = [](const double* w, const double* x, double w6[6]) {
    for (int l = 0; l < 3; ++l)
      for (int mm = l; mm < 3; ++mm) w6[s6(l, mm)] = l == mm ? w[l] * x[l] : w[l] * x[mm] + w[mm] * x[l];
  };
  double Wnn[6], We[2][6], Wen[2][6];
  W(nn, nn, Wnn);
  for (int k = 0; k < 2; ++k) {
    const double e[3] = {V[1 + k], V[4 + k], V[7 + k]};
    W(e, e, We[k]);
    W(e, nn, Wen[k]);
  }
  double num[2] = {0.0, 0.0};
  for (int a = 0; a < 3; ++a)
    for (int bb = a; bb < 3; ++bb) {
      const int ab = s6(a, bb);
      double r1 = 0.0, r2[2] = {0.0, 0.0}, r3[2] = {0.0, 0.0};
      for (int l = 0; l < 6; ++l) {
        const double t = M.c2[ab][l] - mu[a] * M.c1[bb][l] - mu[bb] * M.c1[a][l] + mu[a] * mu[bb] * M.c0[l];
        r1 += t * Wnn[l];
        for (int k = 0; k < 2; ++k) {
          r2[k] += t * Wen[k][l];
          r3[k] += t * We[k][l];
        }
      }
      for (int k = 0; k < 2; ++k) num[k] += We[k][ab] * r1 + 2.0 * Wen[k][ab] * r2[k] + Wnn[ab] * r3[k];
    }
  double tn_inc = 0.0;
  for (int k = 0; k < 2; ++k) {
    const double den = double(n) * (ev[0] - ev[1 + k]);
    tn_inc += num[k] / (den * den);
  }
  o.tn = tn_inc;
  o.tm = (M.c0[0] + M.c0[3] + M.c0[5]) * invn * invn;
  double mw[3];
  for (int a = 0; a < 3; ++a) mw[a] = center[a] + mu[a];
  double axis = -(mw[0] * nn[0] + mw[1] * nn[1] + mw[2] * nn[2]);
  const double sg = axis < 0.0 ? -1.0 : 1.0;
  for (int a = 0; a < 3; ++a) o.nrm[a] = sg * nn[a];
  return o;
}
// A refit's effect on the voxel: every refit initialises it; a plane refit
// adds to the two traces and replaces the normal, a non-plane one clears
// the plane flag.  A voxel's refits apply in walk order.
__device__ inline void refit_apply(const RefitOut& o, uint32_t& flags, double* pl) {
  flags |= kInit;
  if (!o.plane) {
    flags &= ~kPlane;
    return;
  }
  flags |= kPlane;
  pl[6] += o.tn;
  pl[7] += o.tm;
  for (int a = 0; a < 3; ++a) pl[3 + a] = o.nrm[a];
}
__device__ inline void refit(const MapDev& m, Moments& M, int n, uint32_t& flags, double* pl,
                             const double center[3], int& refits) {
  if (n < m.upd_pt) return;
  ++refits;
  refit_apply(refit_eval(m, M, n, center), flags, pl);
}

// Whether this scan's walk of a voxel reaches a refit (updatePlane with
// n >= update_size_thresh, voxel_map.cpp:42-95), from counts alone: an
// uninitialised voxel refits when n reaches the threshold, an initialised
// enabled one when newly_add does, before the walk stops at the freeze
// point (n >= max_point_thresh); a build refits every voxel with enough
// points once at the end.
__device__ inline bool walk_refits(int n0, int newly0, uint32_t flags, int len, int upd, int mx, bool build) {
  if (build) return n0 + len >= upd;
  if ((flags & kInit) && !(flags & kEnable)) return false;
  if (!(flags & kInit)) return len > upd - n0 - 1;
  const int jr = upd - newly0 - 1;                  // walk index of the next refit
  const int f = mx - n0 - 1 > 0 ? mx - n0 - 1 : 0;  // walk index of the freeze point
  return jr < (len < f + 1 ? len : f + 1);
}

// The whole walk of a voxel in this scan from counts alone (the state
// machine of the lane walk in k_replay_heavy, voxel_map.cpp:42-95): the
// walk index of the refit that initialises the voxel, then of the first
// refit of the initialised voxel (one every update_size_thresh points after
// it), of the point that freezes it, how many points the walk adds before a
// frozen voxel stops it, and newly_add after it.  A build adds every point
// and refits once at the end (not listed here).
struct WalkPlan {
  int walk;      // points the walk adds
  int init_j;    // refit that initialises the voxel (-1: none)
  int per_j;     // first refit of the initialised voxel (kNever: none)
  int freeze_j;  // point that freezes the voxel (-1: none)
  int newly;     // newly_add after the walk
  bool stopped;  // the segment's rest reaches a frozen voxel
};
__device__ inline WalkPlan walk_plan(int n0, int newly0, uint32_t flags, int len, int upd, int mx, bool build) {
  WalkPlan p{len, -1, kNever, -1, newly0, false};
  if (build) return p;
  int js = 0;  // first walk index on the initialised path
  if (!(flags & kInit)) {
    const int ji = max(0, upd - n0 - 1);  // n reaches update_size_thresh
    if (ji >= len) return p;
    p.init_j = ji;
    js = ji + 1;
  }
  if (!(flags & kEnable)) {  // frozen once initialised
    p.walk = min(len, js);
    p.stopped = js < len;
    return p;
  }
  const int jf = max(js, mx - n0 - 1);  // n reaches max_point_thresh
  p.walk = min(len, jf + 1);
  p.stopped = jf + 1 < len;
  if (jf < len) p.freeze_j = jf;
  if (p.walk > js) {
    const int j1 = js + max(0, upd - 1 - newly0);
    if (j1 < p.walk) {
      p.per_j = j1;
      p.newly = (p.walk - 1 - j1) % upd;
    } else {
      p.newly = newly0 + (p.walk - js);
    }
  }
  return p;
}

// Replay order: the voxels that refit first (the moment path), then the
// rest, each part by descending segment length so waves walk segments of
// similar length; slots past the segment count sort last.  Lengths are
// clamped below 2^lbits - 1; the class is bit lbits.  Also lists the
// refitting voxels as (vid, start, length, session) records for the
// per-wave replay: all of them from the front of hlist (cnt[5] counts them)
// and, from its back, those with at least long_seg points (cnt[6]; 0 lists
// none).  The two parts cannot meet: a long segment holds two points or
// more.  Both counts are cleared by k_seg_heads.
__global__ void __launch_bounds__(kMapThreads) k_seg_lenkeys(const MapDev* maps, const SessInfo* si, BatchBufs b,
                                                             int lbits, int long_seg) {
  constexpr int kW = kMapThreads / 64;
  __shared__ int woff[2][kW];
  __shared__ int bbase[2];
  const int nseg = b.cnt[0];
  const uint32_t top = uint32_t((uint64_t(1) << lbits) - 1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  VMP_UNIFORM_LOOP(q, valid, b.ntot) {
    uint32_t key = (1u << lbits) | top;
    int heavy = 0, lng = 0, start = 0, len = 0, s = 0;
    int32_t vid = -1;
    if (valid && q < nseg) {
      start = b.seg_start[q];
      len = b.seg_end[q] - start;
      vid = b.seg_vid[q];
      if (vid >= 0) {
        s = b.seg_sess[q];
        const MapDev m = maps[s];
        const VoxState& st = m.vs[vid];
        heavy = walk_refits(st.n, st.newly_add, st.flags, len, m.upd_pt, m.max_pt, si[s].is_build != 0) ? 1 : 0;
      }
      lng = heavy & int(long_seg > 0 && len >= long_seg);
      key = (uint32_t(1 - heavy) << lbits) | (top - min(uint32_t(len), top - 1));
    }
    if (valid) {
      b.len_in[q] = key;
      b.ord_in[q] = q;
      // the merge stage's per-point slots start empty (k_candidates fills them)
      b.cand_in[q] = ~0ull;
      b.cand_spec[q] = 0;
      int2* tn = reinterpret_cast<int2*>(b.tnbr + size_t(q) * 6);
      tn[0] = make_int2(0, 0);
      tn[1] = make_int2(0, 0);
      tn[2] = make_int2(0, 0);
    }
    const uint64_t bh = __ballot(heavy), bl = __ballot(lng);
    if (lane == 0) {
      woff[0][wv] = __popcll(bh);
      woff[1][wv] = __popcll(bl);
    }
    __syncthreads();
    if (threadIdx.x < 2) {  // the block's share of each list, wave by wave
      int run = 0;
      for (int w = 0; w < kW; ++w) {
        const int c = woff[threadIdx.x][w];
        woff[threadIdx.x][w] = run;
        run += c;
      }
      bbase[threadIdx.x] = run ? atomicAdd(&b.cnt[5 + threadIdx.x], run) : 0;
    }
    __syncthreads();
    if (heavy) {
      const uint64_t below = (uint64_t(1) << lane) - 1;
      const int4 rec = make_int4(vid, start, len, s);
      b.hlist[bbase[0] + woff[0][wv] + __popcll(bh & below)] = rec;
      if (lng) b.hlist[b.ntot - 1 - (bbase[1] + woff[1][wv] + __popcll(bl & below))] = rec;
    }
    __syncthreads();
  }
}

// The voxels that refit in this scan (the first cnt[5] of the replay
// order) replay their pending points and then their points of this scan in
// order through the pushPoint state machine, with FP64 moments and refits
// from them, in one of two forms that give the same bits:
//  - one wave per voxel (heavy_wave): lane c keeps moment c (lanes 0-4 also
//    moment 64 + c) and adds the points' terms in order from factors that
//    the wave stages in LDS 64 points at a time; at each refit point (known
//    from counts, walk_plan) it stores a snapshot of the moments, and up to
//    kSnap refits then run side by side, one per lane, applied in walk order.
//    The chain is a few memory round trips and one refit, whatever the
//    segment's length.
//  - one lane per voxel (heavy_lane): a walk per lane, the refits of the
//    wave's lanes taken together in rounds.  For steps with more than
//    wave_max refitting voxels, those whose segment has fewer than long_seg
//    points walk this way, so the FP64 refits run for 64 voxels at once.
constexpr int kHeavyChunk = 4;  // points of a heavy voxel's lane walk loaded together
constexpr int kLightChunk = 8;  // the same for the light walk
constexpr int kSnap = 16;       // refits evaluated together by heavy_wave
constexpr int kAccChunk = 8;    // points whose factors heavy_wave reads together
struct HeavyLds {
  double F[64][kMomF + 1];    // factors of 64 consecutive points of the walk (rows padded
                              // to an odd number of doubles: a lane per row, no bank conflicts)
  double snap[kSnap][kMomC];  // moments at refit points
  int snapn[kSnap];           // point count at each of those refits
};

__device__ inline double readlane_d(double v, int k) {
  const uint64_t u = uint64_t(__double_as_longlong(v));
  const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u)), k));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u >> 32)), k));
  return __longlong_as_double((long long)(uint64_t(lo) | (uint64_t(hi) << 32)));
}

__device__ inline void write_hot_pl(const MapDev& m, int32_t vid, uint32_t flags, const double* pl, int64_t kx,
                                    int64_t ky, int64_t kz) {
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

__device__ inline void heavy_lane(const MapDev& m, const SessInfo& S, const BatchBufs& b, int q, int32_t vid,
                                  int& refits, int& calls, int& dropped) {
  if (vid >= 0) {
    const int start = b.seg_start[q];
    const int end = b.seg_end[q];
    VoxState st = m.vs[vid];
    double* pl = m.plane + size_t(vid) * kPlaneD;
    double* mo = m.mom + size_t(vid) * kMomD;
    double plv[kPlaneD];
    Moments M;
    const int cnt = st.n > 0 ? int(mo[kMomPend]) : 0;
    if (st.n > 0) {
      for (int k = 0; k < kPlaneD; ++k) plv[k] = pl[k];
    } else {  // a new voxel (its records were not cleared at creation)
      for (int k = 0; k < kPlaneD; ++k) plv[k] = 0.0;
    }
    if (st.n - cnt > 0) mom_load(mo, M);
    else mom_zero(M);  // no point has reached the moments yet
    if (cnt > m.pend_cap) atomicOr(&m.ctr->error, 32);
    const float* pd = m.pend + size_t(vid) * m.pend_cap * kPendF;
    const int npd = min(cnt, m.pend_cap);
    for (int k0 = 0; k0 < npd; k0 += kHeavyChunk) {  // a chunk of pending points loaded together
      float v[kHeavyChunk][kPendF];
#pragma unroll
      for (int u = 0; u < kHeavyChunk; ++u) {
        const int k = min(k0 + u, npd - 1);
#pragma unroll
        for (int a = 0; a < kPendF; ++a) v[u][a] = pd[k * kPendF + a];
      }
#pragma unroll
      for (int u = 0; u < kHeavyChunk; ++u)
        if (k0 + u < npd) mom_add(M, v[u], v[u] + 3);
    }
    int64_t kx, ky, kz;
    unpack_key(st.key, kx, ky, kz);
    const double center[3] = {(double(kx) + 0.5) * m.vs64, (double(ky) + 0.5) * m.vs64,
                              (double(kz) + 0.5) * m.vs64};
    uint32_t flags = st.flags;
    int n = st.n, newly = st.newly_add;
    int merge_pos = kNever, freeze_t = -1;
    // Points are replayed in rounds: each lane advances to its next refit
    // (or the end of its segment), then the lanes that reached one refit
    // together, so the wave runs the FP64 refit once per round instead of
    // whenever any lane needs it.  Per lane the order of operations is the
    // reference's: add the point, refit, then the freeze check.
    int pos = start;
    bool active = true;
    while (__any(active)) {
      bool need = false;
      while (active && !need) {
        // the next kHeavyChunk points of the segment, loaded together
        // (contiguous in key order); a refit ends the chunk early
        float4 qa[kHeavyChunk], qb[kHeavyChunk], qd[kHeavyChunk];
#pragma unroll
        for (int u = 0; u < kHeavyChunk; ++u) {
          const size_t p3 = 3 * size_t(min(pos + u, end - 1));
          qa[u] = b.qc[p3];
          qb[u] = b.qc[p3 + 1];
          qd[u] = b.qc[p3 + 2];
        }
#pragma unroll
        for (int u = 0; u < kHeavyChunk; ++u) {
          if (!active || need) break;
          if (pos >= end) {
            active = false;
            break;
          }
          if ((flags & kInit) && !(flags & kEnable) && !S.is_build) {
            if (flags & kPlane) {
              merge_pos = pos - start;
              calls += end - pos;
            } else {
              dropped += end - pos;
            }
            active = false;
            break;
          }
          const int i = __float_as_int(qa[u].w);
          ++pos;
          const float qv[3] = {qa[u].x, qa[u].y, qa[u].z};
          const float C[6] = {qb[u].x, qb[u].y, qb[u].z, qb[u].w, qd[u].x, qd[u].y};
          mom_add(M, qv, C);
          ++n;
          if (S.is_build) continue;
          if (!(flags & kInit)) {
            // updatePlane after every point until the voxel initialises
            need = n >= m.upd_pt;
            continue;
          }
          ++newly;
          if (newly >= m.upd_pt) {
            need = true;
            newly = 0;
          }
          if (n >= m.max_pt) {
            flags &= ~kEnable;
            freeze_t = i;
          }
        }
      }
      if (need) refit(m, M, n, flags, plv, center, refits);
    }
    if (S.is_build) refit(m, M, n, flags, plv, center, refits);
    const double invn = 1.0 / double(n);
    for (int a = 0; a < 3; ++a) plv[a] = fma(M.s1[a], invn, center[a]);
    mom_store(mo, M);
    for (int a = 0; a < 3; ++a) mo[kMomRun + a] = M.s1[a];
    mo[kMomPend] = 0.0;
    for (int k = 0; k < kPlaneD; ++k) pl[k] = plv[k];
    st.flags = flags;
    st.n = n;
    st.newly_add = newly;
    st.freeze_t = freeze_t;
    st.merge_pos = merge_pos;
    m.vs[vid] = st;
    write_hot(m, vid, flags);
  }
}

// The ks refits whose snapshots are in L, lane k evaluating the k-th, then
// applied in order to the wave-uniform flags and plane record.
__device__ inline void heavy_flush(const MapDev& m, HeavyLds& L, int ks, const double center[3], uint32_t& flags,
                                   double plv[kPlaneD]) {
  wave_lds_sync();
  const int lane = threadIdx.x & 63;
  RefitOut o{};
  // the snapshot rows have the layout of Moments: the refit reads them from LDS
  if (lane < ks) o = refit_eval(m, *reinterpret_cast<const Moments*>(L.snap[lane]), L.snapn[lane], center);
  wave_lds_sync();
  for (int k = 0; k < ks; ++k) {
    RefitOut r;
    r.plane = __builtin_amdgcn_readlane(int(o.plane), k) != 0;
    r.tn = readlane_d(o.tn, k);
    r.tm = readlane_d(o.tm, k);
    for (int a = 0; a < 3; ++a) r.nrm[a] = readlane_d(o.nrm[a], k);
    refit_apply(r, flags, plv);
  }
}

// One voxel on one wave; the counts go to lane 0's refits/calls/dropped.
__device__ inline void heavy_wave(const MapDev& m, bool build, const BatchBufs& b, int32_t vid, int start, int len,
                                  HeavyLds& L, int& refits, int& calls, int& dropped) {
  const int lane = threadIdx.x & 63;
  // segment point `lane` and pending point `lane`, loaded with the voxel's
  // state (before the pending count is known) for the first 64 rows
  float4 s0a = make_float4(0.f, 0.f, 0.f, 0.f), s0b = s0a, s0d = s0a;
  if (lane < len) {
    const size_t p3 = 3 * size_t(start + lane);
    s0a = b.qc[p3];
    s0b = b.qc[p3 + 1];
    s0d = b.qc[p3 + 2];
  }
  const float* pd = m.pend + size_t(vid) * m.pend_cap * kPendF;
  float pv0[kPendF];
  for (int a = 0; a < kPendF; ++a) pv0[a] = lane < m.pend_cap ? pd[lane * kPendF + a] : 0.f;
  const VoxState st = m.vs[vid];
  double* pl = m.plane + size_t(vid) * kPlaneD;
  double* mo = m.mom + size_t(vid) * kMomD;
  const int cnt = st.n > 0 ? int(mo[kMomPend]) : 0;
  if (cnt > m.pend_cap && lane == 0) atomicOr(&m.ctr->error, 32);
  const int npd = min(cnt, m.pend_cap);
  // lane c holds moment c and, on lanes 0-4, moment 64 + c
  const bool has_mom = st.n - cnt > 0;  // some point has reached the moments
  const int c2 = 64 + min(lane, kMomC - 65);
  double Ma = has_mom ? mo[lane] : 0.0;
  double Mb = has_mom ? mo[c2] : 0.0;
  const int ia1 = mom_i1(lane), ia2 = mom_i2(lane), ib1 = mom_i1(c2), ib2 = mom_i2(c2);
  double plv[kPlaneD];
  for (int k = 0; k < kPlaneD; ++k) plv[k] = st.n > 0 ? pl[k] : 0.0;  // a new voxel's record is not cleared
  int64_t kx, ky, kz;
  unpack_key(st.key, kx, ky, kz);
  const double center[3] = {(double(kx) + 0.5) * m.vs64, (double(ky) + 0.5) * m.vs64, (double(kz) + 0.5) * m.vs64};
  const WalkPlan P = walk_plan(st.n, st.newly_add, st.flags, len, m.upd_pt, m.max_pt, build);
  uint32_t flags = st.flags;
  // the pending points, then the walk's points of this scan, as rows of F
  // 64 at a time
  const int total = npd + P.walk;
  int next = P.init_j >= 0 ? P.init_j : P.per_j;  // walk index of the next refit
  int ks = 0, nref = 0;
  if (lane < npd) mom_factors(pv0, pv0 + 3, L.F[lane]);
  if (lane < P.walk && npd + lane < 64) {
    const float qv[3] = {s0a.x, s0a.y, s0a.z};
    const float C[6] = {s0b.x, s0b.y, s0b.z, s0b.w, s0d.x, s0d.y};
    mom_factors(qv, C, L.F[npd + lane]);
  }
  for (int p0 = 0; p0 < total; p0 += 64) {
    const int x = p0 + lane;
    if (p0 > 0 && x < total) {
      float qv[3], C[6];
      if (x < npd) {
        for (int a = 0; a < 3; ++a) qv[a] = pd[x * kPendF + a];
        for (int l = 0; l < 6; ++l) C[l] = pd[x * kPendF + 3 + l];
      } else {
        const size_t p3 = 3 * size_t(start + x - npd);
        const float4 qa = b.qc[p3], qb = b.qc[p3 + 1], qd = b.qc[p3 + 2];
        qv[0] = qa.x, qv[1] = qa.y, qv[2] = qa.z;
        C[0] = qb.x, C[1] = qb.y, C[2] = qb.z, C[3] = qb.w, C[4] = qd.x, C[5] = qd.y;
      }
      mom_factors(qv, C, L.F[lane]);
    }
    wave_lds_sync();
    const int np = min(64, total - p0);
    for (int t0 = 0; t0 < np; t0 += kAccChunk) {
      // the refits among the chunk's points must fit beside the snapshots held
      int nc = 0;
      for (int nx = next, j0 = p0 + t0 - npd; nx < j0 + kAccChunk && nx < P.walk;
           nx = nx == P.init_j ? P.per_j : nx + m.upd_pt)
        ++nc;
      if (ks + nc > kSnap) {
        heavy_flush(m, L, ks, center, flags, plv);
        ks = 0;
      }
      double fa[kAccChunk][2], fb[kAccChunk][2];  // the chunk's factors, read together
#pragma unroll
      for (int u = 0; u < kAccChunk; ++u) {
        const double* f = L.F[min(t0 + u, np - 1)];
        fa[u][0] = f[ia1];
        fa[u][1] = f[ia2];
        fb[u][0] = f[ib1];
        fb[u][1] = f[ib2];
      }
#pragma unroll
      for (int u = 0; u < kAccChunk; ++u) {
        if (t0 + u >= np) break;
        Ma = fma(fa[u][0], fa[u][1], Ma);
        Mb = fma(fb[u][0], fb[u][1], Mb);
        const int j = p0 + t0 + u - npd;
        if (j == next) {
          if (st.n + j + 1 >= m.upd_pt) {  // (the lane walk's refit() is a no-op below the threshold)
            L.snap[ks][lane] = Ma;
            if (lane < kMomC - 64) L.snap[ks][64 + lane] = Mb;
            L.snapn[ks] = st.n + j + 1;
            ++ks;
            ++nref;
          }
          next = j == P.init_j ? P.per_j : j + m.upd_pt;
        }
      }
    }
    wave_lds_sync();
  }
  if (ks == kSnap) {
    heavy_flush(m, L, ks, center, flags, plv);
    ks = 0;
  }
  const int n = st.n + P.walk;
  if (build && n >= m.upd_pt) {  // a build refits once with all the points
    L.snap[ks][lane] = Ma;
    if (lane < kMomC - 64) L.snap[ks][64 + lane] = Mb;
    L.snapn[ks] = n;
    ++ks;
    ++nref;
  }
  if (ks) heavy_flush(m, L, ks, center, flags, plv);
  const double invn = 1.0 / double(n);
  for (int a = 0; a < 3; ++a) plv[a] = fma(readlane_d(Ma, a), invn, center[a]);
  if (P.freeze_j >= 0) flags &= ~kEnable;
  int merge_pos = kNever;
  if (P.stopped) {
    if (flags & kPlane) {
      merge_pos = P.walk;
      if (lane == 0) calls += len - P.walk;
    } else if (lane == 0) {
      dropped += len - P.walk;
    }
  }
  const int freeze_t = P.freeze_j >= 0 ? __float_as_int(b.qc[3 * size_t(start + P.freeze_j)].w) : -1;
  mo[lane] = Ma;
  if (lane < kMomC - 64) mo[64 + lane] = Mb;
  if (lane < 3) mo[kMomRun + lane] = Ma;
  if (lane == 0) mo[kMomPend] = 0.0;
  double pv = plv[0];
  for (int k = 1; k < kPlaneD; ++k) pv = lane == k ? plv[k] : pv;
  if (lane < kPlaneD) pl[lane] = pv;
  if (lane == 0) {
    VoxState& vs = m.vs[vid];
    vs.flags = flags;
    vs.n = n;
    vs.newly_add = P.newly;
    vs.freeze_t = freeze_t;
    vs.merge_pos = merge_pos;
    write_hot_pl(m, vid, flags, plv, kx, ky, kz);
    refits += nref;
  }
}

// One wave per voxel of hlist: every refitting voxel when there are at most
// wave_max of them, else the long ones (cnt[6], at hlist's back).  A
// one-session step reads its map's record without waiting for the list.
__global__ void __launch_bounds__(kMapThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_replay_heavy_wave(const MapDev* maps, const SessInfo* si, BatchBufs b, int wave_max) {
  __shared__ HeavyLds lds[kMapThreads / 64];
  constexpr int kWpb = kMapThreads / 64;
  const int nh = b.cnt[5];
  const bool all = nh <= wave_max;
  const int nw = all ? nh : b.cnt[6];
  const int wv = threadIdx.x >> 6;
  const MapDev m0 = maps[0];
  const bool build0 = si[0].is_build != 0;
  for (int u = blockIdx.x * kWpb + wv; u < nw; u += gridDim.x * kWpb) {
    int refits = 0, calls = 0, dropped = 0;
    const int4 r = b.hlist[all ? u : b.ntot - 1 - u];
    MapDev m;
    bool build;
    if (b.nsess == 1) {
      m = m0;
      build = build0;
    } else {
      m = maps[r.w];
      build = si[r.w].is_build != 0;
    }
    heavy_wave(m, build, b, r.x, r.y, r.z, lds[wv], refits, calls, dropped);
    wave_fetch_add(&m.ctr->refits, refits);
    wave_fetch_add(&m.ctr->merge_calls, calls);
    wave_fetch_add(&m.ctr->dropped, dropped);
  }
}

// The rest of the refitting voxels, one per lane.
__global__ void __launch_bounds__(kMapThreads) k_replay_heavy(const MapDev* maps, const SessInfo* si, BatchBufs b,
                                                              int wave_max) {
  const int nh = b.cnt[5];
  const int nw = nh <= wave_max ? nh : b.cnt[6];
  VMP_UNIFORM_LOOP(i, valid, nh - nw) {
    int refits = 0, calls = 0, dropped = 0;
    const int q = valid ? b.ord[nw + i] : 0;
    const int s = valid ? b.seg_sess[q] : 0;
    heavy_lane(maps[s], si[s], b, q, valid ? b.seg_vid[q] : -1, refits, calls, dropped);
    const MapDev& m = maps[s];
    wave_fetch_add(&m.ctr->refits, refits);
    wave_fetch_add(&m.ctr->merge_calls, calls);
    wave_fetch_add(&m.ctr->dropped, dropped);
  }
}

// The other voxels (from cnt[5] on): frozen ones count their points as
// merge() calls or drops; the rest add points without reaching a refit, so
// only the count, the running mean and the freeze state change, and each
// point (relative to the voxel centre, with its world cov) joins the pending
// list that the voxel's next refit folds into the moments.
__global__ void __launch_bounds__(kMapThreads) k_replay_light(const MapDev* maps, const SessInfo* si, BatchBufs b) {
  const int nseg = b.cnt[0], nh = b.cnt[5];
  // a one-session step reads its map's record once, before the voxels
  const MapDev m0 = maps[0];
  const int build0 = si[0].is_build;
  VMP_UNIFORM_LOOP(qi, valid, nseg - nh) {
    int calls = 0, dropped = 0;
    const int q = valid ? b.ord[nh + qi] : 0;
    const int32_t vid = valid ? b.seg_vid[q] : -1;
    const int start = b.seg_start[q], end = b.seg_end[q];
    MapDev m;
    int s, is_build;
    if (b.nsess == 1) {
      s = 0;
      m = m0;
      is_build = build0;
    } else {
      s = valid ? b.seg_sess[q] : 0;
      m = maps[s];
      is_build = si[s].is_build;
    }
    VoxState st = m.vs[vid < 0 ? 0 : vid];
    const bool frozen0 = vid >= 0 && !is_build && (st.flags & kInit) && !(st.flags & kEnable);
    if (frozen0) {
      // frozen before this scan: every point is a merge() call (plane) or dropped
      const int len = end - start;
      if (st.flags & kPlane) {
        m.vs[vid].merge_pos = 0;
        calls += len;
      } else {
        dropped += len;
      }
    } else if (vid >= 0) {
      double* pl = m.plane + size_t(vid) * kPlaneD;
      double nrm[3];  // the plane's normal, kept (zero for a new voxel)
      for (int a = 0; a < 3; ++a) nrm[a] = st.n > 0 ? pl[3 + a] : 0.0;
      int64_t kx, ky, kz;
      unpack_key(st.key, kx, ky, kz);
      const double center[3] = {(double(kx) + 0.5) * m.vs64, (double(ky) + 0.5) * m.vs64,
                                (double(kz) + 0.5) * m.vs64};
      double* mo = m.mom + size_t(vid) * kMomD;
      double s1[3] = {0.0, 0.0, 0.0};
      int cnt = 0;
      if (st.n > 0) {
        for (int a = 0; a < 3; ++a) s1[a] = mo[kMomRun + a];
        cnt = int(mo[kMomPend]);
      }
      float* pd = m.pend + size_t(vid) * m.pend_cap * kPendF;
      uint32_t flags = st.flags;
      int n = st.n, newly = st.newly_add;
      int merge_pos = kNever, freeze_t = -1;
      bool walking = true;
      for (int p0 = start; p0 < end && walking; p0 += kLightChunk) {
       // a chunk of the segment's points loaded together (contiguous in key order)
       float4 qa[kLightChunk], qb[kLightChunk], qd[kLightChunk];
#pragma unroll
       for (int u = 0; u < kLightChunk; ++u) {
         const size_t p3 = 3 * size_t(min(p0 + u, end - 1));
         qa[u] = b.qc[p3];
         qb[u] = b.qc[p3 + 1];
         qd[u] = b.qc[p3 + 2];
       }
#pragma unroll
       for (int u = 0; u < kLightChunk; ++u) {
        const int pos = p0 + u;
        if (pos >= end || !walking) break;
        if ((flags & kInit) && !(flags & kEnable) && !is_build) {
          if (flags & kPlane) {
            merge_pos = pos - start;
            calls += end - pos;
          } else {
            dropped += end - pos;
          }
          walking = false;
          break;
        }
        const int i = __float_as_int(qa[u].w);
        const float qv[3] = {qa[u].x, qa[u].y, qa[u].z};
        const float C[6] = {qb[u].x, qb[u].y, qb[u].z, qb[u].w, qd[u].x, qd[u].y};
        if (cnt < m.pend_cap) {
          for (int a = 0; a < 3; ++a) pd[cnt * kPendF + a] = qv[a];
          for (int a = 0; a < 6; ++a) pd[cnt * kPendF + 3 + a] = C[a];
        } else {
          atomicOr(&m.ctr->error, 32);
        }
        ++cnt;
        for (int a = 0; a < 3; ++a) s1[a] += double(qv[a]);
        ++n;
        if (is_build) continue;
        if (!(flags & kInit)) {
          if (n >= m.upd_pt) atomicOr(&m.ctr->error, 32);  // a refit: the other kernel's voxel
          continue;
        }
        ++newly;
        if (newly >= m.upd_pt) atomicOr(&m.ctr->error, 32);
        if (n >= m.max_pt) {
          flags &= ~kEnable;
          freeze_t = i;
        }
       }
      }
      for (int a = 0; a < 3; ++a) mo[kMomRun + a] = s1[a];
      mo[kMomPend] = double(cnt);
      const double invn = 1.0 / double(n);
      double plv[6];
      for (int a = 0; a < 3; ++a) plv[a] = center[a] + s1[a] * invn;
      for (int a = 0; a < 3; ++a) plv[3 + a] = nrm[a];
      for (int a = 0; a < 3; ++a) pl[a] = plv[a];
      if (st.n == 0)  // a new voxel: its plane record was not cleared at creation
        for (int k = 3; k < kPlaneD; ++k) pl[k] = 0.0;
      st.flags = flags;
      st.n = n;
      st.newly_add = newly;
      st.freeze_t = freeze_t;
      st.merge_pos = merge_pos;
      m.vs[vid] = st;
      write_hot_pl(m, vid, flags, plv, kx, ky, kz);
    }
    wave_fetch_add(&m.ctr->merge_calls, calls);
    wave_fetch_add(&m.ctr->dropped, dropped);
  }
}

}  // namespace vmp

#include "vmp_merge.hpp"

namespace vmp {

// ------------------------------------------------------------------ K9 ----
__global__ void k_finalize_segs(const MapDev* maps, const SessInfo* si, BatchBufs b) {
  const int nseg = b.cnt[0];
  VMP_GRID_LOOP(q, nseg) {
    const int32_t vid = b.seg_vid[q];
    if (vid < 0) continue;
    const int s = b.seg_sess[q];
    const MapDev m = maps[s];
    VoxState& st = m.vs[vid];
    const int last_t = b.sval_out[b.seg_end[q] - 1] - si[s].off;
    const bool calls = st.merge_pos != kNever;
    const bool frozen_plane = (st.flags & kPlane) && !(st.flags & kEnable) && (st.flags & kInit);
    // a dirtying event is resolved by a later (evaluated) merge() call
    const bool dirty = frozen_plane && st.dirty_t > -2 && !(calls && last_t > st.dirty_t);
    if (dirty) st.flags |= kDirty;
    else st.flags &= ~kDirty;
    st.stamp = uint64_t(m.ctr->point_base + last_t);
    b.flags[si[s].off + last_t] = vid + 1;  // LRU append marker (array cleared before)
  }
}

__global__ void k_evict_finalize(const MapDev* maps, const SessInfo* si, int nsess, BatchBufs b) {
  VMP_GRID_LOOP(p, b.ntot) {
    const int s = sess_of(si, nsess, p);
    const MapDev m = maps[s];
    const MapCounters* c = m.ctr;
    const int r = p - si[s].off;
    if (r >= c->nevict) continue;
    const int32_t v = b.ev_vid[p];
    VoxState& st = m.vs[v];
    const int32_t e = hash_find_slot(m.table, m.hmask, st.key);
    if (e >= 0 && m.table[e].vid == v) {
      m.table[e].vid = -1;
      m.table[e].key = kKeyTomb;
    }
    st.flags = 0;
    m.hot[v].flags = 0;
    m.freelist[c->free_top + r] = v;
  }
}

__global__ void k_lru_append(const MapDev* maps, const SessInfo* si, int nsess, BatchBufs b) {
  const int nsel = b.cnt[2];
  VMP_GRID_LOOP(r, nsel) {
    const int g = b.sel[r];
    const int s = sess_of(si, nsess, g);
    const MapDev m = maps[s];
    const int t = g - si[s].off;
    const int32_t vid = b.flags[g] - 1;
    const int64_t ph = (m.ctr->ring_tail + (r - si[s].lru_lo)) % m.qcap;
    m.ring_vid[ph] = vid;
    m.ring_stamp[ph] = uint64_t(m.ctr->point_base + t);
  }
}

// the scan's counter updates: the LRU appends, the points, and the evicted
// voxels that k_evict_finalize pushed above the free-list top (nothing
// after k_evict_finalize reads free_top or tombs; nevict stays 0 in a scan
// without points)
__global__ void k_end_scan(const MapDev* maps, const SessInfo* si, int nsess) {
  if (threadIdx.x != 0 || int(blockIdx.x) >= nsess) return;
  MapCounters* c = maps[blockIdx.x].ctr;
  c->ring_tail += c->nlru;
  c->point_base += si[blockIdx.x].n;
  c->free_top += c->nevict;
  c->tombs += c->nevict;
}

// start of each session's candidates in the sorted candidate list, and
// their total (cnt[3])
// (one workgroup: every session's count is loaded at once, then an
// exclusive scan in LDS, kCsThreads sessions per round)
constexpr int kCsThreads = 256;
__device__ inline void cand_starts_block(const MapDev* maps, int nsess, const BatchBufs& b) {
  __shared__ int sc[kCsThreads];
  __shared__ int carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  for (int s0 = 0; s0 < nsess; s0 += kCsThreads) {
    const int s = s0 + tid;
    const int v = s < nsess ? maps[s].ctr->ncand : 0;
    sc[tid] = v;
    __syncthreads();
    for (int o = 1; o < kCsThreads; o <<= 1) {  // inclusive Hillis-Steele scan
      const int u = tid >= o ? sc[tid - o] : 0;
      __syncthreads();
      sc[tid] += u;
      __syncthreads();
    }
    if (s < nsess) b.cand_start[s] = carry + sc[tid] - v;
    __syncthreads();
    if (tid == kCsThreads - 1) carry += sc[tid];
    __syncthreads();
  }
  if (tid == 0) b.cnt[3] = carry;
}
__global__ void __launch_bounds__(kCsThreads) k_cand_starts(const MapDev* maps, int nsess, BatchBufs b) {
  cand_starts_block(maps, nsess, b);
}
// the same as the tail of the candidate compaction (a block of kCsThreads)
struct TailCandStarts {
  const MapDev* maps;
  int nsess;
  BatchBufs b;
  static constexpr bool kActive = true;
  __device__ void at_tile(int, int, int, const int32_t (&)[kPrimIpt]) const {}
  __device__ void operator()(int) const { cand_starts_block(maps, nsess, b); }
};


// ---------------------------------------------------- housekeeping -----
// hash rebuild: every live voxel is re-inserted into a cleared table
__global__ void k_rehash_collect(MapDev m, int32_t* live, int32_t* nlive) {
  VMP_GRID_LOOP(v, m.pool) {
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
  VMP_GRID_LOOP(r, n) {
    const int32_t v = live[r];
    const int32_t e = hash_find_or_insert(m.table, m.hmask, m.vs[v].key, m.ctr);
    if (e >= 0) m.table[e].vid = v;
  }
}

// LRU ring compaction.  An entry stays when its voxel is alive and still
// carries the entry's stamp (older entries of a re-touched voxel are stale).
// Pass 1 linearises the circular range [head, head + used) into tv/ts and
// flags the survivors; a stable select gives their positions; pass 2 writes
// them back to the front of the ring, so ascending stamp order is kept.
__global__ void k_ring_mark(MapDev m, int64_t head, int64_t used, int32_t* tv, uint64_t* ts, int32_t* keep) {
  for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < used; r += int64_t(gridDim.x) * blockDim.x) {
    const int64_t ph = (head + r) % m.qcap;
    const int32_t v = m.ring_vid[ph];
    const uint64_t s = m.ring_stamp[ph];
    tv[r] = v;
    ts[r] = s;
    keep[r] = (v >= 0 && v < m.pool && (m.vs[v].flags & kAlive) && m.vs[v].stamp == s) ? 1 : 0;
  }
}
__global__ void k_ring_scatter(MapDev m, const int32_t* tv, const uint64_t* ts, const int32_t* sel, const int32_t* cnt) {
  const int n = *cnt;
  VMP_GRID_LOOP(j, n) {
    const int32_t r = sel[j];
    m.ring_vid[j] = tv[r];
    m.ring_stamp[j] = ts[r];
  }
}

}  // namespace vmp
