// Merge stages of the map update: VoxelGrid::merge (voxel_map.cpp:138-186)
// is called for every point that lands in a frozen plane voxel, reads the six
// face neighbours "at that moment" and may rewrite both planes and the
// neighbour's group id, so its calls are ordered by point index.
//
// Which calls can act:  a call is a no-op if the previous evaluated call on
// the same voxel was one and nothing in its closed neighbourhood changed
// since.  Changes are (a) a voxel freezing as a plane (known after the
// per-voxel replay) and (b) merges themselves.  Evictions only remove
// neighbours and new incarnations start unfrozen, so neither can turn an idle
// call into an acting one.
//   k_candidates   candidates = the first call after each (a)-event of the
//                  voxel or its neighbours, plus the first call of voxels
//                  still dirty from earlier scans; also caches each touched
//                  voxel's six neighbours and its candidate times
//   k_spec_eval    every candidate evaluated in parallel on the state before
//                  any merge of this scan: its eligible neighbours at its
//                  time (existence, incarnation, frozen-ness and plane-ness
//                  do not depend on merges) and whether it would merge
//   k_merge_replay one wave per session walks, in point order, the
//                  candidates that would merge plus the calls a real merge
//                  makes necessary: for every voxel whose closed
//                  neighbourhood a merge changed ("tainted"), its next call
//                  and all its later candidates, whose speculative result
//                  assumed the unchanged neighbourhood.
#pragma once

namespace vmp {

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

// first merge() call of segment q strictly after local point t (kNever if none).
// The segment's point times ascend; the search narrows the range by sixteen
// independent samples per step (one memory round trip each) instead of a
// binary search's chain of dependent loads.
__device__ inline int first_call_after(const BatchBufs& b, int q, int merge_pos, int t, int off) {
  const int end = b.seg_end[q];
  int lo = b.seg_start[q] + merge_pos, hi = end;  // answer in [lo, hi], hi = none
  const int tg = t + off;                          // compare global indices
  while (hi - lo > 16) {
    const int step = (hi - lo + 15) >> 4;
    int v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = b.sval_out[min(lo + (k + 1) * step - 1, hi - 1)];
    int k = 0;
    while (k < 16 && v[k] <= tg) ++k;
    if (k == 16) return kNever;  // the last sample is hi - 1
    const int nhi = min(lo + (k + 1) * step - 1, hi - 1);
    lo = k == 0 ? lo : lo + k * step;
    hi = nhi;  // sval_out[nhi] > tg: the answer is at most nhi
    if (lo >= hi) return b.sval_out[hi] - off;
    ++hi;      // search [lo, nhi] inclusive as the half-open [lo, nhi + 1)
  }
  int v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = lo + k < hi ? b.sval_out[lo + k] : 0x7fffffff;
  for (int k = 0; k < 16; ++k)
    if (lo + k < hi && v[k] > tg) return v[k] - off;
  return kNever;
}

__device__ inline int32_t lookup_vid(const MapDev& m, uint64_t key) {
  const int32_t e = hash_find_slot(m.table, m.hmask, key);
  return e < 0 ? -1 : m.table[e].vid;
}

// neighbour j of a voxel if merge() may consider it at time t: it exists,
// is frozen (update_enable false), initialised and a plane; else -1
__device__ inline int32_t eligible_at(const MapDev& m, int32_t cur, int tag, int t) {
  const int32_t nv = vid_at(m, cur, tag, t);
  if (nv < 0) return -1;
  const VoxState& ns = m.vs[nv];
  const bool frozen = !(ns.flags & kEnable) && !(ns.tag == tag && ns.freeze_t > t);
  return (frozen && (ns.flags & kPlane) && (ns.flags & kInit)) ? nv : -1;
}
__device__ inline int32_t nbr_at(const MapDev& m, uint64_t vkey, int j, int tag, int t) {
  if (!key_neighbour_ok(vkey, j)) return -1;
  return eligible_at(m, lookup_vid(m, key_neighbour(vkey, j)), tag, t);
}

// The six face neighbours' voxel ids (-1: none), as lookup_vid gives them;
// the six home-slot entries are loaded together, and only a collision
// continues its probe.
__device__ inline void lookup_nbrs6(const MapDev& m, uint64_t vkey, int32_t nv[6]) {
  uint64_t nk[6];
  uint32_t h[6];
  HashEntry e[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    nk[j] = key_neighbour_ok(vkey, j) ? key_neighbour(vkey, j) : kKeyEmpty;
    h[j] = hash_slot(nk[j], m.hmask);
    e[j] = m.table[h[j]];
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    int32_t v = -1;
    if (nk[j] != kKeyEmpty) {
      if (e[j].key == nk[j]) {
        v = e[j].vid;
      } else if (e[j].key != kKeyEmpty) {
        uint32_t hh = h[j];
        for (uint32_t probe = 1; probe <= m.hmask; ++probe) {
          hh = (hh + 1) & m.hmask;
          const HashEntry ee = m.table[hh];
          if (ee.key == nk[j]) {
            v = ee.vid;
            break;
          }
          if (ee.key == kKeyEmpty) break;
        }
      }
    }
    nv[j] = v;
  }
}

// eligible_at for six voxel ids, their records read in three rounds of six
// loads (the incarnation's creation fields, then the eligibility fields of
// the incarnation at time t) instead of a chain per neighbour
__device__ inline void eligible6(const MapDev& m, const int32_t cur[6], int tag, int t, int32_t out[6]) {
  int32_t v[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const VoxState& r = m.vs[cur[j] < 0 ? 0 : cur[j]];
    const int rt = r.tag, ct = r.create_t, pv = r.prev_vid;
    v[j] = cur[j] < 0 ? -1 : ((rt == tag && ct > t) ? pv : cur[j]);
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const VoxState& r = m.vs[v[j] < 0 ? 0 : v[j]];
    const int rt = r.tag, ev = r.evict_t, fz = r.freeze_t;
    const uint32_t fl = r.flags;
    const bool gone = v[j] < 0 || (rt == tag && ev <= t);
    const bool frozen = !(fl & kEnable) && !(rt == tag && fz > t);
    out[j] = (!gone && frozen && (fl & kPlane) && (fl & kInit)) ? v[j] : -1;
  }
}

// angle and distance gates (voxel_map.cpp:147-155); planes as kPlaneD records
__device__ inline bool merge_gate(const double* pv, const double* pn, double ma, double md) {
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
  return !(norm_distance > ma || axis_distance > md);
}

// trace-weighted fusion (voxel_map.cpp:157-183), written to both records.
// Operator precedence as in the reference: new_mean = tm0*m_n + (tm1*m_v)/(tm0+tm1)
// and likewise for the normal, which is not re-normalised.
__device__ inline void merge_fuse(double* pv, double* pn) {
  const double tn0 = pv[6], tm0 = pv[7], tn1 = pn[6], tm1 = pn[7];
  const double tc0 = tn0 + tm0, tc1 = tn1 + tm1;
  double nm[3], nrm[3];
  for (int k = 0; k < 3; ++k) {
    nm[k] = tm0 * pn[k] + (tm1 * pv[k]) / (tm0 + tm1);
    nrm[k] = tn0 * pn[3 + k] + (tn1 * pv[3 + k]) / (tn0 + tn1);
  }
  const double den = (tc0 + tc1) * (tc0 + tc1);
  const double ntn = (tc0 * tc0 * tn1 + tc1 * tc1 * tn0) / den;
  const double ntm = (tc0 * tc0 * tm1 + tc1 * tc1 * tm0) / den;
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
  pv[6] = pn[6] = ntn;
  pv[7] = pn[7] = ntm;
}

// ------------------------------------------------------------------ K7 ----
__global__ void k_candidates(const MapDev* maps, const SessInfo* si, BatchBufs b) {
  const int nseg = b.cnt[0];
  const MapDev m0 = maps[0];  // a one-session step reads its map's record once
  VMP_UNIFORM_LOOP(q, valid, nseg) {
    const int32_t vid = valid ? b.seg_vid[q] : -1;
    const int s = b.nsess == 1 || !valid ? 0 : b.seg_sess[q];
    MapDev m;
    if (b.nsess == 1) m = m0;
    else m = maps[s];
    MapCounters* c = m.ctr;
    int emitted[7];
    int ne = 0;
    if (valid) {
      b.seg_nct[q] = 0;
      if (vid < 0)
        for (int j = 0; j < 6; ++j) b.seg_nbr[q * 6 + j] = -1;
    }
    if (vid >= 0) {
      const int tag = c->tag, off = si[s].off;
      VoxState& st = m.vs[vid];
      int dirty = (st.flags & kDirty) ? -1 : -2;
      const bool own_freeze = st.freeze_t >= 0 && (st.flags & kPlane);
      if (own_freeze && st.freeze_t > dirty) dirty = st.freeze_t;
      const bool has_calls = st.merge_pos != kNever;
      auto emit = [&](int after) {
        const int t = first_call_after(b, q, st.merge_pos, after, off);
        if (t == kNever) return;
        for (int k = 0; k < ne; ++k)
          if (emitted[k] == t) return;  // one candidate per distinct call
        emitted[ne++] = t;
      };
      if (has_calls && ((st.flags & kDirty) || own_freeze)) emit(-1);
      // the six neighbours' entries, then their records, six loads at a time
      int32_t nvs[6];
      lookup_nbrs6(m, st.key, nvs);
      int ntag[6], nfrz[6];
      uint32_t nflg[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const VoxState& nbs = m.vs[nvs[j] < 0 ? 0 : nvs[j]];
        ntag[j] = nbs.tag;
        nfrz[j] = nbs.freeze_t;
        nflg[j] = nbs.flags;
      }
      for (int j = 0; j < 6; ++j) {
        const int32_t nv = nvs[j];
        b.seg_nbr[q * 6 + j] = nv;
        if (nv < 0) continue;
        if (ntag[j] != tag || nfrz[j] < 0 || !(nflg[j] & kPlane)) continue;
        if (nfrz[j] > dirty) dirty = nfrz[j];
        if (has_calls) emit(nfrz[j]);
      }
      st.dirty_t = dirty;
      // a frozen plane marks its untouched live neighbours dirty for later
      // scans (only their kDirty bit changes: this kernel reads kPlane and
      // kAlive of neighbours, and the own flags of touched voxels only)
      if (st.freeze_t >= 0 && (st.flags & kPlane))
        for (int j = 0; j < 6; ++j)
          if (nvs[j] >= 0 && ntag[j] != tag && (nflg[j] & kAlive)) atomicOr(&m.vs[nvs[j]].flags, kDirty);
    }
    // the session's candidate count, one atomic per wave and session; each
    // candidate goes to its session's offset plus its point time (a call
    // time belongs to one voxel), so the filled slots are in sorted order
    wave_fetch_add(&c->ncand, ne);
    if (ne > 0) {
      const int off = si[s].off, ns_pts = si[s].n;
      for (int k = 0; k < ne; ++k) {
        if (emitted[k] < ns_pts) b.cand_in[off + emitted[k]] = cand_key_batch(s, emitted[k], vid, b.cand_sshift);
        else atomicOr(&c->error, 16);
      }
      // candidate times, ascending (insertion sort of <= 7)
      for (int i = 1; i < ne; ++i) {
        const int v = emitted[i];
        int k = i - 1;
        while (k >= 0 && emitted[k] > v) {
          emitted[k + 1] = emitted[k];
          --k;
        }
        emitted[k + 1] = v;
      }
      for (int k = 0; k < ne; ++k) b.seg_ct[q * 7 + k] = emitted[k];
      b.seg_nct[q] = ne;
    }
  }
}

// the filled slots of the dense candidate array, and their compaction
struct PredCand {
  const uint64_t* c;
  __device__ int32_t operator()(int i) const { return c[i] != ~uint64_t(0) ? 1 : 0; }
};
struct OutCand {
  const uint64_t* in;
  uint64_t* out;
  __device__ void operator()(int k, int i, int32_t) const { out[k] = in[i]; }
};

// ------------------------------------------------------------------ K8 ----
__global__ void k_spec_eval(const MapDev* maps, BatchBufs b) {
  const int ntot = b.cnt[3];
  VMP_GRID_LOOP(p, ntot) {
    const uint64_t cur = b.cand_out[p];
    const int s = int(cur >> b.cand_sshift);
    const MapDev m = maps[s];
    const int tag = m.ctr->tag;
    const int t = cand_t(cand_local(cur, b.cand_sshift));
    const int32_t V = cand_vid(cur);
    const VoxState& vst = m.vs[V];
    const uint64_t gV = vst.group_id;
    double pv[kPlaneD];
    for (int k = 0; k < kPlaneD; ++k) pv[k] = m.plane[size_t(V) * kPlaneD + k];
    const bool cached = vst.tag == tag && vst.seg >= 0;
    // the six neighbours (cached ids of a voxel touched in this scan, else
    // looked up), their eligibility at t, then the gate, each step with the
    // six neighbours' loads together
    int32_t nc[6], nv[6];
    if (cached) {
#pragma unroll
      for (int j = 0; j < 6; ++j) nc[j] = b.seg_nbr[vst.seg * 6 + j];
    } else {
      lookup_nbrs6(m, vst.key, nc);
    }
    eligible6(m, nc, tag, t, nv);
    int spec = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      b.cand_nbr[size_t(p) * 6 + j] = nv[j];
      if (nv[j] >= 0 && m.vs[nv[j]].group_id != gV &&
          merge_gate(pv, m.plane + size_t(nv[j]) * kPlaneD, m.merge_a, m.merge_d))
        spec = 1;
    }
    b.cand_spec[p] = spec;
  }
}

// per-session range of the compacted acting candidates (cnt[4] of them)
// keys of the compacted acting candidates, contiguous (into cand_in, free
// after the candidate sort), so the replay loads a window's keys in one step
__global__ void k_spl_keys(BatchBufs b) {
  const int n = b.cnt[4];
  VMP_GRID_LOOP(k, n) b.cand_in[k] = cand_local(b.cand_out[b.spl[k]], b.cand_sshift);
}

__device__ inline void spl_range(const MapDev* maps, SessInfo* si, int s, const BatchBufs& b, int n) {
  auto lb = [&](int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (b.spl[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const int c0 = b.cand_start[s];
  si[s].spl_lo = lb(c0);
  si[s].spl_hi = lb(c0 + maps[s].ctr->ncand);
}
__global__ void k_spl_ranges(const MapDev* maps, SessInfo* si, int nsess, BatchBufs b) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsess) spl_range(maps, si, s, b, b.cnt[4]);
}
// k_spl_keys and k_spl_ranges as the output and the tail of the compaction
// of the acting candidates
struct OutSpl {
  int32_t* spl;
  uint64_t* keys;
  const uint64_t* cand;
  int sshift;
  __device__ void operator()(int k, int p, int32_t) const {
    spl[k] = p;
    keys[k] = cand_local(cand[p], sshift);
  }
};
struct TailSplRanges {
  const MapDev* maps;
  SessInfo* si;
  int nsess;
  BatchBufs b;
  static constexpr bool kActive = true;
  __device__ void at_tile(int, int, int, const int32_t (&)[kPrimIpt]) const {}
  __device__ void operator()(int n) const {
    for (int s = threadIdx.x; s < nsess; s += blockDim.x) spl_range(maps, si, s, b, n);
  }
};

// The merge-independent part of the taint phase of every acting candidate,
// in parallel ahead of the replay (see BatchBufs::pre_x).
constexpr int kPreItems = 49;  // 7 records x (6 neighbours + itself)
__global__ void k_spl_prefetch(const MapDev* maps, const SessInfo* si, BatchBufs b) {
  const int npre = min(b.cnt[4], b.pre_cap);
  VMP_GRID_LOOP(it, npre * kPreItems) {
    const int k = it / kPreItems, r = it % kPreItems, cj = r / 7, j = r % 7;
    const int pos = b.spl[k];
    const uint64_t cur = b.cand_out[pos];
    const int s = int(cur >> b.cand_sshift);
    const MapDev m = maps[s];
    const int tag = m.ctr->tag, off = si[s].off;
    const int t = cand_t(cand_local(cur, b.cand_sshift));
    const int32_t base = cj == 6 ? cand_vid(cur) : b.cand_nbr[size_t(pos) * 6 + cj];
    int32_t X = -1;
    if (base >= 0) {
      if (j == 6) {
        X = base;
      } else {
        const VoxState& bs = m.vs[base];
        int32_t cur_id;
        if (bs.tag == tag && bs.seg >= 0) cur_id = b.seg_nbr[bs.seg * 6 + j];
        else cur_id = key_neighbour_ok(bs.key, j) ? lookup_vid(m, key_neighbour(bs.key, j)) : -1;
        X = vid_at(m, cur_id, tag, t);
      }
    }
    int xseg = -1, tn = kNever;
    if (X >= 0) {
      const VoxState& xs = m.vs[X];
      if (xs.tag == tag && xs.seg >= 0 && xs.evict_t == kNever) {
        xseg = xs.seg;
        if (xs.merge_pos != kNever) tn = first_call_after(b, xs.seg, xs.merge_pos, t, off);
      }
    }
    b.pre_x[it] = X;
    b.pre_seg[it] = xseg;
    b.pre_tn[it] = tn;
  }
}

constexpr int kMrThreads = 256;  // threads of one session's merge replay
constexpr int kWin = 36;         // window entries (36 x 7 records <= 256 threads)
constexpr int kSPre = 2 * kWin;  // static candidates loaded per window
constexpr int kPoolCap = 2048;   // pending re-evaluations, kept sorted
constexpr int kNewCap = 2048;    // keys created by one window's taint phase
constexpr int kFxCap = 2048;     // deferred taint side effects of one window
constexpr int kTaintPer = 49;    // up to 7 changed voxels x their closed neighbourhood
constexpr int kFwCap = 512;      // first-writer hash: >= 2 x the kWin x 7 voxels one window writes
static_assert(kWin * 7 <= kMrThreads, "one thread per window record");
static_assert(2 * kWin * 7 <= kFwCap, "first-writer hash at most half full");
static_assert(kWin <= 64, "window bookkeeping runs in one wave");

__device__ inline uint32_t fw_hash(int32_t v) { return (uint32_t(v) * 2654435761u) >> (32 - 9); }

__device__ inline int rank_lt(const uint64_t* a, int n, uint64_t v) {  // elements < v
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ inline int rank_le(const uint64_t* a, int n, uint64_t v) {  // elements <= v
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One workgroup per session replays, in point order, the merge() calls that
// act.  Windowed speculation with in-order commit:
//   fill     the next kWin keys of the static list (acting candidates) merged
//            with the sorted pool of re-evaluations, duplicates dropped;
//   records  each entry's voxel and six eligible neighbours (group, plane);
//   evaluate each entry's merge() on its own copy of the current state;
//   taint    for every acting entry, in parallel, the re-evaluations its
//            merges make necessary (new keys and the voxels' dirty/taint
//            marks, held back until the entry commits);
//   commit   in point order.  An entry whose records, or (if it acts) the
//            records its taint phase read, were changed by an earlier commit
//            of the window, or that a new key must precede, stops the commit:
//            it and every later entry return to the pool.
// A committed entry therefore saw exactly the state the serial order gives
// it, and its taint phase read the same records a serial one would.
__global__ void __launch_bounds__(kMrThreads) k_merge_replay(const MapDev* maps, const SessInfo* si, int nsess,
                                                             BatchBufs b) {
  const int s = blockIdx.x;
  if (s >= nsess) return;
  const MapDev m = maps[s];
  MapCounters* c = m.ctr;
  const int tag = c->tag, off = si[s].off;
  const int spl_hi = si[s].spl_hi;
  const int tid = threadIdx.x;
  __shared__ uint64_t pool[2][kPoolCap];
  __shared__ uint64_t skey[kSPre];
  __shared__ int spos[kSPre];
  __shared__ uint64_t wkey[kWin];
  __shared__ int wpos[kWin];
  __shared__ int wspl[kWin];  // index of a static entry in spl (-1: pool entry)
  __shared__ double pl[kWin][7][kPlaneD];
  __shared__ uint64_t grp[kWin][7];
  __shared__ int32_t nbv[kWin][7];
  __shared__ int clos[kWin][8];
  __shared__ int nclos[kWin];
  __shared__ int act[kWin];
  __shared__ int32_t r2[kWin][kTaintPer];
  __shared__ uint64_t nkey[kNewCap];
  __shared__ int nkown[kNewCap];
  __shared__ int32_t fx_v[kFxCap];
  __shared__ int fx_wk[kFxCap];  // owner entry * 2 + kind (0 in-scan taint, 1 dirty flag)
  __shared__ uint64_t gbuf[kNewCap + kWin];
  __shared__ uint64_t sbuf[kNewCap + kWin];
  __shared__ int sh_i[8];
  __shared__ uint64_t mk[2 * kSPre];  // fill: statics and pool entries merged
  __shared__ int mpos[2 * kSPre];
  __shared__ int mspl[2 * kSPre];
  __shared__ int mval[2 * kSPre];
  __shared__ int mpre[2 * kSPre];
  __shared__ int32_t rseg[kWin][7];   // record voxels: segment of this scan (-1 none), key, flags
  __shared__ uint64_t rkey[kWin][7];
  __shared__ uint32_t rflg[kWin][7];
  __shared__ int sh_mg;
  __shared__ int sh_rew;  // statics of this window that are read again next window
  __shared__ uint64_t sh_minkey[kWin];
  __shared__ int sh_conf[kWin];
  __shared__ int32_t fw_key[kFwCap];  // voxels written in this window -> first writing entry
  __shared__ int32_t fw_min[kFwCap];
  int sp = si[s].spl_lo, pb = 0, ph = 0, pn = 0;
  int evals = 0, merges = 0, windows = 0, requeued = 0, max_acting = 0, err = 0;
  uint64_t last = ~0ull;
  while (true) {
    // ---- fill the window: the loaded statics and pool entries merged in
    //      parallel (merge path; statics first on equal keys), duplicates and
    //      keys beyond what is loaded dropped, the first kWin taken
    const int ns = min(kSPre, spl_hi - sp), np = min(kSPre, pn - ph);
    if (tid < ns) {
      skey[tid] = b.cand_in[sp + tid];
      spos[tid] = b.spl[sp + tid];
    }
    __syncthreads();
    const uint64_t* pseg = &pool[pb][ph];
    uint64_t lim = ~0ull;  // keys above an unloaded static or pool entry wait
    if (spl_hi - sp > kSPre) lim = skey[ns - 1];
    if (pn - ph > kSPre && pseg[np - 1] < lim) lim = pseg[np - 1];
    const int M = ns + np;
    if (tid < ns) {
      const int p = tid + rank_lt(pseg, np, skey[tid]);
      mk[p] = skey[tid];
      mpos[p] = spos[tid];
      mspl[p] = sp + tid;
    } else if (tid >= 128 && tid - 128 < np) {
      const int j = tid - 128;
      const int p = j + rank_le(skey, ns, pseg[j]);
      mk[p] = pseg[j];
      mpos[p] = -1;
      mspl[p] = -1;
    }
    __syncthreads();
    // valid: not over the limit, not a duplicate of the previous key; prefix
    // counts from wave ballots (M <= 2 * kSPre = 144 < 3 waves)
    {
      const int lane = tid & 63, wv = tid >> 6;
      const bool in = tid < M;
      const bool over = in && mk[tid] > lim;
      const bool ok = in && !over && mk[tid] != (tid == 0 ? last : mk[tid - 1]);
      const bool stat = in && !over && mpos[tid] >= 0;
      const uint64_t bok = __ballot(ok), bst = __ballot(stat);
      if (lane == 0 && wv < 3) {
        mpre[wv] = __popcll(bok);
        mval[wv] = __popcll(bst);
      }
      if (tid == 0) {
        sh_i[0] = 0x7fffffff;  // first position over the limit
      }
      __syncthreads();
      int base = 0, sbase = 0;
      for (int q = 0; q < wv && q < 3; ++q) {
        base += mpre[q];
        sbase += mval[q];
      }
      const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
      const int pre = base + __popcll(bok & below);
      const int spre = sbase + __popcll(bst & below);
      if (ok && pre < kWin) {
        wkey[pre] = mk[tid];
        wpos[pre] = mpos[tid];
        wspl[pre] = mspl[tid];
      }
      if (over && (tid == 0 || mk[tid - 1] <= lim)) atomicMin(&sh_i[0], tid);
      // the kWin-th taken key ends the consumed range
      if (ok && pre == kWin - 1) {
        sh_i[3] = tid + 1;
        sh_i[4] = spre + (stat ? 1 : 0);
      }
      __syncthreads();
      if (tid == 0) {
        const int total = mpre[0] + mpre[1] + mpre[2];
        const int statics = mval[0] + mval[1] + mval[2];
        int cut, used;
        if (total >= kWin) {
          cut = sh_i[3];
          used = sh_i[4];
        } else {
          cut = sh_i[0] < M ? sh_i[0] : M;  // statics counted only below the limit
          used = statics;
        }
        sh_i[0] = total < kWin ? total : kWin;
        sh_i[1] = used;
        sh_i[2] = ph + (cut - used);
      }
      __syncthreads();
    }
    const int n = sh_i[0];
    const int consumed = sh_i[1] + (sh_i[2] - ph);
    const int used = sh_i[1];
    sp += used;
    ph = sh_i[2];
    if (n == 0) {
      if (consumed == 0) break;  // nothing left
      continue;                  // only duplicates of the last key were loaded
    }
    ++windows;
    // ---- records: thread w*7+j, j = 6 is the voxel itself
    if (tid < n * 7) {
      const int w = tid / 7, j = tid % 7;
      const uint64_t cur = wkey[w];
      const int t = cand_t(cur);
      const int32_t V = cand_vid(cur);
      int32_t nv;
      if (j == 6) {
        nv = V;
      } else if (wpos[w] >= 0) {
        nv = b.cand_nbr[size_t(wpos[w]) * 6 + j];
      } else {  // a re-evaluation: the ids from its first evaluation in this scan, if any
        int32_t* slot = b.tnbr + (size_t(off) + size_t(t)) * 6 + j;
        const int32_t cached = *slot;
        if (cached != 0) {
          nv = cached - 2;
        } else {
          const VoxState& vst = m.vs[V];
          nv = (vst.tag == tag && vst.seg >= 0) ? eligible_at(m, b.seg_nbr[vst.seg * 6 + j], tag, t)
                                                : nbr_at(m, vst.key, j, tag, t);
          *slot = nv + 2;
        }
      }
      nbv[w][j] = nv;
      if (nv >= 0) {
        const VoxState& ns_ = m.vs[nv];
        grp[w][j] = ns_.group_id;
        rseg[w][j] = (ns_.tag == tag && ns_.seg >= 0) ? ns_.seg : -1;
        rkey[w][j] = ns_.key;
        rflg[w][j] = ns_.flags;
        for (int k = 0; k < kPlaneD; ++k) pl[w][j][k] = m.plane[size_t(nv) * kPlaneD + k];
      }
    }
    __syncthreads();
    // ---- merge(): each entry's six neighbours in order against its own
    //      plane, one thread per entry with its seven records in registers
    //      (an entry no neighbour passes changes nothing)
    if (tid < n) {
      const int w = tid;
      double pv[kPlaneD], pn[6][kPlaneD];
#pragma unroll
      for (int k = 0; k < kPlaneD; ++k) pv[k] = pl[w][6][k];
#pragma unroll
      for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int k = 0; k < kPlaneD; ++k) pn[j][k] = pl[w][j][k];
      const uint64_t g6 = grp[w][6];
      int nc = 0;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        if (nbv[w][j] < 0 || grp[w][j] == g6) continue;
        if (!merge_gate(pv, pn[j], m.merge_a, m.merge_d)) continue;
        merge_fuse(pv, pn[j]);
#pragma unroll
        for (int k = 0; k < kPlaneD; ++k) pl[w][j][k] = pn[j][k];
        grp[w][j] = g6;
        clos[w][nc++] = j;
      }
      if (nc > 0) {
#pragma unroll
        for (int k = 0; k < kPlaneD; ++k) pl[w][6][k] = pv[k];
        clos[w][nc++] = 6;
      }
      nclos[w] = nc;
    }
    __syncthreads();
    if (tid < 64) {  // acting entries in order (n <= kWin < 64: one wave)
      const bool a = tid < n && nclos[tid] > 0;
      const uint64_t bm = __ballot(a);
      const uint64_t below = tid == 0 ? 0ull : (~0ull >> (64 - tid));
      if (a) act[__popcll(bm & below)] = tid;
      if (tid == 0) {
        sh_i[3] = __popcll(bm);
        sh_i[4] = 0;  // new keys
        sh_i[5] = 0;  // side effects
      }
    }
    __syncthreads();
    const int na = sh_i[3];
    max_acting = max(max_acting, na);
    // ---- conflicts through the records: the first writer of every voxel the
    //      acting entries change, then for each entry whether one of its
    //      seven records has an earlier writer.  The commit stops at the
    //      first such entry at the latest, so only the acting entries before
    //      it run the taint phase.
    for (int i = tid; i < kFwCap; i += kMrThreads) {
      fw_key[i] = -1;
      fw_min[i] = kWin;
    }
    if (tid < kWin) {
      sh_minkey[tid] = ~0ull;
      sh_conf[tid] = 0;
    }
    __syncthreads();
    // na * 8 can exceed the block (kWin * 8 > kMrThreads): strided
    for (int q = tid; q < na * 8; q += kMrThreads) {
      const int a = q >> 3, ci = q & 7;
      const int k = act[a];
      if (ci < nclos[k]) {
        const int32_t X = nbv[k][clos[k][ci]];
        uint32_t h = fw_hash(X);
        while (true) {  // at most half full: the probe ends
          const int32_t old = atomicCAS(&fw_key[h], -1, X);
          if (old == -1 || old == X) {
            atomicMin(&fw_min[h], k);
            break;
          }
          h = (h + 1) & (kFwCap - 1);
        }
      }
    }
    __syncthreads();
    for (int q = tid; q < n * 7; q += kMrThreads) {
      const int w = q / 7;
      const int32_t v = nbv[w][q % 7];
      if (v < 0) continue;
      uint32_t h = fw_hash(v);
      while (true) {
        const int32_t kk = fw_key[h];
        if (kk == -1) break;
        if (kk == v) {
          if (fw_min[h] < w) sh_conf[w] = 1;
          break;
        }
        h = (h + 1) & (kFwCap - 1);
      }
    }
    __syncthreads();
    if (tid < 64) {  // acting entries before the first record conflict (act ascends)
      const uint64_t cm = __ballot(tid < n && sh_conf[tid]);
      const int pre = cm ? __ffsll((unsigned long long)cm) - 1 : n;
      const uint64_t am = __ballot(tid < na && act[tid] < pre);
      if (tid == 0) sh_i[6] = __popcll(am);
    }
    __syncthreads();
    const int nat = sh_i[6];
    // ---- taint phase of those acting entries: for each changed voxel and
    //      each voxel of its closed neighbourhood X: X re-evaluates its next
    //      call and its later candidates if X changed itself, or if X (a plane
    //      not in the group) passes the gates against a new record.  A voxel
    //      of the scan records its dirty/taint times, one outside is marked
    //      dirty.
    for (int T = tid; T < nat * kTaintPer; T += kMrThreads) {
      const int a = T / kTaintPer, r = T % kTaintPer, ci = r / 7, j = r % 7;
      const int w = act[a];
      const int nc0 = nclos[w];
      int32_t X = -1;
      if (ci < nc0) {
        const int t = cand_t(wkey[w]);
        const int cj = clos[w][ci];
        const int32_t base_v = nbv[w][cj];
        // a static entry's X, its segment and next call come precomputed
        const int pre = (wspl[w] >= 0 && wspl[w] < b.pre_cap) ? wspl[w] * kPreItems + cj * 7 + j : -1;
        int pre_seg = -1, pre_tn = kNever;
        X = base_v;
        if (pre >= 0) {
          X = b.pre_x[pre];
          pre_seg = b.pre_seg[pre];
          pre_tn = b.pre_tn[pre];
        } else if (j < 6) {
          const int bseg = rseg[w][cj];
          const uint64_t bkey = rkey[w][cj];
          int32_t cur_id;
          if (bseg >= 0) cur_id = b.seg_nbr[bseg * 6 + j];
          else cur_id = key_neighbour_ok(bkey, j) ? lookup_vid(m, key_neighbour(bkey, j)) : -1;
          X = vid_at(m, cur_id, tag, t);
        }
        bool need = false;
        if (X >= 0) {
          // what depends only on X (and, for a static entry, on its
          // precomputed segment) is loaded in one round; nothing here is
          // written before the commit
          const VoxState* xv = m.vs + X;
          const uint32_t xflags = xv->flags;
          const uint64_t xkey = xv->key, xgrp = xv->group_id;
          const int32_t xtag = xv->tag, xsg = xv->seg, xev = xv->evict_t, xmp = xv->merge_pos;
          double px[kPlaneD];
          for (int k = 0; k < kPlaneD; ++k) px[k] = m.plane[size_t(X) * kPlaneD + k];
          int32_t nct = 0, ct[7];
          if (pre_seg >= 0) {
            nct = b.seg_nct[pre_seg];
            for (int k = 0; k < 7; ++k) ct[k] = b.seg_ct[pre_seg * 7 + k];
          }
          for (int k = 0; k < nc0; ++k) need |= (nbv[w][clos[w][k]] == X);
          if (!need && (xflags & kPlane) && (xflags & kInit)) {
            int64_t ax, ay, az;
            unpack_key(xkey, ax, ay, az);
            for (int k = 0; k < nc0; ++k) {
              const int cc = clos[w][k];
              int64_t bx, by, bz;
              unpack_key(rkey[w][cc], bx, by, bz);
              const int64_t dist = (ax > bx ? ax - bx : bx - ax) + (ay > by ? ay - by : by - ay) +
                                   (az > bz ? az - bz : bz - az);
              if (dist == 1 && grp[w][cc] != xgrp && merge_gate(px, pl[w][cc], m.merge_a, m.merge_d)) need = true;
            }
          }
          if (need) {
            int xseg = pre_seg, tn = pre_tn;
            if (pre < 0) {
              xseg = (xtag == tag && xsg >= 0 && xev == kNever) ? xsg : -1;
              tn = (xseg >= 0 && xmp != kNever) ? first_call_after(b, xsg, xmp, t, off) : kNever;
              if (xseg >= 0) {
                nct = b.seg_nct[xseg];
                for (int k = 0; k < nct; ++k) ct[k] = b.seg_ct[xseg * 7 + k];
              }
            }
            if (xseg >= 0) {
              uint64_t loc[8];
              int nl = 0;
              if (tn != kNever) loc[nl++] = cand_key(0, tn, X);
              for (int k = 0; k < nct; ++k) {
                const int tc = ct[k];
                if (tc > t) loc[nl++] = cand_key(0, tc, X);
              }
              if (nl > 0) {
                const int at = atomicAdd(&sh_i[4], nl);
                for (int k = 0; k < nl; ++k)
                  if (at + k < kNewCap) {
                    nkey[at + k] = loc[k];
                    nkown[at + k] = w;
                  }
              }
              const int f = atomicAdd(&sh_i[5], 1);
              if (f < kFxCap) {
                fx_v[f] = X;
                fx_wk[f] = w * 2;
              }
            } else if (xflags & kAlive) {
              const int f = atomicAdd(&sh_i[5], 1);
              if (f < kFxCap) {
                fx_v[f] = X;
                fx_wk[f] = w * 2 + 1;
              }
            }
          }
        }
      }
      r2[w][r] = X;
    }
    __syncthreads();
    const int nnew = sh_i[4], nfx = sh_i[5];
    if (nnew > kNewCap || nfx > kFxCap) {
      err |= 16;
      break;
    }
    // ---- commit in point order, all at once: the first entry that read a
    //      voxel an earlier acting entry of the window writes (its seven
    //      records, found above, and if it acts the records of its taint
    //      phase), or that an earlier entry's new key must precede, stops the
    //      commit.  The entries before it commit together: their write sets
    //      are disjoint (an entry that writes a voxel also read it).
    for (int q = tid; q < nnew; q += kMrThreads)
      atomicMin(reinterpret_cast<unsigned long long*>(&sh_minkey[nkown[q]]), (unsigned long long)nkey[q]);
    for (int q = tid; q < nat * kTaintPer; q += kMrThreads) {  // the taint reads
      const int w = act[q / kTaintPer];
      const int32_t v = r2[w][q % kTaintPer];
      if (v < 0) continue;
      uint32_t h = fw_hash(v);
      while (true) {
        const int32_t kk = fw_key[h];
        if (kk == -1) break;
        if (kk == v) {
          if (fw_min[h] < w) sh_conf[w] = 1;
          break;
        }
        h = (h + 1) & (kFwCap - 1);
      }
    }
    __syncthreads();
    if (tid < 64) {  // one wave (n <= kWin < 64): exclusive prefix minimum of the new keys
      uint64_t pm = tid < n ? sh_minkey[tid] : ~0ull;
      for (int o = 1; o < 64; o <<= 1) {
        const uint64_t u = __shfl_up(pm, o, 64);
        if (tid >= o && u < pm) pm = u;
      }
      uint64_t before = __shfl_up(pm, 1, 64);
      if (tid == 0) before = ~0ull;
      const bool stop_here = tid < n && (sh_conf[tid] || before <= wkey[tid]);
      const uint64_t bm = __ballot(stop_here);
      const int stop = bm ? __ffsll((unsigned long long)bm) - 1 : n;
      int mg = (tid < stop && nclos[tid] > 0) ? nclos[tid] - 1 : 0;
      for (int o = 32; o > 0; o >>= 1) mg += __shfl_xor(mg, o, 64);
      if (tid == 0) {
        sh_i[7] = stop;
        sh_mg = mg;
        // the statics from the stop on are read again from the static list
        // (with their precomputed taint data) rather than sent to the pool;
        // consumed statics ascend, so they are those with key >= the stop key
        sh_rew = stop < n ? used - rank_lt(skey, used, wkey[stop]) : 0;
      }
    }
    __syncthreads();
    const int w = sh_i[7];
    sp -= sh_rew;
    merges += sh_mg;
    evals += w;
    for (int q = tid; q < w * 8; q += kMrThreads) {
      const int e = q >> 3, ci = q & 7;
      if (ci >= nclos[e]) continue;
      const int j = clos[e][ci];
      const int32_t X = nbv[e][j];
      for (int k = 0; k < kPlaneD; ++k) m.plane[size_t(X) * kPlaneD + k] = pl[e][j][k];
      m.vs[X].group_id = grp[e][j];
      // an atomic bit set: another entry of this commit may set kDirty on X
      atomicOr(&m.vs[X].flags, kMerged);
      // the observe record from the LDS copies (merges never change the
      // plane bit, which is all the observe pass reads of the flags)
      int64_t kx, ky, kz;
      unpack_key(rkey[e][j], kx, ky, kz);
      VoxHot h;
      h.mx = float(pl[e][j][0] - double(kx) * m.vs64);
      h.my = float(pl[e][j][1] - double(ky) * m.vs64);
      h.mz = float(pl[e][j][2] - double(kz) * m.vs64);
      h.flags = rflg[e][j] | kMerged;
      h.nx = float(pl[e][j][3]);
      h.ny = float(pl[e][j][4]);
      h.nz = float(pl[e][j][5]);
      h.pad = 0.f;
      m.hot[X] = h;
    }
    for (int f = tid; f < nfx; f += kMrThreads) {
      const int e = fx_wk[f] >> 1;
      if (e >= w) continue;
      VoxState& xs = m.vs[fx_v[f]];
      if (fx_wk[f] & 1) {
        atomicOr(&xs.flags, kDirty);
      } else {
        const int t = cand_t(wkey[e]);
        atomicMax(&xs.dirty_t, t);
        atomicMin(&xs.taint_t, t);
      }
    }
    requeued += n - w;
    // ---- pool := pool[ph, pn) merged with the committed entries' new keys and
    //      the entries sent back (duplicates are dropped at the next fill)
    if (tid == 0) sh_i[6] = 0;
    __syncthreads();
    for (int k = tid; k < nnew + n; k += kMrThreads) {
      uint64_t v;
      bool keep;
      if (k < nnew) {
        v = nkey[k];
        keep = nkown[k] < w;
      } else {
        v = wkey[k - nnew];
        keep = k - nnew >= w && wspl[k - nnew] < 0;  // a static sent back is read again
      }
      if (keep) gbuf[atomicAdd(&sh_i[6], 1)] = v;
    }
    __syncthreads();
    const int g = sh_i[6], rest = pn - ph;
    if (g + rest > kPoolCap) {
      err |= 16;
      break;
    }
    for (int k = tid; k < g; k += kMrThreads) {  // rank sort (ties by index)
      const uint64_t v = gbuf[k];
      int rk = 0;
      for (int q = 0; q < g; ++q) rk += (gbuf[q] < v) || (gbuf[q] == v && q < k);
      sbuf[rk] = v;
    }
    __syncthreads();
    const uint64_t* prest = &pool[pb][ph];
    uint64_t* dst = pool[1 - pb];
    for (int k = tid; k < g; k += kMrThreads) dst[k + rank_lt(prest, rest, sbuf[k])] = sbuf[k];
    for (int k = tid; k < rest; k += kMrThreads) dst[k + rank_le(sbuf, g, prest[k])] = prest[k];
    // duplicates of a key sent back must not be dropped against it: the last
    // processed key is the last committed one
    if (w > 0) last = wkey[w - 1];
    pb = 1 - pb;
    ph = 0;
    pn = g + rest;
    __syncthreads();
  }
  if (tid == 0) {
    c->merge_evals += evals;
    c->merges += merges;
    c->merge_windows += windows;
    c->merge_requeued += requeued;
    c->merge_max_acting = max(c->merge_max_acting, max_acting);
    if (err) atomicOr(&c->error, err);
  }
}

}  // namespace vmp
