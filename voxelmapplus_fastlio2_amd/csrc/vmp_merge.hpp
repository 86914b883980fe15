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
//   k_candidates   first call after each (a)-event of the voxel or its
//                  neighbours, plus the first call of voxels still dirty from
//                  earlier scans
//   k_spec_eval    every candidate evaluated in parallel on the state before
//                  any merge of this scan: the eligible neighbours at its time
//                  (existence, incarnation, frozen-ness and plane-ness do not
//                  depend on merges) and whether it would merge
//   k_merge_replay one wave per session walks candidates in point order; a
//                  candidate is evaluated for real only if it would merge or
//                  a merge earlier in the scan touched its closed
//                  neighbourhood ("tainted"); each real merge adds the next
//                  call of every voxel whose neighbourhood it changed.
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

// first merge() call of segment q strictly after local point t (kNever if none)
__device__ inline int first_call_after(const BatchBufs& b, int q, int merge_pos, int t, int off) {
  const int end = b.seg_end[q];
  int lo = b.seg_start[q] + merge_pos, hi = end;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (b.sval_out[mid] - off > t) hi = mid;
    else lo = mid + 1;
  }
  return lo < end ? b.sval_out[lo] - off : kNever;
}

__device__ inline int32_t lookup_vid(const MapDev& m, uint64_t key) {
  const int32_t e = hash_find_slot(m.table, m.hmask, key);
  return e < 0 ? -1 : m.table[e].vid;
}

// neighbour j of a voxel if merge() may consider it at time t: it exists,
// is frozen (update_enable false), initialised and a plane; else -1
__device__ inline int32_t nbr_at(const MapDev& m, uint64_t vkey, int j, int tag, int t) {
  if (!key_neighbour_ok(vkey, j)) return -1;
  const int32_t nv = vid_at(m, lookup_vid(m, key_neighbour(vkey, j)), tag, t);
  if (nv < 0) return -1;
  const VoxState& ns = m.vs[nv];
  const bool frozen = !(ns.flags & kEnable) && !(ns.tag == tag && ns.freeze_t > t);
  return (frozen && (ns.flags & kPlane) && (ns.flags & kInit)) ? nv : -1;
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
  VMP_GRID_LOOP(q, nseg) {
    const int32_t vid = b.seg_vid[q];
    if (vid < 0) continue;
    const int s = b.seg_sess[q];
    const MapDev& m = maps[s];
    MapCounters* c = m.ctr;
    const int tag = c->tag, off = si[s].off, ns_pts = si[s].n;
    VoxState& st = m.vs[vid];
    int dirty = (st.flags & kDirty) ? -1 : -2;
    const bool own_freeze = st.freeze_t >= 0 && (st.flags & kPlane);
    if (own_freeze && st.freeze_t > dirty) dirty = st.freeze_t;
    const bool has_calls = st.merge_pos != kNever;
    int emitted[7];
    int ne = 0;
    auto emit = [&](int after) {
      const int t = first_call_after(b, q, st.merge_pos, after, off);
      if (t == kNever) return;
      for (int k = 0; k < ne; ++k)
        if (emitted[k] == t) return;  // one candidate per distinct call
      emitted[ne++] = t;
      const int slot = atomicAdd(&c->ncand, 1);
      if (slot < ns_pts) b.cand_in[off + slot] = (uint64_t(uint32_t(t)) << 32) | uint32_t(vid);
      else atomicOr(&c->error, 16);
    };
    if (has_calls && ((st.flags & kDirty) || own_freeze)) emit(-1);
    for (int j = 0; j < 6; ++j) {
      if (!key_neighbour_ok(st.key, j)) continue;
      const int32_t nv = lookup_vid(m, key_neighbour(st.key, j));
      if (nv < 0) continue;
      const VoxState& nbs = m.vs[nv];
      if (nbs.tag != tag || nbs.freeze_t < 0 || !(nbs.flags & kPlane)) continue;
      if (nbs.freeze_t > dirty) dirty = nbs.freeze_t;
      if (has_calls) emit(nbs.freeze_t);
    }
    st.dirty_t = dirty;
  }
}

// voxels not touched in this scan whose neighbour froze as a plane
__global__ void k_dirty_untouched(const MapDev* maps, BatchBufs b) {
  const int nseg = b.cnt[0];
  VMP_GRID_LOOP(q, nseg) {
    if (b.seg_vid[q] < 0) continue;
    const MapDev& m = maps[b.seg_sess[q]];
    const int tag = m.ctr->tag;
    const VoxState& st = m.vs[b.seg_vid[q]];
    if (st.freeze_t < 0 || !(st.flags & kPlane)) continue;
    for (int j = 0; j < 6; ++j) {
      if (!key_neighbour_ok(st.key, j)) continue;
      const int32_t nv = lookup_vid(m, key_neighbour(st.key, j));
      if (nv < 0) continue;
      VoxState& nbs = m.vs[nv];
      if (nbs.tag != tag && (nbs.flags & kAlive)) atomicOr(&nbs.flags, kDirty);
    }
  }
}

// ------------------------------------------------------------------ K8 ----
__global__ void k_spec_eval(const MapDev* maps, const SessInfo* si, int nsess, BatchBufs b) {
  VMP_GRID_LOOP(p, b.ntot) {
    const int s = sess_of(si, nsess, p);
    const MapDev& m = maps[s];
    if (p - si[s].off >= m.ctr->ncand) continue;
    const int tag = m.ctr->tag;
    const uint64_t cur = b.cand_out[p];
    const int t = int(cur >> 32);
    const int32_t V = int32_t(cur & 0xffffffffu);
    const uint64_t vkey = m.vs[V].key;
    const uint64_t gV = m.vs[V].group_id;
    double pv[kPlaneD];
    for (int k = 0; k < kPlaneD; ++k) pv[k] = m.plane[size_t(V) * kPlaneD + k];
    int spec = 0;
    for (int j = 0; j < 6; ++j) {
      const int32_t nv = nbr_at(m, vkey, j, tag, t);
      b.cand_nbr[size_t(p) * 6 + j] = nv;
      if (nv >= 0 && !spec && m.vs[nv].group_id != gV &&
          merge_gate(pv, m.plane + size_t(nv) * kPlaneD, m.merge_a, m.merge_d))
        spec = 1;
    }
    b.cand_spec[p] = spec;
  }
}

__device__ inline void heap_push(uint64_t* heap, int hn, uint64_t v) {
  int i = hn;
  while (i > 0) {
    const int p = (i - 1) >> 1;
    if (heap[p] <= v) break;
    heap[i] = heap[p];
    i = p;
  }
  heap[i] = v;
}
__device__ inline void heap_pop(uint64_t* heap, int hn) {
  const uint64_t lastv = heap[hn - 1];
  const int nn = hn - 1;
  int i = 0;
  while (true) {
    const int l = 2 * i + 1, r = l + 1;
    int sm = i;
    uint64_t sv = lastv;
    if (l < nn && heap[l] < sv) {
      sm = l;
      sv = heap[l];
    }
    if (r < nn && heap[r] < sv) {
      sm = r;
      sv = heap[r];
    }
    if (sm == i) break;
    heap[i] = heap[sm];
    i = sm;
  }
  if (nn > 0) heap[i] = lastv;
}

__global__ void __launch_bounds__(64) k_merge_replay(const MapDev* maps, const SessInfo* si, int nsess, BatchBufs b) {
  const int s = blockIdx.x;
  if (s >= nsess) return;
  const MapDev& m = maps[s];
  MapCounters* c = m.ctr;
  const int ncand = c->ncand, tag = c->tag, off = si[s].off;
  const uint64_t* cand = b.cand_out + off;
  const int32_t* cspec = b.cand_spec + off;
  const int32_t* cnbr = b.cand_nbr + size_t(off) * 6;
  __shared__ uint64_t heap[kHeapCap];
  __shared__ double pl[7][kPlaneD];
  __shared__ uint64_t grp[7];
  __shared__ int32_t nbv[7];
  __shared__ int32_t clos[8];
  __shared__ int nclos;
  const int lane = threadIdx.x;
  int hn = 0, ci = 0, evals = 0, merges = 0, err = 0;
  int64_t cbase = -1;  // static candidates [cbase, cbase + 64) held by the lanes
  uint64_t l_key = ~0ull;
  bool l_need = false;
  uint64_t last = ~0ull;
  while (true) {
    // ---- next static candidate that needs a real evaluation
    uint64_t st_key = ~0ull;
    int st_idx = -1;
    while (ci < ncand) {
      if (cbase < 0 || ci >= cbase + 64) {
        cbase = ci;
        const int idx = int(cbase) + lane;
        l_key = ~0ull;
        l_need = false;
        if (idx < ncand) {
          l_key = cand[idx];
          const int t = int(l_key >> 32);
          const VoxState& vst = m.vs[int32_t(l_key & 0xffffffffu)];
          l_need = cspec[idx] || vst.taint_t < t;
        }
      }
      const unsigned long long bal = __ballot(l_need && (cbase + lane >= ci));
      if (bal) {
        const int w = __ffsll(bal) - 1;
        st_key = __shfl(l_key, w, 64);
        st_idx = int(cbase) + w;
        break;
      }
      ci = int(cbase) + 64;
    }
    const uint64_t h_key = hn > 0 ? heap[0] : ~0ull;
    if (st_key == ~0ull && h_key == ~0ull) break;
    uint64_t cur;
    bool from_static;
    if (st_key <= h_key) {
      cur = st_key;
      from_static = true;
      ci = st_idx + 1;
    } else {
      cur = h_key;
      from_static = false;
      __syncthreads();
      if (lane == 0) heap_pop(heap, hn);
      __syncthreads();
      --hn;
    }
    if (cur == last) continue;
    last = cur;
    const int t = int(cur >> 32);
    const int32_t V = int32_t(cur & 0xffffffffu);
    ++evals;
    // ---- stage the six neighbours and the voxel itself in LDS
    if (lane < 6) {
      const int32_t nv = from_static ? cnbr[size_t(st_idx) * 6 + lane] : nbr_at(m, m.vs[V].key, lane, tag, t);
      nbv[lane] = nv;
      if (nv >= 0) {
        grp[lane] = m.vs[nv].group_id;
        for (int k = 0; k < kPlaneD; ++k) pl[lane][k] = m.plane[size_t(nv) * kPlaneD + k];
      }
    } else if (lane == 6) {
      nbv[6] = V;
      grp[6] = m.vs[V].group_id;
      for (int k = 0; k < kPlaneD; ++k) pl[6][k] = m.plane[size_t(V) * kPlaneD + k];
    }
    if (lane == 0) nclos = 0;
    __syncthreads();
    // ---- merge(): the six neighbours in order against the (updated) own plane
    if (lane == 0) {
      for (int j = 0; j < 6; ++j) {
        if (nbv[j] < 0 || grp[j] == grp[6]) continue;
        if (!merge_gate(pl[6], pl[j], m.merge_a, m.merge_d)) continue;
        merge_fuse(pl[6], pl[j]);
        grp[j] = grp[6];
        clos[nclos++] = j;
      }
      if (nclos > 0) clos[nclos++] = 6;
    }
    __syncthreads();
    const int nc0 = nclos;
    if (nc0 == 0) continue;
    // ---- write back the merged records
    if (lane < nc0) {
      const int j = clos[lane];
      const int32_t X = nbv[j];
      for (int k = 0; k < kPlaneD; ++k) m.plane[size_t(X) * kPlaneD + k] = pl[j][k];
      m.vs[X].group_id = grp[j];
      const uint32_t fl = m.vs[X].flags | kMerged;
      m.vs[X].flags = fl;
      write_hot(m, X, fl);
    }
    merges += nc0 - 1;
    __syncthreads();
    // ---- taint the closed neighbourhoods of every changed voxel and queue
    //      the next call of each touched one
    for (int q0 = 0; q0 < nc0 * 7; q0 += 64) {
      const int qq = q0 + lane;
      int32_t X = -1;
      if (qq < nc0 * 7) {
        const int32_t base_v = nbv[clos[qq / 7]];
        const int j = qq % 7;
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
          if (xs.dirty_t < t) xs.dirty_t = t;
          if (xs.taint_t > t) xs.taint_t = t;
          if (xs.merge_pos != kNever) tn = first_call_after(b, xs.seg, xs.merge_pos, t, off);
        } else if (xs.flags & kAlive) {
          atomicOr(&xs.flags, kDirty);
        }
      }
      for (int w = 0; w < 64; ++w) {
        const int tw = __shfl(tn, w, 64);
        const int32_t xw = __shfl(X, w, 64);
        if (tw == kNever) continue;
        if (hn >= kHeapCap) {
          err |= 16;
          continue;
        }
        if (lane == 0) heap_push(heap, hn, (uint64_t(uint32_t(tw)) << 32) | uint32_t(xw));
        ++hn;
      }
    }
    __syncthreads();
    cbase = -1;  // taints changed: re-read the static window
    if (err) break;
  }
  if (lane == 0) {
    c->merge_evals += evals;
    c->merges += merges;
    if (err) atomicOr(&c->error, err);
  }
}

}  // namespace vmp
