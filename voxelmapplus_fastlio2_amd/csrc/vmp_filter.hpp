// Device kernels of the filter side: parameter builds, the motion prior, and
// the IESKF update loop (kf::IESKF::update, ieskf.cpp:125-156), one 64-lane
// block per session with the state staged in LDS.
#pragma once
#include "vmp_common.hpp"
#include "vmp_ieskf.hpp"
#include "vmp_observe.hpp"

namespace vmp {

__device__ inline void load_state(const vmp_state& src, vmp_state& dst, int lane) {
  const double* s = reinterpret_cast<const double*>(&src);
  double* d = reinterpret_cast<double*>(&dst);
  for (int i = lane; i < 36; i += 64) d[i] = s[i];
}

__global__ void k_obs_prep(FilterDev** fs) {
  FilterDev* f = fs[blockIdx.x];
  if (threadIdx.x == 0) make_obs_params(f->x, f->P, f->cfg, f->obs_prm);
}
__global__ void k_upd_prep(FilterDev** fs) {
  FilterDev* f = fs[blockIdx.x];
  if (threadIdx.x == 0) make_upd_params(f->x, f->P, f->cfg, f->upd_prm);
}

struct Diag23 {
  double q[23];
};
// pos += rot * dp ; rot = rot * dR ; P += diag(q)
__global__ void k_apply_motion(FilterDev* f, D33 dR, D3 dp, Diag23 q) {
  if (threadIdx.x != 0) return;
  const D33 rot = d33_from_cm(f->x.rot);
  const D3 a = d33_mulv(rot, dp);
  for (int i = 0; i < 3; ++i) f->x.pos[i] = f->x.pos[i] + a.v[i];
  d33_to_cm(d33_mul(rot, dR), f->x.rot);
  for (int i = 0; i < 23; ++i) f->P[i * 23 + i] += q.q[i];
}

// the same for many sessions: one thread per session
struct MotionRec {
  FilterDev* f;
  double dR[9];  // column-major
  double dp[3];
  double q[23];
};
__global__ void k_apply_motion_batch(const MotionRec* recs, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MotionRec& r = recs[i];
  FilterDev* f = r.f;
  const D33 rot = d33_from_cm(f->x.rot);
  const D3 dp{{r.dp[0], r.dp[1], r.dp[2]}};
  const D3 a = d33_mulv(rot, dp);
  for (int k = 0; k < 3; ++k) f->x.pos[k] = f->x.pos[k] + a.v[k];
  d33_to_cm(d33_mul(rot, d33_from_cm(r.dR)), f->x.rot);
  for (int k = 0; k < 23; ++k) f->P[k * 23 + k] += r.q[k];
}

// zero the per-session observe counters before a pass
__global__ void k_zero_lists(int32_t** lists, int n) {
  const int s = blockIdx.x;
  if (int(threadIdx.x) < n) lists[s][threadIdx.x] = 0;
}

// start of an update: predict_x = x, P^-1, first observe parameters.  The
// factor runs in the first wave's lanes; the inverse's product W^T W is
// spread over all the workgroup's waves.
constexpr int kBeginEndThreads = 256;
__global__ void __launch_bounds__(kBeginEndThreads) k_ieskf_begin(FilterDev** fs, int32_t** obs_ctr) {
  FilterDev* f = fs[blockIdx.x];
  if (threadIdx.x < 8) obs_ctr[blockIdx.x][threadIdx.x] = 0;  // counters of the first observe pass
  __shared__ vmp_state sx;
  const int lane = threadIdx.x;
  load_state(f->x, sx, lane);
  double a[23];  // row `lane` of P (column-major in memory), factored in lanes
#pragma unroll
  for (int c = 0; c < 23; ++c) a[c] = lane < 23 ? f->P[c * 23 + lane] : 0.0;
  double dinv;
  __shared__ double S[23 * 24], Pi[529];
  chol_rows(a, dinv, lane);
  chol_inverse_rows(a, dinv, Pi, lane, S);  // P^-1
  for (int i = lane; i < 529; i += kBeginEndThreads) as_global(f->Pinv)[i] = Pi[i];
  __syncthreads();
  load_state(sx, f->xp, lane);
  if (lane == 0) {
    f->iters = 0;
    f->active = 1;
    make_obs_params(sx, f->P, f->cfg, f->obs_prm);
  }
}

// one IESKF iteration after an observe pass (ieskf.cpp:133-149)
// The session's observe partial rows (one per observe block, jobs[session])
// are reduced here in a fixed order: row b is added by lane b mod 64, then
// the lanes are summed by the wave (H 12x12 column-major and b as the
// standalone reduction writes them).  The observe counters are read, then
// cleared for the next pass.
// Four waves per session: all of them read and add the partial rows (many
// loads in flight); the filter algebra then runs on the first wave's lanes
// (the other waves pass the same barriers and store nothing).
constexpr int kStepThreads = 256;
template <int NJ>
__global__ void __launch_bounds__(kStepThreads) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_ieskf_step(FilterDev** fs, const ObsJob* __restrict__ jobs, int32_t** obs_ctr, int32_t** stat_out) {
  constexpr int NH = NJ * (NJ + 1) / 2;
  constexpr int NACC = NH + NJ;
  constexpr int ROW = obs_row<NJ>();
  FilterDev* f = fs[blockIdx.x];
  if (!f->active) return;  // converged: later passes are no-ops
  __shared__ double Pi[529];
  __shared__ double dl[23];
  __shared__ vmp_state sx, sxp;
  __shared__ JacBlocks jb;
  __shared__ double hobs[156];
  __shared__ double red[ROW];
  __shared__ double gsum[kStepThreads / 32][32];
  // P's first six columns (the blocks the observe parameters take) and the
  // configuration, loaded with the rest up front for the final lane-0 work
  __shared__ double Pc[6 * 23];
  __shared__ vmp_map_cfg scfg;
  const int lane = threadIdx.x;  // lanes of the first wave carry the algebra
  load_state(f->x, sx, lane);
  load_state(f->xp, sxp, lane);
  for (int i = lane; i < 529; i += kStepThreads) Pi[i] = f->Pinv[i];
  if (lane < 6 * 23) Pc[lane] = f->P[lane];
  if (lane == 0) scfg = f->cfg;
  {
    const ObsJob& J = jobs[blockIdx.x];
    const auto part = as_global(J.partials);
    const int nb = J.nblocks;
    if constexpr (obs_stride<NJ>() == 32) {
      // thread t adds column t & 31 of rows g, g + 8, g + 16, ... (g = t >> 5)
      // in that order, kU loads in flight together (each wave load reads two
      // whole rows); then the eight groups are added in order
      constexpr int kU = 16, kG = kStepThreads / 32;
      const int col = lane & 31, grp = lane >> 5;
      double acc = 0.0;
      for (int r0 = grp; r0 < nb; r0 += kG * kU) {
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) v[u] = part[size_t(min(r0 + kG * u, nb - 1)) * 32 + col];
#pragma unroll
        for (int u = 0; u < kU; ++u)
          if (r0 + kG * u < nb) acc += v[u];
      }
      gsum[grp][col] = acc;
      __syncthreads();
      if (lane < ROW) {
        double t = gsum[0][lane];
#pragma unroll
        for (int g = 1; g < kG; ++g) t += gsum[g][lane];
        red[lane] = t;
      }
    } else if (lane < 64) {
      // lane l adds rows l, l + 64, l + 128, ... in that order, then the
      // wave adds the lanes
      double fin[ROW];
#pragma unroll
      for (int k = 0; k < ROW; ++k) fin[k] = 0.0;
      for (int b = lane; b < nb; b += 64) {
#pragma unroll
        for (int k = 0; k < ROW; ++k) fin[k] += part[size_t(b) * ROW + k];
      }
#pragma unroll
      for (int k = 0; k < ROW; ++k) {
        const double v = wave_sum0(fin[k]);
        if (lane == 0) red[k] = v;
      }
    }
  }
  __syncthreads();
  for (int i = lane; i < 144; i += kStepThreads) {  // H from the upper triangle
    const int r = i % 12, c = i / 12;
    double v = 0.0;
    if (r < NJ && c < NJ) {
      const int a = r < c ? r : c, b2 = r < c ? c : r;
      v = red[a * NJ - a * (a - 1) / 2 + (b2 - a)];
    }
    hobs[i] = v;
  }
  if (lane < 12) hobs[144 + lane] = lane < NJ ? red[NH + lane] : 0.0;
  __syncthreads();
  state_minus_par(sx, sxp, dl, lane);
  jac_blocks_par(sx, sxp, dl, jb, lane);
  // Lane r (< 23) builds row r of JtPi = J^T P^-1 and of H_ = JtPi J + [H_obs]
  // in registers (J is the identity outside three small blocks, so a block
  // row or column adds up at most three products), then factors H_ and
  // solves H_ delta = -b_ with the rows in lanes.
  double a[23];
  double br = 0.0;
#pragma unroll
  for (int c = 0; c < 23; ++c) a[c] = 0.0;
  if (lane < 23) {
    const int r = lane;
    double t[23];
    // JtPi[r][c] = sum_k J[k][r] Pi[k][c]; a block row sums its nonzero terms
    // in ascending k (the zero terms of the dense sum add nothing)
    if (r >= 3 && r < 6) {
#pragma unroll
      for (int c = 0; c < 23; ++c) {
        double s = jb.a[0 * 3 + (r - 3)] * Pi[3 * 23 + c];
        s += jb.a[1 * 3 + (r - 3)] * Pi[4 * 23 + c];
        s += jb.a[2 * 3 + (r - 3)] * Pi[5 * 23 + c];
        t[c] = s;
      }
    } else if (r >= 6 && r < 9) {
#pragma unroll
      for (int c = 0; c < 23; ++c) {
        double s = jb.b[0 * 3 + (r - 6)] * Pi[6 * 23 + c];
        s += jb.b[1 * 3 + (r - 6)] * Pi[7 * 23 + c];
        s += jb.b[2 * 3 + (r - 6)] * Pi[8 * 23 + c];
        t[c] = s;
      }
    } else if (r >= 21) {
#pragma unroll
      for (int c = 0; c < 23; ++c) {
        double s = jb.c[0 * 2 + (r - 21)] * Pi[21 * 23 + c];
        s += jb.c[1 * 2 + (r - 21)] * Pi[22 * 23 + c];
        t[c] = s;
      }
    } else {
#pragma unroll
      for (int c = 0; c < 23; ++c) t[c] = Pi[r * 23 + c];
    }
    // H_[r][c] = sum_k JtPi[r][k] J[k][c]
#pragma unroll
    for (int c = 0; c < 23; ++c) {
      double s;
      if (c >= 3 && c < 6) {
        s = t[3] * jb.a[0 * 3 + (c - 3)];
        s += t[4] * jb.a[1 * 3 + (c - 3)];
        s += t[5] * jb.a[2 * 3 + (c - 3)];
      } else if (c >= 6 && c < 9) {
        s = t[6] * jb.b[0 * 3 + (c - 6)];
        s += t[7] * jb.b[1 * 3 + (c - 6)];
        s += t[8] * jb.b[2 * 3 + (c - 6)];
      } else if (c >= 21) {
        s = t[21] * jb.c[0 * 2 + (c - 21)];
        s += t[22] * jb.c[1 * 2 + (c - 21)];
      } else {
        s = t[c];
      }
      if (r < 12 && c < 12) s += hobs[c * 12 + r];
      a[c] = s;
    }
    // b_[r] = JtPi[r] . delta + [b_obs]
    double s = t[0] * dl[0];
#pragma unroll
    for (int k = 1; k < 23; ++k) s += t[k] * dl[k];
    if (r < 12) s += hobs[144 + r];
    br = s;
  }
  double dinv;
  __syncthreads();  // Pi is no longer read: its LDS stages L from here on
  chol_rows(a, dinv, lane);
  const double xr = chol_solve_rows(a, dinv, br, lane, Pi);  // delta = -H_^-1 b_
  if (lane < 23) {
#pragma unroll
    for (int c = 0; c < 23; ++c) f->Hinv[lane * 23 + c] = a[c];  // Cholesky factor of the last H_
  }
  __syncthreads();  // every lane has read dl
  if (lane < 23) dl[lane] = -xr;
  __syncthreads();
  if (lane < 23) f->delta[lane] = dl[lane];
  state_plus_par(sx, dl, lane);
  load_state(sx, f->x, lane);
  if (lane == 0) {
    f->iters += 1;
    double mx = dl[0];
    for (int k = 1; k < 23; ++k) mx = mx > dl[k] ? mx : dl[k];
    if (mx < 0.001 || f->iters >= scfg.max_iter) f->active = 0;  // signed maxCoeff (ieskf.cpp:148)
    make_obs_params(sx, Pc, scfg, f->obs_prm);
    f->obs_prm.active = f->active;
    int32_t* so = stat_out[blockIdx.x];
    int32_t* oc = obs_ctr[blockIdx.x];
    if (f->iters == 1)
      for (int k = 4; k < 8; ++k) so[k] = 0;
    int32_t cnt[4];  // hits, plane hits, effective, fallbacks of this pass
    for (int k = 0; k < 4; ++k) cnt[k] = int32_t(red[NACC + k]);
    for (int k = 0; k < 4; ++k) so[k] = cnt[k];
    so[4] += cnt[0];  // hits over all passes of this update
    so[5] += cnt[1];
    so[6] += cnt[2];
    so[7] |= oc[4];  // hash probe overflow
    for (int k = 0; k < 8; ++k) oc[k] = 0;
  }
}

// P = L H_^-1 L^T (ieskf.cpp:151-155) and the insertion parameters; the
// inverse's product and the two J products on all the workgroup's waves
__global__ void __launch_bounds__(kBeginEndThreads) k_ieskf_end(FilterDev** fs) {
  FilterDev* f = fs[blockIdx.x];
  __shared__ double T[529], Hi[529];
  __shared__ double dl[23];
  __shared__ vmp_state sx, sxp;
  __shared__ JacBlocks jb;
  const int lane = threadIdx.x;
  load_state(f->x, sx, lane);
  load_state(f->xp, sxp, lane);
  // rows of the last H_'s Cholesky factor in lanes (its lower triangle)
  double a[23];
#pragma unroll
  for (int c = 0; c < 23; ++c) a[c] = lane < 23 && c <= lane ? f->Hinv[lane * 23 + c] : 0.0;
  double d = 1.0;
#pragma unroll
  for (int c = 0; c < 23; ++c)
    if (lane == c) d = a[c];
  if (lane < 23) dl[lane] = f->delta[lane];
  __shared__ double S[23 * 24];
  chol_inverse_rows(a, 1.0 / d, Hi, lane, S);  // H_^-1
  __syncthreads();
  jac_blocks_par(sx, sxp, dl, jb, lane);
  // P = L H_^-1 L^T with L = J; J is the identity outside three small
  // blocks, so a block row or column adds up at most three products (in
  // ascending k, as the dense sum would)
  for (int i = lane; i < 529; i += kBeginEndThreads) {  // T = J Hi
    const int r = i / 23, c = i % 23;
    double s;
    if (r >= 3 && r < 6) {
      s = jb.a[(r - 3) * 3 + 0] * Hi[3 * 23 + c];
      s += jb.a[(r - 3) * 3 + 1] * Hi[4 * 23 + c];
      s += jb.a[(r - 3) * 3 + 2] * Hi[5 * 23 + c];
    } else if (r >= 6 && r < 9) {
      s = jb.b[(r - 6) * 3 + 0] * Hi[6 * 23 + c];
      s += jb.b[(r - 6) * 3 + 1] * Hi[7 * 23 + c];
      s += jb.b[(r - 6) * 3 + 2] * Hi[8 * 23 + c];
    } else if (r >= 21) {
      s = jb.c[(r - 21) * 2 + 0] * Hi[21 * 23 + c];
      s += jb.c[(r - 21) * 2 + 1] * Hi[22 * 23 + c];
    } else {
      s = Hi[r * 23 + c];
    }
    T[i] = s;
  }
  __syncthreads();
  for (int i = lane; i < 529; i += kBeginEndThreads) {  // P = T J^T
    const int r = i / 23, c = i % 23;
    double s;
    if (c >= 3 && c < 6) {
      s = T[r * 23 + 3] * jb.a[(c - 3) * 3 + 0];
      s += T[r * 23 + 4] * jb.a[(c - 3) * 3 + 1];
      s += T[r * 23 + 5] * jb.a[(c - 3) * 3 + 2];
    } else if (c >= 6 && c < 9) {
      s = T[r * 23 + 6] * jb.b[(c - 6) * 3 + 0];
      s += T[r * 23 + 7] * jb.b[(c - 6) * 3 + 1];
      s += T[r * 23 + 8] * jb.b[(c - 6) * 3 + 2];
    } else if (c >= 21) {
      s = T[r * 23 + 21] * jb.c[(c - 21) * 2 + 0];
      s += T[r * 23 + 22] * jb.c[(c - 21) * 2 + 1];
    } else {
      s = T[r * 23 + c];
    }
    f->P[c * 23 + r] = s;  // column-major
  }
  __syncthreads();
  if (lane == 0) {
    make_upd_params(sx, f->P, f->cfg, f->upd_prm);
    f->active = 0;
  }
}

}  // namespace vmp
