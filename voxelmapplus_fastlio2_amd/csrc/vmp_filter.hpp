// Device kernels of the filter side: parameter builds, the motion prior, and
// the IESKF update loop (kf::IESKF::update, ieskf.cpp:125-156), one 64-lane
// block per session with the state staged in LDS.
#pragma once
#include "vmp_common.hpp"
#include "vmp_ieskf.hpp"

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

// start of an update: predict_x = x, P^-1, first observe parameters
__global__ void __launch_bounds__(64) k_ieskf_begin(FilterDev** fs) {
  FilterDev* f = fs[blockIdx.x];
  __shared__ double M[529], Inv[529];
  __shared__ vmp_state sx;
  const int lane = threadIdx.x;
  load_state(f->x, sx, lane);
  for (int i = lane; i < 529; i += 64) M[(i % 23) * 23 + i / 23] = f->P[i];  // column-major -> row-major
  __syncthreads();
  chol_factor(M, lane);
  chol_inverse(M, Inv, lane);
  for (int i = lane; i < 529; i += 64) f->Pinv[i] = Inv[i];
  load_state(sx, f->xp, lane);
  if (lane == 0) {
    f->iters = 0;
    f->active = 1;
    make_obs_params(sx, f->P, f->cfg, f->obs_prm);
  }
}

// one IESKF iteration after an observe pass (ieskf.cpp:133-149)
__global__ void __launch_bounds__(64) k_ieskf_step(FilterDev** fs, int32_t** obs_ctr, int32_t** stat_out) {
  FilterDev* f = fs[blockIdx.x];
  if (!f->active) return;
  __shared__ double Pi[529], H[529], JtPi[529];
  __shared__ double dl[23], bb[23], xs[23];
  __shared__ vmp_state sx, sxp;
  __shared__ JacBlocks jb;
  __shared__ double hobs[156];
  const int lane = threadIdx.x;
  load_state(f->x, sx, lane);
  load_state(f->xp, sxp, lane);
  for (int i = lane; i < 529; i += 64) Pi[i] = f->Pinv[i];
  for (int i = lane; i < 156; i += 64) hobs[i] = f->obs[i];
  __syncthreads();
  state_minus_par(sx, sxp, dl, lane);
  jac_blocks_par(sx, sxp, dl, jb, lane);
  // JtPi = J^T P^-1 (J is the identity outside three small blocks)
  for (int i = lane; i < 529; i += 64) {
    const int r = i / 23, c = i % 23;
    double s = jac_at(jb, 0, r) * Pi[c];
    for (int k = 1; k < 23; ++k) s += jac_at(jb, k, r) * Pi[k * 23 + c];
    JtPi[i] = s;
  }
  __syncthreads();
  // H_ = JtPi J + [H_obs] ; b_ = JtPi delta + [b_obs]
  for (int i = lane; i < 529; i += 64) {
    const int r = i / 23, c = i % 23;
    double s = JtPi[r * 23] * jac_at(jb, 0, c);
    for (int k = 1; k < 23; ++k) s += JtPi[r * 23 + k] * jac_at(jb, k, c);
    if (r < 12 && c < 12) s += hobs[c * 12 + r];
    H[i] = s;
  }
  if (lane < 23) {
    double s = JtPi[lane * 23] * dl[0];
    for (int k = 1; k < 23; ++k) s += JtPi[lane * 23 + k] * dl[k];
    if (lane < 12) s += hobs[144 + lane];
    bb[lane] = s;
  }
  __syncthreads();
  chol_factor(H, lane);
  chol_solve_vec(H, bb, xs, lane);  // delta = -H_^-1 b_
  if (lane < 23) dl[lane] = -xs[lane];
  for (int i = lane; i < 529; i += 64) f->Hinv[i] = H[i];  // Cholesky factor of the last H_
  __syncthreads();
  if (lane < 23) f->delta[lane] = dl[lane];
  state_plus_par(sx, dl, lane);
  load_state(sx, f->x, lane);
  if (lane == 0) {
    f->iters += 1;
    double mx = dl[0];
    for (int k = 1; k < 23; ++k) mx = mx > dl[k] ? mx : dl[k];
    if (mx < 0.001 || f->iters >= f->cfg.max_iter) f->active = 0;  // signed maxCoeff (ieskf.cpp:148)
    make_obs_params(sx, f->P, f->cfg, f->obs_prm);
    f->obs_prm.active = f->active;
    int32_t* so = stat_out[blockIdx.x];
    const int32_t* oc = obs_ctr[blockIdx.x];
    if (f->iters == 1)
      for (int k = 4; k < 8; ++k) so[k] = 0;
    for (int k = 0; k < 4; ++k) so[k] = oc[k];
    so[4] += oc[0];  // hits over all passes of this update
    so[5] += oc[1];
    so[6] += oc[2];
    so[7] |= oc[4];  // hash probe overflow
  }
}

// P = L H_^-1 L^T (ieskf.cpp:151-155) and the insertion parameters
__global__ void __launch_bounds__(64) k_ieskf_end(FilterDev** fs) {
  FilterDev* f = fs[blockIdx.x];
  __shared__ double Lc[529], T[529], Hi[529];
  __shared__ double dl[23];
  __shared__ vmp_state sx, sxp;
  __shared__ JacBlocks jb;
  const int lane = threadIdx.x;
  load_state(f->x, sx, lane);
  load_state(f->xp, sxp, lane);
  for (int i = lane; i < 529; i += 64) Lc[i] = f->Hinv[i];
  if (lane < 23) dl[lane] = f->delta[lane];
  __syncthreads();
  chol_inverse(Lc, Hi, lane);
  jac_blocks_par(sx, sxp, dl, jb, lane);
  for (int i = lane; i < 529; i += 64) {
    const int r = i / 23, c = i % 23;
    double s = jac_at(jb, r, 0) * Hi[c];
    for (int k = 1; k < 23; ++k) s += jac_at(jb, r, k) * Hi[k * 23 + c];
    T[i] = s;
  }
  __syncthreads();
  for (int i = lane; i < 529; i += 64) {
    const int r = i / 23, c = i % 23;
    double s = T[r * 23] * jac_at(jb, c, 0);
    for (int k = 1; k < 23; ++k) s += T[r * 23 + k] * jac_at(jb, c, k);
    f->P[c * 23 + r] = s;  // column-major
  }
  __syncthreads();
  if (lane == 0) {
    make_upd_params(sx, f->P, f->cfg, f->upd_prm);
    f->active = 0;
  }
}

}  // namespace vmp
