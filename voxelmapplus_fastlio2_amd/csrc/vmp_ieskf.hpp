// Iterated error-state Kalman filter update on the device (kf::IESKF::update,
// ieskf.cpp:125-156) — one 64-lane wave per session, FP64, 23x23 algebra on
// rows held in lanes (and in LDS).  The measurement function is the observe
// kernel; each iteration is
//   observe -> k_ieskf_step (reduction of the observe partial rows, H_, b_
//              assembly, Cholesky solve, x += dx),
// and once converged (signed max(dx) < 1e-3, ieskf.cpp:148) or after max_iter
// passes the session is marked inactive so later passes are no-ops, which is
// what the reference's `break` does.  P = L H_^-1 L^T follows with L built
// from the last increment (ieskf.cpp:151-155).
#pragma once
#include "vmp_common.hpp"

namespace vmp {

constexpr double kGravity = 9.81;

struct FilterDev {
  vmp_state x, xp;
  double P[529];     // column-major (API layout)
  double Pinv[529];  // row-major
  double Hinv[529];  // row-major, last iteration
  double delta[23];  // last increment
  double obs[156];   // observe output: H (144, column-major), b (12)
  ObsParams obs_prm;
  UpdParams upd_prm;
  vmp_map_cfg cfg;
  int32_t iters;
  int32_t active;
  int32_t pad[2];
};

// ------------------------------------------------------- SO3 in FP64 ------
// Rodrigues through the unit quaternion with a Taylor branch for tiny angles.
__device__ inline D33 so3_exp_d(const double w[3]) {
  const double tsq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double imag, real;
  if (tsq < 1e-20) {
    imag = 0.5 - tsq / 48.0 + tsq * tsq / 3840.0;
    real = 1.0 - tsq / 8.0 + tsq * tsq / 384.0;
  } else {
    const double th = sqrt(tsq), half = 0.5 * th;
    imag = sin(half) / th;
    real = cos(half);
  }
  const double qw = real, qx = imag * w[0], qy = imag * w[1], qz = imag * w[2];
  const double tx = 2 * qx, ty = 2 * qy, tz = 2 * qz;
  const double twx = tx * qw, twy = ty * qw, twz = tz * qw;
  const double txx = tx * qx, txy = ty * qx, txz = tz * qx;
  const double tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
  D33 R;
  R.m[0] = 1 - (tyy + tzz);
  R.m[1] = txy - twz;
  R.m[2] = txz + twy;
  R.m[3] = txy + twz;
  R.m[4] = 1 - (txx + tzz);
  R.m[5] = tyz - twx;
  R.m[6] = txz - twy;
  R.m[7] = tyz + twx;
  R.m[8] = 1 - (txx + tyy);
  return R;
}

// rotation -> quaternion (largest-diagonal branch) -> 2 atan(|v|/w) v/|v|
__device__ inline void so3_log_d(const D33& R, double out[3]) {
  const double* m = R.m;
  double q[4];
  double t = m[0] + m[4] + m[8];
  if (t > 0) {
    t = sqrt(t + 1.0);
    q[3] = 0.5 * t;
    t = 0.5 / t;
    q[0] = (m[7] - m[5]) * t;
    q[1] = (m[2] - m[6]) * t;
    q[2] = (m[3] - m[1]) * t;
  } else {
    int i = 0;
    if (m[4] > m[0]) i = 1;
    if (m[8] > m[i * 4]) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    t = sqrt(m[i * 4] - m[j * 4] - m[k * 4] + 1.0);
    q[i] = 0.5 * t;
    t = 0.5 / t;
    q[3] = (m[k * 3 + j] - m[j * 3 + k]) * t;
    q[j] = (m[j * 3 + i] + m[i * 3 + j]) * t;
    q[k] = (m[k * 3 + i] + m[i * 3 + k]) * t;
  }
  const double qn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int c = 0; c < 4; ++c) q[c] /= qn;
  const double sqn = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  const double w = q[3];
  double f;
  if (sqn < 1e-20) {
    f = 2.0 / w - 2.0 / 3.0 * sqn / (w * w * w);
  } else {
    const double nn = sqrt(sqn);
    if (fabs(w) < 1e-10) f = (w > 0 ? 3.14159265358979323846 : -3.14159265358979323846) / nn;
    else f = 2.0 * atan(nn / w) / nn;
  }
  for (int c = 0; c < 3; ++c) out[c] = f * q[c];
}

__device__ inline D33 hat_d(const double v[3]) {
  D33 h;
  h.m[0] = 0;
  h.m[1] = -v[2];
  h.m[2] = v[1];
  h.m[3] = v[2];
  h.m[4] = 0;
  h.m[5] = -v[0];
  h.m[6] = -v[1];
  h.m[7] = v[0];
  h.m[8] = 0;
  return h;
}

// left Jacobian I + (1-cos)/th^2 W + (th-sin)/th^3 W^2; first-order branch
// for th^2 < 1e-10, where 1 - cos(th) cancels (Sophus' threshold).
__device__ inline D33 left_jac_d(const double w[3]) {
  const double tsq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  const double th = sqrt(tsq);
  const D33 W = hat_d(w);
  D33 J;
  if (tsq < 1e-10) {
    for (int i = 0; i < 9; ++i) J.m[i] = (i % 4 == 0 ? 1.0 : 0.0) + 0.5 * W.m[i];
    return J;
  }
  const D33 W2 = d33_mul(W, W);
  const double a = (1.0 - cos(th)) / (th * th), b = (th - sin(th)) / (th * th * th);
  for (int i = 0; i < 9; ++i) J.m[i] = ((i % 4 == 0 ? 1.0 : 0.0) + a * W.m[i]) + b * W2.m[i];
  return J;
}

// S2 gravity chart (ieskf.cpp:57-83)
__device__ inline void getBx_d(const double g[3], double B[6]) {  // 3x2 row-major
  const double G = kGravity;
  B[0] = -g[1];
  B[1] = -g[2];
  B[2] = G - g[1] * g[1] / (G + g[0]);
  B[3] = -g[2] * g[1] / (G + g[0]);
  B[4] = -g[2] * g[1] / (G + g[0]);
  B[5] = G - g[2] * g[2] / (G + g[0]);
  for (int i = 0; i < 6; ++i) B[i] /= G;
}
__device__ inline void getNx_d(const double g[3], double N[6]) {  // 2x3 row-major: 1/G^2 Bx^T hat(g)
  double B[6];
  getBx_d(g, B);
  const D33 h = hat_d(g);
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = B[r] * h.m[c] + B[2 + r] * h.m[3 + c] + B[4 + r] * h.m[6 + c];
      N[r * 3 + c] = (1.0 / kGravity / kGravity) * s;
    }
}
__device__ inline void getMx_d(const double g[3], const double res[2], double Mx[6]) {  // 3x2
  double B[6];
  getBx_d(g, B);
  const double bu[3] = {B[0] * res[0] + B[1] * res[1], B[2] * res[0] + B[3] * res[1], B[4] * res[0] + B[5] * res[1]};
  const D33 E = so3_exp_d(bu), h = hat_d(g), Jl = left_jac_d(bu);
  const D33 A = d33_mul(d33_mul(E, h), d33_t(Jl));
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 2; ++c) Mx[r * 2 + c] = -(A.m[r * 3] * B[c] + A.m[r * 3 + 1] * B[2 + c] + A.m[r * 3 + 2] * B[4 + c]);
}

// gravity part of x - xp on S2 (ieskf.cpp:34-52)
__device__ inline void s2_minus_d(const double* g, const double* og, double* out) {
  const D33 hg = hat_d(g);
  double hv[3];
  for (int i = 0; i < 3; ++i) hv[i] = hg.m[i * 3] * og[0] + hg.m[i * 3 + 1] * og[1] + hg.m[i * 3 + 2] * og[2];
  const double v_sin = sqrt(hv[0] * hv[0] + hv[1] * hv[1] + hv[2] * hv[2]);
  const double v_cos = g[0] * og[0] + g[1] * og[1] + g[2] * og[2];
  const double theta = atan2(v_sin, v_cos);
  if (v_sin < 1e-11) {
    out[0] = fabs(theta) > 1e-11 ? 3.1415926 : 0.0;
    out[1] = 0.0;
  } else {
    double B[6];
    getBx_d(og, B);
    const D33 ho = hat_d(og);
    double hx[3];
    for (int i = 0; i < 3; ++i) hx[i] = ho.m[i * 3] * g[0] + ho.m[i * 3 + 1] * g[1] + ho.m[i * 3 + 2] * g[2];
    for (int r = 0; r < 2; ++r) out[r] = theta / v_sin * (B[r] * hx[0] + B[2 + r] * hx[1] + B[4 + r] * hx[2]);
  }
}

// x += d (ieskf.cpp:11-21)
__device__ inline void state_plus(vmp_state& x, const double d[23]) {
  for (int i = 0; i < 3; ++i) x.pos[i] += d[i];
  d33_to_cm(d33_mul(d33_from_cm(x.rot), so3_exp_d(d + 3)), x.rot);
  d33_to_cm(d33_mul(d33_from_cm(x.rot_ext), so3_exp_d(d + 6)), x.rot_ext);
  for (int i = 0; i < 3; ++i) {
    x.pos_ext[i] += d[9 + i];
    x.vel[i] += d[12 + i];
    x.bg[i] += d[15 + i];
    x.ba[i] += d[18 + i];
  }
  double B[6];
  getBx_d(x.g, B);
  const double w[3] = {B[0] * d[21] + B[1] * d[22], B[2] * d[21] + B[3] * d[22], B[4] * d[21] + B[5] * d[22]};
  const D33 E = so3_exp_d(w);
  double g2[3];
  for (int i = 0; i < 3; ++i) g2[i] = E.m[i * 3] * x.g[0] + E.m[i * 3 + 1] * x.g[1] + E.m[i * 3 + 2] * x.g[2];
  for (int i = 0; i < 3; ++i) x.g[i] = g2[i];
}

// blocks of J (and L): right Jacobians of the two rotations, S2 block
struct JacBlocks {
  double a[9], b[9], c[4];
};
__device__ inline void jac_blocks(const vmp_state& x, const vmp_state& xp, const double d[23], JacBlocks& J) {
  // right Jacobian = leftJacobian(w)^T
  const D33 A = d33_t(left_jac_d(d + 3)), B = d33_t(left_jac_d(d + 6));
  for (int i = 0; i < 9; ++i) {
    J.a[i] = A.m[i];
    J.b[i] = B.m[i];
  }
  double N[6], M[6];
  getNx_d(x.g, N);
  getMx_d(xp.g, d + 21, M);
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) J.c[r * 2 + c] = N[r * 3] * M[c] + N[r * 3 + 1] * M[2 + c] + N[r * 3 + 2] * M[4 + c];
}
// element (r, c) of the 23x23 J built from the blocks
__device__ inline double jac_at(const JacBlocks& J, int r, int c) {
  if (r >= 3 && r < 6 && c >= 3 && c < 6) return J.a[(r - 3) * 3 + (c - 3)];
  if (r >= 6 && r < 9 && c >= 6 && c < 9) return J.b[(r - 6) * 3 + (c - 6)];
  if (r >= 21 && c >= 21) return J.c[(r - 21) * 2 + (c - 21)];
  return r == c ? 1.0 : 0.0;
}

// ---- Cholesky factorisation, solve and inverse of a symmetric positive
//      definite 23x23 on rows held in lanes ---------------------------------
// Lane i (< 23) holds row i of the matrix in a[0..22]; values cross lanes by
// readlane (uniform index), so nothing goes through LDS and no barrier is
// needed.  The factorisation is right-looking (column k scaled, then the
// trailing lower triangle updated in ascending column order); divisions by a
// diagonal element are products with its reciprocal (kept per lane in dinv).
__device__ inline double rl64(double v, int src) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(x), src);
  const int hi = __builtin_amdgcn_readlane(int(x >> 32), src);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
// in place: row i becomes L[i][0..i], its upper part untouched; dinv = 1/L[i][i]
// (column k of L is read by readlane as it is used: measured faster here
// than staging each column through LDS)
__device__ inline void chol_rows(double (&a)[23], double& dinv, int lane) {
  dinv = 0.0;
#pragma unroll
  for (int k = 0; k < 23; ++k) {
    const double d = sqrt(rl64(a[k], k));
    const double inv = 1.0 / d;
    if (lane == k) {
      a[k] = d;
      dinv = inv;
    } else if (lane > k) {
      a[k] = a[k] * inv;
    }
#pragma unroll
    for (int j = k + 1; j < 23; ++j) {
      const double ljk = rl64(a[k], j);
      if (lane >= j) a[j] -= a[k] * ljk;
    }
  }
}
// x = A^-1 b for A = L L^T from chol_rows; b and x one element per lane.
// Forward: L y = b by rows; backward: L^T x = y by columns, lane j taking
// L[i][j] x_i off y_j as each x_i is found (L staged row-major in Ls, 529
// doubles of LDS).  Called by all threads of a 64-thread workgroup.
__device__ inline double chol_solve_rows(const double (&a)[23], double dinv, double b, int lane, double* Ls) {
  double y = b;
#pragma unroll
  for (int j = 0; j < 23; ++j) {  // L y = b
    if (lane == j) y = y * dinv;
    const double yj = rl64(y, j);
    if (lane > j) y -= a[j] * yj;
  }
  if (lane < 23) {
#pragma unroll
    for (int c = 0; c < 23; ++c) Ls[lane * 23 + c] = a[c];
  }
  __syncthreads();
#pragma unroll
  for (int i = 22; i >= 0; --i) {  // L^T x = y
    const double xi = rl64(y, i) * rl64(dinv, i);
    if (lane == i) y = xi;
    else if (lane < i) y -= Ls[i * 23 + lane] * xi;
  }
  __syncthreads();  // Ls may be reused by the caller
  return y;
}

// A^-1 for A = L L^T from chol_rows (rows of L in the first wave's lanes,
// dinv = 1/L[i][i]): lane c of the first wave forms column c of W = L^-1 by
// forward substitution, then A^-1 = W^T W, whose row r is written by lane c
// of wave r mod (waves) into out[r * 23 + c] (bitwise symmetric: both halves
// sum the same products in the same order).  L, 1/L[i][i] and then W are
// staged in S (23 * 24 doubles of LDS) and read back as broadcasts; out
// should be LDS too (stores through a generic pointer to global memory would
// wait together with those reads).  Called by all threads of the workgroup
// (one or more waves).
__device__ inline void chol_inverse_rows(const double (&a)[23], double dinv, double* out, int lane, double* S) {
  if (lane < 23) {
#pragma unroll
    for (int c = 0; c < 23; ++c) S[lane * 24 + c] = a[c];
    S[lane * 24 + 23] = dinv;
  }
  __syncthreads();
  double w[23];
  if (lane < 64) {
#pragma unroll
    for (int i = 0; i < 23; ++i) {
      double s = lane == i ? 1.0 : 0.0;
#pragma unroll
      for (int j = 0; j < i; ++j) s -= S[i * 24 + j] * w[j];
      w[i] = s * S[i * 24 + 23];
    }
  }
  __syncthreads();
  if (lane < 23) {
#pragma unroll
    for (int k = 0; k < 23; ++k) S[k * 24 + lane] = w[k];
  }
  __syncthreads();
  const int l = lane & 63, nw = int(blockDim.x) >> 6;
  if (l < 23) {
    double wc[23];
#pragma unroll
    for (int k = 0; k < 23; ++k) wc[k] = S[k * 24 + l];
    for (int r = lane >> 6; r < 23; r += nw) {
      double s = S[r] * wc[0];
#pragma unroll
      for (int k = 1; k < 23; ++k) s += S[k * 24 + r] * wc[k];
      out[r * 23 + l] = s;
    }
  }
  __syncthreads();  // S may be reused by the caller
}

// ---- lane-parallel pieces of the state algebra (inputs in LDS) ----------
// x - xp split over three lanes: rotation, extrinsic rotation, gravity chart
__device__ inline void state_minus_par(const vmp_state& x, const vmp_state& o, double* d, int lane) {
  if (lane == 0) {
    so3_log_d(d33_mul(d33_t(d33_from_cm(o.rot)), d33_from_cm(x.rot)), d + 3);
  } else if (lane == 1) {
    so3_log_d(d33_mul(d33_t(d33_from_cm(o.rot_ext)), d33_from_cm(x.rot_ext)), d + 6);
  } else if (lane == 2) {
    s2_minus_d(x.g, o.g, d + 21);
  } else if (lane == 3) {
    for (int i = 0; i < 3; ++i) {
      d[i] = x.pos[i] - o.pos[i];
      d[9 + i] = x.pos_ext[i] - o.pos_ext[i];
      d[12 + i] = x.vel[i] - o.vel[i];
      d[15 + i] = x.bg[i] - o.bg[i];
      d[18 + i] = x.ba[i] - o.ba[i];
    }
  }
  __syncthreads();
}
__device__ inline void jac_blocks_par(const vmp_state& x, const vmp_state& xp, const double* d, JacBlocks& J,
                                      int lane) {
  if (lane == 0) {
    const D33 A = d33_t(left_jac_d(d + 3));
    for (int i = 0; i < 9; ++i) J.a[i] = A.m[i];
  } else if (lane == 1) {
    const D33 B = d33_t(left_jac_d(d + 6));
    for (int i = 0; i < 9; ++i) J.b[i] = B.m[i];
  } else if (lane == 2) {
    double N[6], M[6];
    getNx_d(x.g, N);
    getMx_d(xp.g, d + 21, M);
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) J.c[r * 2 + c] = N[r * 3] * M[c] + N[r * 3 + 1] * M[2 + c] + N[r * 3 + 2] * M[4 + c];
  }
  __syncthreads();
}
__device__ inline void state_plus_par(vmp_state& x, const double* d, int lane) {
  if (lane == 0) {
    d33_to_cm(d33_mul(d33_from_cm(x.rot), so3_exp_d(d + 3)), x.rot);
  } else if (lane == 1) {
    d33_to_cm(d33_mul(d33_from_cm(x.rot_ext), so3_exp_d(d + 6)), x.rot_ext);
  } else if (lane == 2) {
    double B[6];
    getBx_d(x.g, B);
    const double w[3] = {B[0] * d[21] + B[1] * d[22], B[2] * d[21] + B[3] * d[22], B[4] * d[21] + B[5] * d[22]};
    const D33 E = so3_exp_d(w);
    double g2[3];
    for (int i = 0; i < 3; ++i) g2[i] = E.m[i * 3] * x.g[0] + E.m[i * 3 + 1] * x.g[1] + E.m[i * 3 + 2] * x.g[2];
    for (int i = 0; i < 3; ++i) x.g[i] = g2[i];
  } else if (lane == 3) {
    for (int i = 0; i < 3; ++i) {
      x.pos[i] += d[i];
      x.pos_ext[i] += d[9 + i];
      x.vel[i] += d[12 + i];
      x.bg[i] += d[15 + i];
      x.ba[i] += d[18 + i];
    }
  }
  __syncthreads();
}

// Gauss-Jordan inverse with partial pivoting of the 23x23 row-major M (LDS),
// one wave; result in Inv (LDS).  W is a 23x46 LDS workspace, fac 23 doubles.
__device__ inline void gj_inverse(const double* M, double* Inv, double (*W)[47], double* fac, int lane) {
  constexpr int n = 23;
  for (int idx = lane; idx < n * 2 * n; idx += 64) {
    const int r = idx / (2 * n), c = idx % (2 * n);
    W[r][c] = c < n ? M[r * n + c] : (c - n == r ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    double v = (lane >= k && lane < n) ? fabs(W[lane][k]) : -1.0;
    int pi = lane;
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(v, o, 64);
      const int op = __shfl_xor(pi, o, 64);
      if (ov > v || (ov == v && op < pi)) {
        v = ov;
        pi = op;
      }
    }
    if (pi != k) {
      for (int c = lane; c < 2 * n; c += 64) {
        const double t = W[k][c];
        W[k][c] = W[pi][c];
        W[pi][c] = t;
      }
    }
    __syncthreads();
    const double d = W[k][k];
    __syncthreads();
    for (int c = lane; c < 2 * n; c += 64) W[k][c] = W[k][c] / d;
    __syncthreads();
    if (lane < n) fac[lane] = W[lane][k];
    __syncthreads();
    for (int idx = lane; idx < n * 2 * n; idx += 64) {
      const int r = idx / (2 * n), c = idx % (2 * n);
      if (r != k) W[r][c] -= fac[r] * W[k][c];
    }
    __syncthreads();
  }
  for (int idx = lane; idx < n * n; idx += 64) Inv[idx] = W[idx / n][n + idx % n];
  __syncthreads();
}

}  // namespace vmp
