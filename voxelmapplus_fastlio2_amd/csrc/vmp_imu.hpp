// IMU side of the registration step on the device: IMU initialisation
// (LIOBuilder::initializeImu, lio_builder.cpp:28-63), IMU propagation of the
// filter state over one package (IESKF::predict, ieskf.cpp:101-123, driven as
// in undistortCloud, lio_builder.cpp:65-123) and the backward compensation of
// every point into the frame at the scan's end (lio_builder.cpp:121-153).
//
//   k_imu_init      one thread per session: means of the cached samples,
//                   gravity alignment, bg, P
//   k_imu_predict   one 64-lane workgroup per session: the package's IMU
//                   pairs in order; the state step and the small F/G blocks
//                   on lane 0, the 23 rows of P = F P F^T + G Q G^T on lanes
//                   0-22 (only the structurally non-zero products, added in
//                   ascending order, which is what the dense sums give); the
//                   IMU pose table and its running minimum of offsets
//   k_undistort     one lane per point: the pose interval it falls in (a
//                   binary search over the running minimum, which is the
//                   interval the reference's backward walk reaches it in),
//                   then the compensation in FP64; the earliest point also
//                   repeats it for every earlier interval it is later than,
//                   as the reference's walk does
#pragma once
#include "vmp_common.hpp"
#include "vmp_ieskf.hpp"

namespace vmp {

constexpr int kMaxImu = 512;  // IMU samples of one package
constexpr int kPoseD = 22;    // offset, acc 3, gyro 3, vel 3, pos 3, rot 9 (row-major)

// one SyncPackage's IMU side (lio_node.cpp:137-165); the samples are
// interleaved (t, acc xyz, gyro xyz) so a package uploads as one copy of
// its used bytes
constexpr int kImuS = 7;
struct ImuPkg {
  double cloud_start, cloud_end;
  int32_t m;  // IMU samples
  int32_t pad;
  double s[kMaxImu][kImuS];
};
__host__ __device__ inline size_t imu_pkg_bytes(int m) { return offsetof(ImuPkg, s) + sizeof(double) * kImuS * m; }

struct ImuDev {
  // LIODataGroup (lio_builder.h:43-54)
  double last_imu_t, last_imu_acc[3], last_imu_gyro[3];
  double last_acc[3], last_gyro[3];
  double last_cloud_end;
  double gravity_norm;
  double q[4];     // diagonal of Q by block: ng, na, nbg, nba
  int32_t npose;   // entries of the pose table
  int32_t pad;
  ImuPkg pkg;      // the package being processed
  double pose[kMaxImu + 1][kPoseD];
  double pose_min[kMaxImu + 1];  // min of the offsets from entry i to the second to last
};

struct ImuCfgDev {
  double r_il[9];  // row-major
  double p_il[3];
  int32_t gravity_align;
  int32_t n;       // samples
};

// one session of the preprocessing stage
struct PreJob {
  ImuDev* imu;
  FilterDev* f;
  const float* x;  // raw points, ascending time
  const float* y;
  const float* z;
  const float* t;  // ms since cloud_start
  float* ox;       // compensated points
  float* oy;
  float* oz;
  int32_t n;
  int32_t pad;
};

// Quaternion::setFromTwoVectors(a, b).matrix() for a, b not opposite (the
// half-way quaternion), else a half turn about an axis orthogonal to a
__device__ inline D33 rot_two_vectors_d(const double a[3], const double b[3]) {
  const double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  const double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  const double v0[3] = {a[0] / na, a[1] / na, a[2] / na};
  const double v1[3] = {b[0] / nb, b[1] / nb, b[2] / nb};
  const double c = v1[0] * v0[0] + v1[1] * v0[1] + v1[2] * v0[2];
  double qw, qx, qy, qz;
  if (c < -1.0 + 1e-12) {
    double ax[3];
    if (fabs(v0[0]) < 0.9) {
      ax[0] = 0.0;
      ax[1] = -v0[2];
      ax[2] = v0[1];
    } else {
      ax[0] = -v0[2];
      ax[1] = 0.0;
      ax[2] = v0[0];
    }
    const double nn = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    qw = 0.0;
    qx = ax[0] / nn;
    qy = ax[1] / nn;
    qz = ax[2] / nn;
  } else {
    const double axis[3] = {v0[1] * v1[2] - v0[2] * v1[1], v0[2] * v1[0] - v0[0] * v1[2],
                            v0[0] * v1[1] - v0[1] * v1[0]};
    const double sq = sqrt((1.0 + c) * 2.0);
    const double invs = 1.0 / sq;
    qx = axis[0] * invs;
    qy = axis[1] * invs;
    qz = axis[2] * invs;
    qw = sq * 0.5;
  }
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

// initializeImu over the cached samples (lio_builder.cpp:33-61); the caller
// has checked that there are at least imu_init_num of them
__global__ void k_imu_init(ImuDev* imu, FilterDev* f, const double* t, const double* acc, const double* gyro,
                           ImuCfgDev c) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double am[3] = {0.0, 0.0, 0.0}, gm[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < c.n; ++i)
    for (int k = 0; k < 3; ++k) {
      am[k] = am[k] + acc[3 * i + k];
      gm[k] = gm[k] + gyro[3 * i + k];
    }
  const double cnt = double(c.n);
  for (int k = 0; k < 3; ++k) {
    am[k] = am[k] / cnt;
    gm[k] = gm[k] / cnt;
  }
  imu->gravity_norm = sqrt(am[0] * am[0] + am[1] * am[1] + am[2] * am[2]);
  vmp_state& x = f->x;
  d33_to_cm(D33{{c.r_il[0], c.r_il[1], c.r_il[2], c.r_il[3], c.r_il[4], c.r_il[5], c.r_il[6], c.r_il[7], c.r_il[8]}},
            x.rot_ext);
  for (int k = 0; k < 3; ++k) {
    x.pos_ext[k] = c.p_il[k];
    x.bg[k] = gm[k];
  }
  auto init_g = [&](const double* d) {  // State::initG (ieskf.h:44-47)
    const double nn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) x.g[k] = d[k] / nn * kGravity;
  };
  const double neg[3] = {-am[0], -am[1], -am[2]};
  if (c.gravity_align) {
    const double down[3] = {0.0, 0.0, -1.0};
    d33_to_cm(rot_two_vectors_d(neg, down), x.rot);
    init_g(down);
  } else {
    init_g(neg);
  }
  for (int i = 0; i < 529; ++i) f->P[i] = (i % 24 == 0) ? 1.0 : 0.0;
  for (int k = 0; k < 3; ++k) {
    f->P[(6 + k) * 24] = 0.00001;
    f->P[(9 + k) * 24] = 0.00001;
    f->P[(15 + k) * 24] = 0.0001;
    f->P[(18 + k) * 24] = 0.0001;
  }
  f->P[21 * 24] = 0.00001;
  f->P[22 * 24] = 0.00001;
  const int l = c.n - 1;
  imu->last_imu_t = t[l];
  for (int k = 0; k < 3; ++k) {
    imu->last_imu_acc[k] = acc[3 * l + k];
    imu->last_imu_gyro[k] = gyro[3 * l + k];
  }
}

// the small blocks of F and G of one predict step, from the state before it
struct PredBlocks {
  double Rexp[9];  // exp(-(gyro - bg) dt)
  double Jdt[9];   // -Jr((gyro - bg) dt) dt (F and G)
  double Adt[9];   // (-rot hat(acc - ba)) dt
  double Rdt[9];   // (-rot) dt (F and G)
  double Mdt[6];   // getMx() dt, 3x2
  double NM[4];    // getNx() getMx(), 2x2
  double dt;
};

// IESKF::predict on the state in `x` (lane 0): the blocks, then x += delta
__device__ inline void predict_state(vmp_state& x, const double acc[3], const double gyro[3], double dt,
                                     PredBlocks& B) {
  double wdt[3], dvel[3], dpos[3];
  for (int k = 0; k < 3; ++k) wdt[k] = dt * (gyro[k] - x.bg[k]);
  const D33 rot = d33_from_cm(x.rot);
  double ab[3] = {acc[0] - x.ba[0], acc[1] - x.ba[1], acc[2] - x.ba[2]};
  const D3 ra = d33_mulv(rot, D3{{ab[0], ab[1], ab[2]}});
  for (int k = 0; k < 3; ++k) {
    dvel[k] = dt * (ra.v[k] + x.g[k]);
    dpos[k] = dt * x.vel[k];
  }
  const double nw[3] = {(-1.0) * wdt[0], (-1.0) * wdt[1], (-1.0) * wdt[2]};
  const D33 Re = so3_exp_d(nw);
  const D33 Jr = d33_t(left_jac_d(wdt));
  D33 nrot;
  for (int i = 0; i < 9; ++i) nrot.m[i] = (-1.0) * rot.m[i];
  const D33 Ah = d33_mul(nrot, hat_d(ab));
  for (int i = 0; i < 9; ++i) {
    B.Rexp[i] = Re.m[i];
    B.Jdt[i] = ((-1.0) * Jr.m[i]) * dt;
    B.Adt[i] = Ah.m[i] * dt;
    B.Rdt[i] = nrot.m[i] * dt;
  }
  double Bx[6];
  getBx_d(x.g, Bx);
  const D33 hg = hat_d(x.g);
  double Mx[6];  // (-hat(g)) Bx
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 2; ++c) {
      double s = ((-1.0) * hg.m[r * 3]) * Bx[c];
      s += ((-1.0) * hg.m[r * 3 + 1]) * Bx[2 + c];
      s += ((-1.0) * hg.m[r * 3 + 2]) * Bx[4 + c];
      Mx[r * 2 + c] = s;
    }
  for (int i = 0; i < 6; ++i) B.Mdt[i] = Mx[i] * dt;
  double N[6];
  getNx_d(x.g, N);
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) {
      double s = N[r * 3] * Mx[c];
      s += N[r * 3 + 1] * Mx[2 + c];
      s += N[r * 3 + 2] * Mx[4 + c];
      B.NM[r * 2 + c] = s;
    }
  B.dt = dt;
  // x += delta: rot_ext, pos_ext, bg, ba and g take zero increments
  for (int k = 0; k < 3; ++k) {
    x.pos[k] = x.pos[k] + dpos[k];
    x.vel[k] = x.vel[k] + dvel[k];
  }
  d33_to_cm(d33_mul(rot, so3_exp_d(wdt)), x.rot);
}

// column indices of the structurally non-zero entries of row r of F, in
// ascending order (returns the count)
__device__ inline int f_row_cols(int r, int* cols) {
  int n = 0;
  if (r < 3) {
    cols[n++] = r;
    cols[n++] = 12 + r;
  } else if (r < 6) {
    for (int k = 3; k < 6; ++k) cols[n++] = k;
    for (int k = 15; k < 18; ++k) cols[n++] = k;
  } else if (r >= 12 && r < 15) {
    for (int k = 3; k < 6; ++k) cols[n++] = k;
    cols[n++] = r;
    for (int k = 18; k < 23; ++k) cols[n++] = k;
  } else if (r >= 21) {
    cols[n++] = 21;
    cols[n++] = 22;
  } else {
    cols[n++] = r;
  }
  return n;
}
__device__ inline double f_at(const PredBlocks& B, int r, int c) {
  if (r < 3) return c == r ? 1.0 : (c == 12 + r ? B.dt : 0.0);  // (I dt) off the diagonal is 0 dt
  if (r < 6) {
    if (c >= 3 && c < 6) return B.Rexp[(r - 3) * 3 + (c - 3)];
    if (c >= 15 && c < 18) return B.Jdt[(r - 3) * 3 + (c - 15)];
    return 0.0;
  }
  if (r >= 12 && r < 15) {
    if (c >= 3 && c < 6) return B.Adt[(r - 12) * 3 + (c - 3)];
    if (c >= 18 && c < 21) return B.Rdt[(r - 12) * 3 + (c - 18)];
    if (c >= 21) return B.Mdt[(r - 12) * 2 + (c - 21)];
    return c == r ? 1.0 : 0.0;
  }
  if (r >= 21) return c >= 21 ? B.NM[(r - 21) * 2 + (c - 21)] : 0.0;
  return c == r ? 1.0 : 0.0;
}
// G (23x12): rows 3-5 <- cols 0-2 (-Jr dt), 12-14 <- 3-5 (-rot dt),
// 15-17 <- 6-8 (I dt), 18-20 <- 9-11 (I dt)
__device__ inline int g_group(int r) {
  return (r >= 3 && r < 6) ? 0 : (r >= 12 && r < 15) ? 1 : (r >= 15 && r < 18) ? 2 : (r >= 18 && r < 21) ? 3 : -1;
}
__device__ inline double g_at(const PredBlocks& B, int r, int c) {
  const int gr = g_group(r);
  if (gr < 0 || c / 3 != gr) return 0.0;
  const int i = r - (gr == 0 ? 3 : gr == 1 ? 12 : gr == 2 ? 15 : 18), j = c % 3;
  if (gr == 0) return B.Jdt[i * 3 + j];
  if (gr == 1) return B.Rdt[i * 3 + j];
  return (i == j ? 1.0 : 0.0) * B.dt;
}

// one predict step of P on lanes 0..22 (row per lane); P, FP row-major in LDS
__device__ inline void predict_cov(const PredBlocks& B, const double q[4], double* P, double* FP, int lane) {
  int cols[12];
  if (lane < 23) {  // FP = F P
    const int nc = f_row_cols(lane, cols);
    for (int j = 0; j < 23; ++j) {
      double s = 0.0;
      for (int t = 0; t < nc; ++t) s += f_at(B, lane, cols[t]) * P[cols[t] * 23 + j];
      FP[lane * 23 + j] = s;
    }
  }
  __syncthreads();
  if (lane < 23) {  // P = FP F^T + (G Q) G^T
    const int gi = g_group(lane);
    for (int j = 0; j < 23; ++j) {
      const int nc = f_row_cols(j, cols);
      double s = 0.0;
      for (int t = 0; t < nc; ++t) s += FP[lane * 23 + cols[t]] * f_at(B, j, cols[t]);
      double g = 0.0;
      if (gi >= 0 && g_group(j) == gi)
        for (int c = 3 * gi; c < 3 * gi + 3; ++c) g += (g_at(B, lane, c) * q[gi]) * g_at(B, j, c);
      P[lane * 23 + j] = s + g;
    }
  }
  __syncthreads();
}

// undistortCloud's IMU part for one session per workgroup (64 lanes).  The
// biases and gravity do not change while a package is propagated, so every
// step's rotation increment exp((gyro - bg) dt) and right Jacobian depend on
// the IMU data alone: lanes compute them for up to 64 steps at once (the
// transcendental part), then the steps run in order: the state on lane 0
// (products only), the rows of P on lanes 0-22.  exp(-w) is exp(w)^T bit
// for bit (the quaternion form only flips signs), so F's rotation block is
// the transpose of the step's increment.
__global__ void __launch_bounds__(64) k_imu_predict(const PreJob* jobs, int nsess) {
  if (int(blockIdx.x) >= nsess) return;
  const PreJob J = jobs[blockIdx.x];
  ImuDev* im = J.imu;
  FilterDev* f = J.f;
  const int lane = threadIdx.x;
  __shared__ double P[529], FP[529];
  __shared__ PredBlocks B;
  __shared__ vmp_state sx;
  __shared__ double q[4];
  __shared__ double cE[64][9], cJ[64][9], cA[64][3], cG[64][3], cDt[64], cT[64];
  __shared__ double Mx[6], NM[4];
  for (int i = lane; i < 529; i += 64) P[(i % 23) * 23 + i / 23] = f->P[i];  // column-major -> rows
  load_state(f->x, sx, lane);
  if (lane < 4) q[lane] = im->q[lane];
  __syncthreads();
  const int m = im->pkg.m;
  const double lce = im->last_cloud_end;
  // active IMU pairs (head k-1 or last_imu, tail k) are those whose tail is
  // not before the last scan end: a prefix of the ascending samples is skipped
  int k0 = 0;
  while (k0 < m && im->pkg.s[k0][0] < lce) ++k0;
  const int nsteps = (m - k0) + 1;  // the pairs, then the step to the sweep's end
  if (lane == 0) {
    double Bx[6];
    getBx_d(sx.g, Bx);
    const D33 hg = hat_d(sx.g);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 2; ++c) {
        double v = ((-1.0) * hg.m[r * 3]) * Bx[c];
        v += ((-1.0) * hg.m[r * 3 + 1]) * Bx[2 + c];
        v += ((-1.0) * hg.m[r * 3 + 2]) * Bx[4 + c];
        Mx[r * 2 + c] = v;
      }
    double N[6];
    getNx_d(sx.g, N);
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) {
        double v = N[r * 3] * Mx[c];
        v += N[r * 3 + 1] * Mx[2 + c];
        v += N[r * 3 + 2] * Mx[4 + c];
        NM[r * 2 + c] = v;
      }
    double* p = im->pose[0];  // pose 0: the start of the sweep
    const D33 rot = d33_from_cm(sx.rot);
    p[0] = 0.0;
    for (int c = 0; c < 3; ++c) {
      p[1 + c] = im->last_acc[c];
      p[4 + c] = im->last_gyro[c];
      p[7 + c] = sx.vel[c];
      p[10 + c] = sx.pos[c];
    }
    for (int c = 0; c < 9; ++c) p[13 + c] = rot.m[c];
  }
  // input of step st (the last pair's input for the final step; zero when
  // there is no pair: the reference's input is then uninitialised)
  auto step_input = [&](int st, double acc[3], double gyro[3], double& dt, double& tt) {
    const int k = k0 + (st < nsteps - 1 ? st : nsteps - 2);
    if (k < k0) {
      for (int c = 0; c < 3; ++c) acc[c] = gyro[c] = 0.0;
    } else {
      const double* hs = k == 0 ? nullptr : im->pkg.s[k - 1];
      for (int c = 0; c < 3; ++c) {
        const double ha = k == 0 ? im->last_imu_acc[c] : hs[1 + c];
        const double hgy = k == 0 ? im->last_imu_gyro[c] : hs[4 + c];
        gyro[c] = 0.5 * (hgy + im->pkg.s[k][4 + c]);
        acc[c] = 0.5 * (ha + im->pkg.s[k][1 + c]);
        acc[c] = acc[c] * 9.81 / im->gravity_norm;
      }
    }
    if (st < nsteps - 1) {
      const double ht = k == 0 ? im->last_imu_t : im->pkg.s[k - 1][0];
      tt = im->pkg.s[k][0];
      dt = ht < lce ? tt - lce : tt - ht;
    } else {
      const double imu_end = m > 0 ? im->pkg.s[m - 1][0] : im->last_imu_t;
      tt = 0.0;
      dt = im->pkg.cloud_end - imu_end;
    }
  };
  int np = 1;
  for (int c0 = 0; c0 < nsteps; c0 += 64) {
    const int cn = min(64, nsteps - c0);
    if (lane < cn) {  // the steps' rotation increments and right Jacobians, in parallel
      double a[3], g[3], dt, tt;
      step_input(c0 + lane, a, g, dt, tt);
      double wdt[3];
      for (int c = 0; c < 3; ++c) wdt[c] = dt * (g[c] - sx.bg[c]);
      const D33 E = so3_exp_d(wdt);
      const D33 Jr = d33_t(left_jac_d(wdt));
      for (int i = 0; i < 9; ++i) {
        cE[lane][i] = E.m[i];
        cJ[lane][i] = ((-1.0) * Jr.m[i]) * dt;
      }
      for (int c = 0; c < 3; ++c) {
        cA[lane][c] = a[c];
        cG[lane][c] = g[c];
      }
      cDt[lane] = dt;
      cT[lane] = tt;
    }
    __syncthreads();
    for (int j = 0; j < cn; ++j) {
      const bool last = c0 + j == nsteps - 1;
      if (lane == 0) {  // IESKF::predict on the state: blocks from the state before the step, then x += delta
        const double dt = cDt[j];
        const D33 rot = d33_from_cm(sx.rot);
        const double ab[3] = {cA[j][0] - sx.ba[0], cA[j][1] - sx.ba[1], cA[j][2] - sx.ba[2]};
        const D3 ra = d33_mulv(rot, D3{{ab[0], ab[1], ab[2]}});
        D33 nrot;
        for (int i = 0; i < 9; ++i) nrot.m[i] = (-1.0) * rot.m[i];
        const D33 Ah = d33_mul(nrot, hat_d(ab));
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) B.Rexp[r * 3 + c] = cE[j][c * 3 + r];
        for (int i = 0; i < 9; ++i) {
          B.Jdt[i] = cJ[j][i];
          B.Adt[i] = Ah.m[i] * dt;
          B.Rdt[i] = nrot.m[i] * dt;
        }
        for (int i = 0; i < 6; ++i) B.Mdt[i] = Mx[i] * dt;
        for (int i = 0; i < 4; ++i) B.NM[i] = NM[i];
        B.dt = dt;
        for (int c = 0; c < 3; ++c) {
          const double dvel = dt * (ra.v[c] + sx.g[c]);
          const double dpos = dt * sx.vel[c];
          sx.pos[c] = sx.pos[c] + dpos;
          sx.vel[c] = sx.vel[c] + dvel;
        }
        D33 E;
        for (int i = 0; i < 9; ++i) E.m[i] = cE[j][i];
        const D33 rn = d33_mul(rot, E);
        d33_to_cm(rn, sx.rot);
        if (!last) {  // the IMU pose at the pair's tail
          const D3 rb = d33_mulv(rn, D3{{ab[0], ab[1], ab[2]}});
          double* p = im->pose[np];
          p[0] = cT[j] - im->pkg.cloud_start;
          for (int c = 0; c < 3; ++c) {
            im->last_gyro[c] = cG[j][c] - sx.bg[c];
            im->last_acc[c] = rb.v[c] + sx.g[c];
            p[1 + c] = im->last_acc[c];
            p[4 + c] = im->last_gyro[c];
            p[7 + c] = sx.vel[c];
            p[10 + c] = sx.pos[c];
          }
          for (int c = 0; c < 9; ++c) p[13 + c] = rn.m[c];
        }
      }
      __syncthreads();
      predict_cov(B, q, P, FP, lane);
      if (!last) ++np;
    }
  }
  for (int i = lane; i < 529; i += 64) f->P[i] = P[(i % 23) * 23 + i / 23];
  if (lane == 0) {
    f->x = sx;
    if (m > 0) {
      im->last_imu_t = im->pkg.s[m - 1][0];
      for (int c = 0; c < 3; ++c) {
        im->last_imu_acc[c] = im->pkg.s[m - 1][1 + c];
        im->last_imu_gyro[c] = im->pkg.s[m - 1][4 + c];
      }
    }
    im->last_cloud_end = im->pkg.cloud_end;
    im->npose = np;
    // running minimum of the offsets of the interval heads (entries 0..np-2)
    double mn = 1e300;
    for (int i = np - 2; i >= 0; --i) {
      mn = fmin(mn, im->pose[i][0]);
      im->pose_min[i] = mn;
    }
  }
}

// compensation of point p with interval (head h, tail h+1) at time td (s)
__device__ inline void compensate(const ImuDev* im, int h, double td, const D33& Rc, const D3& pc, const D33& Re,
                                  const D3& te, double p[3]) {
  const double* hd = im->pose[h];
  const double* tl = im->pose[h + 1];
  const double dt = td - hd[0];
  const double w[3] = {dt * tl[4], dt * tl[5], dt * tl[6]};
  const D33 E = so3_exp_d(w);
  D33 Hr;
  for (int i = 0; i < 9; ++i) Hr.m[i] = hd[13 + i];
  const D33 Pr = d33_mul(Hr, E);
  double pp[3];
  for (int k = 0; k < 3; ++k) pp[k] = (hd[10 + k] + dt * hd[7 + k]) + dt * (dt * (0.5 * tl[1 + k]));
  const D3 a = d33_mulv(Re, D3{{p[0], p[1], p[2]}});
  const D3 b = d33_mulv(Pr, D3{{a.v[0] + te.v[0], a.v[1] + te.v[1], a.v[2] + te.v[2]}});
  double c[3];
  for (int k = 0; k < 3; ++k) c[k] = (b.v[k] + pp[k]) - pc.v[k];
  const D3 d = d33_mulv(d33_t(Rc), D3{{c[0], c[1], c[2]}});
  const D3 e = d33_mulv(d33_t(Re), D3{{d.v[0] - te.v[0], d.v[1] - te.v[1], d.v[2] - te.v[2]}});
  for (int k = 0; k < 3; ++k) p[k] = e.v[k];
}

// backward compensation (lio_builder.cpp:121-152); grid (x: points, y: sessions)
__global__ void k_undistort(const PreJob* jobs, int nsess) {
  for (int s = blockIdx.y; s < nsess; s += gridDim.y) {
    const PreJob J = jobs[s];
    const ImuDev* im = J.imu;
    const int np = im->npose;
    const vmp_state& x = J.f->x;
    const D33 Rc = d33_from_cm(x.rot), Re = d33_from_cm(x.rot_ext);
    const D3 pc{{x.pos[0], x.pos[1], x.pos[2]}}, te{{x.pos_ext[0], x.pos_ext[1], x.pos_ext[2]}};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < J.n; i += gridDim.x * blockDim.x) {
      const double td = double(J.t[i]) / double(1000);
      double p[3] = {double(J.x[i]), double(J.y[i]), double(J.z[i])};
      // interval: the last head h (<= np - 2) with td > min(offsets h..np-2)
      int lo = 0, hi = np - 2, h = -1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (td > im->pose_min[mid]) {
          h = mid;
          lo = mid + 1;
        } else {
          hi = mid - 1;
        }
      }
      if (h >= 0) {
        compensate(im, h, td, Rc, pc, Re, te, p);
        if (i == 0)  // the walk stays on the first point: every earlier interval it is later than
          for (int h2 = h - 1; h2 >= 0; --h2)
            if (td > im->pose[h2][0]) compensate(im, h2, td, Rc, pc, Re, te, p);
      }
      J.ox[i] = float(p[0]);
      J.oy[i] = float(p[1]);
      J.oz[i] = float(p[2]);
    }
  }
}


// ------------------------------------------------ VoxelGrid scan filter ----
// PCL VoxelGrid (lio_builder.cpp:13-14, 215-223) on the undistorted points of
// several sessions at once.  Keys are (session << 32) | linear leaf index of
// fp32 floor(p * (1 / leaf)) - min index; a stable radix sort groups each
// leaf's points in point order; every leaf becomes the fp32 mean of its
// points, leaves in ascending index order.  A cloud whose leaf count would
// overflow 32-bit indices passes unfiltered (its points keep their order).
struct VgJob {
  const float* x;  // undistorted points
  const float* y;
  const float* z;
  float* ox;       // filtered points
  float* oy;
  float* oz;
  int32_t n, off;  // points, offset of point 0 in the concatenated key list
  float inv;       // 1 / leaf
  int32_t pad;
};
// per session: float min/max as ordered integers (6), pass-through flag, the
// first segment, the leaf count
constexpr int kVgStat = 10;

// fp32 <-> int32 with the same order (negative floats get their magnitude
// bits flipped), for atomic min / max
__device__ inline int32_t f2ord(float v) {
  const int32_t i = __float_as_int(v);
  return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ inline float ord2f(int32_t o) { return __int_as_float(o >= 0 ? o : (o ^ 0x7fffffff)); }

__global__ void k_vg_init(int32_t* st, int nsess) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsess) return;
  for (int a = 0; a < 3; ++a) {
    st[s * kVgStat + a] = 0x7fffffff;
    st[s * kVgStat + 3 + a] = int32_t(0x80000000u);
  }
  for (int k = 6; k < kVgStat; ++k) st[s * kVgStat + k] = 0;
}

__global__ void k_vg_minmax(const VgJob* jobs, int nsess, int32_t* st) {
  for (int s = blockIdx.y; s < nsess; s += gridDim.y) {
    const VgJob J = jobs[s];
    int32_t mn[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
    int32_t mx[3] = {int32_t(0x80000000u), int32_t(0x80000000u), int32_t(0x80000000u)};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < J.n; i += gridDim.x * blockDim.x) {
      const float p[3] = {J.x[i], J.y[i], J.z[i]};
      for (int a = 0; a < 3; ++a) {
        mn[a] = min(mn[a], f2ord(p[a]));
        mx[a] = max(mx[a], f2ord(p[a]));
      }
    }
    for (int a = 0; a < 3; ++a) {
      for (int o = 32; o > 0; o >>= 1) {
        mn[a] = min(mn[a], __shfl_xor(mn[a], o, 64));
        mx[a] = max(mx[a], __shfl_xor(mx[a], o, 64));
      }
    }
    if ((threadIdx.x & 63) == 0)
      for (int a = 0; a < 3; ++a) {
        atomicMin(&st[s * kVgStat + a], mn[a]);
        atomicMax(&st[s * kVgStat + 3 + a], mx[a]);
      }
  }
}

// leaf geometry of a session from its min/max (the same fp32 expressions as
// the filter's): minimum index, multipliers, pass-through when the leaf
// count overflows 32 bits
__device__ inline bool vg_geometry(const int32_t* st, float inv, int minb[3], int mul[3]) {
  float mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = ord2f(st[a]);
    mx[a] = ord2f(st[3 + a]);
  }
  int64_t d = 1;
  for (int a = 0; a < 3; ++a) d *= int64_t((mx[a] - mn[a]) * inv) + 1;
  int divb[3];
  for (int a = 0; a < 3; ++a) {
    minb[a] = int(floorf(mn[a] * inv));
    divb[a] = int(floorf(mx[a] * inv)) - minb[a] + 1;
  }
  mul[0] = 1;
  mul[1] = divb[0];
  mul[2] = divb[0] * divb[1];
  return d > int64_t(2147483647);
}

__global__ void k_vg_keys(const VgJob* jobs, int nsess, const int32_t* st, uint64_t* keys, int32_t* vals) {
  for (int s = blockIdx.y; s < nsess; s += gridDim.y) {
    const VgJob J = jobs[s];
    int minb[3], mul[3];
    const bool pass = vg_geometry(st + s * kVgStat, J.inv, minb, mul);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < J.n; i += gridDim.x * blockDim.x) {
      uint32_t lin;
      if (pass) {
        lin = uint32_t(i);
      } else {
        const float p[3] = {J.x[i], J.y[i], J.z[i]};
        int l = 0;
        for (int a = 0; a < 3; ++a) l += int(floorf(p[a] * J.inv) - float(minb[a])) * mul[a];
        lin = uint32_t(l);
      }
      keys[J.off + i] = (uint64_t(uint32_t(s)) << 32) | uint64_t(lin);
      vals[J.off + i] = i;
    }
  }
}

__global__ void k_vg_heads(const uint64_t* keys, int ntot, int32_t* flags) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ntot; j += gridDim.x * blockDim.x)
    flags[j] = (j == 0 || keys[j] != keys[j - 1]) ? 1 : 0;
}

// per session: its first segment and leaf count (segments start at the
// ascending positions in sel; the session's points occupy [off, off + n))
__global__ void k_vg_ranges(const VgJob* jobs, int nsess, const int32_t* sel, const int32_t* nseg, int32_t* st) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsess) return;
  const int ns = *nseg;
  auto lb = [&](int v) {
    int lo = 0, hi = ns;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (sel[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const int lo = lb(jobs[s].off), hi = lb(jobs[s].off + jobs[s].n);
  st[s * kVgStat + 7] = lo;
  st[s * kVgStat + 8] = hi - lo;
}

__global__ void k_vg_centroid(const VgJob* jobs, const uint64_t* keys, const int32_t* vals, const int32_t* sel,
                              const int32_t* nseg, int ntot, const int32_t* st) {
  const int ns = *nseg;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ns; q += gridDim.x * blockDim.x) {
    const int b = sel[q], e = q + 1 < ns ? sel[q + 1] : ntot;
    const int s = int(keys[b] >> 32);
    const VgJob J = jobs[s];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int k = b; k < e; ++k) {
      const int i = vals[k];
      sx += J.x[i];
      sy += J.y[i];
      sz += J.z[i];
    }
    const float c = float(e - b);
    const int o = q - st[s * kVgStat + 7];
    J.ox[o] = sx / c;
    J.oy[o] = sy / c;
    J.oz[o] = sz / c;
  }
}
}  // namespace vmp
