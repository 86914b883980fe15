// IMU side of the registration step on the device: IMU initialisation
// (LIOBuilder::initializeImu, lio_builder.cpp:28-63), IMU propagation of the
// filter state over one package (IESKF::predict, ieskf.cpp:101-123, driven as
// in undistortCloud, lio_builder.cpp:65-123) and the backward compensation of
// every point into the frame at the scan's end (lio_builder.cpp:121-153).
//
//   k_imu_init      one thread per session: means of the cached samples,
//                   gravity alignment, bg, P
//   k_imu_predict   one 64-lane workgroup per session: the package's IMU
//                   pairs in order (per-step increments across lanes, the
//                   state chain on one lane, P = F P F^T + G Q G^T over the
//                   structurally non-zero products across lanes); the IMU
//                   pose table and its running minimum of offsets
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

// IMU propagation of one package (IESKF::predict per IMU pair, ieskf.cpp:
// 83-111, from lio_builder.cpp:80-119): one four-wave workgroup per session,
// the pairs in chunks of up to 64 steps, each chunk in four phases (A-C on
// the first wave's lanes, D on all four waves).
//   A  lanes: each step's input (mid-point acc and gyro, dt), its rotation
//      increment exp((gyro - bg) dt) and -Jr dt (IMU data and the biases,
//      which do not change while a package is propagated)
//   B  the first wave: the state chain in order (velocity, position,
//      rotation; the rotation's entries cross lanes by shuffles), keeping the
//      rotation before each step and the state after it
//   C  lanes: each step's blocks that depend on the rotation before it,
//      (-rot hat(acc - ba)) dt and (-rot) dt, and the IMU pose table entry
//      at the step's tail
//   D  all waves, step after step: P = F P F^T + G Q G^T.  F is the identity
//      except in rows 0-5, 12-14 and 21-22, so F P is formed for those rows
//      only (its other rows are rows of P) and (F P) F^T for those columns
//      only (its other columns are columns of F P).  Every sum runs over the
//      structurally non-zero terms in ascending index order, the order of
//      the dense product's sums; G Q G^T adds to the diagonal blocks 3-5,
//      12-14, 15-17 and 18-20.
// exp(-w) is exp(w)^T bit for bit (the quaternion form only flips signs),
// so F's rotation block is the transpose of the step's increment.
constexpr int kPredChunk = 64;
constexpr int kFpRows = 11;  // rows of F that differ from the identity's
__device__ inline int fp_slot(int r) { return r < 6 ? r : (r >= 12 && r < 15) ? r - 6 : r >= 21 ? r - 12 : -1; }
__device__ inline int fp_row(int slot) { return slot < 6 ? slot : slot < 9 ? slot + 6 : slot + 12; }
// G's row groups: rows 3-5 (ng), 12-14 (na), 15-17 (nbg), 18-20 (nba)
__device__ inline int g_group(int r) {
  return (r >= 3 && r < 6) ? 0 : (r >= 12 && r < 15) ? 1 : (r >= 15 && r < 18) ? 2 : (r >= 18 && r < 21) ? 3 : -1;
}

struct PredShared {
  double P[529];            // row-major
  double FP[kFpRows * 23];  // rows fp_row(0..10) of F P
  double E[kPredChunk][9];  // per step of the chunk: exp((gyro - bg) dt), row-major
  double J[kPredChunk][9];  // -Jr((gyro - bg) dt) dt (F and G)
  double A[kPredChunk][9];  // (-rot hat(acc - ba)) dt
  double Rd[kPredChunk][9]; // (-rot) dt (F and G)
  double acc[kPredChunk][3], gyro[kPredChunk][3];
  double dt[kPredChunk], tt[kPredChunk];
  double R[kPredChunk + 1][9];  // rotation before each step, and after the last
  double vel[kPredChunk + 1][3], pos[kPredChunk + 1][3];  // after step j at j + 1
  double Mx[6], NM[4], q[4];
  double off[kMaxImu + 1];  // pose table offsets
  vmp_state sx;
};

// phase D for step j of the chunk, P in place, on the block's four waves.
// F P: lane c of every wave holds column c of P (the entries its rows need)
// and forms column c of some of the 11 non-identity rows of F P: wave 0 rows
// 0-2, wave 1 rows 3-5, wave 2 rows 12-13, wave 3 rows 14, 21 and 22.
// (F P) F^T + G Q G^T: lane r of every wave reads row r of F P (a row of P
// for an identity row of F) into registers; after a barrier each wave
// rewrites its columns of row r of P (wave 0 columns 0-2, wave 1 columns 3-5,
// wave 2 columns 12-13, wave 3 columns 14, 21, 22), wave 0 also the identity
// columns of a non-identity row (from F P) and the diagonal of G Q G^T's
// identity blocks.  Every entry is the same expression, summed in the same
// order, whichever wave forms it.
constexpr int kPredThreads = 256;
__device__ inline void predict_cov_step(PredShared& S, int j, int w, int l) {
  double* P = S.P;
  const double dt = S.dt[j];
  const double* E = S.E[j];
  const double* Jd = S.J[j];
  const double* Ad = S.A[j];
  const double* Rd = S.Rd[j];
  double* fp = S.FP;
  if (l < 23) {  // F P
    const int c = l;
    if (w == 0) {  // rows 0-2
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        double s = P[r * 23 + c];
        s += dt * P[(12 + r) * 23 + c];
        fp[r * 23 + c] = s;
      }
    } else if (w == 1) {  // rows 3-5: exp(-w) row i = exp(w) column i
      double p[6];
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        p[t] = P[(3 + t) * 23 + c];
        p[3 + t] = P[(15 + t) * 23 + c];
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double s = E[i] * p[0];
        s += E[3 + i] * p[1];
        s += E[6 + i] * p[2];
        s += Jd[i * 3] * p[3];
        s += Jd[i * 3 + 1] * p[4];
        s += Jd[i * 3 + 2] * p[5];
        fp[(3 + i) * 23 + c] = s;
      }
    } else {  // rows 12-14 (waves 2, 3) and 21-22 (wave 3)
      const double md[6] = {S.Mx[0] * dt, S.Mx[1] * dt, S.Mx[2] * dt, S.Mx[3] * dt, S.Mx[4] * dt, S.Mx[5] * dt};
      double p[3], q[3], u[5];
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        p[t] = P[(3 + t) * 23 + c];
        q[t] = P[(12 + t) * 23 + c];
      }
#pragma unroll
      for (int t = 0; t < 5; ++t) u[t] = P[(18 + t) * 23 + c];
      const int i0 = w == 2 ? 0 : 2, i1 = w == 2 ? 2 : 3;
#pragma unroll
      for (int i = 0; i < 3; ++i) {  // rows 12-14
        if (i < i0 || i >= i1) continue;
        double s = Ad[i * 3] * p[0];
        s += Ad[i * 3 + 1] * p[1];
        s += Ad[i * 3 + 2] * p[2];
        s += q[i];
        s += Rd[i * 3] * u[0];
        s += Rd[i * 3 + 1] * u[1];
        s += Rd[i * 3 + 2] * u[2];
        s += md[i * 2] * u[3];
        s += md[i * 2 + 1] * u[4];
        fp[(6 + i) * 23 + c] = s;
      }
      if (w == 3) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {  // rows 21-22
          double s = S.NM[i * 2] * u[3];
          s += S.NM[i * 2 + 1] * u[4];
          fp[(9 + i) * 23 + c] = s;
        }
      }
    }
  }
  __syncthreads();
  // (F P) F^T + G Q G^T: every wave first reads the entries of row r it needs
  const int r = l;
  const int rs = r < 23 ? fp_slot(r) : -1;
  const double* fr = rs >= 0 ? &S.FP[rs * 23] : &P[(r < 23 ? r : 0) * 23];
  double f[11];
  if (r < 23) {
    if (w == 0) {
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        f[t] = fr[t];
        f[3 + t] = fr[12 + t];
      }
    } else if (w == 1) {
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        f[t] = fr[3 + t];
        f[3 + t] = fr[15 + t];
      }
    } else {
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        f[t] = fr[3 + t];
        f[3 + t] = fr[12 + t];
      }
#pragma unroll
      for (int t = 0; t < 5; ++t) f[6 + t] = fr[18 + t];
    }
  }
  __syncthreads();  // every read of P's row r comes before the writes below
  if (r < 23) {
    double* pr = &P[r * 23];
    const int gr = g_group(r);
    if (w == 0) {  // columns 0-2
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double s = f[c];
        s += f[3 + c] * dt;
        pr[c] = s;
      }
      if (rs >= 0) {  // identity columns of a non-identity row: F P's entries
#pragma unroll
        for (int c = 6; c < 12; ++c) pr[c] = fr[c];
#pragma unroll
        for (int c = 15; c < 21; ++c) pr[c] = fr[c];
      } else if (gr >= 2) {  // identity row and column: only G Q G^T's diagonal
        pr[r] = pr[r] + (dt * S.q[gr]) * dt;
      }
    } else if (w == 1) {  // columns 3-5 (+ G Q G^T's block 3-5)
      double o[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double s = f[0] * E[i];
        s += f[1] * E[3 + i];
        s += f[2] * E[6 + i];
        s += f[3] * Jd[i * 3];
        s += f[4] * Jd[i * 3 + 1];
        s += f[5] * Jd[i * 3 + 2];
        o[i] = s;
      }
      if (gr == 0) {
        const double q = S.q[0];
        const double* Gr = &Jd[(r - 3) * 3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double g = (Gr[0] * q) * Jd[i * 3];
          g += (Gr[1] * q) * Jd[i * 3 + 1];
          g += (Gr[2] * q) * Jd[i * 3 + 2];
          o[i] = o[i] + g;
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) pr[3 + i] = o[i];
    } else {  // columns 12-13 (wave 2), 14, 21, 22 (wave 3) (+ G Q G^T's block 12-14)
      const double md[6] = {S.Mx[0] * dt, S.Mx[1] * dt, S.Mx[2] * dt, S.Mx[3] * dt, S.Mx[4] * dt, S.Mx[5] * dt};
      const int i0 = w == 2 ? 0 : 2, i1 = w == 2 ? 2 : 3;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        if (i < i0 || i >= i1) continue;
        double s = f[0] * Ad[i * 3];
        s += f[1] * Ad[i * 3 + 1];
        s += f[2] * Ad[i * 3 + 2];
        s += f[3 + i];
        s += f[6] * Rd[i * 3];
        s += f[7] * Rd[i * 3 + 1];
        s += f[8] * Rd[i * 3 + 2];
        s += f[9] * md[i * 2];
        s += f[10] * md[i * 2 + 1];
        if (gr == 1) {
          const double q = S.q[1];
          const double* Gr = &Rd[(r - 12) * 3];
          double g = (Gr[0] * q) * Rd[i * 3];
          g += (Gr[1] * q) * Rd[i * 3 + 1];
          g += (Gr[2] * q) * Rd[i * 3 + 2];
          s = s + g;
        }
        pr[12 + i] = s;
      }
      if (w == 3) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          double s = f[9] * S.NM[i * 2];
          s += f[10] * S.NM[i * 2 + 1];
          pr[21 + i] = s;
        }
      }
    }
  }
}

__global__ void __launch_bounds__(kPredThreads) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_imu_predict(const PreJob* jobs, int nsess) {
  if (int(blockIdx.x) >= nsess) return;
  const PreJob Jb = jobs[blockIdx.x];
  ImuDev* im = Jb.imu;
  FilterDev* f = Jb.f;
  const int lane = threadIdx.x;  // phases A-C and the set-up run on the first wave's lanes
  __shared__ PredShared S;
  for (int i = lane; i < 529; i += kPredThreads) S.P[(i % 23) * 23 + i / 23] = f->P[i];  // column-major -> rows
  load_state(f->x, S.sx, lane);
  if (lane < 4) S.q[lane] = im->q[lane];
  __syncthreads();
  const int m = im->pkg.m;
  const double lce = im->last_cloud_end;
  const double cstart = im->pkg.cloud_start;
  // active IMU pairs (head k-1 or last_imu, tail k) are those whose tail is
  // not before the last scan end: a prefix of the ascending samples is skipped
  int k0 = 0;
  while (k0 < m && im->pkg.s[k0][0] < lce) ++k0;
  const int nsteps = (m - k0) + 1;  // the pairs, then the step to the sweep's end
  if (lane == 0) {
    double Bx[6];
    getBx_d(S.sx.g, Bx);
    const D33 hg = hat_d(S.sx.g);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 2; ++c) {
        double v = ((-1.0) * hg.m[r * 3]) * Bx[c];
        v += ((-1.0) * hg.m[r * 3 + 1]) * Bx[2 + c];
        v += ((-1.0) * hg.m[r * 3 + 2]) * Bx[4 + c];
        S.Mx[r * 2 + c] = v;
      }
    double N[6];
    getNx_d(S.sx.g, N);
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) {
        double v = N[r * 3] * S.Mx[c];
        v += N[r * 3 + 1] * S.Mx[2 + c];
        v += N[r * 3 + 2] * S.Mx[4 + c];
        S.NM[r * 2 + c] = v;
      }
    double* p = im->pose[0];  // pose 0: the start of the sweep
    const D33 rot = d33_from_cm(S.sx.rot);
    p[0] = 0.0;
    S.off[0] = 0.0;
    for (int c = 0; c < 3; ++c) {
      p[1 + c] = im->last_acc[c];
      p[4 + c] = im->last_gyro[c];
      p[7 + c] = S.sx.vel[c];
      p[10 + c] = S.sx.pos[c];
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
  int np = 1;  // pose table entries
  for (int c0 = 0; c0 < nsteps; c0 += kPredChunk) {
    const int cn = min(kPredChunk, nsteps - c0);
    if (lane < cn) {  // A
      double a[3], g[3], dt, tt;
      step_input(c0 + lane, a, g, dt, tt);
      double wdt[3];
      for (int c = 0; c < 3; ++c) wdt[c] = dt * (g[c] - S.sx.bg[c]);
      const D33 E = so3_exp_d(wdt);
      const D33 Jr = d33_t(left_jac_d(wdt));
      for (int i = 0; i < 9; ++i) {
        S.E[lane][i] = E.m[i];
        S.J[lane][i] = ((-1.0) * Jr.m[i]) * dt;
      }
      for (int c = 0; c < 3; ++c) {
        S.acc[lane][c] = a[c];
        S.gyro[lane][c] = g[c];
      }
      S.dt[lane] = dt;
      S.tt[lane] = tt;
    }
    __syncthreads();
    {  // B: x += delta step after step, the rotation in lanes 0-8 (entry
       // lane), velocity and position in lanes 0-2; the blocks of a step come
       // from the state before it
      vmp_state& x = S.sx;
      double rv = 0.0, vel = 0.0, pos = 0.0;
      if (lane < 9) rv = x.rot[(lane % 3) * 3 + lane / 3];  // column-major storage
      if (lane < 3) {
        vel = x.vel[lane];
        pos = x.pos[lane];
      }
      if (lane < 64) {  // the first wave alone: the rotation's entries cross lanes by shuffles
        const int r = lane < 9 ? lane / 3 : 0, c = lane % 3, vr = lane < 3 ? lane : 0;
        for (int j = 0; j < cn; ++j) {
          if (lane < 9) S.R[j][lane] = rv;  // the rotation before step j (lane k: entry k, row-major)
          const double Rv0 = __shfl(rv, vr * 3, 64), Rv1 = __shfl(rv, vr * 3 + 1, 64), Rv2 = __shfl(rv, vr * 3 + 2, 64);
          const double Rr0 = __shfl(rv, r * 3, 64), Rr1 = __shfl(rv, r * 3 + 1, 64), Rr2 = __shfl(rv, r * 3 + 2, 64);
          const double* E = S.E[j];
          if (lane < 3) {
            const double dt = S.dt[j];
            const double ab0 = S.acc[j][0] - x.ba[0], ab1 = S.acc[j][1] - x.ba[1], ab2 = S.acc[j][2] - x.ba[2];
            double ra = Rv0 * ab0;
            ra += Rv1 * ab1;
            ra += Rv2 * ab2;
            const double dvel = dt * (ra + x.g[lane]);
            const double dpos = dt * vel;
            pos = pos + dpos;
            vel = vel + dvel;
            S.vel[j + 1][lane] = vel;
            S.pos[j + 1][lane] = pos;
          }
          if (lane < 9) {
            double v = Rr0 * E[c];
            v += Rr1 * E[3 + c];
            v += Rr2 * E[6 + c];
            rv = v;
          }
        }
        if (lane < 9) S.R[cn][lane] = rv;
      }
      __syncthreads();
      if (lane < 9) x.rot[(lane % 3) * 3 + lane / 3] = rv;
      if (lane < 3) {
        x.vel[lane] = vel;
        x.pos[lane] = pos;
      }
    }
    __syncthreads();
    if (lane < cn) {  // C
      const int j = lane;
      const double dt = S.dt[j];
      const double ab[3] = {S.acc[j][0] - S.sx.ba[0], S.acc[j][1] - S.sx.ba[1], S.acc[j][2] - S.sx.ba[2]};
      D33 nrot, rot;
      for (int i = 0; i < 9; ++i) {
        rot.m[i] = S.R[j][i];
        nrot.m[i] = (-1.0) * rot.m[i];
      }
      const D33 Ah = d33_mul(nrot, hat_d(ab));
      for (int i = 0; i < 9; ++i) {
        S.A[j][i] = Ah.m[i] * dt;
        S.Rd[j][i] = nrot.m[i] * dt;
      }
      if (c0 + j != nsteps - 1) {  // the IMU pose at the pair's tail
        D33 rn;
        for (int i = 0; i < 9; ++i) rn.m[i] = S.R[j + 1][i];
        const D3 rb = d33_mulv(rn, D3{{ab[0], ab[1], ab[2]}});
        double* p = im->pose[np + j];
        p[0] = S.tt[j] - cstart;
        S.off[np + j] = p[0];
        double la[3], lg[3];
        for (int c = 0; c < 3; ++c) {
          lg[c] = S.gyro[j][c] - S.sx.bg[c];
          la[c] = rb.v[c] + S.sx.g[c];
          p[1 + c] = la[c];
          p[4 + c] = lg[c];
          p[7 + c] = S.vel[j + 1][c];
          p[10 + c] = S.pos[j + 1][c];
        }
        for (int c = 0; c < 9; ++c) p[13 + c] = rn.m[c];
        if (c0 + j == nsteps - 2)  // the last pair: LIODataGroup's last acc / gyro
          for (int c = 0; c < 3; ++c) {
            im->last_acc[c] = la[c];
            im->last_gyro[c] = lg[c];
          }
      }
    }
    for (int j = 0; j < cn; ++j) {  // D
      __syncthreads();
      predict_cov_step(S, j, lane >> 6, lane & 63);
    }
    np += (c0 + cn == nsteps) ? cn - 1 : cn;
    __syncthreads();
  }
  for (int i = lane; i < 529; i += kPredThreads) f->P[i] = S.P[(i % 23) * 23 + i / 23];
  if (lane == 0) {
    f->x = S.sx;
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
      mn = fmin(mn, S.off[i]);
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
