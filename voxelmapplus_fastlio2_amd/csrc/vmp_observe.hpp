// Observe pass: the measurement function of the IESKF
// (LIOBuilder::sharedUpdateFunc, lio_builder.cpp:250-311) for one or many
// sessions per launch.
//
// Per point (one lane, SoA point stream, grid-stride over PPT points):
//   u   = r_wl p + (p_wl - anchor)   compensated fp32 (hi/lo split of r_wl
//                                    and t, TwoProduct/TwoSum) ~ 1e-12 m
//   key = floor(u / vs) + anchor     fp32 unless within 1e-4 voxel of a face,
//                                    then the reference's fp64 expression
//   hash probe -> voxel id -> 32 B plane record (mean relative to the voxel
//   corner, normal, flags); residual + 3-sigma gate (voxel_map.cpp:258-276);
//   the slot (normal, residual | NaN = invalid) persists across passes and
//   scans exactly like residual_info (stale on a miss).
//   Valid slots contribute w J^T J and w J r (lio_builder.cpp:283-307).
// Per-point math is fp32; per-lane sums, the wave/block reduction and the
// fixed-order cross-block reduction are fp64.  The block decomposition of a
// session depends only on its point count, so a session's H, b are bitwise
// identical alone or inside a batch.
#pragma once
#include "vmp_common.hpp"

namespace vmp {

constexpr int kObsThreads = 256;
constexpr int kObsGroup = 2;  // points per lane in flight together

// length of one block's partial row: the H upper triangle, b, four counters
template <int NJ>
__host__ __device__ constexpr int obs_row() {
  return NJ * (NJ + 1) / 2 + NJ + 4;
}
constexpr int kObsRowMax = 96;  // >= obs_row<12>() = 94
// distance between consecutive partial rows: rows of up to 32 entries are
// padded to 32 (zeros), so a wave reads two whole rows per load, each lane
// the same column every time
template <int NJ>
__host__ __device__ constexpr int obs_stride() {
  return obs_row<NJ>() <= 32 ? 32 : obs_row<NJ>();
}

struct ObsJob {
  const float* x;
  const float* y;
  const float* z;
  float4* slots;
  const HashEntry* table;
  const VoxHot* hot;
  const ObsParams* prm;
  double* partials;       // [nblocks][obs_row]
  int32_t* counters;      // hits, plane_hits, effective, fallback (+ error)
  double* out;            // H (144, column-major) then b (12)
  uint32_t hmask;
  int32_t n;
  int32_t nblocks;
  int32_t block0;
  int32_t ppt;            // points per thread
  int32_t pad;
};

__device__ inline void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Wave-level fp64 sum (64 lanes, result in every lane)
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// DPP lane moves within 16-lane rows (no LDS traffic): quad_perm [1,0,3,2]
// (0xB1), quad_perm [2,3,0,1] (0x4E), row_half_mirror (0x141), row_mirror
// (0x140).  Every source lane lies inside the row, so all lanes are written.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  return __hiloint2double(dpp_i32<CTRL>(__double2hiint(v)), dpp_i32<CTRL>(__double2loint(v)));
}
__device__ __forceinline__ double rl_f64(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}
// 64-lane sums, valid in lane 0: two quad butterflies and the two mirrors
// leave each 16-lane row's sum in all of its lanes (x + y and y + x round the
// same); the row sums of lanes 0, 16, 32, 48 are added in that order.
__device__ inline double wave_sum0(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  return ((rl_f64(v, 0) + rl_f64(v, 16)) + rl_f64(v, 32)) + rl_f64(v, 48);
}
__device__ inline int wave_sum0_i(int v) {
  v += dpp_i32<0xB1>(v);
  v += dpp_i32<0x4E>(v);
  v += dpp_i32<0x141>(v);
  v += dpp_i32<0x140>(v);
  return ((__builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16)) + __builtin_amdgcn_readlane(v, 32)) +
         __builtin_amdgcn_readlane(v, 48);
}

// Jobs and their parameter blocks do not change while a pass runs: reading
// them through the constant address space lets the compiler use scalar loads
// (one per workgroup, no per-lane vector loads inside the point loop).
template <class T>
using kconst = const __attribute__((address_space(4))) T;
template <class T>
__device__ __forceinline__ kconst<T>* as_const(const T* p) {
  return (kconst<T>*)(p);
}
typedef float vmp_f4 __attribute__((ext_vector_type(4)));
typedef float vmp_f3 __attribute__((ext_vector_type(3)));
typedef uint32_t vmp_u3 __attribute__((ext_vector_type(3)));
static_assert(sizeof(vmp_f3) == 16 && sizeof(vmp_u3) == 16, "3-vectors step by 16 bytes");

__device__ inline int find_job(kconst<ObsJob>* jobs, int njobs, int blk) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].block0 <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Linear probe from slot h, whose entry e was already read; each further
// entry's key and id are read in one load.
__device__ inline int32_t hash_resolve(const HashEntry* table, uint32_t mask, uint64_t key, uint32_t h, vmp_u3 e,
                                       int32_t* err) {
  for (uint32_t probe = 0;; ++probe) {
    const uint64_t k = uint64_t(e.x) | (uint64_t(e.y) << 32);
    if (k == key) return int32_t(e.z);
    if (k == kKeyEmpty) return -1;
    if (probe == mask) break;
    h = (h + 1) & mask;
    e = *as_global(reinterpret_cast<const vmp_u3*>(table + h));
  }
  if (err) atomicOr(err, 1);
  return -1;
}

// W32 selects the plain fp32 world-frame arithmetic of the precision study
// (p_w = fl(R) p + fl(t), key = floorf(p_w / vs), residual against the fp32
// world mean); the fourth counter then counts keys that differ from the
// fp64 key.  The registration path uses W32 = false.
template <int NJ, bool W32 = false>
__global__ void __launch_bounds__(kObsThreads) observe_kernel(const ObsJob* __restrict__ jobs_in, int njobs) {
  constexpr int NH = NJ * (NJ + 1) / 2;
  constexpr int NACC = NH + NJ;
  __shared__ double red[kObsThreads / 64][NACC];
  __shared__ int redi[kObsThreads / 64][4];

  kconst<ObsJob>* jobs = as_const(jobs_in);
  const int ji = find_job(jobs, njobs, blockIdx.x);
  kconst<ObsJob>& J = jobs[ji];
  kconst<ObsParams>& P = *as_const(J.prm);
  if (!P.active) return;  // IESKF converged: reference leaves the loop
  const int lb = blockIdx.x - J.block0;
  if (lb >= J.nblocks) return;  // a grid rounded up past the last session's blocks
  const int tid = threadIdx.x;
  const int n = J.n, ppt = J.ppt;
  const auto X = as_global(J.x);
  const auto Y = as_global(J.y);
  const auto Z = as_global(J.z);
  const auto SL = as_global(reinterpret_cast<vmp_f4*>(J.slots));
  const auto HOT = as_global(reinterpret_cast<const vmp_f4*>(J.hot));

  double acc[NACC];
#pragma unroll
  for (int k = 0; k < NACC; ++k) acc[k] = 0.0;
  int c_hit = 0, c_plane = 0, c_eff = 0, c_fb = 0;

  // kObsGroup points per lane are in flight together: their loads, their
  // hash probes and their plane gathers are each issued as one round, then
  // the points are finished (slot store, accumulation) in index order.
  for (int k0 = 0; k0 < ppt; k0 += kObsGroup) {
    int idx[kObsGroup];
    bool live[kObsGroup];
    float px[kObsGroup], py[kObsGroup], pz[kObsGroup];
    bool zfix[kObsGroup];
    vmp_f4 slot[kObsGroup];
#pragma unroll
    for (int g = 0; g < kObsGroup; ++g) {
      idx[g] = (lb * ppt + k0 + g) * kObsThreads + tid;
      live[g] = k0 + g < ppt && idx[g] < n;
      px[g] = py[g] = pz[g] = 0.0f;
      slot[g] = vmp_f4{0.0f, 0.0f, 0.0f, 0.0f};
      if (live[g]) {
        px[g] = X[idx[g]];
        py[g] = Y[idx[g]];
        pz[g] = Z[idx[g]];
        slot[g] = SL[idx[g]];
      }
    }
    if (!live[0]) break;
    float uh[kObsGroup][3], ul[kObsGroup][3], kl[kObsGroup][3], pw[kObsGroup][3];
    int64_t key[kObsGroup][3];
    uint64_t pk[kObsGroup];
    uint32_t hs[kObsGroup];
    vmp_u3 he[kObsGroup];
#pragma unroll
    for (int g = 0; g < kObsGroup; ++g) {
      zfix[g] = (pz[g] == 0.0f);  // calcBodyCov mutation, commons.cpp:20-23
      if (zfix[g]) pz[g] = 0.001f;
      const float p[3] = {px[g], py[g], pz[g]};
      if constexpr (W32) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          pw[g][r] =
              P.rwl32[r * 3] * p[0] + P.rwl32[r * 3 + 1] * p[1] + P.rwl32[r * 3 + 2] * p[2] + float(P.p_wl[r]);
          key[g][r] = int64_t(floorf(pw[g][r] / P.vs));
        }
        const double pd[3] = {double(p[0]), double(p[1]), zfix[g] ? 0.001 : double(p[2])};
        bool differs = false;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          double s = P.r_wl[r * 3] * pd[0];
          s += P.r_wl[r * 3 + 1] * pd[1];
          s += P.r_wl[r * 3 + 2] * pd[2];
          differs |= int64_t(floor((s + P.p_wl[r]) / P.vs64)) != key[g][r];
        }
        c_fb += (live[g] && differs) ? 1 : 0;
      } else {
        // ---- compensated transform into the anchor frame
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          float s = P.th[r], c = P.tl[r];
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const float hi = P.Rh[r * 3 + q];
            const float prod = hi * p[q];
            const float perr = __builtin_fmaf(hi, p[q], -prod);
            float s2, e2;
            two_sum(s, prod, s2, e2);
            s = s2;
            c += perr + e2 + P.Rl[r * 3 + q] * p[q];
          }
          float h2, l2;
          two_sum(s, c, h2, l2);
          uh[g][r] = h2;
          ul[g][r] = l2;
        }
        // ---- key (bit-exact with the reference's fp64 floor(p_w / vs))
        bool near_face = false;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const float f = (uh[g][r] + ul[g][r]) * P.inv_vs;
          kl[g][r] = floorf(f);
          const float fr = f - kl[g][r];
          near_face |= (fr < 1e-4f) || (fr > 1.0f - 1e-4f);
        }
        if (live[g] && near_face) {
          ++c_fb;
          const double pd[3] = {double(p[0]), double(p[1]), zfix[g] ? 0.001 : double(p[2])};
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            double s = P.r_wl[r * 3] * pd[0];
            s += P.r_wl[r * 3 + 1] * pd[1];
            s += P.r_wl[r * 3 + 2] * pd[2];
            key[g][r] = int64_t(floor((s + P.p_wl[r]) / P.vs64));
            kl[g][r] = float(key[g][r] - P.anchor[r]);
          }
        } else {
#pragma unroll
          for (int r = 0; r < 3; ++r) key[g][r] = int64_t(kl[g][r]) + P.anchor[r];
        }
      }
      // first probe of the voxel's hash slot
      const bool in_range = live[g] && key_ok(key[g][0]) && key_ok(key[g][1]) && key_ok(key[g][2]);
      pk[g] = in_range ? pack_key(key[g][0], key[g][1], key[g][2]) : kKeyEmpty;
      // (read unconditionally: slot 0 stands in for points without a key, so
      // the loads stay in flight together; every load's whole destination
      // is used, so no register is reused while it is pending)
      hs[g] = in_range ? hash_slot(pk[g], J.hmask) : 0u;
      he[g] = *as_global(reinterpret_cast<const vmp_u3*>(J.table + hs[g]));
    }
    int32_t vid[kObsGroup];
    vmp_f4 h0[kObsGroup];
    vmp_f3 h1[kObsGroup];
#pragma unroll
    for (int g = 0; g < kObsGroup; ++g)
      vid[g] = pk[g] == kKeyEmpty ? -1 : hash_resolve(J.table, J.hmask, pk[g], hs[g], he[g], J.counters + 4);
#pragma unroll
    for (int g = 0; g < kObsGroup; ++g) {
      // both halves of the 32 B plane record are loaded together (record 0
      // stands in on a miss); the plane test below only selects the result
      const size_t hv = vid[g] >= 0 ? size_t(vid[g]) : 0;
      h0[g] = HOT[2 * hv];
      h1[g] = *as_global(reinterpret_cast<const vmp_f3*>(J.hot) + 2 * hv + 1);
    }
#pragma unroll
    for (int g = 0; g < kObsGroup; ++g) {
      if (!live[g]) continue;
      const float qx0 = px[g], qy0 = py[g], qz0 = pz[g];
      const float r2 = qx0 * qx0 + qy0 * qy0 + qz0 * qz0;
      const float ca = r2 * P.s2;
      const float cb = (P.rv - ca) / r2;
      const float w0 = P.rwl32[0] * qx0 + P.rwl32[1] * qy0 + P.rwl32[2] * qz0;
      const float w1 = P.rwl32[3] * qx0 + P.rwl32[4] * qy0 + P.rwl32[5] * qz0;
      const float w2 = P.rwl32[6] * qx0 + P.rwl32[7] * qy0 + P.rwl32[8] * qz0;
      vmp_f4 sl = slot[g];
      if (vid[g] >= 0) {
        ++c_hit;
        const bool plane = (__float_as_uint(h0[g].w) & kPlane) != 0;
        c_plane += plane ? 1 : 0;
        const float nx = h1[g].x, ny = h1[g].y, nz = h1[g].z;
        float res;
        if constexpr (W32) {
          // the plane mean held as an fp32 world point
          const float mx = float(double(key[g][0]) * P.vs64 + double(h0[g].x));
          const float my = float(double(key[g][1]) * P.vs64 + double(h0[g].y));
          const float mz = float(double(key[g][2]) * P.vs64 + double(h0[g].z));
          res = nx * (pw[g][0] - mx) + ny * (pw[g][1] - my) + nz * (pw[g][2] - mz);
        } else {
          const float qx = (uh[g][0] - kl[g][0] * P.vs) + ul[g][0];
          const float qy = (uh[g][1] - kl[g][1] * P.vs) + ul[g][1];
          const float qz = (uh[g][2] - kl[g][2] * P.vs) + ul[g][2];
          res = nx * (qx - h0[g].x) + ny * (qy - h0[g].y) + nz * (qz - h0[g].z);
        }
        // sigma^2 = n^T (R C_l R^T + [p]x Prr [p]x^T + Ppp) n   (plane_cov term is 0)
        const float nn = nx * nx + ny * ny + nz * nz;
        const float nw = nx * w0 + ny * w1 + nz * w2;
        const float cx = ny * qz0 - nz * qy0, cy = nz * qx0 - nx * qz0, cz = nx * qy0 - ny * qx0;  // n x p
        kconst<float>* A = P.Prr;
        kconst<float>* B = P.Ppp;
        const float qa = cx * (A[0] * cx + A[1] * cy + A[2] * cz) + cy * (A[1] * cx + A[3] * cy + A[4] * cz) +
                         cz * (A[2] * cx + A[4] * cy + A[5] * cz);
        const float qb = nx * (B[0] * nx + B[1] * ny + B[2] * nz) + ny * (B[1] * nx + B[3] * ny + B[4] * nz) +
                         nz * (B[2] * nx + B[4] * ny + B[5] * nz);
        const float sig2 = ca * nn + cb * nw * nw + qa + qb;
        const bool valid = fabsf(res) < 3.0f * sqrtf(sig2);
        const float qnan = __int_as_float(0x7fc00000);
        if (plane) sl = vmp_f4{nx, ny, nz, valid ? res : qnan};
        else sl.w = qnan;
        SL[idx[g]] = sl;
      }
      if (!isnan(sl.w)) {
        ++c_eff;
        const float nx = sl.x, ny = sl.y, nz = sl.z, res = sl.w;
        const float nw = nx * w0 + ny * w1 + nz * w2;
        const float rc = ca * (nx * nx + ny * ny + nz * nz) + cb * nw * nw;
        const float wi = rc < 0.0002f ? 5000.0f : 1.0f / rc;
        // a' = rot^T n ; v = rot_ext p + pos_ext ; J_rot = v x a'
        kconst<float>* Rt = P.rot;
        const float ax = Rt[0] * nx + Rt[3] * ny + Rt[6] * nz;
        const float ay = Rt[1] * nx + Rt[4] * ny + Rt[7] * nz;
        const float az = Rt[2] * nx + Rt[5] * ny + Rt[8] * nz;
        kconst<float>* E = P.ext_R;
        const float vx = E[0] * qx0 + E[1] * qy0 + E[2] * qz0 + P.ext_t[0];
        const float vy = E[3] * qx0 + E[4] * qy0 + E[5] * qz0 + P.ext_t[1];
        const float vz = E[6] * qx0 + E[7] * qy0 + E[8] * qz0 + P.ext_t[2];
        float Jv[NJ];
        Jv[0] = nx;
        Jv[1] = ny;
        Jv[2] = nz;
        Jv[3] = vy * az - vz * ay;
        Jv[4] = vz * ax - vx * az;
        Jv[5] = vx * ay - vy * ax;
        if constexpr (NJ == 12) {
          kconst<float>* W = P.rwl32;
          const float bx = W[0] * nx + W[3] * ny + W[6] * nz;
          const float by = W[1] * nx + W[4] * ny + W[7] * nz;
          const float bz = W[2] * nx + W[5] * ny + W[8] * nz;
          Jv[6] = qy0 * bz - qz0 * by;
          Jv[7] = qz0 * bx - qx0 * bz;
          Jv[8] = qx0 * by - qy0 * bx;
          Jv[9] = ax;
          Jv[10] = ay;
          Jv[11] = az;
        }
        int o = 0;
#pragma unroll
        for (int a = 0; a < NJ; ++a) {
          const double wa = double(wi) * double(Jv[a]);
#pragma unroll
          for (int b2 = a; b2 < NJ; ++b2) acc[o++] += wa * double(Jv[b2]);
        }
#pragma unroll
        for (int a = 0; a < NJ; ++a) acc[NH + a] += double(wi) * double(Jv[a]) * double(res);
      }
    }
  }

  // ---- block reduction (fp64, fixed order) into this block's partial row:
  //      NACC sums, then hits, plane hits, effective, fallbacks.  No fence or
  //      ticket: k_obs_reduce runs after this launch and the kernel boundary
  //      makes the rows visible (an agent-scope fence per block would write
  //      back and invalidate the whole L2 of its XCD).
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int k = 0; k < NACC; ++k) {
    const double s = wave_sum0(acc[k]);
    if (lane == 0) red[wv][k] = s;
  }
  c_hit = wave_sum0_i(c_hit);
  c_plane = wave_sum0_i(c_plane);
  c_eff = wave_sum0_i(c_eff);
  c_fb = wave_sum0_i(c_fb);
  if (lane == 0) {
    redi[wv][0] = c_hit;
    redi[wv][1] = c_plane;
    redi[wv][2] = c_eff;
    redi[wv][3] = c_fb;
  }
  __syncthreads();
  const auto row = as_global(J.partials + size_t(lb) * obs_stride<NJ>());
  if (tid < NACC) {
    double s = red[0][tid];
    for (int w = 1; w < kObsThreads / 64; ++w) s += red[w][tid];
    row[tid] = s;
  }
  if (tid >= obs_row<NJ>() && tid < obs_stride<NJ>()) row[tid] = 0.0;
  if (tid < 4) {
    int s = 0;
    for (int w = 0; w < kObsThreads / 64; ++w) s += redi[w][tid];
    row[NACC + tid] = double(s);
  }
}

// Reduce the partial rows of every job (one workgroup per session) in a fixed
// order: row b is added by thread b mod 256, threads are combined wave by
// wave.  Writes H (12x12 column-major, symmetric), b and the counters.
template <int NJ>
__global__ void __launch_bounds__(kObsThreads) k_obs_reduce(const ObsJob* __restrict__ jobs) {
  constexpr int NH = NJ * (NJ + 1) / 2;
  constexpr int NACC = NH + NJ;
  constexpr int ROW = obs_row<NJ>();
  __shared__ double red[kObsThreads / 64][ROW];
  const ObsJob& J = jobs[blockIdx.x];
  if (!J.prm->active) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double fin[ROW];
#pragma unroll
  for (int k = 0; k < ROW; ++k) fin[k] = 0.0;
  for (int b = tid; b < J.nblocks; b += kObsThreads) {
    const double* src = J.partials + size_t(b) * obs_stride<NJ>();
#pragma unroll
    for (int k = 0; k < ROW; ++k) fin[k] += src[k];
  }
#pragma unroll
  for (int k = 0; k < ROW; ++k) {
    const double s = wave_sum(fin[k]);
    if (lane == 0) red[wv][k] = s;
  }
  __syncthreads();
  if (tid < ROW) {
    double s = red[0][tid];
    for (int w = 1; w < kObsThreads / 64; ++w) s += red[w][tid];
    red[0][tid] = s;
  }
  __syncthreads();
  if (tid < 144) {  // H, 12x12 column-major, symmetric from the upper triangle
    const int r = tid % 12, c = tid / 12;
    double v = 0.0;
    if (r < NJ && c < NJ) {
      const int a = r < c ? r : c, b2 = r < c ? c : r;
      const int o = a * NJ - a * (a - 1) / 2 + (b2 - a);
      v = red[0][o];
    }
    J.out[tid] = v;
  }
  if (tid < 12) J.out[144 + tid] = tid < NJ ? red[0][NH + tid] : 0.0;
  if (tid < 4) J.counters[tid] = int32_t(red[0][NACC + tid]);
}

}  // namespace vmp
