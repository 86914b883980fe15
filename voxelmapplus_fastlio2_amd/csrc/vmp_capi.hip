// C ABI of the MI355X registration path (include/vmp.h): sessions, device
// memory, and the launch sequences of observe / IESKF / map update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "vmp.h"
#include "vmp_common.hpp"
#include "vmp_ieskf.hpp"
#include "vmp_mapupd.hpp"
#include "vmp_observe.hpp"

using namespace vmp;

// ------------------------------------------------------------ kernels -----
__global__ void k_obs_prep(FilterDev* f) {
  if (threadIdx.x == 0) make_obs_params(f->x, f->P, f->cfg, f->obs_prm);
}
__global__ void k_upd_prep(FilterDev* f) {
  if (threadIdx.x == 0) make_upd_params(f->x, f->P, f->cfg, f->upd_prm);
}
struct Diag23 {
  double q[23];
};
__global__ void k_apply_motion(FilterDev* f, D33 dR, D3 dp, Diag23 q) {
  if (threadIdx.x != 0) return;
  // pos += rot * dp ; rot = rot * dR ; P += diag(q)
  const D33 rot = d33_from_cm(f->x.rot);
  const D3 a = d33_mulv(rot, dp);
  for (int i = 0; i < 3; ++i) f->x.pos[i] = f->x.pos[i] + a.v[i];
  d33_to_cm(d33_mul(rot, dR), f->x.rot);
  for (int i = 0; i < 23; ++i) f->P[i * 23 + i] += q.q[i];
}
__global__ void k_zero_i32(int32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// IESKF: start of an update (predict_x = x, P^-1, first observe params)
__global__ void __launch_bounds__(64) k_ieskf_begin(FilterDev** fs) {
  FilterDev* f = fs[blockIdx.x];
  __shared__ double M[529], Inv[529], W[23][47], fac[23];
  const int lane = threadIdx.x;
  for (int i = lane; i < 529; i += 64) M[(i % 23) * 23 + i / 23] = f->P[i];  // column-major -> row-major
  __syncthreads();
  gj_inverse(M, Inv, W, fac, lane);
  for (int i = lane; i < 529; i += 64) f->Pinv[i] = Inv[i];
  if (lane == 0) {
    f->xp = f->x;
    f->iters = 0;
    f->active = 1;
    make_obs_params(f->x, f->P, f->cfg, f->obs_prm);
  }
}

// IESKF iteration after an observe pass (ieskf.cpp:133-149)
__global__ void __launch_bounds__(64) k_ieskf_step(FilterDev** fs, int32_t** obs_ctr, int32_t** stat_out) {
  FilterDev* f = fs[blockIdx.x];
  if (!f->active) return;
  __shared__ double Pi[529], H[529], Inv[529], W[23][47], fac[23];
  __shared__ double dl[23], bb[23], Jd[529];
  __shared__ double JtPi[529];
  const int lane = threadIdx.x;
  if (lane == 0) state_minus(f->x, f->xp, dl);
  for (int i = lane; i < 529; i += 64) Pi[i] = f->Pinv[i];
  __syncthreads();
  __shared__ JacBlocks jb;
  if (lane == 0) jac_blocks(f->x, f->xp, dl, jb);
  __syncthreads();
  for (int i = lane; i < 529; i += 64) Jd[i] = jac_at(jb, i / 23, i % 23);
  __syncthreads();
  // JtPi = J^T P^-1
  for (int i = lane; i < 529; i += 64) {
    const int r = i / 23, c = i % 23;
    double s = Jd[r] * Pi[c];
    for (int k = 1; k < 23; ++k) s += Jd[k * 23 + r] * Pi[k * 23 + c];
    JtPi[i] = s;
  }
  __syncthreads();
  // H_ = JtPi J + [H_obs] ; b_ = JtPi delta + [b_obs]
  for (int i = lane; i < 529; i += 64) {
    const int r = i / 23, c = i % 23;
    double s = JtPi[r * 23] * Jd[c];
    for (int k = 1; k < 23; ++k) s += JtPi[r * 23 + k] * Jd[k * 23 + c];
    if (r < 12 && c < 12) s += f->obs[c * 12 + r];
    H[i] = s;
  }
  if (lane < 23) {
    double s = JtPi[lane * 23] * dl[0];
    for (int k = 1; k < 23; ++k) s += JtPi[lane * 23 + k] * dl[k];
    if (lane < 12) s += f->obs[144 + lane];
    bb[lane] = s;
  }
  __syncthreads();
  gj_inverse(H, Inv, W, fac, lane);
  // delta = -H_^-1 b_
  if (lane < 23) {
    double s = Inv[lane * 23] * bb[0];
    for (int k = 1; k < 23; ++k) s += Inv[lane * 23 + k] * bb[k];
    dl[lane] = -s;
  }
  for (int i = lane; i < 529; i += 64) f->Hinv[i] = Inv[i];
  __syncthreads();
  if (lane == 0) {
    for (int k = 0; k < 23; ++k) f->delta[k] = dl[k];
    state_plus(f->x, dl);
    f->iters += 1;
    double mx = dl[0];
    for (int k = 1; k < 23; ++k) mx = mx > dl[k] ? mx : dl[k];
    if (mx < 0.001 || f->iters >= f->cfg.max_iter) f->active = 0;
    make_obs_params(f->x, f->P, f->cfg, f->obs_prm);
    f->obs_prm.active = f->active;
    int32_t* so = stat_out[blockIdx.x];
    const int32_t* oc = obs_ctr[blockIdx.x];
    for (int k = 0; k < 4; ++k) so[k] = oc[k];
  }
}

// P = L H_^-1 L^T and the insertion parameters of the posterior
__global__ void __launch_bounds__(64) k_ieskf_end(FilterDev** fs) {
  FilterDev* f = fs[blockIdx.x];
  __shared__ double L[529], T[529], Hi[529];
  __shared__ JacBlocks jb;
  const int lane = threadIdx.x;
  if (lane == 0) jac_blocks(f->x, f->xp, f->delta, jb);
  for (int i = lane; i < 529; i += 64) Hi[i] = f->Hinv[i];
  __syncthreads();
  for (int i = lane; i < 529; i += 64) L[i] = jac_at(jb, i / 23, i % 23);
  __syncthreads();
  for (int i = lane; i < 529; i += 64) {
    const int r = i / 23, c = i % 23;
    double s = L[r * 23] * Hi[c];
    for (int k = 1; k < 23; ++k) s += L[r * 23 + k] * Hi[k * 23 + c];
    T[i] = s;
  }
  __syncthreads();
  for (int i = lane; i < 529; i += 64) {
    const int r = i / 23, c = i % 23;
    double s = T[r * 23] * L[c * 23];
    for (int k = 1; k < 23; ++k) s += T[r * 23 + k] * L[c * 23 + k];
    f->P[c * 23 + r] = s;  // column-major
  }
  __syncthreads();
  if (lane == 0) {
    make_upd_params(f->x, f->P, f->cfg, f->upd_prm);
    f->active = 0;
  }
}

// ------------------------------------------------------------- host -------
#define VMP_HIP(call)                                                   \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) {                                             \
      s->err = std::string(#call) + ": " + hipGetErrorString(e_);       \
      return VMP_EDEVICE;                                               \
    }                                                                   \
  } while (0)

struct vmp_ctx {
  int device;
};

struct vmp_session {
  vmp_ctx* ctx = nullptr;
  vmp_map_cfg cfg{};
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void*> allocs;
  // map
  MapDev map{};
  MapCounters* h_ctr = nullptr;  // pinned
  // scan
  float *own_x = nullptr, *own_y = nullptr, *own_z = nullptr;
  const float *cx = nullptr, *cy = nullptr, *cz = nullptr;
  int32_t n = 0;
  float4* slots = nullptr;
  // filter
  FilterDev* f = nullptr;
  FilterDev** f_list = nullptr;  // device array holding f (batch form of one)
  // observe
  ObsJob* d_job = nullptr;
  double* partials = nullptr;
  int32_t* obs_ctr = nullptr;
  int32_t** obs_ctr_list = nullptr;
  int32_t* obs_stat = nullptr;  // last active iteration's counters
  int32_t** obs_stat_list = nullptr;
  unsigned int* ticket = nullptr;
  int max_blocks = 0;
  int ppt = 2;
  // scan scratch
  ScanBufs sb{};
  void* temp = nullptr;
  size_t temp_bytes = 0;
  int32_t* live = nullptr;
  int32_t* nlive = nullptr;
  int32_t* h_obs = nullptr;  // pinned observe counters + stats
  double* h_out = nullptr;   // pinned 156 doubles
  bool has_points = false;
  hipEvent_t ev[4]{};
  int obs_launches = 0;
};

static std::mutex g_mu;

template <class T>
static hipError_t dalloc(vmp_session* s, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count * sizeof(T), 16));
  if (e == hipSuccess) {
    s->allocs.push_back(q);
    *p = static_cast<T*>(q);
  }
  return e;
}

static uint32_t next_pow2(uint64_t v) {
  uint64_t p = 1024;
  while (p < v) p <<= 1;
  return uint32_t(p);
}

static int grid_for(int n, int threads) { return std::max(1, (n + threads - 1) / threads); }

extern "C" const char* vmp_version(void) { return "voxelmapplus_fastlio2_amd 0.1 (gfx950)"; }

extern "C" int vmp_ctx_create(int device, vmp_ctx** out) {
  if (!out) return VMP_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return VMP_EDEVICE;
  vmp_ctx* c = new (std::nothrow) vmp_ctx();
  if (!c) return VMP_ENOMEM;
  c->device = device;
  *out = c;
  return VMP_OK;
}
extern "C" int vmp_ctx_destroy(vmp_ctx* ctx) {
  delete ctx;
  return VMP_OK;
}

struct Sizes {
  int64_t pool, hcap, qcap, maxn, max_blocks;
};
static Sizes sizes_for(const vmp_map_cfg& c) {
  Sizes z;
  z.maxn = c.max_scan_points;
  z.pool = int64_t(c.map_capacity) + c.max_scan_points + 64;
  z.hcap = next_pow2(uint64_t(4) * (uint64_t(c.map_capacity) + uint64_t(c.max_scan_points)));
  z.qcap = 2 * (int64_t(c.map_capacity) + 2 * int64_t(c.max_scan_points)) + 1024;
  z.max_blocks = (c.max_scan_points + kObsThreads - 1) / kObsThreads + 1;
  return z;
}

extern "C" int64_t vmp_session_bytes(const vmp_map_cfg* cfg) {
  if (!cfg) return -1;
  Sizes z = sizes_for(*cfg);
  int64_t b = 0;
  b += z.hcap * int64_t(sizeof(HashEntry));
  b += z.pool * int64_t(sizeof(VoxHot) + sizeof(VoxState) + 8 * kPlaneD + 8 * kMomD + 4 + 4);
  b += z.qcap * 12;
  b += z.maxn * (3 * 4 + 16 + 4 * 20 + 8 * 2);
  b += z.max_blocks * 90 * 8;
  b += int64_t(sizeof(FilterDev)) + (64 << 20) / 16;  // sort scratch estimate
  return b;
}

static int check_cfg(const vmp_map_cfg* c) {
  if (!c) return VMP_EINVAL;
  if (!(c->voxel_size > 0) || c->max_point_thresh < 1 || c->update_size_thresh < 1 || c->map_capacity < 1 ||
      c->max_scan_points < 1 || c->max_iter < 1 || c->max_iter > 64)
    return VMP_EINVAL;
  if (int64_t(c->map_capacity) + c->max_scan_points > (int64_t(1) << 30)) return VMP_EINVAL;
  return VMP_OK;
}

extern "C" int vmp_session_destroy(vmp_session* s) {
  if (!s) return VMP_EINVAL;
  hipSetDevice(s->ctx->device);
  if (s->stream) hipStreamSynchronize(s->stream);
  for (void* p : s->allocs) hipFree(p);
  if (s->h_ctr) hipHostFree(s->h_ctr);
  if (s->h_obs) hipHostFree(s->h_obs);
  if (s->h_out) hipHostFree(s->h_out);
  for (hipEvent_t e : s->ev)
    if (e) hipEventDestroy(e);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
  return VMP_OK;
}

extern "C" int vmp_session_create(vmp_ctx* ctx, const vmp_map_cfg* cfg, vmp_session** out) {
  if (!ctx || !out || check_cfg(cfg) != VMP_OK) return VMP_EINVAL;
  vmp_session* s = new (std::nothrow) vmp_session();
  if (!s) return VMP_ENOMEM;
  s->ctx = ctx;
  s->cfg = *cfg;
  auto fail = [&](int code) {
    vmp_session_destroy(s);
    return code;
  };
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(VMP_EDEVICE);
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return fail(VMP_EDEVICE);
  for (hipEvent_t& e : s->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(VMP_EDEVICE);
  const Sizes z = sizes_for(*cfg);
  const int maxn = int(z.maxn);
  MapDev& m = s->map;
  bool ok = true;
  ok &= dalloc(s, &m.table, size_t(z.hcap)) == hipSuccess;
  ok &= dalloc(s, &m.hot, size_t(z.pool)) == hipSuccess;
  ok &= dalloc(s, &m.vs, size_t(z.pool)) == hipSuccess;
  ok &= dalloc(s, &m.plane, size_t(z.pool) * kPlaneD) == hipSuccess;
  ok &= dalloc(s, &m.mom, size_t(z.pool) * kMomD) == hipSuccess;
  ok &= dalloc(s, &m.freelist, size_t(z.pool)) == hipSuccess;
  ok &= dalloc(s, &m.ring_vid, size_t(z.qcap)) == hipSuccess;
  ok &= dalloc(s, &m.ring_stamp, size_t(z.qcap)) == hipSuccess;
  ok &= dalloc(s, &m.ctr, 1) == hipSuccess;
  ok &= dalloc(s, &s->own_x, maxn) == hipSuccess;
  ok &= dalloc(s, &s->own_y, maxn) == hipSuccess;
  ok &= dalloc(s, &s->own_z, maxn) == hipSuccess;
  ok &= dalloc(s, &s->slots, maxn) == hipSuccess;
  ok &= dalloc(s, &s->f, 1) == hipSuccess;
  ok &= dalloc(s, &s->f_list, 1) == hipSuccess;
  ok &= dalloc(s, &s->d_job, 1) == hipSuccess;
  ok &= dalloc(s, &s->partials, size_t(z.max_blocks) * 90) == hipSuccess;
  ok &= dalloc(s, &s->obs_ctr, 8) == hipSuccess;
  ok &= dalloc(s, &s->obs_ctr_list, 1) == hipSuccess;
  ok &= dalloc(s, &s->obs_stat, 8) == hipSuccess;
  ok &= dalloc(s, &s->obs_stat_list, 1) == hipSuccess;
  ok &= dalloc(s, &s->ticket, 1) == hipSuccess;
  ScanBufs& b = s->sb;
  ok &= dalloc(s, &b.skey_in, maxn) == hipSuccess;
  ok &= dalloc(s, &b.sval_in, maxn) == hipSuccess;
  ok &= dalloc(s, &b.skey_out, maxn) == hipSuccess;
  ok &= dalloc(s, &b.sval_out, maxn) == hipSuccess;
  ok &= dalloc(s, &b.flags, maxn) == hipSuccess;
  ok &= dalloc(s, &b.seg_start, maxn) == hipSuccess;
  ok &= dalloc(s, &b.seg_vid, maxn) == hipSuccess;
  ok &= dalloc(s, &b.seg_entry, maxn) == hipSuccess;
  ok &= dalloc(s, &b.cre_t, maxn) == hipSuccess;
  ok &= dalloc(s, &b.cre_seg, maxn) == hipSuccess;
  ok &= dalloc(s, &b.co_vid, maxn) == hipSuccess;
  ok &= dalloc(s, &b.co_seg, maxn) == hipSuccess;
  ok &= dalloc(s, &b.co_t, maxn) == hipSuccess;
  ok &= dalloc(s, &b.co_prev, maxn) == hipSuccess;
  ok &= dalloc(s, &b.co_group, maxn) == hipSuccess;
  ok &= dalloc(s, &b.ev_vid, maxn) == hipSuccess;
  ok &= dalloc(s, &b.ev_t, maxn) == hipSuccess;
  ok &= dalloc(s, &b.cand_in, maxn) == hipSuccess;
  ok &= dalloc(s, &b.cand_out, maxn) == hipSuccess;
  ok &= dalloc(s, &b.lru_sel, maxn) == hipSuccess;
  ok &= dalloc(s, &s->live, size_t(z.pool)) == hipSuccess;
  ok &= dalloc(s, &s->nlive, 1) == hipSuccess;
  if (!ok) return fail(VMP_ENOMEM);
  if (hipHostMalloc(reinterpret_cast<void**>(&s->h_ctr), sizeof(MapCounters)) != hipSuccess) return fail(VMP_ENOMEM);
  if (hipHostMalloc(reinterpret_cast<void**>(&s->h_obs), 64 * sizeof(int32_t)) != hipSuccess) return fail(VMP_ENOMEM);
  if (hipHostMalloc(reinterpret_cast<void**>(&s->h_out), 160 * sizeof(double)) != hipSuccess) return fail(VMP_ENOMEM);
  m.hmask = uint32_t(z.hcap - 1);
  m.pool = int32_t(z.pool);
  m.qcap = z.qcap;
  m.capacity = cfg->map_capacity;
  m.max_pt = cfg->max_point_thresh;
  m.upd_pt = cfg->update_size_thresh;
  m.plane_thresh = cfg->plane_thresh;
  m.merge_a = cfg->merge_thresh_for_angle;
  m.merge_d = cfg->merge_thresh_for_distance;
  m.vs64 = cfg->voxel_size;
  s->max_blocks = int(z.max_blocks);
  // rocprim scratch: the largest of the sorts and selections at maxn
  size_t t1 = 0, t2 = 0, t3 = 0, t4 = 0;
  rocprim::radix_sort_pairs(nullptr, t1, b.skey_in, b.skey_out, b.sval_in, b.sval_out, size_t(maxn), 0, 32,
                            s->stream);
  rocprim::radix_sort_keys(nullptr, t2, b.cand_in, b.cand_out, size_t(maxn), 0, 64, s->stream);
  rocprim::select(nullptr, t3, rocprim::counting_iterator<int32_t>(0), b.flags, b.seg_start, &m.ctr->nseg,
                  size_t(maxn), s->stream);
  rocprim::select(nullptr, t4, rocprim::counting_iterator<int32_t>(0), b.flags, b.seg_start, &m.ctr->nseg,
                  size_t(std::max<int64_t>(z.qcap, maxn)), s->stream);
  s->temp_bytes = std::max(std::max(t1, t2), std::max(t3, t4)) + 256;
  if (dalloc(s, reinterpret_cast<uint8_t**>(&s->temp), s->temp_bytes) != hipSuccess) return fail(VMP_ENOMEM);
  // initial device state
  const int T = 256;
  hipLaunchKernelGGL(k_rehash_clear, dim3(grid_for(int(std::min<int64_t>(z.hcap, 1 << 20)), T)), dim3(T), 0,
                     s->stream, m);
  MapCounters c0{};
  c0.tag = 0;
  if (hipMemcpyAsync(m.ctr, &c0, sizeof(c0), hipMemcpyHostToDevice, s->stream) != hipSuccess) return fail(VMP_EDEVICE);
  // slots start invalid (NaN residual), voxel records dead
  {
    std::vector<float4> init(size_t(maxn), make_float4(0.f, 0.f, 0.f, __builtin_nanf("")));
    if (hipMemcpyAsync(s->slots, init.data(), sizeof(float4) * size_t(maxn), hipMemcpyHostToDevice, s->stream) !=
        hipSuccess)
      return fail(VMP_EDEVICE);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(VMP_EDEVICE);
  }
  std::vector<VoxState> vs0(size_t(std::min<int64_t>(z.pool, 1 << 16)));
  for (auto& v : vs0) {
    std::memset(&v, 0, sizeof(v));
    v.tag = -1;
  }
  for (int64_t off = 0; off < z.pool; off += int64_t(vs0.size())) {
    const size_t cnt = size_t(std::min<int64_t>(int64_t(vs0.size()), z.pool - off));
    if (hipMemcpy(m.vs + off, vs0.data(), cnt * sizeof(VoxState), hipMemcpyHostToDevice) != hipSuccess)
      return fail(VMP_EDEVICE);
  }
  if (hipMemset(m.hot, 0, sizeof(VoxHot) * size_t(z.pool)) != hipSuccess) return fail(VMP_EDEVICE);
  FilterDev f0;
  std::memset(&f0, 0, sizeof(f0));
  f0.cfg = *cfg;
  for (int i = 0; i < 9; i += 4) {
    f0.x.rot[i] = 1.0;
    f0.x.rot_ext[i] = 1.0;
  }
  f0.x.g[2] = -kGravity;
  for (int i = 0; i < 23; ++i) f0.P[i * 23 + i] = 1.0;
  if (hipMemcpy(s->f, &f0, sizeof(f0), hipMemcpyHostToDevice) != hipSuccess) return fail(VMP_EDEVICE);
  if (hipMemcpy(s->f_list, &s->f, sizeof(FilterDev*), hipMemcpyHostToDevice) != hipSuccess) return fail(VMP_EDEVICE);
  if (hipMemcpy(s->obs_ctr_list, &s->obs_ctr, sizeof(int32_t*), hipMemcpyHostToDevice) != hipSuccess)
    return fail(VMP_EDEVICE);
  if (hipMemcpy(s->obs_stat_list, &s->obs_stat, sizeof(int32_t*), hipMemcpyHostToDevice) != hipSuccess)
    return fail(VMP_EDEVICE);
  if (hipMemset(s->ticket, 0, sizeof(unsigned int)) != hipSuccess) return fail(VMP_EDEVICE);
  if (hipMemset(s->obs_ctr, 0, 8 * sizeof(int32_t)) != hipSuccess) return fail(VMP_EDEVICE);
  if (hipDeviceSynchronize() != hipSuccess) return fail(VMP_EDEVICE);
  *out = s;
  return VMP_OK;
}

extern "C" const char* vmp_last_error(const vmp_session* s) { return s ? s->err.c_str() : "null session"; }

extern "C" int vmp_state_set(vmp_session* s, const vmp_state* x, const double* P) {
  if (!s) return VMP_EINVAL;
  VMP_HIP(hipSetDevice(s->ctx->device));
  if (x) VMP_HIP(hipMemcpyAsync(&s->f->x, x, sizeof(vmp_state), hipMemcpyHostToDevice, s->stream));
  if (P) VMP_HIP(hipMemcpyAsync(s->f->P, P, sizeof(double) * 529, hipMemcpyHostToDevice, s->stream));
  VMP_HIP(hipStreamSynchronize(s->stream));
  return VMP_OK;
}

extern "C" int vmp_state_get(vmp_session* s, vmp_state* x, double* P) {
  if (!s) return VMP_EINVAL;
  VMP_HIP(hipSetDevice(s->ctx->device));
  if (x) VMP_HIP(hipMemcpyAsync(x, &s->f->x, sizeof(vmp_state), hipMemcpyDeviceToHost, s->stream));
  if (P) VMP_HIP(hipMemcpyAsync(P, s->f->P, sizeof(double) * 529, hipMemcpyDeviceToHost, s->stream));
  VMP_HIP(hipStreamSynchronize(s->stream));
  return VMP_OK;
}

static int enqueue_motion(vmp_session* s, const double* dR, const double* dp, const double* q) {
  D33 R = d33_from_cm(dR);
  D3 p{{dp[0], dp[1], dp[2]}};
  Diag23 d;
  for (int i = 0; i < 23; ++i) d.q[i] = q[i];
  hipLaunchKernelGGL(k_apply_motion, dim3(1), dim3(64), 0, s->stream, s->f, R, p, d);
  return VMP_OK;
}

extern "C" int vmp_state_apply_motion(vmp_session* s, const double* dR, const double* dp, const double* qdiag) {
  if (!s || !dR || !dp || !qdiag) return VMP_EINVAL;
  VMP_HIP(hipSetDevice(s->ctx->device));
  enqueue_motion(s, dR, dp, qdiag);
  VMP_HIP(hipStreamSynchronize(s->stream));
  return VMP_OK;
}

extern "C" int vmp_scan_set_points(vmp_session* s, const float* x, const float* y, const float* z, int32_t n) {
  if (!s || n < 0 || n > s->cfg.max_scan_points || (n > 0 && (!x || !y || !z))) {
    if (s) s->err = "scan size exceeds max_scan_points or null input";
    return VMP_EINVAL;
  }
  VMP_HIP(hipSetDevice(s->ctx->device));
  if (n > 0) {
    VMP_HIP(hipMemcpyAsync(s->own_x, x, sizeof(float) * n, hipMemcpyHostToDevice, s->stream));
    VMP_HIP(hipMemcpyAsync(s->own_y, y, sizeof(float) * n, hipMemcpyHostToDevice, s->stream));
    VMP_HIP(hipMemcpyAsync(s->own_z, z, sizeof(float) * n, hipMemcpyHostToDevice, s->stream));
    VMP_HIP(hipStreamSynchronize(s->stream));
  }
  s->cx = s->own_x;
  s->cy = s->own_y;
  s->cz = s->own_z;
  s->n = n;
  s->has_points = true;
  return VMP_OK;
}

extern "C" int vmp_scan_set_points_device(vmp_session* s, const float* dx, const float* dy, const float* dz,
                                          int32_t n) {
  if (!s || n < 0 || n > s->cfg.max_scan_points || (n > 0 && (!dx || !dy || !dz))) {
    if (s) s->err = "scan size exceeds max_scan_points or null input";
    return VMP_EINVAL;
  }
  s->cx = dx;
  s->cy = dy;
  s->cz = dz;
  s->n = n;
  s->has_points = true;
  return VMP_OK;
}

// ---- observe
static void fill_job(vmp_session* s, ObsJob& j, int block0) {
  j.x = s->cx;
  j.y = s->cy;
  j.z = s->cz;
  j.slots = s->slots;
  j.table = s->map.table;
  j.hot = s->map.hot;
  j.prm = &s->f->obs_prm;
  j.partials = s->partials;
  j.counters = s->obs_ctr;
  j.ticket = s->ticket;
  j.out = s->f->obs;
  j.hmask = s->map.hmask;
  j.n = s->n;
  j.ppt = s->ppt;
  j.nblocks = std::max(1, (s->n + kObsThreads * s->ppt - 1) / (kObsThreads * s->ppt));
  j.block0 = block0;
  j.pad = 0;
}

static int enqueue_observe(vmp_session* s) {
  ObsJob j;
  fill_job(s, j, 0);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s->stream, s->obs_ctr, 8);
  if (hipMemcpyAsync(s->d_job, &j, sizeof(j), hipMemcpyHostToDevice, s->stream) != hipSuccess) return VMP_EDEVICE;
  if (s->cfg.estimate_ext)
    hipLaunchKernelGGL(observe_kernel<12>, dim3(j.nblocks), dim3(kObsThreads), 0, s->stream, s->d_job, 1);
  else
    hipLaunchKernelGGL(observe_kernel<6>, dim3(j.nblocks), dim3(kObsThreads), 0, s->stream, s->d_job, 1);
  s->obs_launches++;
  return VMP_OK;
}

extern "C" int vmp_observe(vmp_session* s, const vmp_state* x, double* H, double* b, vmp_obs_stats* st) {
  if (!s || !x) return VMP_EINVAL;
  if (!s->has_points) {
    s->err = "vmp_observe before vmp_scan_set_points";
    return VMP_ESTATE;
  }
  VMP_HIP(hipSetDevice(s->ctx->device));
  VMP_HIP(hipMemcpyAsync(&s->f->x, x, sizeof(vmp_state), hipMemcpyHostToDevice, s->stream));
  hipLaunchKernelGGL(k_obs_prep, dim3(1), dim3(64), 0, s->stream, s->f);
  VMP_HIP(hipEventRecord(s->ev[0], s->stream));
  if (enqueue_observe(s) != VMP_OK) return VMP_EDEVICE;
  VMP_HIP(hipEventRecord(s->ev[1], s->stream));
  VMP_HIP(hipGetLastError());
  VMP_HIP(hipMemcpyAsync(s->h_out, s->f->obs, sizeof(double) * 156, hipMemcpyDeviceToHost, s->stream));
  VMP_HIP(hipMemcpyAsync(s->h_obs, s->obs_ctr, sizeof(int32_t) * 8, hipMemcpyDeviceToHost, s->stream));
  VMP_HIP(hipStreamSynchronize(s->stream));
  if (s->h_obs[4]) {
    s->err = "hash probe overflow";
    return VMP_EDEVICE;
  }
  if (H) std::memcpy(H, s->h_out, sizeof(double) * 144);
  if (b) std::memcpy(b, s->h_out + 144, sizeof(double) * 12);
  if (st) {
    st->points = s->n;
    st->hits = s->h_obs[0];
    st->plane_hits = s->h_obs[1];
    st->effective = s->h_obs[2];
    st->key_fallback = s->h_obs[3];
    st->iterations = 1;
  }
  return VMP_OK;
}

// ---- IESKF update (device resident loop)
static int enqueue_ieskf(vmp_session* s) {
  hipLaunchKernelGGL(k_ieskf_begin, dim3(1), dim3(64), 0, s->stream, s->f_list);
  for (int it = 0; it < s->cfg.max_iter; ++it) {
    if (enqueue_observe(s) != VMP_OK) return VMP_EDEVICE;
    hipLaunchKernelGGL(k_ieskf_step, dim3(1), dim3(64), 0, s->stream, s->f_list, s->obs_ctr_list,
                       s->obs_stat_list);
  }
  hipLaunchKernelGGL(k_ieskf_end, dim3(1), dim3(64), 0, s->stream, s->f_list);
  return VMP_OK;
}

static int read_obs_stats(vmp_session* s, vmp_obs_stats* st) {
  VMP_HIP(hipMemcpyAsync(s->h_obs, s->obs_stat, sizeof(int32_t) * 4, hipMemcpyDeviceToHost, s->stream));
  VMP_HIP(hipMemcpyAsync(s->h_obs + 8, &s->f->iters, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
  VMP_HIP(hipMemcpyAsync(s->h_obs + 12, s->obs_ctr + 4, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
  VMP_HIP(hipStreamSynchronize(s->stream));
  if (s->h_obs[12]) {
    s->err = "hash probe overflow";
    return VMP_EDEVICE;
  }
  if (st) {
    st->points = s->n;
    st->hits = s->h_obs[0];
    st->plane_hits = s->h_obs[1];
    st->effective = s->h_obs[2];
    st->key_fallback = s->h_obs[3];
    st->iterations = s->h_obs[8];
  }
  return VMP_OK;
}

extern "C" int vmp_ieskf_update(vmp_session* s, vmp_obs_stats* st) {
  if (!s) return VMP_EINVAL;
  if (!s->has_points) {
    s->err = "vmp_ieskf_update before vmp_scan_set_points";
    return VMP_ESTATE;
  }
  VMP_HIP(hipSetDevice(s->ctx->device));
  s->obs_launches = 0;
  VMP_HIP(hipEventRecord(s->ev[0], s->stream));
  if (enqueue_ieskf(s) != VMP_OK) return VMP_EDEVICE;
  VMP_HIP(hipEventRecord(s->ev[1], s->stream));
  VMP_HIP(hipGetLastError());
  return read_obs_stats(s, st);
}

// ---- map update
static int enqueue_map_update(vmp_session* s, int is_build) {
  MapDev& m = s->map;
  ScanBufs b = s->sb;
  b.x = s->cx;
  b.y = s->cy;
  b.z = s->cz;
  b.n = s->n;
  b.is_build = is_build;
  b.prm = &s->f->upd_prm;
  const int n = s->n;
  const int T = kMapThreads;
  const int gn = grid_for(n, T);
  hipStream_t st = s->stream;
  hipLaunchKernelGGL(k_begin_scan, dim3(1), dim3(1), 0, st, m);
  if (n == 0) return VMP_OK;
  hipLaunchKernelGGL(k_ins_keys, dim3(gn), dim3(T), 0, st, m, b);
  size_t tb = s->temp_bytes;
  int hb = 1;
  while ((uint64_t(1) << hb) <= uint64_t(m.hmask)) ++hb;
  if (rocprim::radix_sort_pairs(s->temp, tb, b.skey_in, b.skey_out, b.sval_in, b.sval_out, size_t(n), 0, hb, st) !=
      hipSuccess)
    return VMP_EDEVICE;
  hipLaunchKernelGGL(k_seg_heads, dim3(gn), dim3(T), 0, st, b);
  tb = s->temp_bytes;
  if (rocprim::select(s->temp, tb, rocprim::counting_iterator<int32_t>(0), b.flags, b.seg_start, &m.ctr->nseg,
                      size_t(n), st) != hipSuccess)
    return VMP_EDEVICE;
  hipLaunchKernelGGL(k_seg_info, dim3(gn), dim3(T), 0, st, m, b);
  hipLaunchKernelGGL(k_fill_i32, dim3(gn), dim3(T), 0, st, b.flags, 0, n);
  hipLaunchKernelGGL(k_mark_new, dim3(gn), dim3(T), 0, st, m, b);
  tb = s->temp_bytes;
  if (rocprim::select(s->temp, tb, rocprim::counting_iterator<int32_t>(0), b.flags, b.lru_sel, &m.ctr->ncreate_ev,
                      size_t(n), st) != hipSuccess)
    return VMP_EDEVICE;
  hipLaunchKernelGGL(k_cre_split, dim3(gn), dim3(T), 0, st, m, b);
  hipLaunchKernelGGL(k_create_fast, dim3(gn), dim3(T), 0, st, m, b);
  hipLaunchKernelGGL(k_create_fast_ctr, dim3(1), dim3(1), 0, st, m);
  hipLaunchKernelGGL(k_create_evict, dim3(1), dim3(64), 0, st, m, b);
  hipLaunchKernelGGL(k_create_init, dim3(gn), dim3(T), 0, st, m, b);
  hipLaunchKernelGGL(k_replay, dim3(gn), dim3(T), 0, st, m, b);
  if (!is_build) {
    hipLaunchKernelGGL(k_fill_u64, dim3(gn), dim3(T), 0, st, b.cand_in, ~0ull, n);
    hipLaunchKernelGGL(k_candidates, dim3(gn), dim3(T), 0, st, m, b);
    hipLaunchKernelGGL(k_dirty_untouched, dim3(gn), dim3(T), 0, st, m, b);
    tb = s->temp_bytes;
    if (rocprim::radix_sort_keys(s->temp, tb, b.cand_in, b.cand_out, size_t(n), 0, 64, st) != hipSuccess)
      return VMP_EDEVICE;
    hipLaunchKernelGGL(k_merge_replay, dim3(1), dim3(64), 0, st, m, b);
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(gn), dim3(T), 0, st, b.flags, 0, n);
  hipLaunchKernelGGL(k_finalize_segs, dim3(gn), dim3(T), 0, st, m, b);
  hipLaunchKernelGGL(k_evict_finalize, dim3(gn), dim3(T), 0, st, m, b);
  hipLaunchKernelGGL(k_evict_finalize_ctr, dim3(1), dim3(1), 0, st, m);
  tb = s->temp_bytes;
  if (rocprim::select(s->temp, tb, rocprim::counting_iterator<int32_t>(0), b.flags, b.lru_sel, &m.ctr->nlru,
                      size_t(n), st) != hipSuccess)
    return VMP_EDEVICE;
  hipLaunchKernelGGL(k_lru_append, dim3(gn), dim3(T), 0, st, m, b);
  hipLaunchKernelGGL(k_end_scan, dim3(1), dim3(1), 0, st, m, b);
  return VMP_OK;
}

// housekeeping between scans: hash rebuild and LRU ring compaction
static int maintain(vmp_session* s) {
  MapDev& m = s->map;
  const MapCounters& c = *s->h_ctr;
  const int64_t hcap = int64_t(m.hmask) + 1;
  const int T = 256;
  if (int64_t(c.used_entries) + 2 * int64_t(s->cfg.max_scan_points) > hcap / 2) {
    VMP_HIP(hipMemsetAsync(s->nlive, 0, sizeof(int32_t), s->stream));
    hipLaunchKernelGGL(k_rehash_collect, dim3(grid_for(m.pool, T)), dim3(T), 0, s->stream, m, s->live, s->nlive);
    hipLaunchKernelGGL(k_rehash_clear, dim3(grid_for(int(std::min<int64_t>(hcap, 1 << 20)), T)), dim3(T), 0,
                       s->stream, m);
    MapCounters cc = c;
    cc.used_entries = 0;
    cc.tombs = 0;
    VMP_HIP(hipMemcpyAsync(m.ctr, &cc, sizeof(cc), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_rehash_insert, dim3(grid_for(m.pool, T)), dim3(T), 0, s->stream, m, s->live, s->nlive);
    VMP_HIP(hipMemcpyAsync(s->h_ctr, m.ctr, sizeof(MapCounters), hipMemcpyDeviceToHost, s->stream));
    VMP_HIP(hipStreamSynchronize(s->stream));
  }
  const int64_t used = s->h_ctr->ring_tail - s->h_ctr->ring_head;
  if (used + int64_t(s->cfg.max_scan_points) > m.qcap) {
    // compact live entries (ascending stamps preserved) to the front
    const int64_t head = s->h_ctr->ring_head;
    std::vector<int32_t> hv(static_cast<size_t>(used));
    std::vector<uint64_t> hs(static_cast<size_t>(used));
    // copy ring out, filter against voxel state on host (rare path)
    std::vector<int32_t> rv(size_t(m.qcap));
    std::vector<uint64_t> rs(size_t(m.qcap));
    std::vector<VoxState> vs(size_t(m.pool));
    VMP_HIP(hipMemcpy(rv.data(), m.ring_vid, sizeof(int32_t) * size_t(m.qcap), hipMemcpyDeviceToHost));
    VMP_HIP(hipMemcpy(rs.data(), m.ring_stamp, sizeof(uint64_t) * size_t(m.qcap), hipMemcpyDeviceToHost));
    VMP_HIP(hipMemcpy(vs.data(), m.vs, sizeof(VoxState) * size_t(m.pool), hipMemcpyDeviceToHost));
    int64_t k = 0;
    for (int64_t r = 0; r < used; ++r) {
      const int64_t ph = (head + r) % m.qcap;
      const int32_t v = rv[size_t(ph)];
      if ((vs[size_t(v)].flags & kAlive) && vs[size_t(v)].stamp == rs[size_t(ph)]) {
        hv[size_t(k)] = v;
        hs[size_t(k)] = rs[size_t(ph)];
        ++k;
      }
    }
    VMP_HIP(hipMemcpy(m.ring_vid, hv.data(), sizeof(int32_t) * size_t(k), hipMemcpyHostToDevice));
    VMP_HIP(hipMemcpy(m.ring_stamp, hs.data(), sizeof(uint64_t) * size_t(k), hipMemcpyHostToDevice));
    MapCounters cc = *s->h_ctr;
    cc.ring_head = 0;
    cc.ring_tail = k;
    VMP_HIP(hipMemcpy(m.ctr, &cc, sizeof(cc), hipMemcpyHostToDevice));
    *s->h_ctr = cc;
  }
  return VMP_OK;
}

static int finish_map(vmp_session* s, vmp_map_stats* st) {
  VMP_HIP(hipMemcpyAsync(s->h_ctr, s->map.ctr, sizeof(MapCounters), hipMemcpyDeviceToHost, s->stream));
  VMP_HIP(hipStreamSynchronize(s->stream));
  const MapCounters& c = *s->h_ctr;
  if (c.error) {
    char buf[160];
    std::snprintf(buf, sizeof(buf),
                  "map update error 0x%x (1 hash probe, 2 key range, 4 LRU exhausted in one scan, 8 voxel pool, "
                  "16 replay queue)",
                  unsigned(c.error));
    s->err = buf;
    return (c.error & 2) ? VMP_ERANGE : ((c.error & 4) ? VMP_ECAPACITY : VMP_EDEVICE);
  }
  if (st) {
    st->points = s->n;
    st->touched = c.nseg;
    st->created = c.ncre;
    st->evicted = c.nevict;
    st->refits = c.refits;
    st->merge_calls = c.merge_calls;
    st->merge_evals = c.merge_evals;
    st->merges = c.merges;
    st->dropped = c.dropped;
    st->map_size = int32_t(c.size);
  }
  return maintain(s);
}

extern "C" int vmp_map_update(vmp_session* s, int is_build, vmp_map_stats* st) {
  if (!s) return VMP_EINVAL;
  if (!s->has_points) {
    s->err = "vmp_map_update before vmp_scan_set_points";
    return VMP_ESTATE;
  }
  VMP_HIP(hipSetDevice(s->ctx->device));
  if (is_build) {
    VMP_HIP(hipMemcpyAsync(s->h_ctr, s->map.ctr, sizeof(MapCounters), hipMemcpyDeviceToHost, s->stream));
    VMP_HIP(hipStreamSynchronize(s->stream));
    if (s->h_ctr->size != 0) {
      s->err = "build (MAP_INIT) requires an empty map";
      return VMP_ESTATE;
    }
  }
  hipLaunchKernelGGL(k_upd_prep, dim3(1), dim3(64), 0, s->stream, s->f);
  VMP_HIP(hipEventRecord(s->ev[2], s->stream));
  if (enqueue_map_update(s, is_build) != VMP_OK) {
    s->err = "map update launch failed";
    return VMP_EDEVICE;
  }
  VMP_HIP(hipEventRecord(s->ev[3], s->stream));
  VMP_HIP(hipGetLastError());
  return finish_map(s, st);
}

extern "C" int vmp_register_scan(vmp_session* s, vmp_obs_stats* ost, vmp_map_stats* mst) {
  if (!s) return VMP_EINVAL;
  if (!s->has_points) {
    s->err = "vmp_register_scan before vmp_scan_set_points";
    return VMP_ESTATE;
  }
  VMP_HIP(hipSetDevice(s->ctx->device));
  s->obs_launches = 0;
  VMP_HIP(hipEventRecord(s->ev[0], s->stream));
  if (enqueue_ieskf(s) != VMP_OK) return VMP_EDEVICE;
  VMP_HIP(hipEventRecord(s->ev[1], s->stream));
  VMP_HIP(hipEventRecord(s->ev[2], s->stream));
  if (enqueue_map_update(s, 0) != VMP_OK) return VMP_EDEVICE;
  VMP_HIP(hipEventRecord(s->ev[3], s->stream));
  VMP_HIP(hipGetLastError());
  int rc = read_obs_stats(s, ost);
  if (rc != VMP_OK) return rc;
  return finish_map(s, mst);
}

extern "C" int vmp_register_batch(vmp_session** ss, int32_t n, vmp_obs_stats* ost, vmp_map_stats* mst) {
  if (!ss || n < 0) return VMP_EINVAL;
  for (int32_t i = 0; i < n; ++i) {
    int rc = vmp_register_scan(ss[i], ost ? ost + i : nullptr, mst ? mst + i : nullptr);
    if (rc != VMP_OK) return rc;
  }
  return VMP_OK;
}

extern "C" int vmp_session_timing(vmp_session* s, double* observe_ms, double* map_ms, int32_t* observe_launches) {
  if (!s) return VMP_EINVAL;
  float a = 0.f, b = 0.f;
  if (observe_ms) {
    if (hipEventElapsedTime(&a, s->ev[0], s->ev[1]) != hipSuccess) a = -1.f;
    *observe_ms = a;
  }
  if (map_ms) {
    if (hipEventElapsedTime(&b, s->ev[2], s->ev[3]) != hipSuccess) b = -1.f;
    *map_ms = b;
  }
  if (observe_launches) *observe_launches = s->obs_launches;
  return VMP_OK;
}

extern "C" int64_t vmp_map_size(vmp_session* s) {
  if (!s) return -1;
  if (hipMemcpy(s->h_ctr, s->map.ctr, sizeof(MapCounters), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return s->h_ctr->size;
}

extern "C" int64_t vmp_map_export(vmp_session* s, int64_t cap, int64_t* keys, int32_t* ints, uint64_t* group,
                                  double* dbl) {
  if (!s) return -1;
  if (hipSetDevice(s->ctx->device) != hipSuccess) return -1;
  const MapDev& m = s->map;
  std::vector<VoxState> vs(size_t(m.pool));
  std::vector<double> pl(size_t(m.pool) * kPlaneD);
  if (hipMemcpy(vs.data(), m.vs, sizeof(VoxState) * vs.size(), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  if (hipMemcpy(pl.data(), m.plane, sizeof(double) * pl.size(), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  std::vector<int32_t> live;
  for (int32_t v = 0; v < m.pool; ++v)
    if (vs[size_t(v)].flags & kAlive) live.push_back(v);
  std::sort(live.begin(), live.end(), [&](int32_t a, int32_t b) { return vs[size_t(a)].stamp > vs[size_t(b)].stamp; });
  int64_t cnt = 0;
  for (int32_t v : live) {
    if (cnt >= cap) break;
    const VoxState& st = vs[size_t(v)];
    int64_t kx, ky, kz;
    unpack_key(st.key, kx, ky, kz);
    if (keys) {
      keys[3 * cnt] = kx;
      keys[3 * cnt + 1] = ky;
      keys[3 * cnt + 2] = kz;
    }
    if (ints) {
      const bool en = st.flags & kEnable;
      ints[5 * cnt] = st.n;
      ints[5 * cnt + 1] = ((st.flags & kInit) ? 1 : 0) | ((st.flags & kPlane) ? 2 : 0) | (en ? 4 : 0) |
                          ((st.flags & kMerged) ? 8 : 0);
      ints[5 * cnt + 2] = st.newly_add;
      ints[5 * cnt + 3] = en ? st.n : 0;
      ints[5 * cnt + 4] = 0;
    }
    if (group) group[cnt] = st.group_id;
    if (dbl)
      for (int k = 0; k < 8; ++k) dbl[8 * cnt + k] = pl[size_t(v) * kPlaneD + k];
    ++cnt;
  }
  return cnt;
}

extern "C" int vmp_slots_export(vmp_session* s, int32_t n, uint8_t* valid, float* resid, float* norm3) {
  if (!s || n < 0 || n > s->cfg.max_scan_points) return VMP_EINVAL;
  std::vector<float4> sl(static_cast<size_t>(n));
  VMP_HIP(hipMemcpy(sl.data(), s->slots, sizeof(float4) * size_t(n), hipMemcpyDeviceToHost));
  for (int32_t i = 0; i < n; ++i) {
    const bool v = !std::isnan(sl[size_t(i)].w);
    if (valid) valid[i] = v ? 1 : 0;
    if (resid) resid[i] = sl[size_t(i)].w;
    if (norm3) {
      norm3[3 * i] = sl[size_t(i)].x;
      norm3[3 * i + 1] = sl[size_t(i)].y;
      norm3[3 * i + 2] = sl[size_t(i)].z;
    }
  }
  return VMP_OK;
}
