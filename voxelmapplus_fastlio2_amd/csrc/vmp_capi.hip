// C ABI of the MI355X registration path (include/vmp.h): contexts, sessions,
// device memory, and the launch sequence of one registration step for a
// batch of sessions (the single-session calls are batches of one).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "vmp.h"
#include "vmp_common.hpp"
#include "vmp_filter.hpp"
#include "vmp_ieskf.hpp"
#include "vmp_prim.hpp"
#include "vmp_mapupd.hpp"
#include "vmp_observe.hpp"
#include "vmp_imu.hpp"
// entries of the per-tile first-session table: every tile of the largest
// input the library's compaction takes, and one past the last
constexpr int kSfirstEntries = vmp::kPrimMaxItems / vmp::kPrimTile + 2;
static_assert(vmp::kCsThreads == vmp::kPrimThreads, "the candidate starts run as the tail of a compaction tile");

using namespace vmp;

// session initialisation fills
__global__ void k_init_slots(float4* slots, int n) {
  VMP_GRID_LOOP(i, n) slots[i] = make_float4(0.f, 0.f, 0.f, __builtin_nanf(""));
}
__global__ void k_init_vs(VoxState* vs, int n) {
  VMP_GRID_LOOP(i, n) {
    VoxState v;
    std::memset(&v, 0, sizeof(v));
    v.tag = -1;
    vs[i] = v;
  }
}

// Device scratch of vmp_register_scan_noalloc: the map pipeline's per-point
// work arrays, the rocPRIM temporary storage and the job tables of a
// one-session step (and, between steps, the LRU ring compaction's arrays),
// for scans of up to kNoAllocMaxPoints points.  One array per GPU, loaded
// with the code object and held until the process ends; every context on a
// GPU shares it, so no-allocation calls that overlap on one GPU race.  Host
// code reaches it only through the address hipGetSymbolAddress returns for
// the current device (g_scratch below), never through the name.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int32_t kNoAllocMaxPoints = 131072;
constexpr int32_t kNoAllocPreCap = 16384;  // taint-phase prefetch entries (the rest look up directly)
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint8_t* g_scratch[kMaxDevices] = {};

// Input validation: a scan with a NaN or infinite coordinate is rejected
// before any stage reads it (its key would be undefined).  One launch covers
// many point sets; block row y checks set y (grid-stride over sets).
struct PtsRef {
  const float* x;
  const float* y;
  const float* z;
  const float* t;  // optional per-point times, which must also ascend
  int32_t n;
  int32_t pad;
};
// The point sets come by value (up to kCheckInline) or from a device table;
// a bad set r gets bad[r] = gen, the call's generation number, so the flags
// need no clearing between calls.
constexpr int kCheckInline = 8;
struct PtsRefs {
  PtsRef r[kCheckInline];
};
__global__ void k_check_points(PtsRefs inl, const PtsRef* refs, int nref, int32_t* bad, int32_t gen) {
  for (int r = blockIdx.y; r < nref; r += gridDim.y) {
    const PtsRef R = nref <= kCheckInline ? inl.r[r] : refs[r];
    int b = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < R.n; i += gridDim.x * blockDim.x) {
      b |= !(isfinite(R.x[i]) && isfinite(R.y[i]) && isfinite(R.z[i]));
      if (R.t) b |= !isfinite(R.t[i]) || (i > 0 && !(R.t[i] >= R.t[i - 1]));
    }
    if (__ballot(b) != 0ull && (threadIdx.x & 63) == 0) bad[r] = gen;
  }
}

// gather the per-session counters and posterior states of a batch into one
// array (one copy back)
struct BatchOut {
  MapCounters ctr;
  vmp_state x;
  int32_t obs[8];
  int32_t iters;
  int32_t pad[3];
};
__global__ void k_gather(const MapDev* maps, FilterDev** fs, int32_t** obs_stat, BatchOut* out, int nsess,
                         const int32_t* prim_err) {
  const int s = blockIdx.x;
  if (s >= nsess || threadIdx.x != 0) return;
  out[s].ctr = *maps[s].ctr;
  if (*prim_err) out[s].ctr.error |= 64;  // a bounded wait of the compaction / sort ran out
  for (int k = 0; k < 8; ++k) out[s].obs[k] = obs_stat[s][k];
  out[s].iters = fs[s]->iters;
  out[s].x = fs[s]->x;
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

template <class T>
static hipError_t dalloc(std::vector<void*>& allocs, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count * sizeof(T), 16));
  if (e == hipSuccess) {
    allocs.push_back(q);
    *p = static_cast<T*>(q);
  }
  return e;
}

// Batch workspace owned by a context: per-point / per-segment scratch for the
// map pipeline plus the job tables of a batch.  Grows on demand.
struct BatchWs {
  int64_t cap_points = 0;
  int cap_sess = 0;
  bool fixed = false;  // carved from the device scratch: never reallocated
  std::vector<void*> allocs;
  BatchBufs b{};
  void* temp = nullptr;
  size_t temp_bytes = 0;
  // the per-session job tables (maps, session info, observe jobs, filter /
  // counter pointers) sit in one device block with one pinned mirror, so a
  // step uploads them with one copy
  uint8_t* d_tab = nullptr;
  uint8_t* h_tab = nullptr;
  size_t tab_obsstat = 0;  // offset of d_obsstat
  size_t tab_sfirst = 0;   // offset of the last table (d_sfirst)
  // first session of each compaction tile of the step's points (sessions
  // starting in tile t: [sfirst[t], sfirst[t + 1]))
  int32_t* d_sfirst = nullptr;
  int32_t* h_sfirst = nullptr;
  MapDev* d_maps = nullptr;
  SessInfo* d_si = nullptr;
  ObsJob* d_jobs = nullptr;
  FilterDev** d_flist = nullptr;
  int32_t** d_obsctr = nullptr;
  int32_t** d_obsstat = nullptr;
  BatchOut* d_out = nullptr;
  // pinned host mirrors
  MapDev* h_maps = nullptr;
  SessInfo* h_si = nullptr;
  ObsJob* h_jobs = nullptr;
  FilterDev** h_flist = nullptr;
  int32_t** h_obsctr = nullptr;
  int32_t** h_obsstat = nullptr;
  BatchOut* h_out = nullptr;
  MotionRec* d_motion = nullptr;
  MotionRec* h_motion = nullptr;
  PreJob* d_pre = nullptr;  // sessions of a step with a raw package
  PreJob* h_pre = nullptr;
  VgJob* d_vg = nullptr;    // sessions of a step whose scan is filtered
  VgJob* h_vg = nullptr;
  int32_t* d_vgst = nullptr;
  int32_t* h_vgst = nullptr;
  hipEvent_t motion_done = nullptr;  // the last motion table upload finished
  bool motion_pending = false;
  std::vector<void*> host_allocs;

  void release() {
    for (void* p : allocs) hipFree(p);
    allocs.clear();
    for (void* p : host_allocs) hipHostFree(p);
    host_allocs.clear();
    cap_points = 0;
    cap_sess = 0;
  }
};

// Scratch of the device-side LRU ring compaction (grown on demand).
struct RingTmp {
  int64_t cap = 0;
  bool fixed = false;  // carved from the device scratch: never reallocated
  std::vector<void*> allocs;
  int32_t *tv = nullptr, *keep = nullptr, *sel = nullptr, *cnt = nullptr;
  uint64_t* ts = nullptr;
  void* temp = nullptr;
  size_t temp_bytes = 0;
  void release() {
    for (void* p : allocs) hipFree(p);
    allocs.clear();
    cap = 0;
  }
};

struct vmp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  BatchWs ws;
  RingTmp rt;
  hipEvent_t ev[4]{};
  // the IMU packages of a batch of raw scans in one upload: move records and
  // packages in pinned staging, copied whole, then scattered on the device
  // into the sessions' package slots
  uint8_t* h_pkstage = nullptr;
  uint8_t* d_pkstage = nullptr;
  size_t pkstage_cap = 0;
  hipEvent_t pkstage_done = nullptr;
  bool pkstage_pending = false;
  // optional per-launch timing of the observe kernel / map pipeline
  bool timing = false;
  // observe launches back to back between one event pair (the kernel is
  // idempotent: the repetitions rewrite the same slots and partial rows)
  int obs_time_reps = 1;
  int32_t* h_sig = nullptr;  // pinned coherent word a signal kernel sets when the stream reaches it
  int32_t sig_gen = 0;
  int replay_wave_max = 4096;  // refitting voxels per step up to which each replays on its own wave
  int replay_long_seg = 24;    // beyond that, segments of this many points still do (0: none)
  hipEvent_t tev[2 * 64]{};
  int tev_used = 0;
  // per-kernel event pairs of the timed kernels below (while timing is on)
  hipEvent_t kev[2 * 48]{};
  int kev_id[48]{};
  int kev_used = 0;
  double acc_k_ms[16]{};
  int64_t acc_k_n[16]{};
  double acc_obs_ms = 0.0, acc_map_ms = 0.0;
  int64_t acc_obs_n = 0, acc_map_n = 0;
  int64_t acc_obs_points = 0, acc_obs_hits = 0;
  int obs_launches = 0;
  // launch shape of the observe kernel: points per thread; obs_ppt > 0
  // overrides the per-size-bucket table that vmp_autotune_observe fills
  int obs_ppt = 0;
  int ppt_tab[4] = {2, 2, 2, 2};
  // arithmetic of the observe kernel: 0 = compensated fp32 in the anchor
  // frame (registration), 1 = plain fp32 world frame (precision study arm)
  int obs_arith = 0;
  // the filter stage of a step (IESKF begin, max_iter x (observe, reduce,
  // step), end) captured as one hipGraph per (sessions, observe grid, ...)
  bool use_graph = true;
  struct FilterGraph {
    hipGraphExec_t exec = nullptr;
    int n = 0, grid = 0, iters = 0, ext = 0, arith = 0;
    uint64_t ws_gen = 0;
    uint64_t last_use = 0;
  };
  FilterGraph graphs[8];
  // the map stage of a step captured as one hipGraph per shape (points,
  // sessions, hash / session / time / length bits, build flag), once a shape
  // recurs on consecutive steps (scans of a fixed size); a shape whose
  // capture holds anything but kernel nodes runs eagerly from then on
  bool use_map_graph = true;
  struct MapGraph {
    hipGraphExec_t exec = nullptr;
    int64_t key[3] = {-1, -1, -1};
    uint64_t ws_gen = 0;
    uint64_t last_use = 0;
  };
  MapGraph mgraphs[4];
  int64_t mkey_prev[3] = {-1, -1, -1};
  int64_t mkey_bad[3] = {-1, -1, -1};
  int64_t map_graph_launches = 0, map_graph_builds = 0;
  uint64_t graph_clock = 0;
  uint64_t ws_gen = 1;  // bumped when the batch workspace is reallocated
  int64_t graph_launches = 0, graph_builds = 0;
  // a batch submitted without waiting (vmp_register_batch_submit)
  bool fl = false;
  int fl_mode = 0;
  std::vector<vmp_session*> fl_ss;
  std::vector<int32_t> fl_n;
  std::vector<vmp_obs_stats> res_ost;  // results of the last submitted batch, once finished
  std::vector<vmp_map_stats> res_mst;
  int res_rc = 0;
  bool res_valid = false;  // those results have not been collected by vmp_register_batch_wait
  // validation of device point buffers (k_check_points), run on `stream`
  // when a step starts (no batch is in flight then)
  bool check_points = true;
  int32_t check_gen = 0;  // generation of the last point check (bad-set flags)
  // vmp_register_scan_noalloc: a workspace and ring scratch carved from the
  // device's scratch array at context creation (na_ready once that
  // succeeded); the call swaps them in for its duration
  BatchWs ws_na;
  RingTmp rt_na;
  bool na_ready = false;
  int vcap = 0;
  PtsRef* h_refs = nullptr;
  PtsRef* d_refs = nullptr;
  int32_t* h_bad = nullptr;
  int32_t* d_bad = nullptr;
};

// Points per thread of the observe kernel for a scan of n points.  It depends
// only on n (never on the batch), so a session's block decomposition, and
// with it the summation order of H and b, is the same alone and in a batch
// (for one table: sessions compared bit by bit must share it).
static int ppt_bucket(int n) { return n < 8192 ? 0 : (n < 32768 ? 1 : (n < 131072 ? 2 : 3)); }
static int obs_ppt_for(const vmp_ctx* c, int n) {
  return c->obs_ppt > 0 ? c->obs_ppt : c->ppt_tab[ppt_bucket(n)];
}

// kernels timed one by one while per-launch timing is on, in the order
// vmp_ctx_kernel_times reports them
enum {
  kTkInsKeys = 0,
  kTkCreateEvict,
  kTkReplayHeavy,
  kTkReplayLight,
  kTkMergeReplay,
  kTkCandidates,
  kTkRingLive,
  kTkIeskfStep,
  kTkUndistort,
  kTkImuPredict,
  kTkSpecEval,
  kTkCount
};
template <class F>
static void timed_launch(vmp_ctx* c, int id, F&& launch) {
  if (!c->timing || c->kev_used >= 48 || !c->kev[0]) {
    launch();
    return;
  }
  const int k = c->kev_used++;
  hipEventRecord(c->kev[2 * k], c->stream);
  launch();
  hipEventRecord(c->kev[2 * k + 1], c->stream);
  c->kev_id[k] = id;
}

struct vmp_session {
  vmp_ctx* ctx = nullptr;
  vmp_map_cfg cfg{};
  std::string err;
  std::vector<void*> allocs;
  MapDev map{};
  MapCounters h_ctr{};  // last counters read back
  float *own_x = nullptr, *own_y = nullptr, *own_z = nullptr;
  const float *cx = nullptr, *cy = nullptr, *cz = nullptr;
  int32_t n = 0;
  float4* slots = nullptr;
  FilterDev* f = nullptr;
  double* partials = nullptr;
  int32_t* obs_ctr = nullptr;
  int32_t* obs_stat = nullptr;
  int32_t* live = nullptr;
  int32_t* nlive = nullptr;
  bool has_points = false;
  bool needs_check = false;  // device buffers bound, checked when the next step starts
  // IMU side: device state, the raw package bound for the next step
  ImuDev* imu = nullptr;
  float* own_t = nullptr;
  bool imu_inited = false;
  std::vector<double> init_t, init_acc, init_gyro;  // IMU_INIT cache (host)
  bool raw_pending = false;
  const float *rx = nullptr, *ry = nullptr, *rz = nullptr, *rt = nullptr;
  int32_t raw_n = 0;      // points of the pending raw scan
  int32_t cloud_n = 0;    // undistorted points of the last processed package
  ImuPkg* stage = nullptr;  // pinned staging of one package's IMU samples
  float *own_fx = nullptr, *own_fy = nullptr, *own_fz = nullptr;  // VoxelGrid output (scan_resolution > 0)
  hipEvent_t stage_done = nullptr;
  bool stage_pending = false;
};

static uint32_t next_pow2(uint64_t v) {
  uint64_t p = 1024;
  while (p < v) p <<= 1;
  return uint32_t(p);
}
static int grid_for(int64_t n, int threads) {
  return int(std::max<int64_t>(1, std::min<int64_t>((n + threads - 1) / threads, 65535 * 8)));
}

extern "C" const char* vmp_version(void) { return "voxelmapplus_fastlio2_amd 0.2 (gfx950)"; }

// Sets the context's pinned word to v once the stream reaches it (the
// stream's earlier work, device-to-host copies included, is then complete).
__global__ void k_signal(int32_t* h, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(h, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
// Waits for the context's stream: a signal kernel and a bounded spin on the
// pinned word (a round trip about 3 us shorter than hipStreamSynchronize).
// The word changes only once every earlier operation of the stream has
// completed (a fault stops the stream before it), so a seen signal ends the
// wait; otherwise, after 2 ms, hipStreamSynchronize waits for the rest and
// reports any fault.
static hipError_t stream_wait(vmp_ctx* c) {
  if (c->h_sig) {
    const int32_t want = ++c->sig_gen;
    hipLaunchKernelGGL(k_signal, dim3(1), dim3(64), 0, c->stream, c->h_sig, want);
    if (hipGetLastError() == hipSuccess) {
      const auto t0 = std::chrono::steady_clock::now();
      while (std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(2))
        if (__atomic_load_n(c->h_sig, __ATOMIC_ACQUIRE) == want) return hipSuccess;
    }
  }
  return hipStreamSynchronize(c->stream);
}

static void noalloc_setup(vmp_ctx* c);

extern "C" int vmp_ctx_create(int device, vmp_ctx** out) {
  if (!out) return VMP_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return VMP_EDEVICE;
  vmp_ctx* c = new (std::nothrow) vmp_ctx();
  if (!c) return VMP_ENOMEM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return VMP_EDEVICE;
  }

  for (hipEvent_t& e : c->ev) hipEventCreate(&e);
  void* hs = nullptr;
  if (hipHostMalloc(&hs, 256, hipHostMallocCoherent) == hipSuccess) {
    c->h_sig = static_cast<int32_t*>(hs);
    *c->h_sig = 0;
  }
  noalloc_setup(c);
  *out = c;
  return VMP_OK;
}

static int settle(vmp_ctx* c);

extern "C" int vmp_ctx_destroy(vmp_ctx* ctx) {
  if (!ctx) return VMP_EINVAL;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    settle(ctx);
  }
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  for (auto& g : ctx->graphs)
    if (g.exec) hipGraphExecDestroy(g.exec);
  for (auto& g : ctx->mgraphs)
    if (g.exec) hipGraphExecDestroy(g.exec);
  ctx->ws.release();
  if (ctx->ws.motion_done) hipEventDestroy(ctx->ws.motion_done);
  ctx->ws_na.release();  // its pinned mirrors (the device arrays are the scratch's)
  if (ctx->ws_na.motion_done) hipEventDestroy(ctx->ws_na.motion_done);
  ctx->rt.release();
  for (hipEvent_t e : ctx->ev)
    if (e) hipEventDestroy(e);
  for (hipEvent_t e : ctx->kev)
    if (e) hipEventDestroy(e);
  for (hipEvent_t e : ctx->tev)
    if (e) hipEventDestroy(e);
  if (ctx->h_pkstage) hipHostFree(ctx->h_pkstage);
  if (ctx->h_sig) hipHostFree(ctx->h_sig);
  if (ctx->d_pkstage) hipFree(ctx->d_pkstage);
  if (ctx->pkstage_done) hipEventDestroy(ctx->pkstage_done);
  if (ctx->h_refs) hipHostFree(ctx->h_refs);
  if (ctx->h_bad) hipHostFree(ctx->h_bad);
  if (ctx->d_refs) hipFree(ctx->d_refs);
  if (ctx->d_bad) hipFree(ctx->d_bad);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return VMP_OK;
}

// Check device point sets for non-finite coordinates (and times out of
// order) on the context's stream (c->mu held, no batch in flight).  Returns
// the index of the first bad set, -1 if all are good, or -2 on a device
// error.
// room for nref point sets in the check's tables (at least 64)
static bool grow_checks(vmp_ctx* c, int nref) {
  if (nref <= c->vcap) return true;
  hipStreamSynchronize(c->stream);
  if (c->h_refs) hipHostFree(c->h_refs);
  if (c->h_bad) hipHostFree(c->h_bad);
  if (c->d_refs) hipFree(c->d_refs);
  if (c->d_bad) hipFree(c->d_bad);
  c->h_refs = nullptr;
  c->h_bad = nullptr;
  c->d_refs = nullptr;
  c->d_bad = nullptr;
  c->vcap = 0;
  const int cap = std::max(nref, 64);
  if (hipHostMalloc(&c->h_refs, sizeof(PtsRef) * cap) != hipSuccess ||
      hipHostMalloc(&c->h_bad, sizeof(int32_t) * cap) != hipSuccess ||
      hipMalloc(&c->d_refs, sizeof(PtsRef) * cap) != hipSuccess ||
      hipMalloc(&c->d_bad, sizeof(int32_t) * cap) != hipSuccess ||
      hipMemset(c->d_bad, 0, sizeof(int32_t) * cap) != hipSuccess)  // no stale generation matches
    return false;
  c->vcap = cap;
  return true;
}

static int check_device_points(vmp_ctx* c, const PtsRef* refs, int nref) {
  if (nref <= 0 || !c->check_points) return -1;
  if (!grow_checks(c, nref)) return -2;
  int64_t most = 0;
  PtsRefs inl{};
  for (int i = 0; i < nref; ++i) {
    c->h_refs[i] = refs[i];
    if (i < kCheckInline) inl.r[i] = refs[i];
    most = std::max<int64_t>(most, refs[i].n);
  }
  hipStream_t st = c->stream;
  if (nref > kCheckInline &&
      hipMemcpyAsync(c->d_refs, c->h_refs, sizeof(PtsRef) * nref, hipMemcpyHostToDevice, st) != hipSuccess)
    return -2;
  const int32_t gen = ++c->check_gen;
  const int gx = int(std::max<int64_t>(1, std::min<int64_t>((most + 255) / 256, 1024)));
  hipLaunchKernelGGL(k_check_points, dim3(gx, std::min(nref, 65535)), dim3(256), 0, st, inl, c->d_refs, nref,
                     c->d_bad, gen);
  if (hipMemcpyAsync(c->h_bad, c->d_bad, sizeof(int32_t) * nref, hipMemcpyDeviceToHost, st) != hipSuccess ||
      stream_wait(c) != hipSuccess)
    return -2;
  for (int i = 0; i < nref; ++i)
    if (c->h_bad[i] == gen) return i;
  return -1;
}

static bool host_points_finite(const float* x, const float* y, const float* z, int32_t n) {
  for (int32_t i = 0; i < n; ++i)
    if (!(std::isfinite(x[i]) && std::isfinite(y[i]) && std::isfinite(z[i]))) return false;
  return true;
}

static int settle(vmp_ctx* c);

// Lays out the batch workspace for n points and ns sessions: `dev(&p,
// count)` provides each device array, `host(&p, bytes)` each pinned host
// mirror; both return false on failure.  The rocPRIM temporary storage is
// sized for the largest sort or compaction at n and provided last.
template <class Dev, class Host>
static bool ws_layout(vmp_ctx* c, BatchWs& w, size_t n, int ns, size_t pre_cap, Dev&& dev, Host&& host) {
  BatchBufs& b = w.b;
  bool ok = true;
  ok = ok && dev(&b.skey_in, n) && dev(&b.sval_in, n) && dev(&b.skey_out, n) && dev(&b.sval_out, n);
  ok = ok && dev(&b.flags, n) && dev(&b.sel, n) && dev(&b.cnt, size_t(8));
  ok = ok && dev(&b.seg_start, n) && dev(&b.seg_end, n) && dev(&b.seg_sess, n) && dev(&b.seg_vid, n) &&
       dev(&b.seg_entry, n);
  ok = ok && dev(&b.cre_t, n) && dev(&b.cre_seg, n) && dev(&b.co, n) && dev(&b.ev_vid, n);
  ok = ok && dev(&b.cand_in, n) && dev(&b.cand_out, n) && dev(&b.cand_nbr, n * 6) && dev(&b.tnbr, n * 6) &&
       dev(&b.cand_spec, n) &&
       dev(&b.cand_start, size_t(ns)) && dev(&b.spl, n);
  ok = ok && dev(&b.seg_ct, n * 7) && dev(&b.seg_nct, n) && dev(&b.seg_nbr, n * 6);
  ok = ok && dev(&b.len_in, n) && dev(&b.len_out, n) && dev(&b.ord_in, n) && dev(&b.ord, n);
  ok = ok && dev(&b.lv, n * 2) && dev(&b.lv2, n * 2) && dev(&b.lv_cnt, (n * 2 >> 6) + size_t(ns) + 2) &&
       dev(&b.lv_n, size_t(ns));
  ok = ok && dev(&b.prim, prim_ws_entries(int64_t(n))) && dev(&b.qc, n * 3);
  // the refitting-voxel list of the replay lives in cand_nbr's memory (24 B
  // per point), which k_spec_eval writes only after the replays
  if (ok) b.hlist = reinterpret_cast<int4*>(b.cand_nbr);
  // taint-phase prefetch of up to pre_cap acting candidates per batch (the
  // rest take the replay's own lookups)
  b.pre_cap = int32_t(std::min(n, pre_cap));
  ok = ok && dev(&b.pre_x, size_t(b.pre_cap) * kPreItems) && dev(&b.pre_seg, size_t(b.pre_cap) * kPreItems) &&
       dev(&b.pre_tn, size_t(b.pre_cap) * kPreItems);
  const size_t S = size_t(ns);
  size_t tab = 0;
  auto tab_at = [&](size_t bytes) {
    const size_t o = tab;
    tab = (tab + bytes + 255) & ~size_t(255);
    return o;
  };
  const size_t o_maps = tab_at(sizeof(MapDev) * S), o_si = tab_at(sizeof(SessInfo) * S);
  const size_t o_jobs = tab_at(sizeof(ObsJob) * S), o_fl = tab_at(sizeof(FilterDev*) * S);
  const size_t o_oc = tab_at(sizeof(int32_t*) * S), o_os = tab_at(sizeof(int32_t*) * S);
  const size_t o_sf = tab_at(sizeof(int32_t) * kSfirstEntries);
  ok = ok && dev(&w.d_tab, tab) && host(&w.h_tab, tab);
  if (ok) {
    w.tab_obsstat = o_os;
    w.tab_sfirst = o_sf;
    w.d_sfirst = reinterpret_cast<int32_t*>(w.d_tab + o_sf);
    w.h_sfirst = reinterpret_cast<int32_t*>(w.h_tab + o_sf);
    w.d_maps = reinterpret_cast<MapDev*>(w.d_tab + o_maps);
    w.d_si = reinterpret_cast<SessInfo*>(w.d_tab + o_si);
    w.d_jobs = reinterpret_cast<ObsJob*>(w.d_tab + o_jobs);
    w.d_flist = reinterpret_cast<FilterDev**>(w.d_tab + o_fl);
    w.d_obsctr = reinterpret_cast<int32_t**>(w.d_tab + o_oc);
    w.d_obsstat = reinterpret_cast<int32_t**>(w.d_tab + o_os);
    w.h_maps = reinterpret_cast<MapDev*>(w.h_tab + o_maps);
    w.h_si = reinterpret_cast<SessInfo*>(w.h_tab + o_si);
    w.h_jobs = reinterpret_cast<ObsJob*>(w.h_tab + o_jobs);
    w.h_flist = reinterpret_cast<FilterDev**>(w.h_tab + o_fl);
    w.h_obsctr = reinterpret_cast<int32_t**>(w.h_tab + o_oc);
    w.h_obsstat = reinterpret_cast<int32_t**>(w.h_tab + o_os);
  }
  ok = ok && dev(&w.d_out, S) && dev(&w.d_motion, S) && dev(&w.d_pre, S) && dev(&w.d_vg, S) &&
       dev(&w.d_vgst, S * kVgStat);
  ok = ok && host(&w.h_out, sizeof(BatchOut) * S) && host(&w.h_motion, sizeof(MotionRec) * S) &&
       host(&w.h_pre, sizeof(PreJob) * S) && host(&w.h_vg, sizeof(VgJob) * S) &&
       host(&w.h_vgst, sizeof(int32_t) * kVgStat * S);
  if (ok && !w.motion_done) ok = hipEventCreateWithFlags(&w.motion_done, hipEventDisableTiming) == hipSuccess;
  if (!ok) return false;
  size_t t[4] = {0, 0, 0, 0};
  rocprim::radix_sort_pairs(nullptr, t[0], b.skey_in, b.skey_out, b.sval_in, b.sval_out, n, 0, 64, c->stream);
  rocprim::radix_sort_keys(nullptr, t[1], b.cand_in, b.cand_out, n, 0, 64, c->stream);
  rocprim::select(nullptr, t[2], rocprim::counting_iterator<int32_t>(0), b.flags, b.sel, b.cnt, n, c->stream);
  rocprim::radix_sort_pairs(nullptr, t[3], b.len_in, b.len_out, b.ord_in, b.ord, n, 0, 32, c->stream);
  w.temp_bytes = *std::max_element(t, t + 4) + 256;
  uint8_t* temp = nullptr;
  if (!dev(&temp, w.temp_bytes)) return false;
  w.temp = temp;
  return true;
}

// grow the batch workspace to hold `points` points and `nsess` sessions
static int ws_reserve(vmp_ctx* c, int64_t points, int nsess, bool exact = false) {
  BatchWs& w = c->ws;
  if (points <= w.cap_points && nsess <= w.cap_sess) return VMP_OK;
  if (w.fixed) return VMP_ENOMEM;  // the no-allocation workspace never grows
  settle(c);  // a batch in flight still reads the old workspace
  hipStreamSynchronize(c->stream);
  c->ws_gen++;  // captured filter graphs hold the old workspace pointers
  // growth keeps a quarter of headroom, so scans that vary in size do not
  // reallocate (and recapture the filter graphs) step after step
  const int64_t grown = (points > w.cap_points && !exact) ? points + points / 4 : points;
  const int64_t np = std::max<int64_t>(grown, w.cap_points);
  const int ns = std::max(nsess, w.cap_sess);
  w.release();
  auto& A = w.allocs;
  auto dev = [&](auto** p, size_t count) { return dalloc(A, p, count) == hipSuccess; };
  auto host = [&](auto** p, size_t bytes) {
    void* q = nullptr;
    if (hipHostMalloc(&q, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
    w.host_allocs.push_back(q);
    *p = static_cast<std::remove_reference_t<decltype(**p)>*>(q);
    return true;
  };
  w.motion_pending = false;
  if (!ws_layout(c, w, size_t(np), ns, size_t(1) << 18, dev, host) ||
      hipMemsetAsync(w.b.prim, 0, prim_ws_entries(np) * sizeof(int32_t), c->stream) != hipSuccess) {
    w.release();
    return VMP_ENOMEM;
  }
  w.cap_points = np;
  w.cap_sess = ns;
  return VMP_OK;
}

// Bump placement inside the device scratch (256-byte aligned pieces).
struct ScratchCarver {
  uint8_t* base;
  size_t cap;
  size_t used = 0;
  template <class T>
  bool operator()(T** p, size_t count) {
    const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    const size_t off = (used + 255) & ~size_t(255);
    if (off + bytes > cap) return false;
    *p = reinterpret_cast<T*>(base + off);
    used = off + bytes;
    return true;
  }
};

// Set-up of vmp_register_scan_noalloc for a new context (its device
// current): the device's scratch address, once per device, then the
// one-session workspace for kNoAllocMaxPoints points carved from it (its
// pinned host mirrors and the other host-side tables are allocated here, so
// the entry point allocates nothing), and the LRU ring compaction's arrays
// over the same bytes (they are used only between steps, after the step's
// kernels have finished).  A failure leaves na_ready false; the entry point
// then reports it.
static void noalloc_setup(vmp_ctx* c) {
  uint8_t* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device >= kMaxDevices) return;
    if (!g_scratch[c->device]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return;
      g_scratch[c->device] = static_cast<uint8_t*>(p);
    }
    base = g_scratch[c->device];
  }
  BatchWs& w = c->ws_na;
  ScratchCarver carve{base, kScratchBytes};
  auto host = [&](auto** p, size_t bytes) {
    void* q = nullptr;
    if (hipHostMalloc(&q, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
    w.host_allocs.push_back(q);
    *p = static_cast<std::remove_reference_t<decltype(**p)>*>(q);
    return true;
  };
  if (!ws_layout(c, w, size_t(kNoAllocMaxPoints), 1, size_t(kNoAllocPreCap), carve, host)) {
    w.release();
    return;
  }
  // the single-pass primitives' control words start from zero
  if (hipMemset(w.b.prim, 0, prim_ws_entries(kNoAllocMaxPoints) * sizeof(int32_t)) != hipSuccess) {
    w.release();
    return;
  }
  w.cap_points = kNoAllocMaxPoints;
  w.cap_sess = 1;
  w.fixed = true;
  // ring compaction: 20 bytes per entry plus the compaction's temporary
  // storage, sized for as many entries as the scratch holds
  RingTmp& t = c->rt_na;
  size_t n = kScratchBytes / 20;
  for (;;) {
    size_t tb = 0;
    if (rocprim::select(nullptr, tb, rocprim::counting_iterator<int32_t>(0), t.keep, t.sel, t.cnt, n, c->stream) !=
        hipSuccess)
      return;
    ScratchCarver rc{base, kScratchBytes};
    uint8_t* temp = nullptr;
    if (rc(&t.tv, n) && rc(&t.ts, n) && rc(&t.keep, n) && rc(&t.sel, n) && rc(&t.cnt, size_t(1)) &&
        rc(&temp, tb + 256)) {
      t.temp = temp;
      t.temp_bytes = tb + 256;
      break;
    }
    n -= n / 16;
  }
  t.cap = int64_t(n);
  t.fixed = true;
  // host-side tables a one-session step would otherwise grow on first use
  if (!grow_checks(c, 64)) return;
  c->fl_ss.reserve(64);
  c->fl_n.reserve(64);
  c->res_ost.reserve(64);
  c->res_mst.reserve(64);
  c->na_ready = true;
}

struct Sizes {
  int64_t pool, hcap, qcap, maxn, max_blocks;
};
static Sizes sizes_for(const vmp_map_cfg& c) {
  Sizes z;
  z.maxn = c.max_scan_points;
  z.pool = int64_t(c.map_capacity) + c.max_scan_points + 64;
  z.hcap = next_pow2(uint64_t(2) * (uint64_t(c.map_capacity) + uint64_t(c.max_scan_points)));
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
  b += z.pool * int64_t(std::max(cfg->update_size_thresh - 1, 1)) * kPendF * 4;
  b += z.qcap * 12;
  b += z.maxn * (3 * 4 + 16);
  b += z.max_blocks * kObsRowMax * 8;
  b += int64_t(sizeof(FilterDev));
  b += int64_t(sizeof(ImuDev)) + z.maxn * 4;
  if (cfg->scan_resolution > 0.0) b += z.maxn * 12;
  return b;
}

static int check_cfg(const vmp_map_cfg* c) {
  if (!c) return VMP_EINVAL;
  if (!(c->voxel_size > 0) || c->max_point_thresh < 1 || c->update_size_thresh < 1 || c->map_capacity < 1 ||
      c->max_scan_points < 1 || c->max_iter < 1 || c->max_iter > 64)
    return VMP_EINVAL;
  if (int64_t(c->map_capacity) + c->max_scan_points > (int64_t(1) << 25)) return VMP_EINVAL;
  if (c->max_scan_points > (1 << 24)) return VMP_EINVAL;
  return VMP_OK;
}

extern "C" int vmp_session_destroy(vmp_session* s) {
  if (!s) return VMP_EINVAL;
  if (s->ctx) {
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    for (vmp_session* q : s->ctx->fl_ss)
      if (q == s) {
        settle(s->ctx);
        break;
      }
  }
  hipSetDevice(s->ctx->device);
  hipStreamSynchronize(s->ctx->stream);
  for (void* p : s->allocs) hipFree(p);
  if (s->stage) hipHostFree(s->stage);
  if (s->stage_done) hipEventDestroy(s->stage_done);
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
  const Sizes z = sizes_for(*cfg);
  const int maxn = int(z.maxn);
  MapDev& m = s->map;
  auto& A = s->allocs;
  bool ok = true;
  ok &= dalloc(A, &m.table, size_t(z.hcap)) == hipSuccess;
  ok &= dalloc(A, &m.hot, size_t(z.pool)) == hipSuccess;
  ok &= dalloc(A, &m.vs, size_t(z.pool)) == hipSuccess;
  ok &= dalloc(A, &m.plane, size_t(z.pool) * kPlaneD) == hipSuccess;
  ok &= dalloc(A, &m.mom, size_t(z.pool) * kMomD) == hipSuccess;
  m.pend_cap = std::max(cfg->update_size_thresh - 1, 0);
  ok &= dalloc(A, &m.pend, size_t(z.pool) * size_t(std::max(m.pend_cap, 1)) * kPendF) == hipSuccess;
  ok &= dalloc(A, &m.freelist, size_t(z.pool)) == hipSuccess;
  ok &= dalloc(A, &m.ring_vid, size_t(z.qcap)) == hipSuccess;
  ok &= dalloc(A, &m.ring_stamp, size_t(z.qcap)) == hipSuccess;
  ok &= dalloc(A, &m.ctr, 1) == hipSuccess;
  ok &= dalloc(A, &s->own_x, maxn) == hipSuccess;
  ok &= dalloc(A, &s->own_y, maxn) == hipSuccess;
  ok &= dalloc(A, &s->own_z, maxn) == hipSuccess;
  ok &= dalloc(A, &s->slots, maxn) == hipSuccess;
  ok &= dalloc(A, &s->f, 1) == hipSuccess;
  ok &= dalloc(A, &s->partials, size_t(z.max_blocks) * kObsRowMax) == hipSuccess;
  ok &= dalloc(A, &s->obs_ctr, 8) == hipSuccess;
  ok &= dalloc(A, &s->obs_stat, 8) == hipSuccess;
  ok &= dalloc(A, &s->live, size_t(z.pool)) == hipSuccess;
  ok &= dalloc(A, &s->nlive, 1) == hipSuccess;
  ok &= dalloc(A, &s->imu, 1) == hipSuccess;
  ok &= dalloc(A, &s->own_t, maxn) == hipSuccess;
  if (cfg->scan_resolution > 0.0) {
    ok &= dalloc(A, &s->own_fx, maxn) == hipSuccess;
    ok &= dalloc(A, &s->own_fy, maxn) == hipSuccess;
    ok &= dalloc(A, &s->own_fz, maxn) == hipSuccess;
  }
  ok &= hipHostMalloc(&s->stage, sizeof(ImuPkg)) == hipSuccess;
  ok &= hipEventCreateWithFlags(&s->stage_done, hipEventDisableTiming) == hipSuccess;
  if (!ok) return fail(VMP_ENOMEM);
  m.hmask = uint32_t(z.hcap - 1);
  m.hbits = 0;
  while ((int64_t(1) << m.hbits) < z.hcap) ++m.hbits;
  m.pool = int32_t(z.pool);
  m.qcap = z.qcap;
  m.capacity = cfg->map_capacity;
  m.max_pt = cfg->max_point_thresh;
  m.upd_pt = cfg->update_size_thresh;
  m.plane_thresh = cfg->plane_thresh;
  m.merge_a = cfg->merge_thresh_for_angle;
  m.merge_d = cfg->merge_thresh_for_distance;
  m.vs64 = cfg->voxel_size;
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_rehash_clear, dim3(grid_for(std::min<int64_t>(z.hcap, 1 << 20), 256)), dim3(256), 0, st, m);
  MapCounters c0{};
  if (hipMemcpyAsync(m.ctr, &c0, sizeof(c0), hipMemcpyHostToDevice, st) != hipSuccess) return fail(VMP_EDEVICE);
  // residual slots start invalid (NaN residual), voxel records empty; all
  // stream-ordered so creating a session does not wait for other streams
  hipLaunchKernelGGL(k_init_slots, dim3(grid_for(maxn, 256)), dim3(256), 0, st, s->slots, maxn);
  hipLaunchKernelGGL(k_init_vs, dim3(grid_for(z.pool, 256)), dim3(256), 0, st, m.vs, int32_t(z.pool));
  if (hipMemsetAsync(m.hot, 0, sizeof(VoxHot) * size_t(z.pool), st) != hipSuccess) return fail(VMP_EDEVICE);
  FilterDev f0;
  std::memset(&f0, 0, sizeof(f0));
  f0.cfg = *cfg;
  for (int i = 0; i < 9; i += 4) {
    f0.x.rot[i] = 1.0;
    f0.x.rot_ext[i] = 1.0;
  }
  f0.x.g[2] = -kGravity;
  for (int i = 0; i < 23; ++i) f0.P[i * 23 + i] = 1.0;
  if (hipMemcpyAsync(s->f, &f0, sizeof(f0), hipMemcpyHostToDevice, st) != hipSuccess) return fail(VMP_EDEVICE);
  if (hipMemsetAsync(s->obs_ctr, 0, 8 * sizeof(int32_t), st) != hipSuccess) return fail(VMP_EDEVICE);
  if (hipMemsetAsync(s->imu, 0, sizeof(ImuDev), st) != hipSuccess) return fail(VMP_EDEVICE);
  if (hipMemsetAsync(s->obs_stat, 0, 8 * sizeof(int32_t), st) != hipSuccess) return fail(VMP_EDEVICE);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return fail(VMP_EDEVICE);
  *out = s;
  return VMP_OK;
}

extern "C" const char* vmp_last_error(const vmp_session* s) { return s ? s->err.c_str() : "null session"; }

extern "C" int vmp_state_set(vmp_session* s, const vmp_state* x, const double* P) {
  if (!s) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(s->ctx->mu);
  VMP_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  if (x) VMP_HIP(hipMemcpyAsync(&s->f->x, x, sizeof(vmp_state), hipMemcpyHostToDevice, st));
  if (P) VMP_HIP(hipMemcpyAsync(s->f->P, P, sizeof(double) * 529, hipMemcpyHostToDevice, st));
  VMP_HIP(hipStreamSynchronize(st));
  return VMP_OK;
}

extern "C" int vmp_state_get(vmp_session* s, vmp_state* x, double* P) {
  if (!s) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(s->ctx->mu);
  VMP_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  if (x) VMP_HIP(hipMemcpyAsync(x, &s->f->x, sizeof(vmp_state), hipMemcpyDeviceToHost, st));
  if (P) VMP_HIP(hipMemcpyAsync(P, s->f->P, sizeof(double) * 529, hipMemcpyDeviceToHost, st));
  VMP_HIP(hipStreamSynchronize(st));
  return VMP_OK;
}

extern "C" int vmp_state_apply_motion(vmp_session* s, const double* dR, const double* dp, const double* qdiag) {
  if (!s || !dR || !dp || !qdiag) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(s->ctx->mu);
  VMP_HIP(hipSetDevice(s->ctx->device));
  const D33 R = d33_from_cm(dR);
  const D3 p{{dp[0], dp[1], dp[2]}};
  Diag23 d;
  for (int i = 0; i < 23; ++i) d.q[i] = qdiag[i];
  // ordered before the next call on this context's stream
  hipLaunchKernelGGL(k_apply_motion, dim3(1), dim3(64), 0, s->ctx->stream, s->f, R, p, d);
  VMP_HIP(hipGetLastError());
  return VMP_OK;
}

// motion prior of n sessions on context c's stream (c->mu held by the caller);
// qdiag holds 23 values per session, `qstride` apart (0: one shared vector)
static int enqueue_motion(vmp_ctx* c, vmp_session** ss, int n, const double* dR, const double* dp,
                          const double* qdiag, int qstride) {
  vmp_session* s = ss[0];  // carries the error message
  VMP_HIP(hipSetDevice(c->device));
  if (ws_reserve(c, 1, n) != VMP_OK) return VMP_ENOMEM;
  BatchWs& w = c->ws;
  // the pinned table is reused: wait until the previous upload has been read
  if (w.motion_pending) VMP_HIP(hipEventSynchronize(w.motion_done));
  for (int i = 0; i < n; ++i) {
    MotionRec& r = w.h_motion[i];
    r.f = ss[i]->f;
    std::memcpy(r.dR, dR + 9 * size_t(i), sizeof(r.dR));
    std::memcpy(r.dp, dp + 3 * size_t(i), sizeof(r.dp));
    std::memcpy(r.q, qdiag + size_t(qstride) * size_t(i), sizeof(r.q));
  }
  VMP_HIP(hipMemcpyAsync(w.d_motion, w.h_motion, sizeof(MotionRec) * n, hipMemcpyHostToDevice, c->stream));
  VMP_HIP(hipEventRecord(w.motion_done, c->stream));
  w.motion_pending = true;
  hipLaunchKernelGGL(k_apply_motion_batch, dim3((n + 63) / 64), dim3(64), 0, c->stream, w.d_motion, n);
  VMP_HIP(hipGetLastError());
  return VMP_OK;
}

extern "C" int vmp_state_apply_motion_batch(vmp_session** ss, int32_t n, const double* dR, const double* dp,
                                            const double* qdiag) {
  if (!ss || n <= 0 || !dR || !dp || !qdiag) return VMP_EINVAL;
  vmp_session* s = ss[0];  // carries the error message
  vmp_ctx* c = s ? s->ctx : nullptr;
  if (!c) return VMP_EINVAL;
  for (int i = 0; i < n; ++i)
    if (!ss[i] || ss[i]->ctx != c) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  return enqueue_motion(c, ss, n, dR, dp, qdiag, 23);
}

extern "C" int vmp_scan_set_points(vmp_session* s, const float* x, const float* y, const float* z, int32_t n) {
  if (!s || n < 0 || n > s->cfg.max_scan_points || (n > 0 && (!x || !y || !z))) {
    if (s) s->err = "scan size exceeds max_scan_points or null input";
    return VMP_EINVAL;
  }
  if (!host_points_finite(x, y, z, n)) {
    s->err = "scan holds a non-finite coordinate";
    return VMP_EINVAL;
  }
  std::lock_guard<std::mutex> lk(s->ctx->mu);
  VMP_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  if (n > 0) {
    VMP_HIP(hipMemcpyAsync(s->own_x, x, sizeof(float) * n, hipMemcpyHostToDevice, st));
    VMP_HIP(hipMemcpyAsync(s->own_y, y, sizeof(float) * n, hipMemcpyHostToDevice, st));
    VMP_HIP(hipMemcpyAsync(s->own_z, z, sizeof(float) * n, hipMemcpyHostToDevice, st));
    VMP_HIP(hipStreamSynchronize(st));
  }
  s->cx = s->own_x;
  s->cy = s->own_y;
  s->cz = s->own_z;
  s->n = n;
  s->has_points = true;
  return VMP_OK;
}

// bind device points without validation (the caller has checked them)
static int bind_points_device(vmp_session* s, const float* dx, const float* dy, const float* dz, int32_t n) {
  if (!s || n < 0 || n > s->cfg.max_scan_points || (n > 0 && (!dx || !dy || !dz))) {
    if (s) s->err = "scan size exceeds max_scan_points or null input";
    return VMP_EINVAL;
  }
  s->cx = dx;
  s->cy = dy;
  s->cz = dz;
  s->n = n;
  s->has_points = true;
  s->raw_pending = false;
  return VMP_OK;
}

extern "C" int vmp_scan_set_points_device(vmp_session* s, const float* dx, const float* dy, const float* dz,
                                          int32_t n) {
  if (!s || n < 0 || n > s->cfg.max_scan_points || (n > 0 && (!dx || !dy || !dz))) {
    if (s) s->err = "scan size exceeds max_scan_points or null input";
    return VMP_EINVAL;
  }
  const int rc = bind_points_device(s, dx, dy, dz, n);
  if (rc == VMP_OK) {
    s->rt = nullptr;
    s->needs_check = s->ctx->check_points;
  }
  return rc;
}

extern "C" int vmp_scan_set_points_device_batch(vmp_session** ss, int32_t n, const uint64_t* dx, const uint64_t* dy,
                                                const uint64_t* dz, const int32_t* counts) {
  if (!ss || n <= 0 || !dx || !dy || !dz || !counts || !ss[0]) return VMP_EINVAL;
  for (int i = 0; i < n; ++i)
    if (!ss[i] || ss[i]->ctx->device != ss[0]->ctx->device || counts[i] < 0 ||
        counts[i] > ss[i]->cfg.max_scan_points)
      return VMP_EINVAL;
  for (int i = 0; i < n; ++i) {
    const int rc = vmp_scan_set_points_device(ss[i], reinterpret_cast<const float*>(dx[i]),
                                              reinterpret_cast<const float*>(dy[i]),
                                              reinterpret_cast<const float*>(dz[i]), counts[i]);
    if (rc != VMP_OK) return rc;
  }
  return VMP_OK;
}

// ----------------------------------------------------------- IMU side -----
extern "C" int vmp_imu_init(vmp_session* s, const vmp_imu_cfg* cfg, const double* t, const double* acc,
                            const double* gyro, int32_t m, double cloud_end_time, int32_t* initialized) {
  if (!s || !cfg || m < 0 || (m > 0 && (!t || !acc || !gyro)) || cfg->imu_init_num < 1) return VMP_EINVAL;
  if (initialized) *initialized = s->imu_inited ? 1 : 0;
  if (s->imu_inited) return VMP_OK;
  for (int32_t i = 0; i < m; ++i) {
    if (!std::isfinite(t[i])) return VMP_EINVAL;
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(acc[3 * i + k]) || !std::isfinite(gyro[3 * i + k])) return VMP_EINVAL;
  }
  s->init_t.insert(s->init_t.end(), t, t + m);
  s->init_acc.insert(s->init_acc.end(), acc, acc + 3 * size_t(m));
  s->init_gyro.insert(s->init_gyro.end(), gyro, gyro + 3 * size_t(m));
  const int total = int(s->init_t.size());
  if (total < cfg->imu_init_num || m == 0) return VMP_OK;  // lio_builder.cpp:31-32
  vmp_ctx* c = s->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  VMP_HIP(hipSetDevice(c->device));
  settle(c);
  double* buf = nullptr;
  VMP_HIP(hipMalloc(&buf, sizeof(double) * 7 * size_t(total)));
  hipStream_t st = c->stream;
  int rc = VMP_OK;
  if (hipMemcpyAsync(buf, s->init_t.data(), sizeof(double) * total, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(buf + total, s->init_acc.data(), sizeof(double) * 3 * total, hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      hipMemcpyAsync(buf + 4 * total, s->init_gyro.data(), sizeof(double) * 3 * total, hipMemcpyHostToDevice, st) !=
          hipSuccess)
    rc = VMP_EDEVICE;
  ImuCfgDev d{};
  const D33 ril = d33_from_cm(cfg->r_il);
  for (int i = 0; i < 9; ++i) d.r_il[i] = ril.m[i];
  for (int i = 0; i < 3; ++i) d.p_il[i] = cfg->p_il[i];
  d.gravity_align = cfg->gravity_align;
  d.n = total;
  // Q (lio_builder.cpp:9-12) and the IMU_INIT -> MAP_INIT bookkeeping
  // (lio_builder.cpp:181-182): the scan end of this package
  const double q[4] = {cfg->ng, cfg->na, cfg->nbg, cfg->nba};
  if (rc == VMP_OK) {
    hipLaunchKernelGGL(k_imu_init, dim3(1), dim3(64), 0, st, s->imu, s->f, buf, buf + total, buf + 4 * total, d);
    if (hipMemcpyAsync(s->imu->q, q, sizeof(q), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(&s->imu->last_cloud_end, &cloud_end_time, sizeof(double), hipMemcpyHostToDevice, st) !=
            hipSuccess)
      rc = VMP_EDEVICE;
  }
  if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess) rc = VMP_EDEVICE;
  hipFree(buf);
  if (rc != VMP_OK) {
    s->err = "IMU initialisation failed on the device";
    return rc;
  }
  s->imu_inited = true;
  s->init_t.clear();
  s->init_acc.clear();
  s->init_gyro.clear();
  if (initialized) *initialized = 1;
  return VMP_OK;
}

// stage a package's IMU samples (c->mu held); the copy is ordered on the
// context's stream before the step that reads them
static int stage_imu(vmp_session* s, const double* it, const double* iacc, const double* igyro, int32_t m,
                     double cloud_start, double cloud_end) {
  if (s->stage_pending) VMP_HIP(hipEventSynchronize(s->stage_done));
  ImuPkg& p = *s->stage;
  p.cloud_start = cloud_start;
  p.cloud_end = cloud_end;
  p.m = m;
  p.pad = 0;
  for (int32_t i = 0; i < m; ++i) {
    p.s[i][0] = it[i];
    for (int k = 0; k < 3; ++k) {
      p.s[i][1 + k] = iacc[3 * i + k];
      p.s[i][4 + k] = igyro[3 * i + k];
    }
  }
  VMP_HIP(hipMemcpyAsync(&s->imu->pkg, s->stage, imu_pkg_bytes(m), hipMemcpyHostToDevice, s->ctx->stream));
  VMP_HIP(hipEventRecord(s->stage_done, s->ctx->stream));
  s->stage_pending = true;
  return VMP_OK;
}

// one package's place in the batch staging: its session's slot, its offset
// in the staging and its length (a multiple of 8 bytes)
struct PkgMove {
  ImuPkg* dst;
  int64_t src_off;
  int32_t bytes;
  int32_t pad;
};
__global__ void k_scatter_pkgs(const uint8_t* stage) {
  const PkgMove mv = reinterpret_cast<const PkgMove*>(stage)[blockIdx.x];
  const uint2* src = reinterpret_cast<const uint2*>(stage + mv.src_off);
  uint2* dst = reinterpret_cast<uint2*>(mv.dst);
  for (int i = threadIdx.x; i < mv.bytes / 8; i += blockDim.x) dst[i] = src[i];
}

// Stage the IMU packages of n sessions of one context with one copy and one
// scatter launch (c->mu held); ioff[i] is session i's first sample.
static int stage_imu_batch(vmp_ctx* c, vmp_session** ss, int n, const int64_t* ioff, const double* it,
                           const double* iacc, const double* igyro, const int32_t* m, const double* cstart,
                           const double* cend) {
  vmp_session* s = ss[0];  // a failure is reported on the batch's first session
  const size_t head = (sizeof(PkgMove) * size_t(n) + 255) & ~size_t(255);
  size_t total = head;
  for (int i = 0; i < n; ++i) total += imu_pkg_bytes(m[i]);
  hipStream_t st = c->stream;
  if (c->pkstage_pending) VMP_HIP(hipEventSynchronize(c->pkstage_done));  // its copy has read the staging
  c->pkstage_pending = false;
  if (total > c->pkstage_cap) {
    // the last scatter may still read the device staging
    VMP_HIP(hipStreamSynchronize(st));
    if (c->h_pkstage) hipHostFree(c->h_pkstage);
    if (c->d_pkstage) hipFree(c->d_pkstage);
    c->h_pkstage = nullptr;
    c->d_pkstage = nullptr;
    c->pkstage_cap = 0;
    const size_t cap = total + total / 4;
    if (hipHostMalloc(&c->h_pkstage, cap) != hipSuccess || hipMalloc(&c->d_pkstage, cap) != hipSuccess) {
      if (c->h_pkstage) hipHostFree(c->h_pkstage);
      c->h_pkstage = nullptr;
      c->d_pkstage = nullptr;
      return VMP_ENOMEM;
    }
    c->pkstage_cap = cap;
  }
  if (!c->pkstage_done) VMP_HIP(hipEventCreateWithFlags(&c->pkstage_done, hipEventDisableTiming));
  PkgMove* mv = reinterpret_cast<PkgMove*>(c->h_pkstage);
  size_t o = head;
  for (int i = 0; i < n; ++i) {
    ImuPkg& p = *reinterpret_cast<ImuPkg*>(c->h_pkstage + o);
    p.cloud_start = cstart[i];
    p.cloud_end = cend[i];
    p.m = m[i];
    p.pad = 0;
    const int64_t b = ioff[i];
    for (int32_t k = 0; k < m[i]; ++k) {
      p.s[k][0] = it[b + k];
      for (int a = 0; a < 3; ++a) {
        p.s[k][1 + a] = iacc[3 * (b + k) + a];
        p.s[k][4 + a] = igyro[3 * (b + k) + a];
      }
    }
    mv[i] = PkgMove{&ss[i]->imu->pkg, int64_t(o), int32_t(imu_pkg_bytes(m[i])), 0};
    o += imu_pkg_bytes(m[i]);
  }
  VMP_HIP(hipMemcpyAsync(c->d_pkstage, c->h_pkstage, total, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_scatter_pkgs, dim3(n), dim3(256), 0, st, c->d_pkstage);
  VMP_HIP(hipEventRecord(c->pkstage_done, st));
  c->pkstage_pending = true;
  return hipGetLastError() == hipSuccess ? VMP_OK : VMP_EDEVICE;
}

static int check_raw_args(vmp_session* s, int32_t n, const double* it, const double* iacc, const double* igyro,
                          int32_t m, double cloud_start, double cloud_end) {
  if (!s->imu_inited) {
    s->err = "raw package before IMU initialisation (vmp_imu_init)";
    return VMP_ESTATE;
  }
  if (n < 0 || n > s->cfg.max_scan_points || m < 1 || m > kMaxImu || !it || !iacc || !igyro ||
      !std::isfinite(cloud_start) || !std::isfinite(cloud_end)) {
    s->err = "raw package: point count, 1..512 IMU samples or cloud times out of range";
    return VMP_EINVAL;
  }
  for (int32_t i = 0; i < m; ++i) {
    bool ok = std::isfinite(it[i]) && (i == 0 || it[i] >= it[i - 1]);
    for (int k = 0; k < 3; ++k) ok = ok && std::isfinite(iacc[3 * i + k]) && std::isfinite(igyro[3 * i + k]);
    if (!ok) {
      s->err = "raw package: non-finite or unordered IMU sample";
      return VMP_EINVAL;
    }
  }
  return VMP_OK;
}

extern "C" int vmp_scan_set_raw(vmp_session* s, const float* x, const float* y, const float* z, const float* t_ms,
                                int32_t n, const double* it, const double* iacc, const double* igyro, int32_t m,
                                double cloud_start, double cloud_end) {
  if (!s) return VMP_EINVAL;
  int rc = check_raw_args(s, n, it, iacc, igyro, m, cloud_start, cloud_end);
  if (rc != VMP_OK) return rc;
  if (n > 0 && (!x || !y || !z || !t_ms)) return VMP_EINVAL;
  if (!host_points_finite(x, y, z, n)) {
    s->err = "scan holds a non-finite coordinate";
    return VMP_EINVAL;
  }
  for (int32_t i = 0; i < n; ++i)
    if (!std::isfinite(t_ms[i]) || (i > 0 && !(t_ms[i] >= t_ms[i - 1]))) {
      s->err = "raw scan: point times must be finite and ascending";
      return VMP_EINVAL;
    }
  std::lock_guard<std::mutex> lk(s->ctx->mu);
  VMP_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  if (n > 0) {
    VMP_HIP(hipMemcpyAsync(s->own_x, x, sizeof(float) * n, hipMemcpyHostToDevice, st));
    VMP_HIP(hipMemcpyAsync(s->own_y, y, sizeof(float) * n, hipMemcpyHostToDevice, st));
    VMP_HIP(hipMemcpyAsync(s->own_z, z, sizeof(float) * n, hipMemcpyHostToDevice, st));
    VMP_HIP(hipMemcpyAsync(s->own_t, t_ms, sizeof(float) * n, hipMemcpyHostToDevice, st));
  }
  rc = stage_imu(s, it, iacc, igyro, m, cloud_start, cloud_end);
  if (rc != VMP_OK) return rc;
  VMP_HIP(hipStreamSynchronize(st));
  s->rx = s->own_x;
  s->ry = s->own_y;
  s->rz = s->own_z;
  s->rt = s->own_t;
  s->raw_n = n;
  s->raw_pending = true;
  s->needs_check = false;
  s->cx = s->own_x;
  s->cy = s->own_y;
  s->cz = s->own_z;
  s->n = n;
  s->has_points = true;
  return VMP_OK;
}

extern "C" int vmp_scan_set_raw_device(vmp_session* s, const float* dx, const float* dy, const float* dz,
                                       const float* dt_ms, int32_t n, const double* it, const double* iacc,
                                       const double* igyro, int32_t m, double cloud_start, double cloud_end) {
  if (!s) return VMP_EINVAL;
  int rc = check_raw_args(s, n, it, iacc, igyro, m, cloud_start, cloud_end);
  if (rc != VMP_OK) return rc;
  if (n > 0 && (!dx || !dy || !dz || !dt_ms)) return VMP_EINVAL;
  vmp_ctx* c = s->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  VMP_HIP(hipSetDevice(c->device));
  rc = stage_imu(s, it, iacc, igyro, m, cloud_start, cloud_end);
  if (rc != VMP_OK) return rc;
  s->needs_check = c->check_points;
  s->rx = dx;
  s->ry = dy;
  s->rz = dz;
  s->rt = dt_ms;
  s->raw_n = n;
  s->raw_pending = true;
  s->cx = s->own_x;
  s->cy = s->own_y;
  s->cz = s->own_z;
  s->n = n;
  s->has_points = true;
  return VMP_OK;
}

extern "C" int vmp_scan_set_raw_device_batch(vmp_session** ss, int32_t n, const uint64_t* dx, const uint64_t* dy,
                                             const uint64_t* dz, const uint64_t* dt, const int32_t* counts,
                                             const double* imu_t, const double* imu_acc, const double* imu_gyro,
                                             const int32_t* imu_counts, const double* cloud_start,
                                             const double* cloud_end) {
  if (!ss || n <= 0 || !dx || !dy || !dz || !dt || !counts || !imu_t || !imu_acc || !imu_gyro || !imu_counts ||
      !cloud_start || !cloud_end || !ss[0])
    return VMP_EINVAL;
  vmp_ctx* c = ss[0]->ctx;
  std::vector<PtsRef> refs(static_cast<size_t>(n));
  std::vector<int64_t> ioff(static_cast<size_t>(n) + 1, 0);
  for (int i = 0; i < n; ++i) {
    if (!ss[i] || ss[i]->ctx != c) return VMP_EINVAL;
    const int64_t o = ioff[size_t(i)];
    const int rc = check_raw_args(ss[i], counts[i], imu_t + o, imu_acc + 3 * o, imu_gyro + 3 * o, imu_counts[i],
                                  cloud_start[i], cloud_end[i]);
    if (rc != VMP_OK) return rc;
    ioff[size_t(i) + 1] = o + imu_counts[i];
    refs[size_t(i)] = PtsRef{reinterpret_cast<const float*>(dx[i]), reinterpret_cast<const float*>(dy[i]),
                             reinterpret_cast<const float*>(dz[i]), reinterpret_cast<const float*>(dt[i]), counts[i],
                             0};
  }
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return VMP_EDEVICE;
  const int rc = stage_imu_batch(c, ss, n, ioff.data(), imu_t, imu_acc, imu_gyro, imu_counts, cloud_start, cloud_end);
  if (rc != VMP_OK) return rc;
  for (int i = 0; i < n; ++i) {
    vmp_session* s = ss[i];
    s->rx = refs[size_t(i)].x;
    s->ry = refs[size_t(i)].y;
    s->rz = refs[size_t(i)].z;
    s->rt = refs[size_t(i)].t;
    s->raw_n = counts[i];
    s->raw_pending = true;
    s->needs_check = c->check_points;
    s->cx = s->own_x;
    s->cy = s->own_y;
    s->cz = s->own_z;
    s->n = counts[i];
    s->has_points = true;
  }
  return VMP_OK;
}

extern "C" int32_t vmp_scan_export(vmp_session* s, int32_t cap, float* x, float* y, float* z) {
  if (!s || cap < 0) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(s->ctx->mu);
  if (hipSetDevice(s->ctx->device) != hipSuccess) return VMP_EDEVICE;
  settle(s->ctx);
  const int32_t n = s->has_points ? s->n : 0;
  const int32_t k = std::min(cap, n);
  hipStream_t st = s->ctx->stream;
  if (k > 0) {
    if ((x && hipMemcpyAsync(x, s->cx, sizeof(float) * k, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        (y && hipMemcpyAsync(y, s->cy, sizeof(float) * k, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        (z && hipMemcpyAsync(z, s->cz, sizeof(float) * k, hipMemcpyDeviceToHost, st) != hipSuccess))
      return VMP_EDEVICE;
  }
  if (hipStreamSynchronize(st) != hipSuccess) return VMP_EDEVICE;
  return n;
}

// lidarToBody: fp32 transform with the extrinsic, in PCL's order
__global__ void k_to_body(const FilterDev* f, const float* x, const float* y, const float* z, int n, float* ox,
                          float* oy, float* oz) {
  __shared__ float m[9], t[3];
  if (threadIdx.x == 0) {
    const D33 r = d33_from_cm(f->x.rot_ext);
    for (int i = 0; i < 9; ++i) m[i] = float(r.m[i]);
    for (int i = 0; i < 3; ++i) t[i] = float(f->x.pos_ext[i]);
  }
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float o[3];
    ins_transform(m, t, x[i], y[i], z[i], o);
    ox[i] = o[0];
    oy[i] = o[1];
    oz[i] = o[2];
  }
}

extern "C" int32_t vmp_cloud_body_export(vmp_session* s, int32_t cap, float* x, float* y, float* z) {
  if (!s || cap < 0) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(s->ctx->mu);
  if (hipSetDevice(s->ctx->device) != hipSuccess) return VMP_EDEVICE;
  settle(s->ctx);
  const int32_t n = s->cloud_n;
  const int32_t k = std::min(cap, n);
  if (k <= 0) return n;
  hipStream_t st = s->ctx->stream;
  float* buf = nullptr;
  if (hipMalloc(&buf, sizeof(float) * 3 * size_t(k)) != hipSuccess) return VMP_ENOMEM;
  hipLaunchKernelGGL(k_to_body, dim3(std::max(1, std::min((k + 255) / 256, 1024))), dim3(256), 0, st, s->f,
                     s->own_x, s->own_y, s->own_z, k, buf, buf + k, buf + 2 * k);
  int rc = n;
  if ((x && hipMemcpyAsync(x, buf, sizeof(float) * k, hipMemcpyDeviceToHost, st) != hipSuccess) ||
      (y && hipMemcpyAsync(y, buf + k, sizeof(float) * k, hipMemcpyDeviceToHost, st) != hipSuccess) ||
      (z && hipMemcpyAsync(z, buf + 2 * k, sizeof(float) * k, hipMemcpyDeviceToHost, st) != hipSuccess) ||
      hipStreamSynchronize(st) != hipSuccess)
    rc = VMP_EDEVICE;
  hipFree(buf);
  return rc;
}

// Eigen::Quaterniond(const Matrix3d&): Shepperd's method (largest of the
// trace and the diagonal picks the branch)
static void quat_from_rot(const D33& R, double q[4]) {
  const double* m = R.m;
  const double t = m[0] + m[4] + m[8];
  double w, v[3];
  if (t > 0) {
    double r = std::sqrt(t + 1.0);
    w = 0.5 * r;
    r = 0.5 / r;
    v[0] = (m[7] - m[5]) * r;
    v[1] = (m[2] - m[6]) * r;
    v[2] = (m[3] - m[1]) * r;
  } else {
    int i = 0;
    if (m[4] > m[0]) i = 1;
    if (m[8] > m[i * 4]) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    double r = std::sqrt(m[i * 4] - m[j * 4] - m[k * 4] + 1.0);
    v[i] = 0.5 * r;
    r = 0.5 / r;
    w = (m[k * 3 + j] - m[j * 3 + k]) * r;
    v[j] = (m[j * 3 + i] + m[i * 3 + j]) * r;
    v[k] = (m[k * 3 + i] + m[i * 3 + k]) * r;
  }
  q[0] = w;
  q[1] = v[0];
  q[2] = v[1];
  q[3] = v[2];
}

extern "C" int vmp_odom_export(vmp_session* s, double pos[3], double quat_wxyz[4], double cov6x6[36]) {
  if (!s) return VMP_EINVAL;
  vmp_state x;
  std::vector<double> P(529);
  const int rc = vmp_state_get(s, &x, P.data());
  if (rc != VMP_OK) return rc;
  if (pos)
    for (int i = 0; i < 3; ++i) pos[i] = x.pos[i];
  if (quat_wxyz) quat_from_rot(d33_from_cm(x.rot), quat_wxyz);
  if (cov6x6)
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) cov6x6[r * 6 + c] = P[size_t(c) * 23 + size_t(r)];
  return VMP_OK;
}

// ------------------------------------------------------- step pipeline -----
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
  j.out = s->f->obs;
  j.hmask = s->map.hmask;
  j.n = s->n;
  j.ppt = obs_ppt_for(s->ctx, s->n);
  j.nblocks = std::max(1, (s->n + kObsThreads * j.ppt - 1) / (kObsThreads * j.ppt));
  j.block0 = block0;
  j.pad = 0;
}

enum { kStepIeskf = 1, kStepMap = 2, kStepBuild = 4, kStepObserveOnce = 8 };

// Host-side job tables for a batch, uploaded in one copy each.
static int stage_batch(vmp_ctx* c, vmp_session** ss, int n, int64_t* ntot_out, int* nblocks_out) {
  BatchWs& w = c->ws;
  int64_t off = 0;
  int blocks = 0;
  for (int i = 0; i < n; ++i) {
    vmp_session* s = ss[i];
    w.h_maps[i] = s->map;
    SessInfo& S = w.h_si[i];
    std::memset(&S, 0, sizeof(S));
    S.x = s->cx;
    S.y = s->cy;
    S.z = s->cz;
    S.prm = &s->f->upd_prm;
    S.n = s->n;
    S.off = int32_t(off);
    off += s->n;
    fill_job(s, w.h_jobs[i], blocks);
    blocks += w.h_jobs[i].nblocks;
    w.h_flist[i] = s->f;
    w.h_obsctr[i] = s->obs_ctr;
    w.h_obsstat[i] = s->obs_stat;
  }
  *ntot_out = off;
  *nblocks_out = blocks;
  // the sessions starting in each compaction tile (only when the library's
  // compaction runs: at most kPrimMaxItems points)
  if (off <= kPrimMaxItems) {
    const int tiles = prim_tiles(off);
    int sidx = 0;
    for (int t = 0; t <= tiles; ++t) {
      const int64_t lo = int64_t(t) * kPrimTile;
      while (sidx < n && w.h_si[sidx].off < lo) ++sidx;
      w.h_sfirst[t] = sidx;
    }
  }
  return VMP_OK;
}

// Instantiated filter-stage graph for this step's shape, captured on first
// use.  The observe grid is rounded up (to a multiple of 1/8 of its power of
// two) so that one graph serves steps whose block totals differ a little;
// the surplus blocks return at once.  The job tables the kernels read are
// uploaded before each launch, so a replay sees the step's own sessions.
template <class Stage>
static hipGraphExec_t filter_graph(vmp_ctx* c, int n, int nblocks, int iters, bool ext, Stage&& stage) {
  int p2 = 1;
  while (p2 < nblocks) p2 <<= 1;
  const int quant = std::max(1, p2 / 8);
  const int grid = (nblocks + quant - 1) / quant * quant;
  vmp_ctx::FilterGraph* hit = nullptr;
  vmp_ctx::FilterGraph* victim = &c->graphs[0];
  for (auto& g : c->graphs) {
    if (g.exec && g.ws_gen == c->ws_gen && g.n == n && g.grid == grid && g.iters == iters && g.ext == int(ext) &&
        g.arith == c->obs_arith) {
      hit = &g;
      break;
    }
    if (!g.exec || g.last_use < victim->last_use) victim = &g;
  }
  if (!hit) {
    hipStream_t st = c->stream;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    const int saved = c->obs_launches;
    stage(grid);
    c->obs_launches = saved;
    hipGraph_t graph = nullptr;
    if (hipStreamEndCapture(st, &graph) != hipSuccess || !graph) return nullptr;
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess || !exec) return nullptr;
    if (victim->exec) hipGraphExecDestroy(victim->exec);
    *victim = vmp_ctx::FilterGraph{};
    victim->exec = exec;
    victim->n = n;
    victim->grid = grid;
    victim->iters = iters;
    victim->ext = int(ext);
    victim->arith = c->obs_arith;
    victim->ws_gen = c->ws_gen;
    hit = victim;
    c->graph_builds++;
  }
  hit->last_use = ++c->graph_clock;
  return hit->exec;
}

// Instantiated map-stage graph for this step's shape, or nullptr (then the
// caller launches the stage itself).  A shape is captured the second time in
// a row it occurs, so steps whose sizes change every time (batches of uneven
// scans) never pay for a capture.  Only graphs made of kernel nodes are
// kept: a memset node (which a library routine may record) has given wrong
// results on replays here.
template <class Stage>
static hipGraphExec_t map_graph(vmp_ctx* c, const int64_t key[3], Stage&& stage) {
  auto same = [&](const int64_t* a) { return a[0] == key[0] && a[1] == key[1] && a[2] == key[2]; };
  vmp_ctx::MapGraph* victim = &c->mgraphs[0];
  for (auto& g : c->mgraphs) {
    if (g.exec && g.ws_gen == c->ws_gen && same(g.key)) {
      g.last_use = ++c->graph_clock;
      return g.exec;
    }
    if (!g.exec || g.last_use < victim->last_use) victim = &g;
  }
  if (same(c->mkey_bad)) return nullptr;
  if (!same(c->mkey_prev)) {
    for (int k = 0; k < 3; ++k) c->mkey_prev[k] = key[k];
    return nullptr;
  }
  hipStream_t st = c->stream;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  const int saved_kev = c->kev_used;
  const int rc = stage();
  c->kev_used = saved_kev;
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(st, &graph);
  bool ok = rc == VMP_OK && ce == hipSuccess && graph;
  if (ok) {
    size_t nn = 0;
    ok = hipGraphGetNodes(graph, nullptr, &nn) == hipSuccess && nn > 0;
    std::vector<hipGraphNode_t> nodes(ok ? nn : 0);
    if (ok) ok = hipGraphGetNodes(graph, nodes.data(), &nn) == hipSuccess;
    for (size_t i = 0; ok && i < nn; ++i) {
      hipGraphNodeType t;
      ok = hipGraphNodeGetType(nodes[i], &t) == hipSuccess && t == hipGraphNodeTypeKernel;
    }
  }
  hipGraphExec_t exec = nullptr;
  if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess && exec;
  if (graph) hipGraphDestroy(graph);
  hipGetLastError();  // a failed capture leaves nothing behind for the eager launches
  if (!ok) {  // this shape runs eagerly from now on
    for (int k = 0; k < 3; ++k) c->mkey_bad[k] = key[k];
    return nullptr;
  }
  if (victim->exec) hipGraphExecDestroy(victim->exec);
  *victim = vmp_ctx::MapGraph{};
  victim->exec = exec;
  for (int k = 0; k < 3; ++k) victim->key[k] = key[k];
  victim->ws_gen = c->ws_gen;
  victim->last_use = ++c->graph_clock;
  c->map_graph_builds++;
  return exec;
}

static int enqueue_step(vmp_ctx* c, vmp_session** ss, int n, int mode) {
  if (n > kMaxBatchSessions) {
    ss[0]->err = "more than 32768 sessions in one batch";
    return VMP_EINVAL;
  }
  BatchWs& w = c->ws;
  hipStream_t st = c->stream;
  int64_t ntot64 = 0;
  int nblocks = 0;
  stage_batch(c, ss, n, &ntot64, &nblocks);
  const int ntot = int(ntot64);
  for (int i = 0; i < n; ++i) w.h_si[i].is_build = (mode & kStepBuild) ? 1 : 0;
  // the job tables in one copy: from the first table through the last
  // one's n entries (the unused tails of the earlier tables ride along)
  const size_t tab_bytes = ntot <= kPrimMaxItems
                               ? w.tab_sfirst + sizeof(int32_t) * size_t(prim_tiles(ntot) + 1)
                               : w.tab_obsstat + sizeof(int32_t*) * n;
  if (hipMemcpyAsync(w.d_tab, w.h_tab, tab_bytes, hipMemcpyHostToDevice, st) != hipSuccess)
    return VMP_EDEVICE;
  const bool ext = ss[0]->cfg.estimate_ext != 0;
  // zero: clear the sessions' observe counters first (inside the filter stage
  // k_ieskf_begin and k_ieskf_step leave them cleared instead); reduce: add the
  // partial rows here (inside the filter stage k_ieskf_step adds them)
  auto observe = [&](int grid, bool zero, bool reduce) {
    if (zero) hipLaunchKernelGGL(k_zero_lists, dim3(n), dim3(64), 0, st, w.d_obsctr, 8);
    const bool tm = c->timing && c->tev_used + 2 <= 2 * 64;
    const bool w32 = c->obs_arith == 1;
    auto launch = [&]() {
      if (ext && w32)
        hipLaunchKernelGGL((observe_kernel<12, true>), dim3(grid), dim3(kObsThreads), 0, st, w.d_jobs, n);
      else if (ext)
        hipLaunchKernelGGL(observe_kernel<12>, dim3(grid), dim3(kObsThreads), 0, st, w.d_jobs, n);
      else if (w32)
        hipLaunchKernelGGL((observe_kernel<6, true>), dim3(grid), dim3(kObsThreads), 0, st, w.d_jobs, n);
      else
        hipLaunchKernelGGL(observe_kernel<6>, dim3(grid), dim3(kObsThreads), 0, st, w.d_jobs, n);
    };
    if (tm) hipEventRecord(c->tev[c->tev_used], st);
    const int reps = tm ? c->obs_time_reps : 1;
    for (int r = 0; r < reps; ++r) launch();
    if (tm) {
      hipEventRecord(c->tev[c->tev_used + 1], st);
      c->tev_used += 2;
      if (reps > 1) {  // the counters of this pass once more, from zero
        hipLaunchKernelGGL(k_zero_lists, dim3(n), dim3(64), 0, st, w.d_obsctr, 8);
        launch();
      }
    }
    if (reduce && ext)
      hipLaunchKernelGGL(k_obs_reduce<12>, dim3(n), dim3(kObsThreads), 0, st, w.d_jobs);
    else if (reduce)
      hipLaunchKernelGGL(k_obs_reduce<6>, dim3(n), dim3(kObsThreads), 0, st, w.d_jobs);
    c->obs_launches++;
  };
  if (mode & kStepObserveOnce) {
    hipLaunchKernelGGL(k_obs_prep, dim3(n), dim3(64), 0, st, w.d_flist);
    observe(nblocks, true, true);
    return VMP_OK;
  }
  hipEventRecord(c->ev[0], st);
  if (mode & kStepIeskf) {
    int max_iter = 1;
    for (int i = 0; i < n; ++i) max_iter = std::max(max_iter, ss[i]->cfg.max_iter);
    auto filter_stage = [&](int grid) {
      hipLaunchKernelGGL(k_ieskf_begin, dim3(n), dim3(kBeginEndThreads), 0, st, w.d_flist, w.d_obsctr);
      for (int it = 0; it < max_iter; ++it) {
        observe(grid, false, false);
        timed_launch(c, kTkIeskfStep, [&] {
          if (ext)
            hipLaunchKernelGGL(k_ieskf_step<12>, dim3(n), dim3(kStepThreads), 0, st, w.d_flist, w.d_jobs, w.d_obsctr,
                               w.d_obsstat);
          else
            hipLaunchKernelGGL(k_ieskf_step<6>, dim3(n), dim3(kStepThreads), 0, st, w.d_flist, w.d_jobs, w.d_obsctr,
                               w.d_obsstat);
        });
      }
      hipLaunchKernelGGL(k_ieskf_end, dim3(n), dim3(kBeginEndThreads), 0, st, w.d_flist);
    };
    // per-launch timing (vmp_session_set_timing) records events between the
    // launches, so it takes the launch-by-launch path
    hipGraphExec_t exec = (c->use_graph && !c->timing) ? filter_graph(c, n, nblocks, max_iter, ext, filter_stage)
                                                       : nullptr;
    if (exec) {
      if (hipGraphLaunch(exec, st) != hipSuccess) return VMP_EDEVICE;
      c->graph_launches++;
      c->obs_launches += max_iter;
    } else {
      filter_stage(nblocks);
    }
  }
  hipEventRecord(c->ev[1], st);
  hipEventRecord(c->ev[2], st);
  if (mode & kStepMap) {
    if (!(mode & kStepIeskf)) hipLaunchKernelGGL(k_upd_prep, dim3(n), dim3(64), 0, st, w.d_flist);
    BatchBufs b = w.b;
    b.ntot = ntot;
    b.nsess = n;
    const MapDev* M = w.d_maps;
    SessInfo* S = w.d_si;
    const int T = kMapThreads;
    const int gn = grid_for(ntot, T);
    int hbits = 1;
    for (int i = 0; i < n; ++i) hbits = std::max(hbits, int(ss[i]->map.hbits));
    int sbits = 0;
    while ((1 << sbits) < n) ++sbits;
    b.hbits = hbits;
    int tbits = 1;  // point times of the largest scan fit in tbits bits
    for (int i = 0; i < n; ++i)
      while ((int64_t(1) << tbits) < int64_t(ss[i]->n)) ++tbits;
    b.cand_sshift = 25 + tbits;
    // the stage's launches; `lbits` etc. depend only on the step's shape
    int lbits = 1;  // replay order by segment length, capped at 6 bits
    while ((int64_t(1) << lbits) <= int64_t(ntot) && lbits < 6) ++lbits;
    // long segments sort first only below the length clamp (2^lbits - 2)
    const int long_seg = c->replay_long_seg < (1 << lbits) - 1 ? c->replay_long_seg : 0;
    const int wave_max = c->replay_wave_max;
    // every session's counters are the ones its last step returned (a step
    // on this context starts only after the previous one finished), and a
    // scan creates at most one voxel per point: when every session's size
    // plus its points fits its capacity, no session can evict in this scan,
    // all take the fast creation path and the eviction kernels (which would
    // return at once) are left out
    bool all_fast = true;
    for (int i = 0; i < n; ++i)
      all_fast &= int64_t(ss[i]->h_ctr.size) + int64_t(ss[i]->n) <= int64_t(ss[i]->map.capacity);
    // batches up to kPrimMaxItems points use this library's compaction and
    // radix sort (kernels only, so the stage can be captured); larger ones
    // rocPRIM's device-wide routines
    const bool prim = ntot <= kPrimMaxItems;
    auto select_flags = [&](int32_t* flags, int32_t* out, int32_t* count) -> bool {
      if (prim) {
        select_stable(st, ntot, PredFlag{flags}, OutIndex{out}, b.prim, count);
        return true;
      }
      size_t tb = w.temp_bytes;
      return rocprim::select(w.temp, tb, rocprim::counting_iterator<int32_t>(0), flags, out, count, size_t(ntot),
                             st) == hipSuccess;
    };
    auto sort_pairs = [&](auto* kin, int32_t* vin, auto* kout, int32_t* vout, int bits) -> bool {
      if (prim) {
        radix_sort_stable(st, kin, vin, kout, vout, ntot, bits, b.prim);
        return true;
      }
      size_t tb = w.temp_bytes;
      return rocprim::radix_sort_pairs(w.temp, tb, kin, kout, vin, vout, size_t(ntot), 0, bits, st) == hipSuccess;
    };
    auto map_stage = [&]() -> int {
      hipLaunchKernelGGL(k_begin_scan, dim3(n), dim3(1), 0, st, M, n);
      if (ntot > 0) {
        timed_launch(c, kTkInsKeys, [&] { hipLaunchKernelGGL(k_ins_keys, dim3(gn), dim3(T), 0, st, M, S, n, b); });
        // one stable radix sort over (session, hash slot): every key's points
        // become a contiguous segment in point order, sessions stay in order
        if (!sort_pairs(b.skey_in, b.sval_in, b.skey_out, b.sval_out, hbits + sbits)) return VMP_EDEVICE;
        hipLaunchKernelGGL(k_seg_heads, dim3(gn), dim3(T), 0, st, M, S, n, b);
        // the library's compaction also clears the flags it reads, computes
        // the per-session ranges in its last tile and (creation events)
        // splits each event into its time and segment
        if (prim) {
          select_stable(st, ntot, PredFlagClear{b.flags}, OutIndex{b.sel}, b.prim, b.cnt + 0,
                        TailRanges{M, S, n, b.sel, 0, w.d_sfirst, ntot});
        } else {
          if (!select_flags(b.flags, b.sel, b.cnt + 0)) return VMP_EDEVICE;
          hipLaunchKernelGGL(k_ranges, dim3(grid_for(n, 64)), dim3(64), 0, st, M, S, n, b.sel, b.cnt + 0, 0);
          hipLaunchKernelGGL(k_fill_i32, dim3(gn), dim3(T), 0, st, b.flags, 0, ntot);
        }
        hipLaunchKernelGGL(k_seg_info, dim3(gn), dim3(T), 0, st, M, S, n, b);
        if (prim) {
          select_stable(st, ntot, PredFlagClear{b.flags}, OutCre{b.sel, b.cre_t, b.cre_seg, S, n}, b.prim,
                        b.cnt + 1, TailRanges{M, S, n, b.sel, 1, w.d_sfirst, ntot});
        } else {
          if (!select_flags(b.flags, b.sel, b.cnt + 1)) return VMP_EDEVICE;
          hipLaunchKernelGGL(k_ranges, dim3(grid_for(n, 64)), dim3(64), 0, st, M, S, n, b.sel, b.cnt + 1, 1);
          hipLaunchKernelGGL(k_cre_split, dim3(gn), dim3(T), 0, st, S, n, b);
        }
        hipLaunchKernelGGL(k_create_fast, dim3(gn), dim3(T), 0, st, M, S, n, b);
        if (!all_fast) {
          timed_launch(c, kTkRingLive, [&] {
            hipLaunchKernelGGL(k_ring_live, dim3(kLvBlocks, n), dim3(64 * kLvWaves), 0, st, M, S, n, b);
          });
          hipLaunchKernelGGL(k_ring_pack, dim3(n), dim3(kLpThreads), 0, st, M, S, n, b);
          timed_launch(c, kTkCreateEvict,
                       [&] { hipLaunchKernelGGL(k_create_evict, dim3(n), dim3(128), 0, st, M, S, n, b); });
        }
        hipLaunchKernelGGL(k_create_init, dim3(gn), dim3(T), 0, st, M, S, n, b);
        // replay order by segment length in 6 bits (one radix pass; segments
        // of 62 points and more share the first key): the order only spreads
        // the work evenly over waves, every segment replays on its own
        hipLaunchKernelGGL(k_seg_lenkeys, dim3(gn), dim3(T), 0, st, M, S, b, lbits, long_seg);
        if (!sort_pairs(b.len_in, b.ord_in, b.len_out, b.ord, lbits + 1)) return VMP_EDEVICE;
        // the moment path: a voxel per wave up to wave_max refitting voxels
        // (a grid of up to 4096 waves strides over them), else for the long
        // segments only and 64 voxels per wave for the rest; the light path
        // covers the voxels that do not refit
        timed_launch(c, kTkReplayHeavy, [&] {
          hipLaunchKernelGGL(k_replay_heavy_wave, dim3(std::min(gn, 1024)), dim3(T), 0, st, M, S, b, wave_max);
          hipLaunchKernelGGL(k_replay_heavy, dim3(std::min(gn, 256)), dim3(T), 0, st, M, S, b, wave_max);
        });
        timed_launch(c, kTkReplayLight, [&] { hipLaunchKernelGGL(k_replay_light, dim3(gn), dim3(T), 0, st, M, S, b); });
        if (!(mode & kStepBuild)) {
          timed_launch(c, kTkCandidates, [&] { hipLaunchKernelGGL(k_candidates, dim3(gn), dim3(T), 0, st, M, S, b); });
          // the call at point time t is made by the voxel that received point
          // t, so each candidate sits at its session offset plus t: the
          // non-empty slots in order are the list sorted on (session, time)
          if (prim) {
            select_stable(st, ntot, PredCand{b.cand_in}, OutCand{b.cand_in, b.cand_out}, b.prim, b.cnt + 7,
                          TailCandStarts{M, n, b});
          } else {
            size_t tb = w.temp_bytes;
            if (rocprim::radix_sort_keys(w.temp, tb, b.cand_in, b.cand_out, size_t(ntot), 25, b.cand_sshift + sbits,
                                         st) != hipSuccess)
              return VMP_EDEVICE;
            hipLaunchKernelGGL(k_cand_starts, dim3(1), dim3(kCsThreads), 0, st, M, n, b);
          }
          timed_launch(c, kTkSpecEval, [&] { hipLaunchKernelGGL(k_spec_eval, dim3(gn), dim3(T), 0, st, M, b); });
          if (prim) {
            select_stable(st, ntot, PredFlag{b.cand_spec}, OutSpl{b.spl, b.cand_in, b.cand_out, b.cand_sshift},
                          b.prim, b.cnt + 4, TailSplRanges{M, S, n, b});
          } else {
            if (!select_flags(b.cand_spec, b.spl, b.cnt + 4)) return VMP_EDEVICE;
            hipLaunchKernelGGL(k_spl_ranges, dim3(grid_for(n, 64)), dim3(64), 0, st, M, S, n, b);
            hipLaunchKernelGGL(k_spl_keys, dim3(gn), dim3(T), 0, st, b);
          }
          hipLaunchKernelGGL(k_spl_prefetch, dim3(std::min(gn * kPreItems, 4096)), dim3(T), 0, st, M, S, b);
          timed_launch(c, kTkMergeReplay,
                       [&] { hipLaunchKernelGGL(k_merge_replay, dim3(n), dim3(kMrThreads), 0, st, M, S, n, b); });
        }
        // (the library's compaction of the creation events left the flags cleared)
        if (!prim) hipLaunchKernelGGL(k_fill_i32, dim3(gn), dim3(T), 0, st, b.flags, 0, ntot);
        hipLaunchKernelGGL(k_finalize_segs, dim3(gn), dim3(T), 0, st, M, S, b);
        hipLaunchKernelGGL(k_evict_finalize, dim3(gn), dim3(T), 0, st, M, S, n, b);
        if (prim) {
          select_stable(st, ntot, PredFlag{b.flags}, OutIndex{b.sel}, b.prim, b.cnt + 2,
                        TailRanges{M, S, n, b.sel, 2, w.d_sfirst, ntot});
        } else {
          if (!select_flags(b.flags, b.sel, b.cnt + 2)) return VMP_EDEVICE;
          hipLaunchKernelGGL(k_ranges, dim3(grid_for(n, 64)), dim3(64), 0, st, M, S, n, b.sel, b.cnt + 2, 2);
        }
        hipLaunchKernelGGL(k_lru_append, dim3(gn), dim3(T), 0, st, M, S, n, b);
      }
      hipLaunchKernelGGL(k_end_scan, dim3(n), dim3(1), 0, st, M, S, n);
  
      return VMP_OK;
    };
    const int64_t key[3] = {int64_t(ntot), int64_t(n),
                            int64_t(hbits) | (int64_t(sbits) << 8) | (int64_t(tbits) << 16) |
                                (int64_t(lbits) << 24) | (int64_t(all_fast) << 32) |
                                (int64_t(mode & kStepBuild) << 32) |
                                (int64_t(long_seg) << 36) | (int64_t(wave_max) << 42)};
    hipGraphExec_t mexec = (c->use_graph && c->use_map_graph && !c->timing) ? map_graph(c, key, map_stage) : nullptr;
    if (mexec) {
      if (hipGraphLaunch(mexec, st) != hipSuccess) return VMP_EDEVICE;
      c->map_graph_launches++;
    } else {
      const int rc = map_stage();
      if (rc != VMP_OK) return rc;
    }
  }
  hipEventRecord(c->ev[3], st);
  hipLaunchKernelGGL(k_gather, dim3(n), dim3(64), 0, st, w.d_maps, w.d_flist, w.d_obsstat, w.d_out, n,
                     &prim_ws(w.b.prim).ctl->err);
  if (hipMemcpyAsync(w.h_out, w.d_out, sizeof(BatchOut) * n, hipMemcpyDeviceToHost, st) != hipSuccess)
    return VMP_EDEVICE;
  return hipGetLastError() == hipSuccess ? VMP_OK : VMP_EDEVICE;
}

// housekeeping between scans: hash rebuild and LRU ring compaction (rare)
static int maintain(vmp_ctx* ex, vmp_session* s, int32_t* did) {
  *did = 0;
  MapDev& m = s->map;
  const MapCounters& c = s->h_ctr;
  hipStream_t st = ex->stream;
  const int64_t hcap = int64_t(m.hmask) + 1;
  const int T = 256;
  if (int64_t(c.used_entries) + 2 * int64_t(s->cfg.max_scan_points) > hcap / 2) {
    VMP_HIP(hipMemsetAsync(s->nlive, 0, sizeof(int32_t), st));
    hipLaunchKernelGGL(k_rehash_collect, dim3(grid_for(m.pool, T)), dim3(T), 0, st, m, s->live, s->nlive);
    hipLaunchKernelGGL(k_rehash_clear, dim3(grid_for(std::min<int64_t>(hcap, 1 << 20), T)), dim3(T), 0, st, m);
    MapCounters cc = c;
    cc.used_entries = 0;
    cc.tombs = 0;
    VMP_HIP(hipMemcpyAsync(m.ctr, &cc, sizeof(cc), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_rehash_insert, dim3(grid_for(m.pool, T)), dim3(T), 0, st, m, s->live, s->nlive);
    VMP_HIP(hipMemcpyAsync(&s->h_ctr, m.ctr, sizeof(MapCounters), hipMemcpyDeviceToHost, st));
    VMP_HIP(hipStreamSynchronize(st));
    *did |= 1;
  }
  const int64_t used = s->h_ctr.ring_tail - s->h_ctr.ring_head;
  if (used + int64_t(s->cfg.max_scan_points) > m.qcap) {
    RingTmp& t = ex->rt;
    if (t.cap < used && t.fixed) {
      s->err = "LRU ring compaction: the ring exceeds the no-allocation scratch";
      return VMP_ECAPACITY;
    }
    if (t.cap < used) {
      VMP_HIP(hipStreamSynchronize(st));
      t.release();
      const size_t n = size_t(used);
      bool ok = dalloc(t.allocs, &t.tv, n) == hipSuccess && dalloc(t.allocs, &t.ts, n) == hipSuccess &&
                dalloc(t.allocs, &t.keep, n) == hipSuccess && dalloc(t.allocs, &t.sel, n) == hipSuccess &&
                dalloc(t.allocs, &t.cnt, 1) == hipSuccess;
      t.temp_bytes = 0;
      if (ok)
        ok = rocprim::select(nullptr, t.temp_bytes, rocprim::counting_iterator<int32_t>(0), t.keep, t.sel, t.cnt, n,
                             st) == hipSuccess;
      if (ok) ok = dalloc(t.allocs, reinterpret_cast<uint8_t**>(&t.temp), t.temp_bytes + 256) == hipSuccess;
      if (!ok) {
        t.release();
        s->err = "LRU ring compaction: device allocation failed";
        return VMP_ENOMEM;
      }
      t.temp_bytes += 256;
      t.cap = used;
    }
    hipLaunchKernelGGL(k_ring_mark, dim3(grid_for(used, T)), dim3(T), 0, st, m, s->h_ctr.ring_head, used, t.tv, t.ts,
                       t.keep);
    size_t tb = t.temp_bytes;
    VMP_HIP(rocprim::select(t.temp, tb, rocprim::counting_iterator<int32_t>(0), t.keep, t.sel, t.cnt, size_t(used), st));
    hipLaunchKernelGGL(k_ring_scatter, dim3(grid_for(used, T)), dim3(T), 0, st, m, t.tv, t.ts, t.sel, t.cnt);
    int32_t kept = 0;
    VMP_HIP(hipMemcpyAsync(&kept, t.cnt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    VMP_HIP(hipStreamSynchronize(st));
    MapCounters cc = s->h_ctr;
    cc.ring_head = 0;
    cc.ring_tail = int64_t(kept);
    VMP_HIP(hipMemcpyAsync(m.ctr, &cc, sizeof(cc), hipMemcpyHostToDevice, st));
    VMP_HIP(hipStreamSynchronize(st));
    s->h_ctr = cc;
    *did |= 2;
    if (t.fixed) {
      // these arrays share the device scratch with the workspace of
      // vmp_register_scan_noalloc, whose primitives' control words must
      // start from zero at its next step
      VMP_HIP(hipMemsetAsync(ex->ws.b.prim, 0, prim_ws_entries(ex->ws.cap_points) * sizeof(int32_t), st));
    }
  }
  return VMP_OK;
}

static void collect_kernel_timing(vmp_ctx* c) {
  for (int k = 0; k < c->kev_used; ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->kev[2 * k], c->kev[2 * k + 1]) == hipSuccess) {
      c->acc_k_ms[c->kev_id[k]] += ms;
      c->acc_k_n[c->kev_id[k]] += 1;
    }
  }
  c->kev_used = 0;
}

static void collect_timing(vmp_ctx* c, int active_max) {
  for (int k = 0; k + 1 < c->tev_used && k / 2 < active_max; k += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->tev[k], c->tev[k + 1]) == hipSuccess && ms > 0.f) {
      c->acc_obs_ms += ms / c->obs_time_reps;
      c->acc_obs_n += 1;
    }
  }
  c->tev_used = 0;
}

// run one step for a batch on the stream and workspace of context `c` (the
// sessions may belong to other contexts of the same device; the caller makes
// sure no session is in two batches at once) and distribute the results
static int step_finish(vmp_ctx* c, vmp_obs_stats* ost, vmp_map_stats* mst);

// the deferred check of device buffers bound since the last step (c->mu
// held, nothing in flight): a session with a non-finite coordinate or time,
// or times out of order, fails the call before anything is launched, and its
// binding is dropped
static int check_bound_points(vmp_ctx* c, vmp_session** ss, int n) {
  // tables on the stack for small batches (the no-allocation entry point
  // passes one session), on the heap beyond
  PtsRef refs_small[16];
  int who_small[16];
  std::vector<PtsRef> refs_big;
  std::vector<int> who_big;
  PtsRef* refs = refs_small;
  int* who = who_small;
  if (n > 16) {
    refs_big.resize(size_t(n));
    who_big.resize(size_t(n));
    refs = refs_big.data();
    who = who_big.data();
  }
  int nr = 0;
  for (int i = 0; i < n; ++i) {
    vmp_session* q = ss[i];
    if (!q->needs_check) continue;
    if (q->raw_pending) refs[nr] = PtsRef{q->rx, q->ry, q->rz, q->rt, q->raw_n, 0};
    else refs[nr] = PtsRef{q->cx, q->cy, q->cz, nullptr, q->n, 0};
    who[nr++] = i;
  }
  if (nr == 0) return VMP_OK;
  const int bad = check_device_points(c, refs, nr);
  if (bad == -2) {
    ss[0]->err = "point validation failed on the device";
    return VMP_EDEVICE;
  }
  if (bad >= 0) {
    vmp_session* q = ss[who[bad]];
    q->err = "scan holds a non-finite coordinate or time, or point times out of order";
    q->has_points = false;
    q->raw_pending = false;
    q->needs_check = false;
    return VMP_EINVAL;
  }
  for (int k = 0; k < nr; ++k) ss[who[k]]->needs_check = false;
  return VMP_OK;
}

// undistortCloud of the sessions that carry a raw package (IMU propagation of
// x and P, then the points compensated into the sweep's end frame) and, for
// registered scans with scan_resolution > 0, the VoxelGrid filter, whose
// output counts are read back before the step is sized (c->mu held, the
// workspace reserved for the raw points)
static int prepare_scans(vmp_ctx* c, vmp_session** ss, int n, bool is_build) {
  BatchWs& w = c->ws;
  hipStream_t st = c->stream;
  int npre = 0;
  int64_t most = 0;
  for (int i = 0; i < n; ++i) {
    vmp_session* q = ss[i];
    if (!q->raw_pending) continue;
    w.h_pre[npre++] = PreJob{q->imu, q->f, q->rx, q->ry, q->rz, q->rt, q->own_x, q->own_y, q->own_z, q->raw_n, 0};
    most = std::max<int64_t>(most, q->raw_n);
  }
  if (npre == 0) return VMP_OK;
  if (hipMemcpyAsync(w.d_pre, w.h_pre, sizeof(PreJob) * npre, hipMemcpyHostToDevice, st) != hipSuccess)
    return VMP_EDEVICE;
  timed_launch(c, kTkImuPredict, [&] { hipLaunchKernelGGL(k_imu_predict, dim3(npre), dim3(kPredThreads), 0, st, w.d_pre, npre); });
  const int gx = int(std::max<int64_t>(1, std::min<int64_t>((most + 255) / 256, 1024)));
  timed_launch(c, kTkUndistort,
               [&] { hipLaunchKernelGGL(k_undistort, dim3(gx, std::min(npre, 65535)), dim3(256), 0, st, w.d_pre, npre); });
  // the filter (never at MAP_INIT, lio_builder.cpp:185-211); the filtered
  // sessions listed on the stack for small batches, on the heap beyond
  vmp_session* fs_small[16];
  std::vector<vmp_session*> fs_big;
  vmp_session** fs = fs_small;
  if (n > 16) {
    fs_big.resize(size_t(n));
    fs = fs_big.data();
  }
  int nf = 0;
  int64_t off = 0;
  most = 0;
  for (int i = 0; i < n && !is_build; ++i) {
    vmp_session* q = ss[i];
    if (!q->raw_pending || !(q->cfg.scan_resolution > 0.0) || !q->own_fx) continue;
    const float leaf = float(q->cfg.scan_resolution);
    w.h_vg[nf] = VgJob{q->own_x, q->own_y, q->own_z, q->own_fx, q->own_fy, q->own_fz, q->raw_n, int32_t(off),
                       1.0f / leaf, 0};
    off += q->raw_n;
    most = std::max<int64_t>(most, q->raw_n);
    fs[nf++] = q;
  }
  for (int i = 0; i < n; ++i)
    if (ss[i]->raw_pending) {
      ss[i]->raw_pending = false;
      ss[i]->cloud_n = ss[i]->raw_n;
    }
  if (nf == 0) return VMP_OK;
  BatchBufs& b = w.b;
  const int ntot = int(off);
  if (hipMemcpyAsync(w.d_vg, w.h_vg, sizeof(VgJob) * nf, hipMemcpyHostToDevice, st) != hipSuccess) return VMP_EDEVICE;
  const int gy = std::min(nf, 65535);
  const int gxp = int(std::max<int64_t>(1, std::min<int64_t>((most + 255) / 256, 256)));
  hipLaunchKernelGGL(k_vg_init, dim3((nf + 63) / 64), dim3(64), 0, st, w.d_vgst, nf);
  hipLaunchKernelGGL(k_vg_minmax, dim3(gxp, gy), dim3(256), 0, st, w.d_vg, nf, w.d_vgst);
  hipLaunchKernelGGL(k_vg_keys, dim3(gxp, gy), dim3(256), 0, st, w.d_vg, nf, w.d_vgst, b.skey_in, b.sval_in);
  int sbits = 0;
  while ((1 << sbits) < nf) ++sbits;
  size_t tb = w.temp_bytes;
  if (ntot > 0) {
    const bool prim = ntot <= kPrimMaxItems;
    if (prim) {
      radix_sort_stable(st, b.skey_in, b.sval_in, b.skey_out, b.sval_out, ntot, 32 + sbits, b.prim);
    } else if (rocprim::radix_sort_pairs(w.temp, tb, b.skey_in, b.skey_out, b.sval_in, b.sval_out, size_t(ntot), 0,
                                         32 + sbits, st) != hipSuccess) {
      return VMP_EDEVICE;
    }
    const int gn = grid_for(ntot, 256);
    hipLaunchKernelGGL(k_vg_heads, dim3(gn), dim3(256), 0, st, b.skey_out, ntot, b.flags);
    tb = w.temp_bytes;
    if (prim) {
      select_stable(st, ntot, PredFlag{b.flags}, OutIndex{b.sel}, b.prim, b.cnt + 6);
    } else if (rocprim::select(w.temp, tb, rocprim::counting_iterator<int32_t>(0), b.flags, b.sel, b.cnt + 6,
                               size_t(ntot), st) != hipSuccess) {
      return VMP_EDEVICE;
    }
    hipLaunchKernelGGL(k_vg_ranges, dim3((nf + 63) / 64), dim3(64), 0, st, w.d_vg, nf, b.sel, b.cnt + 6, w.d_vgst);
    hipLaunchKernelGGL(k_vg_centroid, dim3(gn), dim3(256), 0, st, w.d_vg, b.skey_out, b.sval_out, b.sel, b.cnt + 6,
                       ntot, w.d_vgst);
  }
  if (hipMemcpyAsync(w.h_vgst, w.d_vgst, sizeof(int32_t) * kVgStat * nf, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return VMP_EDEVICE;
  for (int k = 0; k < nf; ++k) {
    vmp_session* q = fs[k];
    q->n = ntot > 0 ? w.h_vgst[k * kVgStat + 8] : 0;
    q->cx = q->own_fx;
    q->cy = q->own_fy;
    q->cz = q->own_fz;
  }
  return VMP_OK;
}

// validate and launch a batch on the stream and workspace of context `c` (the
// sessions may belong to other contexts of the same device; the caller makes
// sure no session is in two batches at once); c->mu held, no batch in flight
static int step_begin(vmp_ctx* c, vmp_session** ss, int n, int mode) {
  vmp_session* s = ss[0];
  if (n > kMaxBatchSessions) {
    s->err = "more than 32768 sessions in one batch";
    return VMP_EINVAL;
  }
  for (int i = 0; i < n; ++i) {
    if (!ss[i] || ss[i]->ctx->device != c->device) return VMP_EINVAL;
    if (!ss[i]->has_points) {
      ss[i]->err = "registration call before vmp_scan_set_points";
      return VMP_ESTATE;
    }
    if (ss[i]->cfg.estimate_ext != s->cfg.estimate_ext) {
      ss[i]->err = "a batch must share estimate_ext";
      return VMP_EINVAL;
    }
  }
  VMP_HIP(hipSetDevice(c->device));
  int rc = check_bound_points(c, ss, n);
  if (rc != VMP_OK) return rc;
  int64_t ntot = 0;
  for (int i = 0; i < n; ++i) ntot += ss[i]->n;
  if (ntot > (int64_t(1) << 30)) return VMP_EINVAL;
  rc = ws_reserve(c, std::max<int64_t>(ntot, 1), n);
  if (rc != VMP_OK) {
    s->err = "batch workspace allocation failed";
    return rc;
  }
  if (mode & kStepBuild) {
    for (int i = 0; i < n; ++i) {
      VMP_HIP(hipMemcpyAsync(&ss[i]->h_ctr, ss[i]->map.ctr, sizeof(MapCounters), hipMemcpyDeviceToHost, c->stream));
    }
    VMP_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i)
      if (ss[i]->h_ctr.size != 0) {
        ss[i]->err = "build (MAP_INIT) requires an empty map";
        return VMP_ESTATE;
      }
  }
  rc = prepare_scans(c, ss, n, (mode & kStepBuild) != 0);
  if (rc != VMP_OK) {
    s->err = "undistortion / scan filter failed";
    return rc;
  }
  c->obs_launches = 0;
  rc = enqueue_step(c, ss, n, mode);
  if (rc != VMP_OK) {
    s->err = "launch failed";
    return rc;
  }
  c->fl = true;
  c->fl_mode = mode;
  c->fl_ss.assign(ss, ss + n);
  c->fl_n.resize(size_t(n));
  for (int i = 0; i < n; ++i) c->fl_n[size_t(i)] = ss[i]->n;
  return VMP_OK;
}

// wait for the batch in flight on c, distribute its results (ost / mst may be
// null) and run the sessions' housekeeping; c->mu held
static int step_finish(vmp_ctx* c, vmp_obs_stats* ost, vmp_map_stats* mst) {
  if (!c->fl) return VMP_OK;
  c->fl = false;
  vmp_session** ss = c->fl_ss.data();
  const int n = int(c->fl_ss.size());
  const int mode = c->fl_mode;
  vmp_session* s = ss[0];
  int rc = VMP_OK;
  VMP_HIP(stream_wait(c));
  int max_iters = 0;
  int64_t pts_iters = 0, hits = 0;
  for (int i = 0; i < n; ++i) {
    const BatchOut& o = c->ws.h_out[i];
    max_iters = std::max(max_iters, o.iters);
    pts_iters += int64_t(c->fl_n[size_t(i)]) * o.iters;
    hits += o.obs[4];
  }
  if (c->timing) {
    collect_timing(c, (mode & kStepIeskf) ? max_iters : 0);
    collect_kernel_timing(c);
    c->acc_obs_points += pts_iters;
    c->acc_obs_hits += hits;
    if (mode & kStepMap) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) {
        c->acc_map_ms += ms;
        c->acc_map_n += 1;
      }
    }
  }
  int first_err = VMP_OK;
  for (int i = 0; i < n; ++i) {
    vmp_session* si = ss[i];
    const BatchOut& o = c->ws.h_out[i];
    si->h_ctr = o.ctr;
    if ((mode & kStepIeskf) && o.obs[7]) {
      si->err = "hash probe overflow";
      if (first_err == VMP_OK) first_err = VMP_EDEVICE;
    }
    if (ost) {
      ost[i].points = c->fl_n[size_t(i)];
      ost[i].hits = o.obs[0];
      ost[i].plane_hits = o.obs[1];
      ost[i].effective = o.obs[2];
      ost[i].key_fallback = o.obs[3];
      ost[i].iterations = o.iters;
    }
    if (mode & kStepMap) {
      const MapCounters& m = o.ctr;
      if (m.error) {
        char buf[200];
        std::snprintf(buf, sizeof(buf),
                      "map update error 0x%x (1 hash probe, 2 key range, 4 LRU exhausted in one scan, 8 voxel "
                      "pool, 16 replay queue, 32 replay class, 64 compaction / sort wait)",
                      unsigned(m.error));
        si->err = buf;
        if (first_err == VMP_OK) first_err = (m.error & 2) ? VMP_ERANGE : ((m.error & 4) ? VMP_ECAPACITY : VMP_EDEVICE);
        continue;
      }
      if (mst) {
        mst[i].points = c->fl_n[size_t(i)];
        mst[i].touched = m.nseg;
        mst[i].created = m.ncre;
        mst[i].evicted = m.nevict;
        mst[i].refits = m.refits;
        mst[i].merge_calls = m.merge_calls;
        mst[i].merge_evals = m.merge_evals;
        mst[i].merge_windows = m.merge_windows;
        mst[i].merge_requeued = m.merge_requeued;
        mst[i].merge_max_acting = m.merge_max_acting;
        mst[i].merges = m.merges;
        mst[i].dropped = m.dropped;
        mst[i].map_size = int32_t(m.size);
      }
      int32_t did = 0;
      rc = maintain(c, si, &did);
      if (mst) mst[i].maintenance = did;
      if (rc != VMP_OK && first_err == VMP_OK) first_err = rc;
    }
  }
  return first_err;
}

// complete a submitted batch, keeping its results for vmp_register_batch_wait
// (c->mu held)
static int settle(vmp_ctx* c) {
  if (!c->fl) return VMP_OK;
  const size_t n = c->fl_ss.size();
  c->res_ost.assign(n, vmp_obs_stats{});
  c->res_mst.assign(n, vmp_map_stats{});
  c->res_rc = step_finish(c, c->res_ost.data(), c->res_mst.data());
  c->res_valid = true;
  return c->res_rc;
}

// run one batch step on context c and distribute the results
static int run_step_on(vmp_ctx* c, vmp_session** ss, int n, int mode, vmp_obs_stats* ost, vmp_map_stats* mst) {
  std::lock_guard<std::mutex> lk(c->mu);
  int rc = settle(c);  // a submitted batch completes first
  if (rc != VMP_OK) return rc;
  rc = step_begin(c, ss, n, mode);
  if (rc != VMP_OK) return rc;
  return step_finish(c, ost, mst);
}

static int run_step(vmp_session** ss, int n, int mode, vmp_obs_stats* ost, vmp_map_stats* mst) {
  if (!ss || n <= 0 || !ss[0]) return VMP_EINVAL;
  vmp_ctx* c = ss[0]->ctx;
  for (int i = 0; i < n; ++i)
    if (!ss[i] || ss[i]->ctx != c) return VMP_EINVAL;
  return run_step_on(c, ss, n, mode, ost, mst);
}

extern "C" int vmp_observe(vmp_session* s, const vmp_state* x, double* H, double* b, vmp_obs_stats* st) {
  if (!s || !x) return VMP_EINVAL;
  if (!s->has_points) {
    s->err = "vmp_observe before vmp_scan_set_points";
    return VMP_ESTATE;
  }
  vmp_ctx* c = s->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  VMP_HIP(hipSetDevice(c->device));
  settle(c);
  vmp_session* one[1] = {s};
  const int crc = check_bound_points(c, one, 1);
  if (crc != VMP_OK) return crc;
  if (ws_reserve(c, std::max(s->n, 1), 1) != VMP_OK) return VMP_ENOMEM;
  VMP_HIP(hipMemcpyAsync(&s->f->x, x, sizeof(vmp_state), hipMemcpyHostToDevice, c->stream));
  if (prepare_scans(c, one, 1, false) != VMP_OK) return VMP_EDEVICE;
  if (enqueue_step(c, one, 1, kStepObserveOnce) != VMP_OK) return VMP_EDEVICE;
  double out[156];
  int32_t ctr[8];
  VMP_HIP(hipMemcpyAsync(out, s->f->obs, sizeof(out), hipMemcpyDeviceToHost, c->stream));
  VMP_HIP(hipMemcpyAsync(ctr, s->obs_ctr, sizeof(ctr), hipMemcpyDeviceToHost, c->stream));
  VMP_HIP(hipStreamSynchronize(c->stream));
  if (ctr[4]) {
    s->err = "hash probe overflow";
    return VMP_EDEVICE;
  }
  if (H) std::memcpy(H, out, sizeof(double) * 144);
  if (b) std::memcpy(b, out + 144, sizeof(double) * 12);
  if (st) {
    st->points = s->n;
    st->hits = ctr[0];
    st->plane_hits = ctr[1];
    st->effective = ctr[2];
    st->key_fallback = ctr[3];
    st->iterations = 1;
  }
  return VMP_OK;
}

// Observe-pass autotuner.  For each size bucket (< 8k, < 32k, < 128k, larger
// points) that holds one of the sessions, in turn, every candidate (1, 2, 4,
// 8 points per thread) is timed on the sessions' current scans and states
// as one observe launch over all of them (HIP events around the observe
// kernel only, median of reps after one untimed launch), and the fastest is
// kept in the context's table before the next bucket is tuned.  The
// residual slots the passes write are restored, so the sessions' next
// registrations are the same as without tuning.  c->mu held.
static int autotune_locked(vmp_ctx* c, vmp_session** ss, int n, int reps, int32_t* table, float* cand_ms) {
  if (hipSetDevice(c->device) != hipSuccess) return VMP_EDEVICE;
  for (int i = 0; i < n; ++i)
    if (ss[i]->raw_pending) {  // its timing passes would undistort the package again and again
      ss[i]->err = "vmp_autotune_observe with a raw package pending";
      return VMP_ESTATE;
    }
  settle(c);
  {
    const int crc = check_bound_points(c, ss, n);
    if (crc != VMP_OK) return crc;
  }
  int64_t ntot = 0;
  for (int i = 0; i < n; ++i) ntot += ss[i]->n;
  if (ws_reserve(c, std::max<int64_t>(ntot, 1), n) != VMP_OK) return VMP_ENOMEM;
  hipStream_t st = c->stream;
  float4* saved = nullptr;
  if (hipMalloc(&saved, sizeof(float4) * size_t(std::max<int64_t>(ntot, 1))) != hipSuccess) return VMP_ENOMEM;
  int rc = VMP_OK;
  int64_t o = 0;
  for (int i = 0; i < n && rc == VMP_OK; ++i) {
    if (hipMemcpyAsync(saved + o, ss[i]->slots, sizeof(float4) * size_t(ss[i]->n), hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      rc = VMP_EDEVICE;
    o += ss[i]->n;
  }
  if (rc == VMP_OK && !c->tev[0])
    for (hipEvent_t& e : c->tev)
      if (hipEventCreate(&e) != hipSuccess) rc = VMP_EDEVICE;
  const bool keep_timing = c->timing;
  const int keep_ppt = c->obs_ppt, keep_reps = c->obs_time_reps;
  c->timing = true;
  c->obs_ppt = 0;
  c->obs_time_reps = 1;
  const int cands[4] = {1, 2, 4, 8};
  for (int bk = 0; bk < 4 && rc == VMP_OK; ++bk) {
    bool present = false;
    for (int i = 0; i < n; ++i) present |= ppt_bucket(ss[i]->n) == bk;
    if (!present) continue;
    float med[4] = {0, 0, 0, 0};
    for (int ci = 0; ci < 4 && rc == VMP_OK; ++ci) {
      c->ppt_tab[bk] = cands[ci];
      std::vector<float> ms;
      for (int r = 0; r <= reps && rc == VMP_OK; ++r) {
        c->tev_used = 0;
        if (enqueue_step(c, ss, n, kStepObserveOnce) != VMP_OK || hipStreamSynchronize(st) != hipSuccess) {
          rc = VMP_EDEVICE;
          break;
        }
        float t = 0.f;
        if (r > 0 && c->tev_used >= 2 && hipEventElapsedTime(&t, c->tev[0], c->tev[1]) == hipSuccess) ms.push_back(t);
      }
      if (!ms.empty()) {
        std::sort(ms.begin(), ms.end());
        med[ci] = ms[ms.size() / 2];
      }
    }
    int best = 0;
    for (int ci = 1; ci < 4; ++ci)
      if (med[ci] > 0.f && (med[best] <= 0.f || med[ci] < med[best])) best = ci;
    c->ppt_tab[bk] = cands[best];
    if (cand_ms)
      for (int ci = 0; ci < 4; ++ci) cand_ms[bk * 4 + ci] = med[ci];
  }
  c->tev_used = 0;
  c->timing = keep_timing;
  c->obs_ppt = keep_ppt;
  c->obs_time_reps = keep_reps;
  o = 0;
  for (int i = 0; i < n; ++i) {
    if (hipMemcpyAsync(ss[i]->slots, saved + o, sizeof(float4) * size_t(ss[i]->n), hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      rc = VMP_EDEVICE;
    o += ss[i]->n;
  }
  if (hipStreamSynchronize(st) != hipSuccess) rc = VMP_EDEVICE;
  hipFree(saved);
  if (table)
    for (int k = 0; k < 4; ++k) table[k] = c->ppt_tab[k];
  return rc;
}

extern "C" int vmp_autotune_observe_batch(vmp_session** ss, int32_t n, int32_t reps, int32_t* table, float* cand_ms) {
  if (!ss || n < 1 || reps < 1 || reps > 1000 || !ss[0]) return VMP_EINVAL;
  vmp_ctx* c = ss[0]->ctx;
  for (int i = 0; i < n; ++i) {
    if (!ss[i] || ss[i]->ctx != c) return VMP_EINVAL;
    if (!ss[i]->has_points || ss[i]->n < 1) {
      ss[i]->err = "vmp_autotune_observe before vmp_scan_set_points";
      return VMP_ESTATE;
    }
  }
  std::lock_guard<std::mutex> lk(c->mu);
  return autotune_locked(c, ss, n, reps, table, cand_ms);
}

extern "C" int vmp_autotune_observe(vmp_session* s, int32_t reps, int32_t* best_ppt, float* cand_ms) {
  if (!s || reps < 1 || reps > 1000) return VMP_EINVAL;
  if (!s->has_points || s->n < 1) {
    s->err = "vmp_autotune_observe before vmp_scan_set_points";
    return VMP_ESTATE;
  }
  vmp_ctx* c = s->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  vmp_session* one[1] = {s};
  int32_t table[4];
  float all[16];
  const int rc = autotune_locked(c, one, 1, reps, table, all);
  if (rc != VMP_OK) return rc;
  const int bk = ppt_bucket(s->n);
  if (best_ppt) *best_ppt = table[bk];
  if (cand_ms)
    for (int ci = 0; ci < 4; ++ci) cand_ms[ci] = all[bk * 4 + ci];
  return VMP_OK;
}

extern "C" int vmp_ieskf_update(vmp_session* s, vmp_obs_stats* st) {
  if (!s) return VMP_EINVAL;
  vmp_session* one[1] = {s};
  return run_step(one, 1, kStepIeskf, st, nullptr);
}

extern "C" int vmp_map_update(vmp_session* s, int is_build, vmp_map_stats* st) {
  if (!s) return VMP_EINVAL;
  vmp_session* one[1] = {s};
  return run_step(one, 1, kStepMap | (is_build ? kStepBuild : 0), nullptr, st);
}

extern "C" int vmp_register_scan(vmp_session* s, vmp_obs_stats* ost, vmp_map_stats* mst) {
  if (!s) return VMP_EINVAL;
  vmp_session* one[1] = {s};
  return run_step(one, 1, kStepIeskf | kStepMap, ost, mst);
}

// vmp_register_scan with the context's workspace and ring scratch swapped for
// the ones carved from the device scratch, the filter stage launched kernel
// by kernel (instantiating a graph allocates), and every size checked
// against the scratch before anything is launched.
extern "C" int vmp_register_scan_noalloc(vmp_session* s, vmp_obs_stats* ost, vmp_map_stats* mst) {
  if (!s) return VMP_EINVAL;
  vmp_ctx* c = s->ctx;
  if (!c->na_ready) {
    s->err = "no-allocation scratch unavailable on this device (set-up at context creation failed)";
    return VMP_EDEVICE;
  }
  const int64_t np = s->raw_pending ? int64_t(s->raw_n) : int64_t(s->n);
  if (np > kNoAllocMaxPoints) {
    s->err = "scan larger than the no-allocation limit of 131072 points";
    return VMP_EINVAL;
  }
  if (s->map.qcap > c->rt_na.cap) {
    s->err = "LRU ring (map_capacity, max_scan_points) larger than the no-allocation scratch holds";
    return VMP_EINVAL;
  }
  std::lock_guard<std::mutex> lk(c->mu);
  int rc = settle(c);  // a submitted batch completes first (on the regular workspace)
  if (rc != VMP_OK) return rc;
  struct Swap {
    vmp_ctx* c;
    bool graph;
    explicit Swap(vmp_ctx* cc) : c(cc), graph(cc->use_graph) {
      std::swap(c->ws, c->ws_na);
      std::swap(c->rt, c->rt_na);
      c->use_graph = false;
    }
    ~Swap() {
      std::swap(c->ws, c->ws_na);
      std::swap(c->rt, c->rt_na);
      c->use_graph = graph;
    }
  } swap(c);
  vmp_session* one[1] = {s};
  rc = step_begin(c, one, 1, kStepIeskf | kStepMap);
  if (rc != VMP_OK) return rc;
  return step_finish(c, ost, mst);
}

extern "C" int vmp_register_batch_submit(vmp_session** ss, int32_t n) {
  if (!ss || n <= 0 || !ss[0]) return VMP_EINVAL;
  vmp_ctx* c = ss[0]->ctx;
  for (int i = 0; i < n; ++i)
    if (!ss[i] || ss[i]->ctx != c) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  const int prev = settle(c);
  const int rc = step_begin(c, ss, n, kStepIeskf | kStepMap);
  return rc != VMP_OK ? rc : prev;
}

extern "C" int vmp_register_batch_wait(vmp_ctx* c, int32_t n, vmp_obs_stats* ost, vmp_map_stats* mst) {
  if (!c) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  settle(c);
  // only a submitted batch leaves results here, and they are handed out once:
  // a wait with nothing submitted since the last one reports that instead of
  // returning an older batch's statistics
  if (!c->res_valid) return VMP_ESTATE;
  c->res_valid = false;
  const int m = std::min<int>(n, int(c->res_mst.size()));
  for (int i = 0; i < m; ++i) {
    if (ost) ost[i] = c->res_ost[size_t(i)];
    if (mst) mst[i] = c->res_mst[size_t(i)];
  }
  return c->res_rc;
}

extern "C" int vmp_register_batch(vmp_session** ss, int32_t n, vmp_obs_stats* ost, vmp_map_stats* mst) {
  if (!ss || n < 0) return VMP_EINVAL;
  if (n == 0) return VMP_OK;
  return run_step(ss, n, kStepIeskf | kStepMap, ost, mst);
}

extern "C" int vmp_session_timing(vmp_session* s, double* observe_ms, double* map_ms, int32_t* observe_launches) {
  if (!s) return VMP_EINVAL;
  vmp_ctx* c = s->ctx;
  float a = 0.f, b = 0.f;
  if (observe_ms) {
    if (hipEventElapsedTime(&a, c->ev[0], c->ev[1]) != hipSuccess) a = -1.f;
    *observe_ms = a;
  }
  if (map_ms) {
    if (hipEventElapsedTime(&b, c->ev[2], c->ev[3]) != hipSuccess) b = -1.f;
    *map_ms = b;
  }
  if (observe_launches) *observe_launches = c->obs_launches;
  return VMP_OK;
}

extern "C" int vmp_ctx_reserve(vmp_ctx* c, int64_t points, int32_t sessions) {
  if (!c || points < 1 || points > (int64_t(1) << 30) || sessions < 1) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return VMP_EDEVICE;
  return ws_reserve(c, points, sessions, true);
}

extern "C" int vmp_ctx_set_param(vmp_ctx* c, const char* name, int64_t value) {
  if (!c || !name) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  if (std::strcmp(name, "obs_ppt") == 0) {
    if (value < 0 || value > 64) return VMP_EINVAL;
    c->obs_ppt = int(value);
    return VMP_OK;
  }
  if (std::strncmp(name, "obs_ppt_bucket", 14) == 0 && name[14] >= '0' && name[14] <= '3' && name[15] == 0) {
    if (value < 1 || value > 64) return VMP_EINVAL;
    c->ppt_tab[name[14] - '0'] = int(value);
    return VMP_OK;
  }
  if (std::strcmp(name, "filter_graph") == 0) {
    if (value < 0 || value > 1) return VMP_EINVAL;
    c->use_graph = value != 0;
    return VMP_OK;
  }
  if (std::strcmp(name, "map_graph") == 0) {
    if (value < 0 || value > 1) return VMP_EINVAL;
    c->use_map_graph = value != 0;
    return VMP_OK;
  }
  if (std::strcmp(name, "obs_arith") == 0) {
    if (value < 0 || value > 1) return VMP_EINVAL;
    c->obs_arith = int(value);
    return VMP_OK;
  }
  if (std::strcmp(name, "check_points") == 0) {
    if (value < 0 || value > 1) return VMP_EINVAL;
    c->check_points = value != 0;
    return VMP_OK;
  }
  if (std::strcmp(name, "obs_time_reps") == 0) {
    if (value < 1 || value > 64) return VMP_EINVAL;
    c->obs_time_reps = int(value);
    return VMP_OK;
  }
  if (std::strcmp(name, "replay_wave_max") == 0) {
    if (value < 0 || value > (int64_t(1) << 20)) return VMP_EINVAL;
    c->replay_wave_max = int(value);
    return VMP_OK;
  }
  if (std::strcmp(name, "replay_long_seg") == 0) {
    if (value != 0 && (value < 2 || value > 62)) return VMP_EINVAL;
    c->replay_long_seg = int(value);
    return VMP_OK;
  }
  return VMP_EINVAL;
}

extern "C" int vmp_ctx_graph_stats(vmp_ctx* c, int64_t* out) {
  if (!c || !out) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  out[0] = c->graph_launches;
  out[1] = c->graph_builds;
  out[2] = c->map_graph_launches;
  out[3] = c->map_graph_builds;
  return VMP_OK;
}

extern "C" int vmp_session_set_timing(vmp_session* s, int on) {
  if (!s) return VMP_EINVAL;
  vmp_ctx* c = s->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  VMP_HIP(hipSetDevice(c->device));
  if (on && !c->tev[0])
    for (hipEvent_t& e : c->tev) VMP_HIP(hipEventCreate(&e));
  if (on && !c->kev[0])
    for (hipEvent_t& e : c->kev) VMP_HIP(hipEventCreate(&e));
  c->kev_used = 0;
  for (int k = 0; k < 16; ++k) {
    c->acc_k_ms[k] = 0.0;
    c->acc_k_n[k] = 0;
  }
  c->timing = on != 0;
  c->tev_used = 0;
  c->acc_obs_ms = c->acc_map_ms = 0.0;
  c->acc_obs_n = c->acc_map_n = 0;
  c->acc_obs_points = c->acc_obs_hits = 0;
  return VMP_OK;
}

extern "C" int vmp_session_timing_totals(vmp_session* s, double* obs_ms, int64_t* obs_launches, double* map_ms,
                                         int64_t* map_calls, int64_t* obs_points, int64_t* obs_hits) {
  if (!s) return VMP_EINVAL;
  const vmp_ctx* c = s->ctx;
  if (obs_ms) *obs_ms = c->acc_obs_ms;
  if (obs_launches) *obs_launches = c->acc_obs_n;
  if (map_ms) *map_ms = c->acc_map_ms;
  if (map_calls) *map_calls = c->acc_map_n;
  if (obs_points) *obs_points = c->acc_obs_points;
  if (obs_hits) *obs_hits = c->acc_obs_hits;
  return VMP_OK;
}

extern "C" int32_t vmp_ctx_kernel_times(vmp_ctx* c, int32_t cap, double* total_ms, int64_t* launches) {
  if (!c || cap < 0) return VMP_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  for (int k = 0; k < kTkCount && k < cap; ++k) {
    if (total_ms) total_ms[k] = c->acc_k_ms[k];
    if (launches) launches[k] = c->acc_k_n[k];
  }
  return kTkCount;
}

extern "C" int64_t vmp_map_size(vmp_session* s) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->ctx->mu);
  MapCounters c;
  if (hipMemcpy(&c, s->map.ctr, sizeof(MapCounters), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return c.size;
}

extern "C" int64_t vmp_map_export(vmp_session* s, int64_t cap, int64_t* keys, int32_t* ints, uint64_t* group,
                                  double* dbl) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->ctx->mu);
  if (hipSetDevice(s->ctx->device) != hipSuccess) return -1;
  const MapDev& m = s->map;
  std::vector<VoxState> vs(static_cast<size_t>(m.pool));
  std::vector<double> pl(static_cast<size_t>(m.pool) * kPlaneD);
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
  std::lock_guard<std::mutex> lk(s->ctx->mu);
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

// ------------------------------------------------------------ serving -----
#include "vmp_serve.hpp"
