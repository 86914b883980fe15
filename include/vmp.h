/*
 * vmp.h — C ABI of the MI355X registration path for voxel_plus
 * (VoxelMapPlus + FAST-LIO2 IESKF).  One context per process/GPU, one session
 * per LiDAR sequence.  A session owns, in HBM: the voxel hash map with its LRU
 * order, the per-point residual slots, and the filter state (x, P).
 *
 * Conventions
 *   - every 3x3 / 12x12 / 23x23 matrix is column-major (Eigen's default), so
 *     an Eigen object's .data() can be passed directly;
 *   - status codes: 0 ok, < 0 error (see VMP_E*); vmp_last_error() explains;
 *   - host pointers are only read/written during the call;
 *   - calls on one session must not overlap; distinct sessions may be driven
 *     from distinct host threads.
 *
 * Reference entry points each function replaces (paths relative to the
 * reference root, voxel_plus/src/map_builder/):
 *   vmp_scan_set_points   residual-slot prep, lio_builder.cpp:225-229
 *   vmp_observe           kf::measure_func = LIOBuilder::sharedUpdateFunc,
 *                         ieskf.h:74, lio_builder.cpp:250-311
 *   vmp_map_update        VoxelMap::update / build, voxel_map.h:114-116,
 *                         voxel_map.cpp:200-256, with the insertion points and
 *                         covariances of lio_builder.cpp:155-163,185-211,231-246
 *   vmp_ieskf_update      kf::IESKF::update, ieskf.cpp:125-156, with
 *                         vmp_observe as its measurement function
 *   vmp_register_scan     lio_builder.cpp:224-246 (LIO_MAPPING branch after
 *                         undistortion): set points, update, map update
 */
#ifndef VMP_H
#define VMP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VMP_OK = 0,
  VMP_EINVAL = -1,     /* bad argument / size beyond the session's limits */
  VMP_ENOMEM = -2,     /* device allocation failed */
  VMP_EDEVICE = -3,    /* HIP runtime error */
  VMP_ERANGE = -4,     /* voxel key outside +-2^20 voxels per axis */
  VMP_ECAPACITY = -5,  /* one scan touches more voxels than the LRU can evict
                          in scan order (map_capacity too small for the scan) */
  VMP_ESTATE = -6      /* call sequence error (e.g. observe before points) */
};

/* LIOConfig subset on the path (lio_builder.h:17-42) plus device limits. */
typedef struct {
  double voxel_size;                /* 0.5 */
  double plane_thresh;              /* 0.01 */
  double merge_thresh_for_angle;    /* 0.2 (mapping.yaml:16) */
  double merge_thresh_for_distance; /* 0.04 */
  double ranging_cov;               /* 0.04 */
  double angle_cov;                 /* 0.1 (degrees) */
  int32_t max_point_thresh;         /* 100 */
  int32_t update_size_thresh;       /* 10 */
  int32_t map_capacity;             /* LRU capacity in voxels */
  int32_t max_scan_points;          /* residual slots = largest scan */
  int32_t max_iter;                 /* IESKF iterations, 5 */
  int32_t estimate_ext;             /* 0 */
  double scan_resolution;           /* VoxelGrid leaf of registered scans
                                       (lio_builder.cpp:13-14,215-223); <= 0:
                                       none (the yaml's 0.1; 0 in the configs
                                       this path is measured on) */
} vmp_map_cfg;

/* IMU side of LIOConfig (lio_builder.h:19-28). */
typedef struct {
  double na, ng, nba, nbg;  /* noise of Q's blocks: 0.01, 0.01, 1e-4, 1e-4 */
  int32_t imu_init_num;     /* samples averaged at IMU_INIT: 20 */
  int32_t gravity_align;    /* 1 */
  double r_il[9];           /* LiDAR -> IMU rotation, column-major: I */
  double p_il[3];           /* LiDAR -> IMU translation: 0 */
} vmp_imu_cfg;

/* kf::State (ieskf.h:31-62); rot and rot_ext column-major. */
typedef struct {
  double pos[3];
  double rot[9];
  double rot_ext[9];
  double pos_ext[3];
  double vel[3];
  double bg[3];
  double ba[3];
  double g[3];
} vmp_state;

typedef struct {
  int32_t points;      /* N */
  int32_t hits;        /* lookups that found a voxel */
  int32_t plane_hits;  /* ... whose voxel is a plane */
  int32_t effective;   /* valid slots entering H (incl. stale ones) */
  int32_t key_fallback;/* points re-keyed in FP64 near a voxel face */
  int32_t iterations;  /* IESKF iterations run (vmp_ieskf_update) */
} vmp_obs_stats;

typedef struct {
  int32_t points;
  int32_t touched;     /* distinct voxels touched */
  int32_t created;
  int32_t evicted;
  int32_t refits;
  int32_t merge_calls;
  int32_t merge_evals; /* merge() calls actually evaluated on device */
  int32_t merges;
  int32_t dropped;
  int32_t map_size;
  int32_t maintenance; /* housekeeping after this scan: 1 = hash table
                          rebuilt, 2 = LRU ring compacted (bit set) */
  int32_t merge_windows;  /* evaluation windows of the merge replay */
  int32_t merge_requeued; /* window entries re-evaluated after a conflict */
  int32_t merge_max_acting; /* most merge() calls that acted in one window */
} vmp_map_stats;

typedef struct vmp_ctx vmp_ctx;
typedef struct vmp_session vmp_session;

const char* vmp_version(void);

int vmp_ctx_create(int device, vmp_ctx** out);
int vmp_ctx_destroy(vmp_ctx* ctx);

int vmp_session_create(vmp_ctx* ctx, const vmp_map_cfg* cfg, vmp_session** out);
int vmp_session_destroy(vmp_session* s);
/* bytes of HBM the session holds (for memory-based admission) */
int64_t vmp_session_bytes(const vmp_map_cfg* cfg);
const char* vmp_last_error(const vmp_session* s);

/* Filter state resident on the device (x and P, P column-major 23x23). */
int vmp_state_set(vmp_session* s, const vmp_state* x, const double* P);
int vmp_state_get(vmp_session* s, vmp_state* x, double* P);
/* Motion prior between scans (stand-in for IMU propagation, which is outside
 * this path): pos += rot*dp; rot = rot*dR; P += diag(qdiag[23]). */
int vmp_state_apply_motion(vmp_session* s, const double* dR, const double* dp, const double* qdiag);
/* The same for n sessions of one context in one launch: dR 9n (column-major
 * per session), dp 3n, qdiag 23n. */
int vmp_state_apply_motion_batch(vmp_session** s, int32_t n, const double* dR, const double* dp, const double* qdiag);

/* Scan points in the LiDAR frame, SoA fp32, scan (time) order.  Host
 * pointers are copied; _device takes pointers that already live in HBM (the
 * buffers must hold the points from the registration call that uses them
 * until it returns).  A scan holding a NaN or infinite coordinate is
 * rejected with VMP_EINVAL (the reference would index a voxel from an
 * undefined key): host buffers when they are bound; device buffers when the
 * next registration call starts, by one small kernel on the context's
 * stream, before anything is launched, so the failed call changes nothing
 * and the binding is dropped (context parameter "check_points"). */
int vmp_scan_set_points(vmp_session* s, const float* x, const float* y, const float* z, int32_t n);
int vmp_scan_set_points_device(vmp_session* s, const float* dx, const float* dy, const float* dz, int32_t n);
/* The same for n sessions in one call: device addresses and point counts. */
int vmp_scan_set_points_device_batch(vmp_session** s, int32_t n, const uint64_t* dx, const uint64_t* dy,
                                     const uint64_t* dz, const int32_t* counts);

/* IMU_INIT (LIOBuilder::initializeImu, lio_builder.cpp:28-63, with
 * lio_builder.cpp:177-184): the package's samples (t seconds, acc and gyro 3
 * per sample, host arrays) join the session's cache; once it holds
 * imu_init_num samples the device sets rot (gravity alignment), g, bg,
 * rot_ext, pos_ext and P, and *initialized = 1.  The scan of an IMU_INIT
 * package is not used (as in the reference). */
int vmp_imu_init(vmp_session* s, const vmp_imu_cfg* cfg, const double* t, const double* acc, const double* gyro,
                 int32_t m, double cloud_end_time, int32_t* initialized);
/* A raw SyncPackage (lio_node.cpp:137-165) for the next registration call:
 * points in the LiDAR frame at their own times, t_ms per point in ms since
 * cloud_start (ascending), and the package's 1..512 IMU samples (host
 * arrays).  The next vmp_map_update(build) / vmp_register_scan / batch call
 * first runs undistortCloud on the device (lio_builder.cpp:65-153: IMU
 * propagation of x and P over the samples, then every point compensated
 * into the frame at cloud_end) and, with scan_resolution > 0 and not at
 * MAP_INIT, the VoxelGrid filter.  Replaces vmp_scan_set_points and the
 * motion prior for that call.  _device takes points and times already in
 * HBM (read during the call that processes them). */
int vmp_scan_set_raw(vmp_session* s, const float* x, const float* y, const float* z, const float* t_ms, int32_t n,
                     const double* imu_t, const double* imu_acc, const double* imu_gyro, int32_t m,
                     double cloud_start, double cloud_end);
int vmp_scan_set_raw_device(vmp_session* s, const float* dx, const float* dy, const float* dz, const float* dt_ms,
                            int32_t n, const double* imu_t, const double* imu_acc, const double* imu_gyro, int32_t m,
                            double cloud_start, double cloud_end);
/* The same for n sessions of one context: device addresses of points and
 * times, point counts, the sessions' IMU samples concatenated (imu_counts
 * per session) and their cloud times; one check of all point sets. */
int vmp_scan_set_raw_device_batch(vmp_session** s, int32_t n, const uint64_t* dx, const uint64_t* dy,
                                  const uint64_t* dz, const uint64_t* dt_ms, const int32_t* counts,
                                  const double* imu_t, const double* imu_acc, const double* imu_gyro,
                                  const int32_t* imu_counts, const double* cloud_start, const double* cloud_end);
/* The points the last registration used (undistorted and filtered), SoA;
 * returns their count (at most cap are copied). */
int32_t vmp_scan_export(vmp_session* s, int32_t cap, float* x, float* y, float* z);
/* The LiDAR cloud of the last processed package, undistorted and before
 * filtering, in the body (IMU) frame: lidarToBody's fp32 transform with
 * rot_ext, pos_ext (lio_builder.cpp:165-173), the cloud PointCloudWithOdom
 * carries (lio_node.cpp:176-181,200-214).  Returns the count. */
int32_t vmp_cloud_body_export(vmp_session* s, int32_t cap, float* x, float* y, float* z);
/* PointCloudWithOdom's pose (interface/msg/PointCloudWithOdom.msg:1-3,
 * lio_node.cpp:219-227): position, orientation quaternion (w, x, y, z) of
 * rot, and the 6x6 covariance P[0:6, 0:6] row-major (position, then
 * rotation error). */
int vmp_odom_export(vmp_session* s, double pos[3], double quat_wxyz[4], double cov6x6[36]);

/* One measurement pass at state x with the session's P as prior:
 * H (12x12) and b (12) exactly as SharedState (ieskf.h:22-29). */
int vmp_observe(vmp_session* s, const vmp_state* x, double* H, double* b, vmp_obs_stats* st);

/* Insert the current scan with the session's (posterior) state.
 * is_build != 0 runs VoxelMap::build semantics (MAP_INIT, empty map only). */
int vmp_map_update(vmp_session* s, int is_build, vmp_map_stats* st);

/* IESKF update on the device with vmp_observe as measurement function. */
int vmp_ieskf_update(vmp_session* s, vmp_obs_stats* st);

/* set points (device) + IESKF update + map update, no host round trip */
int vmp_register_scan(vmp_session* s, vmp_obs_stats* ost, vmp_map_stats* mst);

/* vmp_register_scan that allocates nothing (no counterpart in the
 * reference): for callers that cannot allocate device memory in the call
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The step's work arrays live in a 64 MiB array per GPU that the
 * library's code object holds until the process ends, set up by
 * vmp_ctx_create.  Limits: the bound scan (raw package or points) holds at
 * most 131072 points, and the session's LRU ring, 2 (map_capacity + 2
 * max_scan_points) + 1024 entries, fits the scratch (about 3.2M entries:
 * map_capacity up to about 1.3M at max_scan_points 131072); beyond either
 * the call returns VMP_EINVAL before launching anything.  The filter stage
 * is launched kernel by kernel (no hipGraph).  Results are bitwise those of
 * vmp_register_scan.  Every context on a GPU shares its scratch, whose
 * contents the next call overwrites: no-allocation calls that overlap on one
 * GPU (from contexts on different host threads) race. */
int vmp_register_scan_noalloc(vmp_session* s, vmp_obs_stats* ost, vmp_map_stats* mst);

/* Map export, LRU order (most recent first).  Any output may be NULL.
 * keys[3i]; ints[5i] = {n, flags(1 init,2 plane,4 update_enable,8 merged),
 * newly_add_point, points held, 0}; group[i]; dbl[8i] = {mean xyz, norm xyz,
 * trace of normal-cov block, trace of centre-cov block}.  Returns count. */
int64_t vmp_map_export(vmp_session* s, int64_t cap, int64_t* keys, int32_t* ints, uint64_t* group, double* dbl);
int64_t vmp_map_size(vmp_session* s);
/* Residual slots: valid flag, residual, normal (3 per point). */
int vmp_slots_export(vmp_session* s, int32_t n, uint8_t* valid, float* resid, float* norm3);

/* Batched forms: sessions advance together, one launch per stage (at most
 * 32768 sessions per batch). */
int vmp_register_batch(vmp_session** s, int32_t n, vmp_obs_stats* ost, vmp_map_stats* mst);
/* The same without waiting: the batch is launched and the call returns; the
 * next submit (or any call that needs the context) first completes it, so the
 * host can prepare the next scans while the device works.  Returns the status
 * of the batch it completed (or of this launch).  vmp_register_batch_wait
 * completes the batch in flight and copies the statistics of the last
 * submitted batch (n entries) with its status; each submitted batch's
 * statistics are handed out once, and a wait with no submitted batch since
 * the previous wait returns VMP_ESTATE (synchronous calls leave none). */
int vmp_register_batch_submit(vmp_session** s, int32_t n);
int vmp_register_batch_wait(vmp_ctx* ctx, int32_t n, vmp_obs_stats* ost, vmp_map_stats* mst);

/* Size the context's batch workspace for batches of up to `points` points
 * in total over up to `sessions` sessions, so no later batch allocates (no
 * reference counterpart; without it the workspace grows on demand, with a
 * quarter of headroom). */
int vmp_ctx_reserve(vmp_ctx* ctx, int64_t points, int32_t sessions);

/* Launch-shape parameters of a context (no reference counterpart):
 *   "obs_ppt"  points per thread of the observe kernel, 1..64; 0 = taken
 *              from the size-bucket table (the default)
 *   "obs_ppt_bucket0".."obs_ppt_bucket3"  table entries for scans of < 8k,
 *              < 32k, < 128k and more points (default 2; set by
 *              vmp_autotune_observe)
 *   "filter_graph" 1 = the filter stage of a step (IESKF begin, the
 *              iterations' observe/reduce/step launches, end) replays a
 *              captured hipGraph (the default), 0 = launched one by one
 *              (also while per-launch timing is on).
 *   "map_graph" 1 = the map stage of a step replays a captured hipGraph
 *              once the step's shape (points, sessions, hash bits, build or
 *              not) recurs on consecutive steps (the default), 0 = always
 *              launched one by one (also while timing is on or when
 *              "filter_graph" is 0).
 *   "obs_arith" arithmetic of the observe pass: 0 = compensated fp32 in a
 *              frame anchored at the scan's pose voxel (the default),
 *              1 = plain fp32 world frame, a precision-study arm whose
 *              fallback counter counts keys that differ from the fp64 key.
 *   "check_points" 1 = device point buffers are checked for non-finite
 *              coordinates (and raw scans for ascending times) when the
 *              call that uses them starts (the default), 0 = the caller
 *              guarantees them (host buffers are always checked).
 *   "obs_time_reps" 1..64: while per-launch timing is on, each observe
 *              launch runs this many times back to back between one event
 *              pair (the kernel is idempotent), then once more for the
 *              pass's counters; the timing reports the mean (default 1).
 *   "replay_wave_max" 0..2^20: a map update with at most this many
 *              refitting voxels replays each of them on its own wave
 *              (default 4096); with more, only the voxels of
 *              "replay_long_seg" points or more do, and the others walk one
 *              per lane.  Both forms give the same bits.
 *   "replay_long_seg" 2..62, or 0 for none (default 24).
 * Applies to later calls. */
int vmp_ctx_set_param(vmp_ctx* ctx, const char* name, int64_t value);
/* Graph use of a context so far (no reference counterpart): out[0] filter
 * stage replays, out[1] filter graphs captured, out[2] map stage replays,
 * out[3] map graphs captured. */
int vmp_ctx_graph_stats(vmp_ctx* ctx, int64_t* out);

/* Autotuner of the observe launch shape (no reference counterpart): times
 * the observe kernel on the session's current scan and state for 1, 2, 4
 * and 8 points per thread (HIP events around the kernel, median of reps
 * launches), stores the fastest in the context's table for the scan's size
 * bucket and returns it (cand_ms: the four medians, may be NULL).  The
 * residual slots are restored afterwards. */
int vmp_autotune_observe(vmp_session* s, int32_t reps, int32_t* best_ppt, float* cand_ms);
/* The same for a batch of sessions of one context, tuned as the batch runs:
 * bucket by bucket (< 8k, < 32k, < 128k, larger points), each candidate is
 * timed as one observe launch over all the sessions.  table: the resulting
 * four entries; cand_ms: 4 x 4 medians (bucket-major, 0 for buckets no
 * session falls in); either may be NULL. */
int vmp_autotune_observe_batch(vmp_session** s, int32_t n, int32_t reps, int32_t* table, float* cand_ms);

/* Device events around the last call's kernels (for roofline timing). */
int vmp_session_timing(vmp_session* s, double* observe_ms, double* map_ms, int32_t* observe_launches);
/* Per-launch HIP-event timing of the observe kernel and of the map-update
 * pipeline, accumulated until the next vmp_session_set_timing call.
 * No reference counterpart (instrumentation). */
int vmp_session_set_timing(vmp_session* s, int on);
int vmp_session_timing_totals(vmp_session* s, double* obs_ms, int64_t* obs_launches, double* map_ms,
                              int64_t* map_calls, int64_t* obs_points, int64_t* obs_hits);
/* Per-kernel HIP-event totals over the same period (instrumentation, no
 * reference counterpart): milliseconds and launches of, in this order,
 * k_ins_keys, k_create_evict, k_replay_heavy, k_replay_light,
 * k_merge_replay, k_candidates, k_ring_live, k_ieskf_step, k_undistort,
 * k_imu_predict, k_spec_eval.  Returns the number of kernels (up to cap
 * entries are written). */
int32_t vmp_ctx_kernel_times(vmp_ctx* ctx, int32_t cap, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
