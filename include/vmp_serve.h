/*
 * vmp_serve.h — continuous-batching server over the registration path
 * (include/vmp.h).  No counterpart in the reference, which registers one
 * sequence per process (lio_node.cpp:98-160 drives one LIOBuilder); this is
 * the multi-session form of that loop.
 *
 * A server owns up to `nsessions` sequences.  Their voxel maps and filter
 * states stay resident in HBM while the total stays under `hbm_budget`; when
 * a session that is not resident needs the device and the budget is full,
 * the least recently used idle session is evicted (its map is dropped, its
 * last posterior state is kept on the host) and the new one is admitted.  An
 * evicted session's next scan starts a new map (MAP_INIT) at that state.
 *
 * vmp_server_run replays a list of requests (scans) in real time: a request
 * becomes eligible at its arrival time; each of `executors` workers (one HIP
 * stream each) repeatedly takes every eligible request whose session is idle
 * (at most one per session, in arrival order, up to the batch limits) and
 * runs them as one batch.  Requests that arrive while batches are on the
 * device join the next batch; batches of different workers overlap on the
 * GPU.  Each result is the same as the session's scans registered alone.
 */
#ifndef VMP_SERVE_H
#define VMP_SERVE_H
#include <stdint.h>

#include "vmp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t device;
  int32_t executors;          /* concurrent batches (streams), >= 1 */
  int32_t max_batch_sessions; /* sessions per batch */
  int64_t max_batch_points;   /* points per batch */
  int64_t hbm_budget;         /* bytes for resident sessions */
} vmp_server_cfg;

typedef struct {
  /* set by the caller */
  int32_t session;            /* 0 .. nsessions-1 */
  int32_t n;                  /* points */
  const float* x;             /* device pointers, SoA, LiDAR frame */
  const float* y;
  const float* z;
  double dR[9];               /* motion prior since the session's previous */
  double dp[3];               /* scan (column-major dR); unused on its first */
  double arrival;             /* seconds after the start of the run */
  /* set by the server */
  double start;               /* batch launch, seconds after the start */
  double done;                /* batch completion */
  int32_t status;             /* 0 or a VMP_E* code */
  int32_t kind;               /* 1 MAP_INIT, 2 registered, 3 MAP_INIT after eviction */
  int32_t batch;              /* batch id */
  int32_t batch_sessions;     /* sessions in that batch */
  int32_t iterations;         /* IESKF iterations (kind 2) */
  int32_t pad;
  double rot[9];              /* posterior pose: rot (column-major), pos */
  double pos[3];
} vmp_request;

typedef struct vmp_server vmp_server;

int vmp_server_create(const vmp_server_cfg* cfg, const vmp_map_cfg* session_cfg, int32_t nsessions,
                      vmp_server** out);
int vmp_server_destroy(vmp_server* sv);
/* Initial state of a session (the MAP_INIT pose and P). Required before its
 * first request. */
int vmp_server_session_init(vmp_server* sv, int32_t id, const vmp_state* x0, const double* P0);
/* Replay `n` requests sorted by arrival; qdiag (23) is the motion-prior noise.
 * Returns 0 or the first error; per-request status is in reqs[i].status. */
int vmp_server_run(vmp_server* sv, vmp_request* reqs, int32_t n, const double* qdiag);
/* The resident session of `id` (NULL if evicted or never admitted); valid
 * until the next run. */
vmp_session* vmp_server_session(vmp_server* sv, int32_t id);
/* Counters of the last run: batches, requests, evictions, admissions,
 * resident sessions, resident bytes, largest batch (sessions), executors. */
int vmp_server_stats(vmp_server* sv, int64_t out[8]);
const char* vmp_server_last_error(const vmp_server* sv);

#ifdef __cplusplus
}
#endif
#endif
