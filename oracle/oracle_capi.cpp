// ORACLE — test infrastructure only.  extern "C" surface over the FP64
// restatement so pytest (ctypes) and bench.py's cpu_baseline leg can drive it.
// Layouts: state = 36 doubles {pos[3], rot[9], rot_ext[9], pos_ext[3], vel[3],
// bg[3], ba[3], g[3]}, 3x3 and 23x23 matrices column-major (Eigen layout).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vmp_oracle.hpp"

using namespace orc;

extern "C" {

typedef struct {
  double voxel_size, plane_thresh, merge_thresh_for_angle, merge_thresh_for_distance, ranging_cov, angle_cov;
  int32_t max_point_thresh, update_size_thresh, map_capacity, max_scan_points, max_iter, estimate_ext;
  double scan_resolution;
} orc_cfg;

static void m3_in(const double* cm, M3& m) {
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) m(r, c) = cm[c * 3 + r];
}
static void m3_out(const M3& m, double* cm) {
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) cm[c * 3 + r] = m(r, c);
}
static void v3_in(const double* p, V3& v) { v = v3(p[0], p[1], p[2]); }
static void v3_out(const V3& v, double* p) {
  p[0] = v(0, 0);
  p[1] = v(1, 0);
  p[2] = v(2, 0);
}
static void state_in(const double* s, State& x) {
  v3_in(s + 0, x.pos);
  m3_in(s + 3, x.rot);
  m3_in(s + 12, x.rot_ext);
  v3_in(s + 21, x.pos_ext);
  v3_in(s + 24, x.vel);
  v3_in(s + 27, x.bg);
  v3_in(s + 30, x.ba);
  v3_in(s + 33, x.g);
}
static void state_out(const State& x, double* s) {
  v3_out(x.pos, s + 0);
  m3_out(x.rot, s + 3);
  m3_out(x.rot_ext, s + 12);
  v3_out(x.pos_ext, s + 21);
  v3_out(x.vel, s + 24);
  v3_out(x.bg, s + 27);
  v3_out(x.ba, s + 30);
  v3_out(x.g, s + 33);
}

void* orc_create(const orc_cfg* c) {
  Config cfg;
  cfg.voxel_size = c->voxel_size;
  cfg.plane_thresh = c->plane_thresh;
  cfg.merge_thresh_for_angle = c->merge_thresh_for_angle;
  cfg.merge_thresh_for_distance = c->merge_thresh_for_distance;
  cfg.ranging_cov = c->ranging_cov;
  cfg.angle_cov = c->angle_cov;
  cfg.max_point_thresh = c->max_point_thresh;
  cfg.update_size_thresh = c->update_size_thresh;
  cfg.map_capacity = c->map_capacity;
  cfg.max_iter = c->max_iter;
  cfg.estimate_ext = c->estimate_ext;
  cfg.scan_resolution = c->scan_resolution;
  return new Registrar(cfg);
}
void orc_destroy(void* h) { delete static_cast<Registrar*>(h); }

void orc_set_state(void* h, const double* st, const double* P) {
  Registrar* r = static_cast<Registrar*>(h);
  if (st) state_in(st, r->kf.x);
  if (P)
    for (int c = 0; c < 23; ++c)
      for (int i = 0; i < 23; ++i) r->kf.P(i, c) = P[c * 23 + i];
}
void orc_get_state(void* h, double* st, double* P) {
  Registrar* r = static_cast<Registrar*>(h);
  if (st) state_out(r->kf.x, st);
  if (P)
    for (int c = 0; c < 23; ++c)
      for (int i = 0; i < 23; ++i) P[c * 23 + i] = r->kf.P(i, c);
}

int orc_map_init(void* h, const float* x, const float* y, const float* z, int n) {
  static_cast<Registrar*>(h)->map_init(x, y, z, n);
  return 0;
}
int orc_set_scan(void* h, const float* x, const float* y, const float* z, int n) {
  static_cast<Registrar*>(h)->set_scan(x, y, z, n);
  return 0;
}
// One sharedUpdateFunc pass at the given state (prior P = the registrar's P).
int orc_observe(void* h, const double* st, double* H, double* b, int* effect) {
  Registrar* r = static_cast<Registrar*>(h);
  State s;
  state_in(st, s);
  SharedState sd;
  r->observe(s, sd);
  for (int c = 0; c < 12; ++c) {
    for (int i = 0; i < 12; ++i) H[c * 12 + i] = sd.H(i, c);
    b[c] = sd.b(c, 0);
  }
  if (effect) *effect = r->effect_num;
  return 0;
}
int orc_ieskf_update(void* h, int* iters) {
  Registrar* r = static_cast<Registrar*>(h);
  r->kf.update();
  if (iters) *iters = r->kf.last_iters;
  return 0;
}
int orc_map_update(void* h) {
  static_cast<Registrar*>(h)->map_update();
  return 0;
}
// set_scan + IESKF update + map update: one registered scan.
int orc_register(void* h, const float* x, const float* y, const float* z, int n, int* iters) {
  Registrar* r = static_cast<Registrar*>(h);
  r->set_scan(x, y, z, n);
  r->kf.update();
  if (iters) *iters = r->kf.last_iters;
  r->map_update();
  return 0;
}

// Motion prior standing in for IMU propagation between scans:
// pos += rot * dp; rot = rot * dR; P += diag(qdiag).
void orc_apply_motion(void* h, const double* dR, const double* dp, const double* qdiag) {
  Registrar* r = static_cast<Registrar*>(h);
  M3 D;
  m3_in(dR, D);
  V3 p;
  v3_in(dp, p);
  r->kf.x.pos = r->kf.x.pos + r->kf.x.rot * p;
  r->kf.x.rot = r->kf.x.rot * D;
  for (int i = 0; i < 23; ++i) r->kf.P(i, i) += qdiag[i];
}

int64_t orc_map_size(void* h) { return int64_t(static_cast<Registrar*>(h)->map->cache.size()); }

// Voxels in LRU order (front = most recently touched).
// ints[5*i]: n, flags (1 init, 2 plane, 4 update_enable, 8 merged), newly_add_point, temp size, 0
// dbl[8*i]: mean xyz, norm xyz, trace(cov[0:3,0:3]), trace(cov[3:6,3:6])
// Per voxel in LRU order: eigenvalues (ascending) and major axis of the last
// refit (6 doubles), for conditioning-aware normal comparisons.
int64_t orc_map_fit(void* h, int64_t cap, double* out) {
  Registrar* r = static_cast<Registrar*>(h);
  int64_t i = 0;
  for (const VoxelKey& k : r->map->cache) {
    if (i >= cap) break;
    const VoxelGrid& g = *r->map->featmap.at(k);
    for (int a = 0; a < 3; ++a) {
      out[6 * i + a] = g.fit_evals[a];
      out[6 * i + 3 + a] = g.fit_major(a, 0);
    }
    ++i;
  }
  return i;
}

int64_t orc_map_export(void* h, int64_t cap, int64_t* keys, int32_t* ints, uint64_t* group, double* dbl) {
  Registrar* r = static_cast<Registrar*>(h);
  int64_t i = 0;
  for (const VoxelKey& k : r->map->cache) {
    if (i >= cap) break;
    const VoxelGrid& g = *r->map->featmap.at(k);
    if (keys) {
      keys[3 * i] = k.x;
      keys[3 * i + 1] = k.y;
      keys[3 * i + 2] = k.z;
    }
    if (ints) {
      ints[5 * i] = g.plane->n;
      ints[5 * i + 1] = (g.is_init ? 1 : 0) | (g.is_plane ? 2 : 0) | (g.update_enable ? 4 : 0) | (g.merged ? 8 : 0);
      ints[5 * i + 2] = g.newly_add_point;
      ints[5 * i + 3] = int32_t(g.temp_points.size());
      ints[5 * i + 4] = 0;
    }
    if (group) group[i] = g.group_id;
    if (dbl) {
      const M6& c = g.plane->cov;
      for (int a = 0; a < 3; ++a) {
        dbl[8 * i + a] = g.plane->mean(a, 0);
        dbl[8 * i + 3 + a] = g.plane->norm(a, 0);
      }
      dbl[8 * i + 6] = c(0, 0) + c(1, 1) + c(2, 2);
      dbl[8 * i + 7] = c(3, 3) + c(4, 4) + c(5, 5);
    }
    ++i;
  }
  return i;
}

int orc_slots_export(void* h, int n, uint8_t* valid, double* resid, double* norm3) {
  Registrar* r = static_cast<Registrar*>(h);
  if (n > int(r->residual_info.size())) n = int(r->residual_info.size());
  for (int i = 0; i < n; ++i) {
    const ResidualData& d = r->residual_info[i];
    if (valid) valid[i] = d.is_valid ? 1 : 0;
    if (resid) resid[i] = d.residual;
    if (norm3) v3_out(d.plane_norm, norm3 + 3 * i);
  }
  return n;
}

void orc_stats(void* h, int64_t* out) {
  const MapStats& s = static_cast<Registrar*>(h)->map->stats;
  out[0] = s.new_voxels;
  out[1] = s.evicted;
  out[2] = s.refits;
  out[3] = s.merge_calls;
  out[4] = s.merges;
  out[5] = s.dropped;
}

// Direct map access for brute-force tests: world points + covs (3x3 col-major).
int orc_map_push(void* h, const double* pts, const double* covs, int n, int is_build) {
  Registrar* r = static_cast<Registrar*>(h);
  std::vector<PointWithCov> pv(n);
  for (int i = 0; i < n; ++i) {
    v3_in(pts + 3 * i, pv[i].point);
    m3_in(covs + 9 * i, pv[i].cov);
  }
  if (is_build) r->map->build(pv);
  else r->map->update(pv);
  return 0;
}

// The insertion transform at the registrar's current state (fp32 out).
void orc_lidar_to_world_f32(void* h, const float* x, const float* y, const float* z, int n, float* ox, float* oy,
                            float* oz) {
  Registrar* r = static_cast<Registrar*>(h);
  std::vector<float> wx, wy, wz;
  lidar_to_world_f32(r->kf.x, x, y, z, n, wx, wy, wz);
  for (int i = 0; i < n; ++i) {
    ox[i] = wx[i];
    oy[i] = wy[i];
    oz[i] = wz[i];
  }
}

void orc_key(void* h, const double* p, int64_t* key) {
  V3 v;
  v3_in(p, v);
  VoxelKey k = static_cast<Registrar*>(h)->map->index(v);
  key[0] = k.x;
  key[1] = k.y;
  key[2] = k.z;
}

// ---- IMU side of LIOBuilder ----
static std::vector<ImuData> imus_in(const double* t, const double* acc, const double* gyro, int m) {
  std::vector<ImuData> v(static_cast<size_t>(m));
  for (int i = 0; i < m; ++i) {
    v3_in(acc + 3 * i, v[size_t(i)].acc);
    v3_in(gyro + 3 * i, v[size_t(i)].gyro);
    v[size_t(i)].t = t[i];
  }
  return v;
}
// noise (na, ng, nba, nbg), imu_init_num, gravity_align, r_il (column-major), p_il
void orc_imu_config(void* h, const double* noise4, int init_num, int gravity_align, const double* r_il,
                    const double* p_il) {
  Registrar* r = static_cast<Registrar*>(h);
  ImuConfig c;
  c.na = noise4[0];
  c.ng = noise4[1];
  c.nba = noise4[2];
  c.nbg = noise4[3];
  c.imu_init_num = init_num;
  c.gravity_align = gravity_align != 0;
  m3_in(r_il, c.r_il);
  v3_in(p_il, c.p_il);
  r->set_imu_config(c);
}
void orc_set_scan_resolution(void* h, double leaf) { static_cast<Registrar*>(h)->cfg.scan_resolution = leaf; }
// One SyncPackage (lio_builder.cpp:175-248): returns the status after it
// (0 IMU_INIT, 1 MAP_INIT, 2 LIO_MAPPING); iters gets the IESKF iterations.
int orc_process(void* h, const float* x, const float* y, const float* z, const float* t_ms, int n, const double* it,
                const double* iacc, const double* igyro, int m, double cloud_start, double cloud_end, int* iters) {
  Registrar* r = static_cast<Registrar*>(h);
  r->kf.last_iters = 0;
  const LioStatus st = r->process(std::vector<float>(x, x + n), std::vector<float>(y, y + n),
                                  std::vector<float>(z, z + n), std::vector<float>(t_ms, t_ms + n),
                                  imus_in(it, iacc, igyro, m), cloud_start, cloud_end);
  if (iters) *iters = r->kf.last_iters;
  return int(st);
}
// undistortCloud alone (points in/out, sorted by time on return)
void orc_undistort(void* h, float* x, float* y, float* z, float* t_ms, int n, const double* it, const double* iacc,
                   const double* igyro, int m, double cloud_start, double cloud_end) {
  Registrar* r = static_cast<Registrar*>(h);
  std::vector<float> a(x, x + n), b(y, y + n), c(z, z + n), d(t_ms, t_ms + n);
  r->undistort(a, b, c, d, imus_in(it, iacc, igyro, m), cloud_start, cloud_end);
  for (int i = 0; i < n; ++i) {
    x[i] = a[size_t(i)];
    y[i] = b[size_t(i)];
    z[i] = c[size_t(i)];
    t_ms[i] = d[size_t(i)];
  }
}
int orc_initialize_imu(void* h, const double* it, const double* iacc, const double* igyro, int m) {
  return static_cast<Registrar*>(h)->initialize_imu(imus_in(it, iacc, igyro, m)) ? 1 : 0;
}
void orc_predict(void* h, const double* acc, const double* gyro, double dt) {
  Registrar* r = static_cast<Registrar*>(h);
  V3 a, g;
  v3_in(acc, a);
  v3_in(gyro, g);
  r->kf.predict(a, g, dt, r->Q);
}
// IMU bookkeeping: last_acc (3), last_gyro (3), last_cloud_end_time,
// gravity_norm, last_imu time, status
void orc_imu_state(void* h, double* out10) {
  Registrar* r = static_cast<Registrar*>(h);
  v3_out(r->last_acc, out10);
  v3_out(r->last_gyro, out10 + 3);
  out10[6] = r->last_cloud_end_time;
  out10[7] = r->gravity_norm;
  out10[8] = r->last_imu.t;
  out10[9] = double(int(r->status));
}
// IMU pose table of the last undistortion: per entry offset, acc(3),
// gyro(3), vel(3), pos(3), rot(9, column-major) = 22 doubles
int orc_imu_poses(void* h, int cap, double* out) {
  Registrar* r = static_cast<Registrar*>(h);
  const int n = int(r->imu_poses.size());
  for (int i = 0; i < n && i < cap; ++i) {
    const ImuPose& q = r->imu_poses[size_t(i)];
    double* o = out + 22 * i;
    o[0] = q.offset;
    v3_out(q.acc, o + 1);
    v3_out(q.gyro, o + 4);
    v3_out(q.vel, o + 7);
    v3_out(q.pos, o + 10);
    m3_out(q.rot, o + 13);
  }
  return n;
}

// points of the last registered scan (after undistortion and downsampling)
int orc_scan_points(void* h, int cap, float* x, float* y, float* z) {
  Registrar* r = static_cast<Registrar*>(h);
  const int n = std::min(cap, r->size);
  for (int i = 0; i < n; ++i) {
    if (x) x[i] = r->cx[size_t(i)];
    if (y) y[i] = r->cy[size_t(i)];
    if (z) z[i] = r->cz[size_t(i)];
  }
  return r->size;
}
int orc_voxel_grid(const float* x, const float* y, const float* z, int n, float leaf, float* ox, float* oy, float* oz) {
  std::vector<float> a, b, c;
  voxel_grid_filter(x, y, z, n, leaf, a, b, c);
  for (size_t i = 0; i < a.size(); ++i) {
    ox[i] = a[i];
    oy[i] = b[i];
    oz[i] = c[i];
  }
  return int(a.size());
}
void orc_rot_from_two_vectors(const double* a, const double* b, double* R) {
  V3 va, vb;
  v3_in(a, va);
  v3_in(b, vb);
  m3_out(rot_from_two_vectors(va, vb), R);
}

// ---- math exposed for pinning against numpy / scipy ----
void orc_eigh3(const double* A, double* evals, double* evecs) {  // column-major in/out
  M3 m, V;
  m3_in(A, m);
  eigh3(m, evals, V);
  m3_out(V, evecs);
}
int orc_invert(const double* A, double* Ainv, int n) { return invert(A, Ainv, n) ? 0 : -1; }  // row-major
void orc_so3_exp(const double* w, double* R) {
  V3 v;
  v3_in(w, v);
  m3_out(so3_exp(v), R);
}
void orc_so3_log(const double* R, double* w) {
  M3 m;
  m3_in(R, m);
  v3_out(so3_log(m), w);
}
void orc_left_jacobian(const double* w, double* J) {
  V3 v;
  v3_in(w, v);
  m3_out(left_jacobian(v), J);
}
void orc_calc_body_cov(double* p, double ri, double di, double* cov) {
  V3 v;
  v3_in(p, v);
  M3 c;
  calc_body_cov(v, ri, di, c);
  v3_out(v, p);
  m3_out(c, cov);
}
void orc_boxplus(const double* st, const double* d23, double* out) {
  State x;
  state_in(st, x);
  V23 d;
  for (int i = 0; i < 23; ++i) d(i, 0) = d23[i];
  x.plus(d);
  state_out(x, out);
}
void orc_boxminus(const double* a, const double* b, double* d23) {
  State x, y;
  state_in(a, x);
  state_in(b, y);
  V23 d = x.minus(y);
  for (int i = 0; i < 23; ++i) d23[i] = d(i, 0);
}

}  // extern "C"
