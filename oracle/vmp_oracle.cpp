// ORACLE — test infrastructure only (see vmp_oracle.hpp).  FP64 restatement of
// the reference registration path; every function cites what it follows.
#include "vmp_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace orc {

M3 hat(const V3& v) {  // Sophus::SO3d::hat
  M3 m = M3::zero();
  m(0, 1) = -v(2, 0);
  m(0, 2) = v(1, 0);
  m(1, 0) = v(2, 0);
  m(1, 2) = -v(0, 0);
  m(2, 0) = -v(1, 0);
  m(2, 1) = v(0, 0);
  return m;
}

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 (Golub & Van Loan
// 4th ed. 8.5.2).  Replaces Eigen::SelfAdjointEigenSolver (voxel_map.cpp:104):
// eigenvalues ascending, eigenvectors in the columns of `evecs`.
void eigh3(const M3& Ain, double evals[3], M3& V) {
  double A[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = 0.5 * (Ain(i, j) + Ain(j, i));
  V = M3::ident();
  const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (off <= 1e-36 * dia || off == 0.0) break;
    for (int r = 0; r < 3; ++r) {
      int p = P[r], q = Q[r];
      double apq = A[p][q];
      if (apq == 0.0) continue;
      double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
      double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
      double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
      for (int k = 0; k < 3; ++k) {
        double akp = A[k][p], akq = A[k][q];
        A[k][p] = c * akp - s * akq;
        A[k][q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; ++k) {
        double apk = A[p][k], aqk = A[q][k];
        A[p][k] = c * apk - s * aqk;
        A[q][k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; ++k) {
        double vkp = V(k, p), vkq = V(k, q);
        V(k, p) = c * vkp - s * vkq;
        V(k, q) = s * vkp + c * vkq;
      }
    }
  }
  int idx[3] = {0, 1, 2};
  double d[3] = {A[0][0], A[1][1], A[2][2]};
  std::sort(idx, idx + 3, [&](int a, int b) { return d[a] < d[b]; });
  M3 Vs;
  for (int c = 0; c < 3; ++c) {
    evals[c] = d[idx[c]];
    for (int r = 0; r < 3; ++r) Vs(r, c) = V(r, idx[c]);
  }
  V = Vs;
}

// Dense inverse by LU with partial pivoting (replaces Eigen's .inverse() on
// Matrix23d, ieskf.cpp:141-142,145,155).  Row-major n x n.
bool invert(const double* Ain, double* Ainv, int n) {
  std::vector<double> LU(Ain, Ain + size_t(n) * n);
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int k = 0; k < n; ++k) {
    int piv = k;
    double best = std::fabs(LU[size_t(k) * n + k]);
    for (int i = k + 1; i < n; ++i) {
      double v = std::fabs(LU[size_t(i) * n + k]);
      if (v > best) { best = v; piv = i; }
    }
    if (best == 0.0) return false;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(LU[size_t(k) * n + j], LU[size_t(piv) * n + j]);
      std::swap(perm[k], perm[piv]);
    }
    double dkk = LU[size_t(k) * n + k];
    for (int i = k + 1; i < n; ++i) {
      double l = LU[size_t(i) * n + k] / dkk;
      LU[size_t(i) * n + k] = l;
      for (int j = k + 1; j < n; ++j) LU[size_t(i) * n + j] -= l * LU[size_t(k) * n + j];
    }
  }
  std::vector<double> col(n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) col[i] = (perm[i] == c) ? 1.0 : 0.0;
    for (int i = 0; i < n; ++i) {
      double s = col[i];
      for (int j = 0; j < i; ++j) s -= LU[size_t(i) * n + j] * col[j];
      col[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = col[i];
      for (int j = i + 1; j < n; ++j) s -= LU[size_t(i) * n + j] * col[j];
      col[i] = s / LU[size_t(i) * n + i];
    }
    for (int i = 0; i < n; ++i) Ainv[size_t(i) * n + c] = col[i];
  }
  return true;
}

// SO3 exponential through the unit quaternion (Rodrigues), with the Taylor
// branch below theta = 1e-10 (Sophus::SO3d::exp, used at ieskf.cpp:14-15,20).
M3 so3_exp(const V3& w) {
  double tsq = dot(w, w);
  double imag, real;
  if (tsq < 1e-20) {
    imag = 0.5 - tsq / 48.0 + tsq * tsq / 3840.0;
    real = 1.0 - tsq / 8.0 + tsq * tsq / 384.0;
  } else {
    double th = std::sqrt(tsq), half = 0.5 * th;
    imag = std::sin(half) / th;
    real = std::cos(half);
  }
  double qw = real, qx = imag * w(0, 0), qy = imag * w(1, 0), qz = imag * w(2, 0);
  double tx = 2 * qx, ty = 2 * qy, tz = 2 * qz;
  double twx = tx * qw, twy = ty * qw, twz = tz * qw;
  double txx = tx * qx, txy = ty * qx, txz = tz * qx;
  double tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
  M3 R;
  R(0, 0) = 1 - (tyy + tzz);
  R(0, 1) = txy - twz;
  R(0, 2) = txz + twy;
  R(1, 0) = txy + twz;
  R(1, 1) = 1 - (txx + tzz);
  R(1, 2) = tyz - twx;
  R(2, 0) = txz - twy;
  R(2, 1) = tyz + twx;
  R(2, 2) = 1 - (txx + tyy);
  return R;
}

// SO3 logarithm: rotation matrix -> unit quaternion (Shepperd's method) ->
// 2*atan2 form (Sophus::SO3d::log, used at ieskf.cpp:26-27).
V3 so3_log(const M3& m) {
  double q[4];  // x y z w
  double t = m(0, 0) + m(1, 1) + m(2, 2);
  if (t > 0) {
    t = std::sqrt(t + 1.0);
    q[3] = 0.5 * t;
    t = 0.5 / t;
    q[0] = (m(2, 1) - m(1, 2)) * t;
    q[1] = (m(0, 2) - m(2, 0)) * t;
    q[2] = (m(1, 0) - m(0, 1)) * t;
  } else {
    int i = 0;
    if (m(1, 1) > m(0, 0)) i = 1;
    if (m(2, 2) > m(i, i)) i = 2;
    int j = (i + 1) % 3, k = (j + 1) % 3;
    t = std::sqrt(m(i, i) - m(j, j) - m(k, k) + 1.0);
    q[i] = 0.5 * t;
    t = 0.5 / t;
    q[3] = (m(k, j) - m(j, k)) * t;
    q[j] = (m(j, i) + m(i, j)) * t;
    q[k] = (m(k, i) + m(i, k)) * t;
  }
  double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (double& c : q) c /= qn;
  double sqn = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  double w = q[3], f;
  if (sqn < 1e-20) {
    f = 2.0 / w - 2.0 / 3.0 * sqn / (w * w * w);
  } else {
    double nn = std::sqrt(sqn);
    if (std::fabs(w) < 1e-10) f = (w > 0 ? M_PI : -M_PI) / nn;
    else f = 2.0 * std::atan(nn / w) / nn;
  }
  return v3(f * q[0], f * q[1], f * q[2]);
}

// Left Jacobian of SO3 (Sophus::SO3d::leftJacobian; ieskf.cpp:5-8 takes its
// transpose as the right Jacobian).  First-order branch for theta^2 < 1e-10
// (Sophus' Constants<double>::epsilon()), where 1 - cos(theta) cancels.
M3 left_jacobian(const V3& w) {
  double tsq = dot(w, w);
  M3 O = hat(w);
  M3 I = M3::ident();
  if (tsq < 1e-10) return I + 0.5 * O;
  double th = std::sqrt(tsq);
  M3 O2 = O * O;
  return (I + ((1.0 - std::cos(th)) / (th * th)) * O) + ((th - std::sin(th)) / (th * th * th)) * O2;
}

// commons.cpp:18-45.  Mutates pb[2] == 0 -> 0.001 like the reference.
void calc_body_cov(V3& pb, double range_inc, double degree_inc, M3& cov) {
  if (pb(2, 0) == 0) pb(2, 0) = 0.001;
  double range = std::sqrt(pb(0, 0) * pb(0, 0) + pb(1, 0) * pb(1, 0) + pb(2, 0) * pb(2, 0));
  double range_var = range_inc * range_inc;
  double s = std::pow(std::sin(degree_inc * M_PI / 180.0), 2);
  Mat<2, 2> direction_var = Mat<2, 2>::zero();
  direction_var(0, 0) = s;
  direction_var(1, 1) = s;
  V3 d = pb;
  {
    double z = dot(d, d);
    if (z > 0) {
      double n = std::sqrt(z);
      for (int i = 0; i < 3; ++i) d(i, 0) /= n;
    }
  }
  M3 dh = M3::zero();
  dh(0, 1) = -d(2, 0);
  dh(0, 2) = d(1, 0);
  dh(1, 0) = d(2, 0);
  dh(1, 2) = -d(0, 0);
  dh(2, 0) = -d(1, 0);
  dh(2, 1) = d(0, 0);
  V3 b1 = v3(1, 1, -(d(0, 0) + d(1, 0)) / d(2, 0));
  {
    double n = std::sqrt(dot(b1, b1));
    for (int i = 0; i < 3; ++i) b1(i, 0) /= n;
  }
  V3 b2 = v3(b1(1, 0) * d(2, 0) - b1(2, 0) * d(1, 0), b1(2, 0) * d(0, 0) - b1(0, 0) * d(2, 0),
             b1(0, 0) * d(1, 0) - b1(1, 0) * d(0, 0));
  {
    double n = std::sqrt(dot(b2, b2));
    for (int i = 0; i < 3; ++i) b2(i, 0) /= n;
  }
  Mat<3, 2> N;
  for (int i = 0; i < 3; ++i) {
    N(i, 0) = b1(i, 0);
    N(i, 1) = b2(i, 0);
  }
  Mat<3, 2> A = (range * dh) * N;
  cov = ((range_var * d) * tr(d)) + (A * direction_var) * tr(A);
}

// ---------------- VoxelGrid (voxel_map.cpp:11-186) ----------------
VoxelGrid::VoxelGrid(int max_pt, int upd_pt, double pt, VoxelKey pos, VoxelMap* m)
    : max_point_thresh(max_pt), update_point_thresh(upd_pt), plane_thresh(pt), position(pos), map(m) {
  group_id = map->group_counter++;
  temp_points.reserve(max_point_thresh);
  plane = std::make_shared<Plane>();
}

void VoxelGrid::addToPlane(const PointWithCov& pv) {  // voxel_map.cpp:29-34
  for (int i = 0; i < 3; ++i)
    plane->mean(i, 0) = plane->mean(i, 0) + (pv.point(i, 0) - plane->mean(i, 0)) / (plane->n + 1.0);
  plane->ppt = plane->ppt + pv.point * tr(pv.point);
  plane->n += 1;
}

void VoxelGrid::addPoint(const PointWithCov& pv) {  // voxel_map.cpp:36-40
  addToPlane(pv);
  temp_points.push_back(pv);
}

void VoxelGrid::pushPoint(const PointWithCov& pv) {  // voxel_map.cpp:42-95
  if (!is_init) {
    addToPlane(pv);
    temp_points.push_back(pv);
    updatePlane();
    return;
  }
  if (is_plane && !update_enable) {
    merge();
    return;
  }
  if (!update_enable) {  // frozen non-plane voxel drops the point
    map->stats.dropped++;
    return;
  }
  addToPlane(pv);
  temp_points.push_back(pv);
  newly_add_point++;
  if (newly_add_point >= update_point_thresh) {
    updatePlane();
    newly_add_point = 0;
  }
  if (int(temp_points.size()) >= max_point_thresh) {
    update_enable = false;
    std::vector<PointWithCov>().swap(temp_points);
  }
}

void VoxelGrid::updatePlane() {  // voxel_map.cpp:97-136
  if (plane->n < update_point_thresh) return;
  is_init = true;
  map->stats.refits++;
  M3 cov;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      cov(i, j) = plane->ppt(i, j) / double(plane->n) - plane->mean(i, 0) * plane->mean(j, 0);
  double evals[3];
  M3 evecs;
  eigh3(cov, evals, evecs);
  for (int i = 0; i < 3; ++i) fit_evals[i] = evals[i];
  fit_major = v3(evecs(0, 2), evecs(1, 2), evecs(2, 2));
  if (evals[0] > plane_thresh) {
    is_plane = false;
    return;
  }
  is_plane = true;
  M3 J_Q = (1.0 / double(plane->n)) * M3::ident();
  V3 pn = v3(evecs(0, 0), evecs(1, 0), evecs(2, 0));
  for (const PointWithCov& pv : temp_points) {
    Mat<6, 3> J;
    M3 F = M3::zero();
    for (int m = 1; m < 3; ++m) {
      V3 em = v3(evecs(0, m), evecs(1, m), evecs(2, m));
      double den = double(plane->n) * (evals[0] - evals[m]);
      Mat<1, 3> row = tr(pv.point - plane->mean);
      for (int k = 0; k < 3; ++k) row(0, k) = row(0, k) / den;
      M3 S = em * tr(pn) + pn * tr(em);
      Mat<1, 3> Fm = row * S;
      for (int k = 0; k < 3; ++k) F(m, k) = Fm(0, k);
    }
    M3 top = evecs * F;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        J(i, j) = top(i, j);
        J(3 + i, j) = J_Q(i, j);
      }
    plane->cov = plane->cov + (J * pv.cov) * tr(J);
  }
  double axis_distance = -dot(plane->mean, pn);
  if (axis_distance < 0.0) pn = -1.0 * pn;
  plane->norm = pn;
  center = plane->mean;
}

void VoxelGrid::merge() {  // voxel_map.cpp:138-186
  map->stats.merge_calls++;
  const VoxelKey near[6] = {{position.x - 1, position.y, position.z}, {position.x, position.y - 1, position.z},
                            {position.x, position.y, position.z - 1}, {position.x + 1, position.y, position.z},
                            {position.x, position.y + 1, position.z}, {position.x, position.y, position.z + 1}};
  for (const VoxelKey& k : near) {
    auto it = map->featmap.find(k);
    if (it == map->featmap.end()) continue;
    std::shared_ptr<VoxelGrid> nn = it->second;
    if (nn->group_id == group_id || nn->update_enable || !nn->is_plane) continue;
    double norm_distance = 1.0 - dot(nn->plane->norm, plane->norm);
    double axis_distance = std::fabs(dot(nn->plane->norm, nn->plane->mean) - dot(plane->norm, plane->mean));
    if (norm_distance > map->merge_thresh_for_angle || axis_distance > map->merge_thresh_for_distance) continue;
    const M6& c0 = plane->cov;
    const M6& c1 = nn->plane->cov;
    double tn0 = c0(0, 0) + c0(1, 1) + c0(2, 2), tm0 = c0(3, 3) + c0(4, 4) + c0(5, 5);
    double tn1 = c1(0, 0) + c1(1, 1) + c1(2, 2), tm1 = c1(3, 3) + c1(4, 4) + c1(5, 5);
    double tc0 = tn0 + tm0, tc1 = tn1 + tm1;
    // operator precedence as written in the reference: only the second term is divided
    V3 new_mean;
    V3 t1 = tm1 * plane->mean;
    V3 n1 = tn1 * plane->norm;
    for (int i = 0; i < 3; ++i) new_mean(i, 0) = tm0 * nn->plane->mean(i, 0) + t1(i, 0) / (tm0 + tm1);
    V3 new_norm;
    for (int i = 0; i < 3; ++i) new_norm(i, 0) = tn0 * nn->plane->norm(i, 0) + n1(i, 0) / (tn0 + tn1);
    M6 new_cov;
    double den = (tc0 + tc1) * (tc0 + tc1);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) new_cov(i, j) = (tc0 * tc0 * c1(i, j) + tc1 * tc1 * c0(i, j)) / den;
    nn->group_id = group_id;
    merged = true;
    nn->merged = true;
    map->stats.merges++;
    if (-dot(new_mean, new_norm) < 0.0) new_norm = -1.0 * new_norm;
    plane->mean = new_mean;
    plane->norm = new_norm;
    plane->cov = new_cov;
    nn->plane->mean = new_mean;
    nn->plane->norm = new_norm;
    nn->plane->cov = new_cov;
  }
}

// ---------------- VoxelMap (voxel_map.cpp:188-276) ----------------
VoxelMap::VoxelMap(int max_pt, int upd_pt, double pt, double vs, int cap, double ma, double md)
    : max_point_thresh(max_pt), update_point_thresh(upd_pt), plane_thresh(pt), voxel_size(vs), capacity(cap),
      merge_thresh_for_angle(ma), merge_thresh_for_distance(md) {}

VoxelKey VoxelMap::index(const V3& p) const {  // voxel_map.cpp:194-198
  return VoxelKey{int64_t(std::floor(p(0, 0) / voxel_size)), int64_t(std::floor(p(1, 0) / voxel_size)),
                  int64_t(std::floor(p(2, 0) / voxel_size))};
}

void VoxelMap::touch(const VoxelKey& k) {  // LRU handling shared by build/update
  auto it = featmap.find(k);
  if (it == featmap.end()) {
    auto g = std::make_shared<VoxelGrid>(max_point_thresh, update_point_thresh, plane_thresh, k, this);
    featmap[k] = g;
    cache.push_front(k);
    g->cache_it = cache.begin();
    stats.new_voxels++;
    if (int64_t(cache.size()) > int64_t(capacity)) {
      featmap.erase(cache.back());
      cache.pop_back();
      stats.evicted++;
    }
  } else {
    cache.splice(cache.begin(), cache, it->second->cache_it);
  }
}

void VoxelMap::build(std::vector<PointWithCov>& pvs) {  // voxel_map.cpp:200-230
  for (PointWithCov& pv : pvs) {
    VoxelKey k = index(pv.point);
    touch(k);
    featmap[k]->addPoint(pv);
  }
  for (auto& kv : featmap) kv.second->updatePlane();
}

void VoxelMap::update(std::vector<PointWithCov>& pvs) {  // voxel_map.cpp:232-256
  for (PointWithCov& pv : pvs) {
    VoxelKey k = index(pv.point);
    touch(k);
    featmap[k]->pushPoint(pv);
  }
}

bool VoxelMap::buildResidual(ResidualData& data, const std::shared_ptr<VoxelGrid>& g) {  // voxel_map.cpp:258-276
  data.is_valid = false;
  if (g->is_plane) {
    V3 p2m = data.point_world - g->plane->mean;
    data.plane_norm = g->plane->norm;
    data.plane_mean = g->plane->mean;
    data.residual = dot(data.plane_norm, p2m);
    Mat<1, 6> J_nq;
    for (int i = 0; i < 3; ++i) {
      J_nq(0, i) = p2m(i, 0);
      J_nq(0, 3 + i) = -data.plane_norm(i, 0);
    }
    double sigma_l = ((J_nq * data.plane_cov) * tr(J_nq))(0, 0);
    sigma_l += ((tr(data.plane_norm) * data.cov_world) * data.plane_norm)(0, 0);
    if (std::fabs(data.residual) < 3.0 * std::sqrt(sigma_l)) data.is_valid = true;
  }
  return data.is_valid;
}

// ---------------- State / IESKF (ieskf.cpp:11-156) ----------------
static V3 seg3(const V23& d, int o) { return v3(d(o, 0), d(o + 1, 0), d(o + 2, 0)); }

void State::plus(const V23& d) {  // ieskf.cpp:11-21
  pos = pos + seg3(d, 0);
  rot = rot * so3_exp(seg3(d, 3));
  rot_ext = rot_ext * so3_exp(seg3(d, 6));
  pos_ext = pos_ext + seg3(d, 9);
  vel = vel + seg3(d, 12);
  bg = bg + seg3(d, 15);
  ba = ba + seg3(d, 18);
  Mat<2, 1> r;
  r(0, 0) = d(21, 0);
  r(1, 0) = d(22, 0);
  g = so3_exp(getBx() * r) * g;
}

V23 State::minus(const State& o) const {  // ieskf.cpp:23-55
  V23 d = V23::zero();
  for (int i = 0; i < 3; ++i) d(i, 0) = pos(i, 0) - o.pos(i, 0);
  V3 r = so3_log(tr(o.rot) * rot);
  V3 re = so3_log(tr(o.rot_ext) * rot_ext);
  for (int i = 0; i < 3; ++i) {
    d(3 + i, 0) = r(i, 0);
    d(6 + i, 0) = re(i, 0);
    d(9 + i, 0) = pos_ext(i, 0) - o.pos_ext(i, 0);
    d(12 + i, 0) = vel(i, 0) - o.vel(i, 0);
    d(15 + i, 0) = bg(i, 0) - o.bg(i, 0);
    d(18 + i, 0) = ba(i, 0) - o.ba(i, 0);
  }
  V3 hg = hat(g) * o.g;
  double v_sin = std::sqrt(dot(hg, hg));
  double v_cos = dot(g, o.g);
  double theta = std::atan2(v_sin, v_cos);
  Mat<2, 1> res;
  if (v_sin < 1e-11) {
    if (std::fabs(theta) > 1e-11) {
      res(0, 0) = 3.1415926;
      res(1, 0) = 0;
    } else {
      res(0, 0) = 0;
      res(1, 0) = 0;
    }
  } else {
    res = (theta / v_sin) * ((tr(o.getBx()) * hat(o.g)) * g);
  }
  d(21, 0) = res(0, 0);
  d(22, 0) = res(1, 0);
  return d;
}

Mat<3, 2> State::getBx() const {  // ieskf.cpp:57-65
  Mat<3, 2> r;
  double g0 = g(0, 0), g1 = g(1, 0), g2 = g(2, 0);
  r(0, 0) = -g1;
  r(0, 1) = -g2;
  r(1, 0) = GRAVITY - g1 * g1 / (GRAVITY + g0);
  r(1, 1) = -g2 * g1 / (GRAVITY + g0);
  r(2, 0) = -g2 * g1 / (GRAVITY + g0);
  r(2, 1) = GRAVITY - g2 * g2 / (GRAVITY + g0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) r(i, j) /= GRAVITY;
  return r;
}

Mat<3, 2> State::getMx() const {  // ieskf.cpp:79-83
  return ((-1.0) * hat(g)) * getBx();
}

Mat<3, 2> State::getMx(const Mat<2, 1>& res) const {  // ieskf.cpp:73-78
  Mat<3, 2> bx = getBx();
  V3 bu = bx * res;
  return ((((-1.0) * so3_exp(bu)) * hat(g)) * tr(left_jacobian(bu))) * bx;
}

Mat<2, 3> State::getNx() const {  // ieskf.cpp:80-83
  return (1.0 / GRAVITY / GRAVITY) * (tr(getBx()) * hat(g));
}

static M23 inv23(const M23& A) {
  M23 out;
  invert(&A.a[0][0], &out.a[0][0], 23);
  return out;
}

static M23 jac23(const State& x, const State& xp, const V23& delta) {
  M23 J = M23::ident();
  M3 a = right_jacobian(seg3(delta, 3)), b = right_jacobian(seg3(delta, 6));
  Mat<2, 1> r;
  r(0, 0) = delta(21, 0);
  r(1, 0) = delta(22, 0);
  Mat<2, 2> c = x.getNx() * xp.getMx(r);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      J(3 + i, 3 + j) = a(i, j);
      J(6 + i, 6 + j) = b(i, j);
    }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) J(21 + i, 21 + j) = c(i, j);
  return J;
}

void IESKF::update() {  // ieskf.cpp:125-156
  State predict_x = x;
  SharedState sd;
  sd.iter_num = 0;
  V23 delta = V23::zero();
  M23 H_ = M23::zero();
  M23 Pinv = inv23(P);
  last_iters = 0;
  for (size_t i = 0; i < max_iter; ++i) {
    func(x, sd);
    delta = x.minus(predict_x);
    M23 J = jac23(x, predict_x, delta);
    M23 JtPi = tr(J) * Pinv;
    V23 b_ = JtPi * delta;
    H_ = JtPi * J;
    for (int r = 0; r < 12; ++r) {
      for (int c = 0; c < 12; ++c) H_(r, c) += sd.H(r, c);
      b_(r, 0) += sd.b(r, 0);
    }
    delta = (-1.0) * (inv23(H_) * b_);
    x.plus(delta);
    sd.iter_num += 1;
    last_iters++;
    double mx = delta(0, 0);
    for (int k = 1; k < 23; ++k) mx = std::max(mx, delta(k, 0));
    if (mx < eps) break;  // signed maxCoeff, as in the reference
  }
  M23 L = jac23(x, predict_x, delta);
  P = (L * inv23(H_)) * tr(L);
}

// ---------------- registration driver ----------------
Registrar::Registrar(const Config& c) : cfg(c) {
  map = std::make_shared<VoxelMap>(c.max_point_thresh, c.update_size_thresh, c.plane_thresh, c.voxel_size,
                                   c.map_capacity, c.merge_thresh_for_angle, c.merge_thresh_for_distance);
  kf.max_iter = size_t(c.max_iter);
  kf.func = [this](State& s, SharedState& d) { observe(s, d); };
}

// PCL transformPointCloud with an fp32 Matrix4f (lio_builder.cpp:155-163).
void lidar_to_world_f32(const State& x, const float* px, const float* py, const float* pz, int n,
                               std::vector<float>& ox, std::vector<float>& oy, std::vector<float>& oz) {
  M3 r = x.rot * x.rot_ext;
  V3 t = (x.rot * x.pos_ext) + x.pos;
  float m[3][3], tt[3];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) m[i][j] = float(r(i, j));
    tt[i] = float(t(i, 0));
  }
  ox.resize(n);
  oy.resize(n);
  oz.resize(n);
  for (int i = 0; i < n; ++i) {
    float o[3];
    for (int r2 = 0; r2 < 3; ++r2) {
      float a = m[r2][2] * pz[i] + tt[r2];
      float b = m[r2][1] * py[i] + a;
      o[r2] = m[r2][0] * px[i] + b;
    }
    ox[i] = o[0];
    oy[i] = o[1];
    oz[i] = o[2];
  }
}

static M3 block3(const M23& P, int o) {
  M3 b;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) b(i, j) = P(o + i, o + j);
  return b;
}

void Registrar::map_init(const float* x, const float* y, const float* z, int n) {  // lio_builder.cpp:185-211
  std::vector<float> wx, wy, wz;
  lidar_to_world_f32(kf.x, x, y, z, n, wx, wy, wz);
  M3 r_wl = kf.x.rot * kf.x.rot_ext;
  M3 Prr = block3(kf.P, 3), Ppp = block3(kf.P, 0);
  std::vector<PointWithCov> pv_list;
  pv_list.reserve(n);
  for (int i = 0; i < n; ++i) {
    PointWithCov pv;
    pv.point = v3(double(wx[i]), double(wy[i]), double(wz[i]));
    V3 pb = v3(double(x[i]), double(y[i]), double(z[i]));
    M3 pc;
    calc_body_cov(pb, cfg.ranging_cov, cfg.angle_cov, pc);
    M3 cm = hat(pb);
    pv.cov = (((r_wl * pc) * tr(r_wl)) + ((cm * Prr) * tr(cm))) + Ppp;
    pv_list.push_back(pv);
  }
  map->build(pv_list);
}

void Registrar::set_scan(const float* x, const float* y, const float* z, int n) {  // lio_builder.cpp:215-229
  cx.assign(x, x + n);
  cy.assign(y, y + n);
  cz.assign(z, z + n);
  size = n;
  if (int(residual_info.size()) < n) residual_info.resize(n);  // per-index slots persist
  for (int i = 0; i < n; ++i) {
    residual_info[i].point_lidar = v3(double(x[i]), double(y[i]), double(z[i]));
    calc_body_cov(residual_info[i].point_lidar, cfg.ranging_cov, cfg.angle_cov, residual_info[i].cov_lidar);
  }
}

void Registrar::observe(State& state, SharedState& sd) {  // lio_builder.cpp:250-311
  M3 r_wl = state.rot * state.rot_ext;
  V3 p_wl = (state.rot * state.pos_ext) + state.pos;
  M3 Prr = block3(kf.P, 3), Ppp = block3(kf.P, 0);
  for (int i = 0; i < size; ++i) {
    ResidualData& rd = residual_info[i];
    rd.point_world = (r_wl * rd.point_lidar) + p_wl;
    M3 cm = hat(rd.point_lidar);
    rd.cov_world = (((r_wl * rd.cov_lidar) * tr(r_wl)) + ((cm * Prr) * tr(cm))) + Ppp;
    VoxelKey k = map->index(rd.point_world);
    auto it = map->featmap.find(k);
    if (it != map->featmap.end()) map->buildResidual(rd, it->second);
  }
  sd.H = Mat<12, 12>::zero();
  sd.b = Mat<12, 1>::zero();
  effect_num = 0;
  for (int i = 0; i < size; ++i) {
    const ResidualData& rd = residual_info[i];
    if (!rd.is_valid) continue;
    effect_num++;
    Mat<1, 12> J = Mat<1, 12>::zero();
    V3 pn = rd.plane_norm;
    Mat<1, 6> Jn;
    V3 dpm = rd.point_world - rd.plane_mean;
    for (int k = 0; k < 3; ++k) {
      Jn(0, k) = dpm(k, 0);
      Jn(0, 3 + k) = -pn(k, 0);
    }
    double r_cov = ((Jn * rd.plane_cov) * tr(Jn))(0, 0);
    r_cov += ((((tr(pn) * r_wl) * rd.cov_lidar) * tr(r_wl)) * pn)(0, 0);
    double r_info = r_cov < 0.0002 ? 5000 : 1.0 / r_cov;
    Mat<1, 3> jr = ((-1.0 * tr(pn)) * state.rot) * hat((state.rot_ext * rd.point_lidar) + state.pos_ext);
    for (int k = 0; k < 3; ++k) {
      J(0, k) = pn(k, 0);
      J(0, 3 + k) = jr(0, k);
    }
    if (cfg.estimate_ext) {
      Mat<1, 3> je = ((-1.0 * tr(pn)) * r_wl) * hat(rd.point_lidar);
      Mat<1, 3> jp = tr(pn) * state.rot;
      for (int k = 0; k < 3; ++k) {
        J(0, 6 + k) = je(0, k);
        J(0, 9 + k) = jp(0, k);
      }
    }
    for (int a = 0; a < 12; ++a) {
      double ja = J(0, a) * r_info;
      for (int c = 0; c < 12; ++c) sd.H(a, c) += ja * J(0, c);
      sd.b(a, 0) += ja * rd.residual;
    }
  }
}

void Registrar::map_update() {  // lio_builder.cpp:231-246
  std::vector<float> wx, wy, wz;
  lidar_to_world_f32(kf.x, cx.data(), cy.data(), cz.data(), size, wx, wy, wz);
  M3 r_wl = kf.x.rot * kf.x.rot_ext;
  M3 Prr = block3(kf.P, 3), Ppp = block3(kf.P, 0);
  std::vector<PointWithCov> pv_list;
  pv_list.reserve(size);
  for (int i = 0; i < size; ++i) {
    PointWithCov pv;
    pv.point = v3(double(wx[i]), double(wy[i]), double(wz[i]));
    M3 cm = hat(residual_info[i].point_lidar);
    pv.cov = (((r_wl * residual_info[i].cov_lidar) * tr(r_wl)) + ((cm * Prr) * tr(cm))) + Ppp;
    pv_list.push_back(pv);
  }
  map->update(pv_list);
}


// ---------------- IMU propagation (ieskf.cpp:101-123) ----------------
// F and G are built from the state before the step; x += delta with the
// 24-vector whose rot_ext, pos_ext, bias and gravity parts are zero, so
// those stay exactly as they were.
void IESKF::predict(const V3& acc, const V3& gyro, double dt, const Mat<12, 12>& Q) {
  const V3 wb = gyro - x.bg;
  const V3 wdt = dt * wb;                               // (gyro - bg) * dt
  const V3 dpos = dt * x.vel;                           // vel * dt
  const V3 dvel = dt * ((x.rot * (acc - x.ba)) + x.g);  // (rot (acc - ba) + g) * dt
  M23 F = M23::ident();
  Mat<23, 12> G = Mat<23, 12>::zero();
  const M3 Jr = right_jacobian(wdt);
  const M3 Rexp = so3_exp((-1.0) * wdt);
  const M3 Ahat = ((-1.0) * x.rot) * hat(acc - x.ba);
  const Mat<3, 2> Mx = x.getMx();
  const Mat<2, 2> NM = x.getNx() * Mx;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      F(i, 12 + j) = (i == j ? 1.0 : 0.0) * dt;
      F(3 + i, 3 + j) = Rexp(i, j);
      F(3 + i, 15 + j) = ((-1.0) * Jr(i, j)) * dt;
      F(12 + i, 3 + j) = Ahat(i, j) * dt;
      F(12 + i, 18 + j) = ((-1.0) * x.rot(i, j)) * dt;
      G(3 + i, j) = ((-1.0) * Jr(i, j)) * dt;
      G(12 + i, 3 + j) = ((-1.0) * x.rot(i, j)) * dt;
      G(15 + i, 6 + j) = (i == j ? 1.0 : 0.0) * dt;
      G(18 + i, 9 + j) = (i == j ? 1.0 : 0.0) * dt;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) F(12 + i, 21 + j) = Mx(i, j) * dt;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) F(21 + i, 21 + j) = NM(i, j);
  x.pos = x.pos + dpos;
  x.rot = x.rot * so3_exp(wdt);
  x.rot_ext = x.rot_ext * so3_exp(V3::zero());
  x.vel = x.vel + dvel;
  x.g = so3_exp(V3::zero()) * x.g;
  P = (((F * P) * tr(F)) + ((G * Q) * tr(G)));
}

// Eigen's Quaternion::setFromTwoVectors for a, b not opposite: the quaternion
// half-way between them, then the rotation matrix of that quaternion.
M3 rot_from_two_vectors(const V3& a, const V3& b) {
  auto unit = [](const V3& v) {
    const double n = std::sqrt(dot(v, v));
    return v3(v(0, 0) / n, v(1, 0) / n, v(2, 0) / n);
  };
  const V3 v0 = unit(a), v1 = unit(b);
  const double c = dot(v1, v0);
  double qw, qx, qy, qz;
  if (c < -1.0 + 1e-12) {  // opposite: a half turn about an axis orthogonal to a
    V3 ax = std::fabs(v0(0, 0)) < 0.9 ? v3(0.0, -v0(2, 0), v0(1, 0)) : v3(-v0(2, 0), 0.0, v0(0, 0));
    ax = unit(ax);
    qw = 0.0;
    qx = ax(0, 0);
    qy = ax(1, 0);
    qz = ax(2, 0);
  } else {
    const V3 axis = v3(v0(1, 0) * v1(2, 0) - v0(2, 0) * v1(1, 0), v0(2, 0) * v1(0, 0) - v0(0, 0) * v1(2, 0),
                       v0(0, 0) * v1(1, 0) - v0(1, 0) * v1(0, 0));
    const double sq = std::sqrt((1.0 + c) * 2.0);
    const double invs = 1.0 / sq;
    qx = axis(0, 0) * invs;
    qy = axis(1, 0) * invs;
    qz = axis(2, 0) * invs;
    qw = sq * 0.5;
  }
  const double tx = 2 * qx, ty = 2 * qy, tz = 2 * qz;
  const double twx = tx * qw, twy = ty * qw, twz = tz * qw;
  const double txx = tx * qx, txy = ty * qx, txz = tz * qx;
  const double tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
  M3 R;
  R(0, 0) = 1 - (tyy + tzz);
  R(0, 1) = txy - twz;
  R(0, 2) = txz + twy;
  R(1, 0) = txy + twz;
  R(1, 1) = 1 - (txx + tzz);
  R(1, 2) = tyz - twx;
  R(2, 0) = txz - twy;
  R(2, 1) = tyz + twx;
  R(2, 2) = 1 - (txx + tyy);
  return R;
}

void voxel_grid_filter(const float* x, const float* y, const float* z, int n, float leaf, std::vector<float>& ox,
                       std::vector<float>& oy, std::vector<float>& oz) {
  ox.clear();
  oy.clear();
  oz.clear();
  if (n <= 0) return;
  const float* P[3] = {x, y, z};
  float mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = P[a][0];
    mx[a] = P[a][0];
    for (int i = 1; i < n; ++i) {
      mn[a] = std::min(mn[a], P[a][i]);
      mx[a] = std::max(mx[a], P[a][i]);
    }
  }
  const float inv = 1.0f / leaf;
  // index range check: too many leaves for 32-bit indices -> the cloud
  // passes unfiltered
  int64_t dxyz = 1;
  for (int a = 0; a < 3; ++a) dxyz *= int64_t((mx[a] - mn[a]) * inv) + 1;
  if (dxyz > int64_t(2147483647)) {
    ox.assign(x, x + n);
    oy.assign(y, y + n);
    oz.assign(z, z + n);
    return;
  }
  int minb[3], maxb[3], divb[3], mul[3];
  for (int a = 0; a < 3; ++a) {
    minb[a] = int(std::floor(mn[a] * inv));
    maxb[a] = int(std::floor(mx[a] * inv));
    divb[a] = maxb[a] - minb[a] + 1;
  }
  mul[0] = 1;
  mul[1] = divb[0];
  mul[2] = divb[0] * divb[1];
  std::vector<std::pair<int, int>> idx(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    int lin = 0;
    for (int a = 0; a < 3; ++a) {
      const int ijk = int(std::floor(P[a][i] * inv) - float(minb[a]));
      lin += ijk * mul[a];
    }
    idx[size_t(i)] = {lin, i};
  }
  std::stable_sort(idx.begin(), idx.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
    return a.first < b.first;
  });
  size_t q = 0;
  while (q < idx.size()) {
    size_t e = q + 1;
    while (e < idx.size() && idx[e].first == idx[q].first) ++e;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (size_t k = q; k < e; ++k) {
      sx += x[idx[k].second];
      sy += y[idx[k].second];
      sz += z[idx[k].second];
    }
    const float cnt = float(e - q);
    ox.push_back(sx / cnt);
    oy.push_back(sy / cnt);
    oz.push_back(sz / cnt);
    q = e;
  }
}

// ---------------- LIOBuilder IMU side (lio_builder.cpp:5-153) ----------------
void Registrar::set_imu_config(const ImuConfig& c) {  // lio_builder.cpp:9-12
  icfg = c;
  Q = Mat<12, 12>::ident();
  for (int i = 0; i < 3; ++i) {
    Q(i, i) = c.ng;
    Q(3 + i, 3 + i) = c.na;
    Q(6 + i, 6 + i) = c.nbg;
    Q(9 + i, 9 + i) = c.nba;
  }
}

bool Registrar::initialize_imu(const std::vector<ImuData>& imus) {  // lio_builder.cpp:28-63
  imu_cache.insert(imu_cache.end(), imus.begin(), imus.end());
  if (int(imu_cache.size()) < icfg.imu_init_num) return false;
  V3 acc_mean = V3::zero(), gyro_mean = V3::zero();
  for (const ImuData& d : imu_cache) {
    acc_mean = acc_mean + d.acc;
    gyro_mean = gyro_mean + d.gyro;
  }
  const double cnt = double(imu_cache.size());
  for (int i = 0; i < 3; ++i) {
    acc_mean(i, 0) = acc_mean(i, 0) / cnt;
    gyro_mean(i, 0) = gyro_mean(i, 0) / cnt;
  }
  gravity_norm = std::sqrt(dot(acc_mean, acc_mean));
  kf.x.rot_ext = icfg.r_il;
  kf.x.pos_ext = icfg.p_il;
  kf.x.bg = gyro_mean;
  auto init_g = [&](const V3& dir) {  // State::initG (ieskf.h:44-47)
    const double nn = std::sqrt(dot(dir, dir));
    kf.x.g = v3(dir(0, 0) / nn * GRAVITY, dir(1, 0) / nn * GRAVITY, dir(2, 0) / nn * GRAVITY);
  };
  if (icfg.gravity_align) {
    kf.x.rot = rot_from_two_vectors((-1.0) * acc_mean, v3(0.0, 0.0, -1.0));
    init_g(v3(0.0, 0.0, -1.0));
  } else {
    init_g((-1.0) * acc_mean);
  }
  kf.P = M23::ident();
  for (int i = 0; i < 3; ++i) {
    kf.P(6 + i, 6 + i) = 0.00001;
    kf.P(9 + i, 9 + i) = 0.00001;
    kf.P(15 + i, 15 + i) = 0.0001;
    kf.P(18 + i, 18 + i) = 0.0001;
  }
  kf.P(21, 21) = 0.00001;
  kf.P(22, 22) = 0.00001;
  last_imu = imus.back();
  return true;
}

void Registrar::undistort(std::vector<float>& x, std::vector<float>& y, std::vector<float>& z,
                          std::vector<float>& t_ms, const std::vector<ImuData>& imus, double cloud_start,
                          double cloud_end) {  // lio_builder.cpp:65-153
  imu_cache.clear();
  imu_cache.push_back(last_imu);
  imu_cache.insert(imu_cache.end(), imus.begin(), imus.end());
  const double imu_time_end = imu_cache.back().t;
  const size_t n = x.size();
  {  // points by time (the generator's times are unique, so any sort agrees)
    std::vector<size_t> ord(n);
    for (size_t i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return t_ms[a] < t_ms[b]; });
    std::vector<float> a(n), b(n), c(n), d(n);
    for (size_t i = 0; i < n; ++i) {
      a[i] = x[ord[i]];
      b[i] = y[ord[i]];
      c[i] = z[ord[i]];
      d[i] = t_ms[ord[i]];
    }
    x.swap(a);
    y.swap(b);
    z.swap(c);
    t_ms.swap(d);
  }
  imu_poses.clear();
  imu_poses.push_back(ImuPose{0.0, last_acc, last_gyro, kf.x.vel, kf.x.pos, kf.x.rot});
  V3 acc_val = V3::zero(), gyro_val = V3::zero();
  // deviation: the reference predicts its final step with an uninitialised
  // input when no IMU pair falls after the last scan end; zero here
  V3 inp_acc = V3::zero(), inp_gyro = V3::zero();
  double dt = 0.0;
  for (size_t k = 0; k + 1 < imu_cache.size(); ++k) {
    const ImuData& head = imu_cache[k];
    const ImuData& tail = imu_cache[k + 1];
    if (tail.t < last_cloud_end_time) continue;
    gyro_val = 0.5 * (head.gyro + tail.gyro);
    acc_val = 0.5 * (head.acc + tail.acc);
    for (int i = 0; i < 3; ++i) acc_val(i, 0) = acc_val(i, 0) * 9.81 / gravity_norm;
    if (head.t < last_cloud_end_time) dt = tail.t - last_cloud_end_time;
    else dt = tail.t - head.t;
    inp_acc = acc_val;
    inp_gyro = gyro_val;
    kf.predict(inp_acc, inp_gyro, dt, Q);
    last_gyro = gyro_val - kf.x.bg;
    last_acc = (kf.x.rot * (acc_val - kf.x.ba)) + kf.x.g;
    const double offset = tail.t - cloud_start;
    imu_poses.push_back(ImuPose{offset, last_acc, last_gyro, kf.x.vel, kf.x.pos, kf.x.rot});
  }
  dt = cloud_end - imu_time_end;
  kf.predict(inp_acc, inp_gyro, dt, Q);
  last_imu = imus.back();
  last_cloud_end_time = cloud_end;
  const M3 cur_rot = kf.x.rot, cur_rot_ext = kf.x.rot_ext;
  const V3 cur_pos = kf.x.pos, cur_pos_ext = kf.x.pos_ext;
  if (n == 0) return;
  size_t ip = n - 1;
  for (size_t kp = imu_poses.size() - 1; kp > 0; --kp) {
    const ImuPose& head = imu_poses[kp - 1];
    const ImuPose& tail = imu_poses[kp];
    for (; double(t_ms[ip]) / double(1000) > head.offset; ip--) {
      dt = double(t_ms[ip]) / double(1000) - head.offset;
      const V3 point = v3(double(x[ip]), double(y[ip]), double(z[ip]));
      const M3 point_rot = head.rot * so3_exp(dt * tail.gyro);
      const V3 point_pos = (head.pos + dt * head.vel) + dt * (dt * (0.5 * tail.acc));
      const V3 pc = tr(cur_rot_ext) *
                    ((tr(cur_rot) * (((point_rot * ((cur_rot_ext * point) + cur_pos_ext)) + point_pos) - cur_pos)) -
                     cur_pos_ext);
      x[ip] = float(pc(0, 0));
      y[ip] = float(pc(1, 0));
      z[ip] = float(pc(2, 0));
      if (ip == 0) break;
    }
  }
}

LioStatus Registrar::process(std::vector<float> x, std::vector<float> y, std::vector<float> z, std::vector<float> t_ms,
                             const std::vector<ImuData>& imus, double cloud_start, double cloud_end) {
  if (status == LioStatus::IMU_INIT) {  // lio_builder.cpp:177-184
    if (initialize_imu(imus)) {
      status = LioStatus::MAP_INIT;
      last_cloud_end_time = cloud_end;
    }
  } else if (status == LioStatus::MAP_INIT) {  // lio_builder.cpp:185-211
    undistort(x, y, z, t_ms, imus, cloud_start, cloud_end);
    map_init(x.data(), y.data(), z.data(), int(x.size()));
    status = LioStatus::LIO_MAPPING;
  } else {  // lio_builder.cpp:212-247
    undistort(x, y, z, t_ms, imus, cloud_start, cloud_end);
    std::vector<float> fx, fy, fz;
    if (cfg.scan_resolution > 0.0) {
      voxel_grid_filter(x.data(), y.data(), z.data(), int(x.size()), float(cfg.scan_resolution), fx, fy, fz);
    } else {
      fx = x;
      fy = y;
      fz = z;
    }
    set_scan(fx.data(), fy.data(), fz.data(), int(fx.size()));
    kf.update();
    map_update();
  }
  return status;
}

}  // namespace orc
