// ORACLE — test infrastructure only.  A CPU, FP64, Eigen-free restatement of
// the reference's per-scan registration path (voxel_plus), used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  The
// product (voxelmapplus_fastlio2_amd) never links or calls this code.
//
// Followed files (paths relative to the reference root):
//   voxel_plus/src/map_builder/voxel_map.{h,cpp}   whole
//   voxel_plus/src/map_builder/ieskf.{h,cpp}       State ops, update()
//   voxel_plus/src/map_builder/lio_builder.cpp:5-26, 175-311
//   voxel_plus/src/map_builder/commons.cpp:18-45   calcBodyCov
// Quirks reproduced on purpose (see the quirk ledger in the design notes):
//   plane_cov in the residual slots is never written (stays 0); stale slots on
//   a map miss; merge() precedence bug and non-renormalised normal;
//   never-reset Plane::cov; signed maxCoeff convergence; fixed 5 iterations;
//   calcBodyCov's z==0 -> 0.001 mutation; fp32 insertion transform.
// Deviations: residual slots sized to the scan (not a fixed 10000);
//   group-id counter and merge thresholds are per map, not process-global.
// Unpinned choices (no reference test pins them; the reference cannot be built
// here because Eigen/Sophus/PCL/ROS are absent): summation order of small
// products is left-to-right over the inner index; PCL's fp32 transform is
// x' = m0*x + (m1*y + (m2*z + t)); DEG2RAD uses pi/180; SO3 exp/log follow
// Rodrigues with Taylor branches below 1e-10 rad (exp, log) and a first-order
// left Jacobian below theta^2 = 1e-10 (Sophus' epsilon); 23x23 inverse is LU with
// partial pivoting; 3x3 eigenvectors come from cyclic Jacobi.
#pragma once
#include <cstdint>
#include <functional>
#include <list>
#include <memory>
#include <unordered_map>
#include <vector>

namespace orc {

template <int R, int C>
struct Mat {
  double a[R][C];
  double& operator()(int i, int j) { return a[i][j]; }
  double operator()(int i, int j) const { return a[i][j]; }
  static Mat zero() {
    Mat m;
    for (int i = 0; i < R; ++i)
      for (int j = 0; j < C; ++j) m.a[i][j] = 0.0;
    return m;
  }
  static Mat ident() {
    Mat m = zero();
    for (int i = 0; i < (R < C ? R : C); ++i) m.a[i][i] = 1.0;
    return m;
  }
};
using V3 = Mat<3, 1>;
using M3 = Mat<3, 3>;
using M6 = Mat<6, 6>;
using M23 = Mat<23, 23>;
using V23 = Mat<23, 1>;

template <int R, int K, int C>
inline Mat<R, C> operator*(const Mat<R, K>& A, const Mat<K, C>& B) {
  Mat<R, C> out;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) {
      double s = A.a[i][0] * B.a[0][j];
      for (int k = 1; k < K; ++k) s += A.a[i][k] * B.a[k][j];
      out.a[i][j] = s;
    }
  return out;
}
template <int R, int C>
inline Mat<R, C> operator+(const Mat<R, C>& A, const Mat<R, C>& B) {
  Mat<R, C> o;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) o.a[i][j] = A.a[i][j] + B.a[i][j];
  return o;
}
template <int R, int C>
inline Mat<R, C> operator-(const Mat<R, C>& A, const Mat<R, C>& B) {
  Mat<R, C> o;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) o.a[i][j] = A.a[i][j] - B.a[i][j];
  return o;
}
template <int R, int C>
inline Mat<R, C> operator*(double s, const Mat<R, C>& A) {
  Mat<R, C> o;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) o.a[i][j] = s * A.a[i][j];
  return o;
}
template <int R, int C>
inline Mat<C, R> tr(const Mat<R, C>& A) {
  Mat<C, R> o;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) o.a[j][i] = A.a[i][j];
  return o;
}
inline V3 v3(double x, double y, double z) {
  V3 v;
  v.a[0][0] = x;
  v.a[1][0] = y;
  v.a[2][0] = z;
  return v;
}
inline double dot(const V3& a, const V3& b) {
  return a.a[0][0] * b.a[0][0] + a.a[1][0] * b.a[1][0] + a.a[2][0] * b.a[2][0];
}
M3 hat(const V3& v);

// ---- math (own implementations, named methods) ----
void eigh3(const M3& A, double evals[3], M3& evecs);  // cyclic Jacobi, ascending
bool invert(const double* A, double* Ainv, int n);     // LU, partial pivoting
M3 so3_exp(const V3& w);
V3 so3_log(const M3& R);
M3 left_jacobian(const V3& w);
inline M3 right_jacobian(const V3& w) { return tr(left_jacobian(w)); }  // ieskf.cpp:5-8
void calc_body_cov(V3& pb, double range_inc, double degree_inc, M3& cov);

// ---- voxel map (voxel_map.h:17-128) ----
struct VoxelKey {
  int64_t x, y, z;
  bool operator==(const VoxelKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct KeyHasher {  // voxel_map.h:27-33 (bucket placement only)
  size_t operator()(const VoxelKey& k) const {
    const int64_t P = 116101, N = 10000000000LL;
    return size_t((((k.z) * P) % N + (k.y)) * P % N + (k.x));
  }
};
struct PointWithCov {
  V3 point;
  M3 cov;
};
struct Plane {
  V3 mean = V3::zero();
  M3 ppt = M3::zero();
  V3 norm = V3::zero();
  M6 cov = M6::zero();
  int n = 0;
};
struct ResidualData {
  V3 point_lidar = V3::zero(), point_world = V3::zero(), plane_mean = V3::zero(), plane_norm = V3::zero();
  M6 plane_cov = M6::zero();  // never written (stays zero)
  M3 cov_lidar = M3::zero(), cov_world = M3::zero();
  bool is_valid = false;
  double residual = 0.0;
};

struct MapStats {
  int64_t new_voxels = 0, evicted = 0, refits = 0, merge_calls = 0, merges = 0, dropped = 0;
};

class VoxelMap;
class VoxelGrid {
 public:
  VoxelGrid(int max_pt, int upd_pt, double plane_thresh, VoxelKey pos, VoxelMap* map);
  void updatePlane();
  void addToPlane(const PointWithCov& pv);
  void addPoint(const PointWithCov& pv);
  void pushPoint(const PointWithCov& pv);
  void merge();

  int max_point_thresh, update_point_thresh;
  double plane_thresh;
  bool is_init = false, is_plane = false, update_enable = true, merged = false;
  int newly_add_point = 0;
  uint64_t group_id;
  std::vector<PointWithCov> temp_points;
  VoxelKey position;
  VoxelMap* map;
  std::shared_ptr<Plane> plane;
  V3 center = V3::zero();
  std::list<VoxelKey>::iterator cache_it;
  // eigenvalues and major axis of the last refit's covariance: test
  // diagnostics only (the conditioning of the fitted normal), no part of
  // the restated algorithm
  double fit_evals[3] = {0.0, 0.0, 0.0};
  V3 fit_major = V3::zero();
};

class VoxelMap {
 public:
  VoxelMap(int max_pt, int upd_pt, double plane_thresh, double voxel_size, int capacity, double merge_a,
           double merge_d);
  VoxelKey index(const V3& p) const;
  void build(std::vector<PointWithCov>& pvs);
  void update(std::vector<PointWithCov>& pvs);
  bool buildResidual(ResidualData& d, const std::shared_ptr<VoxelGrid>& g);

  int max_point_thresh, update_point_thresh;
  double plane_thresh, voxel_size;
  std::unordered_map<VoxelKey, std::shared_ptr<VoxelGrid>, KeyHasher> featmap;
  std::list<VoxelKey> cache;
  int capacity;
  uint64_t group_counter = 0;  // VoxelGrid::count, per map
  double merge_thresh_for_angle, merge_thresh_for_distance;
  MapStats stats;

 private:
  void touch(const VoxelKey& k);
};

// ---- IESKF (ieskf.h:22-110) ----
constexpr double GRAVITY = 9.81;
struct State {
  V3 pos = V3::zero();
  M3 rot = M3::ident();
  M3 rot_ext = M3::ident();
  V3 pos_ext = V3::zero();
  V3 vel = V3::zero();
  V3 bg = V3::zero();
  V3 ba = V3::zero();
  V3 g = v3(0.0, 0.0, -GRAVITY);
  void plus(const V23& d);        // operator+=(Vector23d)
  V23 minus(const State& o) const;  // operator-
  Mat<3, 2> getBx() const;
  Mat<3, 2> getMx() const;  // -hat(g) Bx (ieskf.cpp:79-83)
  Mat<3, 2> getMx(const Mat<2, 1>& res) const;
  Mat<2, 3> getNx() const;
};
struct SharedState {
  Mat<12, 12> H;
  Mat<12, 1> b;
  size_t iter_num = 0;
};
class IESKF {
 public:
  State x;
  M23 P = M23::ident();
  std::function<void(State&, SharedState&)> func;
  size_t max_iter = 5;
  double eps = 0.001;
  int last_iters = 0;
  void update();
  // IMU propagation (ieskf.cpp:101-123); Q is the 12x12 noise of LIODataGroup
  void predict(const V3& acc, const V3& gyro, double dt, const Mat<12, 12>& Q);
};

// ---- IMU side of LIOBuilder (lio_builder.cpp:28-153) ----
struct ImuData {  // commons.h:12-20
  V3 acc, gyro;
  double t;
};
struct ImuPose {  // commons.h:30-43
  double offset;
  V3 acc, gyro, vel, pos;
  M3 rot;
};
struct ImuConfig {  // LIOConfig fields of the IMU side (lio_builder.h:19-28)
  double na = 0.01, ng = 0.01, nba = 0.0001, nbg = 0.0001;
  int imu_init_num = 20;
  bool gravity_align = true;
  M3 r_il = M3::ident();
  V3 p_il = V3::zero();
};
// Eigen::Quaterniond::FromTwoVectors(a, b).matrix(): the rotation by the
// smallest angle that takes direction a to direction b (half-way quaternion
// w = s/2, v = (a x b)/s, s = sqrt(2 (1 + a.b)) on the normalised vectors).
M3 rot_from_two_vectors(const V3& a, const V3& b);

// PCL VoxelGrid<PointXYZINormal> (lio_builder.cpp:13-14, 215-223) on xyz:
// leaf indices from fp32 floor(p * (1/leaf)) minus the cloud's minimum
// index, points sorted by the linear index (ties in input order), every
// occupied leaf emitted as the fp32 mean of its points in that order.
void voxel_grid_filter(const float* x, const float* y, const float* z, int n, float leaf, std::vector<float>& ox,
                       std::vector<float>& oy, std::vector<float>& oz);

// PCL transformPointCloud with the fp32 Matrix4f of lidarToWorld
// (lio_builder.cpp:155-163): x' = m0*x + (m1*y + (m2*z + t)) in fp32.
void lidar_to_world_f32(const State& x, const float* px, const float* py, const float* pz, int n,
                        std::vector<float>& ox, std::vector<float>& oy, std::vector<float>& oz);

// ---- registration driver (LIOBuilder restricted to the path) ----
struct Config {
  double voxel_size = 0.5;
  int max_point_thresh = 100;
  int update_size_thresh = 10;
  double plane_thresh = 0.01;
  double merge_thresh_for_angle = 0.2;
  double merge_thresh_for_distance = 0.04;
  double ranging_cov = 0.04;
  double angle_cov = 0.1;
  int map_capacity = 100000;
  int max_iter = 5;
  int estimate_ext = 0;
  double scan_resolution = 0.0;  // VoxelGrid leaf of LIO_MAPPING scans (<= 0: none)
};

enum class LioStatus { IMU_INIT, MAP_INIT, LIO_MAPPING };  // lio_builder.h:10-16

class Registrar {
 public:
  explicit Registrar(const Config& c);
  // MAP_INIT (lio_builder.cpp:185-211): fp32 world points, cov with current P
  void map_init(const float* x, const float* y, const float* z, int n);
  // residual-slot prep (lio_builder.cpp:225-229)
  void set_scan(const float* x, const float* y, const float* z, int n);
  void observe(State& s, SharedState& d);  // sharedUpdateFunc (lio_builder.cpp:250-311)
  void map_update();                        // lio_builder.cpp:231-246
  // IMU side (lio_builder.cpp:28-153)
  void set_imu_config(const ImuConfig& c);                 // lio_builder.cpp:9-12
  bool initialize_imu(const std::vector<ImuData>& imus);   // lio_builder.cpp:28-63
  // predict over the package's IMU samples, then backward compensation of
  // the points (sorted by time, t_ms per point) into the frame at cloud_end
  void undistort(std::vector<float>& x, std::vector<float>& y, std::vector<float>& z, std::vector<float>& t_ms,
                 const std::vector<ImuData>& imus, double cloud_start, double cloud_end);
  // one SyncPackage through the state machine of LIOBuilder::process
  // (lio_builder.cpp:175-248); returns the status after it
  LioStatus process(std::vector<float> x, std::vector<float> y, std::vector<float> z, std::vector<float> t_ms,
                    const std::vector<ImuData>& imus, double cloud_start, double cloud_end);
  Config cfg;
  ImuConfig icfg;
  Mat<12, 12> Q = Mat<12, 12>::ident();
  std::vector<ImuData> imu_cache;
  std::vector<ImuPose> imu_poses;
  ImuData last_imu{};
  V3 last_acc = V3::zero(), last_gyro = V3::zero();
  double last_cloud_end_time = 0.0;
  double gravity_norm = 0.0;
  LioStatus status = LioStatus::IMU_INIT;
  IESKF kf;
  std::shared_ptr<VoxelMap> map;
  std::vector<ResidualData> residual_info;
  std::vector<float> cx, cy, cz;  // current scan (lidar frame, unmutated)
  int size = 0;
  int effect_num = 0;
};

}  // namespace orc
