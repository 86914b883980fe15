// Seeded synthetic LiDAR sequences (room / street) for the registration path.
// Deterministic: every random draw is a counter-based hash of (seed, stream,
// index), so results do not depend on thread count or call order.
#include "vmp_synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kScanDt = 0.1;        // 10 Hz scans
constexpr int kSubSteps = 20;          // Simpson sub-steps per scan interval
constexpr double kMaxRange = 40.0;
constexpr double kMinRange = 0.1;
constexpr double kRangeSigma = 0.02;

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(mix64(mix64(seed) ^ a) ^ (b * 0xD6E8FEB86659FD93ull));
}
inline double u01(uint64_t h) { return double(h >> 11) * (1.0 / 9007199254740992.0); }
inline double gauss(uint64_t h1, uint64_t h2) {
  double u1 = 1.0 - u01(h1);  // (0, 1]
  double u2 = u01(h2);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

struct Rng {  // sequential stream for scene layout
  uint64_t s;
  double uni() { s = mix64(s); return u01(s); }
  double uni(double a, double b) { return a + (b - a) * uni(); }
};

struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }

struct M3 { double m[3][3]; };  // row-major internally
inline V3 mul(const M3& R, V3 v) {
  return {R.m[0][0] * v.x + R.m[0][1] * v.y + R.m[0][2] * v.z,
          R.m[1][0] * v.x + R.m[1][1] * v.y + R.m[1][2] * v.z,
          R.m[2][0] * v.x + R.m[2][1] * v.y + R.m[2][2] * v.z};
}
inline M3 mulm(const M3& A, const M3& B) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
  return C;
}
inline M3 transpose(const M3& A) {
  M3 T;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T.m[i][j] = A.m[j][i];
  return T;
}
M3 rpy(double roll, double pitch, double yaw) {
  double cr = std::cos(roll), sr = std::sin(roll);
  double cp = std::cos(pitch), sp = std::sin(pitch);
  double cy = std::cos(yaw), sy = std::sin(yaw);
  M3 Rz{{{cy, -sy, 0}, {sy, cy, 0}, {0, 0, 1}}};
  M3 Ry{{{cp, 0, sp}, {0, 1, 0}, {-sp, 0, cp}}};
  M3 Rx{{{1, 0, 0}, {0, cr, -sr}, {0, sr, cr}}};
  return mulm(Rz, mulm(Ry, Rx));
}
M3 expso3(V3 w) {  // Rodrigues
  double th = std::sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
  M3 K{{{0, -w.z, w.y}, {w.z, 0, -w.x}, {-w.y, w.x, 0}}};
  double a, b;
  if (th < 1e-8) { a = 1.0; b = 0.5; }
  else { a = std::sin(th) / th; b = (1.0 - std::cos(th)) / (th * th); }
  M3 K2 = mulm(K, K);
  M3 R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R.m[i][j] = (i == j ? 1.0 : 0.0) + a * K.m[i][j] + b * K2.m[i][j];
  return R;
}

V3 logso3(const M3& R) {  // rotation vector of a rotation matrix
  const V3 v{R.m[2][1] - R.m[1][2], R.m[0][2] - R.m[2][0], R.m[1][0] - R.m[0][1]};
  double c = 0.5 * (R.m[0][0] + R.m[1][1] + R.m[2][2] - 1.0);
  c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
  const double th = std::acos(c);
  const double f = th < 1e-8 ? 0.5 : th / (2.0 * std::sin(th));
  return f * v;
}

struct OBox {  // box with yaw about its vertical axis
  V3 c;        // centre
  V3 h;        // half extents
  double cy, sy;
};

struct Pole { double x, y, r, z0, z1; };

bool slab(V3 o, V3 d, V3 lo, V3 hi, double& tnear) {
  double t0 = -1e30, t1 = 1e30;
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  const double l[3] = {lo.x, lo.y, lo.z}, u[3] = {hi.x, hi.y, hi.z};
  for (int a = 0; a < 3; ++a) {
    if (std::fabs(dd[a]) < 1e-15) {
      if (oo[a] < l[a] || oo[a] > u[a]) return false;
    } else {
      double ta = (l[a] - oo[a]) / dd[a], tb = (u[a] - oo[a]) / dd[a];
      if (ta > tb) std::swap(ta, tb);
      if (ta > t0) t0 = ta;
      if (tb < t1) t1 = tb;
      if (t0 > t1) return false;
    }
  }
  if (t1 < 1e-6) return false;
  tnear = t0 > 1e-6 ? t0 : t1;  // inside the box -> exit point
  return true;
}

}  // namespace

// IMU-mode clock: the sensor rests for kRest seconds (IMU initialisation
// needs a static start), then its speed ramps smoothly to the trajectory's
// 1 m/s over kRamp seconds; the path itself is the trajectory's, traversed
// at path parameter tau(t).
constexpr double kRest = 0.5;
constexpr double kRamp = 1.0;
constexpr double kImuDt = 0.005;  // 200 Hz

struct vmp_synth {
  int scene;
  uint64_t seed;
  int n;
  int imu_mode = 0;
  V3 offset;
  std::vector<V3> bg, ba;  // IMU biases per sample (random walks)
  std::vector<OBox> boxes;
  std::vector<OBox> facades;
  std::vector<Pole> poles;
  std::vector<V3> xy_table;  // integrated planar position at each scan end

  double heading(double t) const {
    if (scene == VMP_SCENE_ROOM) return 0.2 * t + 0.15 * std::sin(0.5 * t);
    return 0.05 * std::sin(0.3 * t);
  }
  V3 start() const {
    if (scene == VMP_SCENE_ROOM) return {0.0, -5.0, 0.0};
    return {-90.0, 0.0, 0.0};
  }
  double ground(double x, double y) const { return 0.02 * x + 0.005 * y; }

  void build_trajectory(int max_scans) {
    xy_table.resize(size_t(max_scans) + 1);
    V3 p = start();
    xy_table[0] = p;
    double h = kScanDt / kSubSteps;
    for (int k = 0; k < max_scans; ++k) {
      double t0 = k * kScanDt;
      for (int s = 0; s < kSubSteps; ++s) {
        double a = t0 + s * h, m = a + 0.5 * h, b = a + h;
        double cx = (std::cos(heading(a)) + 4.0 * std::cos(heading(m)) + std::cos(heading(b))) / 6.0;
        double cy = (std::sin(heading(a)) + 4.0 * std::sin(heading(m)) + std::sin(heading(b))) / 6.0;
        p.x += h * cx;  // 1 m/s
        p.y += h * cy;
      }
      xy_table[k + 1] = p;
    }
  }

  // path parameter and its first two derivatives at clock time t
  void tau_at(double t, double& tau, double& s, double& sd) const {
    if (!imu_mode) {
      tau = t;
      s = 1.0;
      sd = 0.0;
      return;
    }
    if (t <= kRest) {
      tau = 0.0;
      s = 0.0;
      sd = 0.0;
      return;
    }
    const double u = (t - kRest) / kRamp;
    if (u >= 1.0) {
      tau = 0.5 * kRamp + (t - kRest - kRamp);  // the ramp covers half its length
      s = 1.0;
      sd = 0.0;
      return;
    }
    // speed 3u^2 - 2u^3 (smoothstep); its integral over [0, u] times kRamp
    tau = kRamp * (u * u * u - 0.5 * u * u * u * u);
    s = 3.0 * u * u - 2.0 * u * u * u;
    sd = (6.0 * u - 6.0 * u * u) / kRamp;
  }
  double heading_rate(double tau) const {
    if (scene == VMP_SCENE_ROOM) return 0.2 + 0.075 * std::cos(0.5 * tau);
    return 0.015 * std::cos(0.3 * tau);
  }
  // planar position at path parameter tau (table entries every 0.1, then
  // Simpson over the remainder)
  V3 xy_at(double tau) const {
    int k0 = int(std::floor(tau / kScanDt));
    if (k0 < 0) k0 = 0;
    if (size_t(k0) >= xy_table.size()) k0 = int(xy_table.size()) - 1;
    V3 p = xy_table[size_t(k0)];
    const double a0 = k0 * kScanDt, len = tau - a0;
    if (len > 0.0) {
      const int ns = 8;
      const double h = len / ns;
      for (int q = 0; q < ns; ++q) {
        const double a = a0 + q * h, m = a + 0.5 * h, b = a + h;
        p.x += h * (std::cos(heading(a)) + 4.0 * std::cos(heading(m)) + std::cos(heading(b))) / 6.0;
        p.y += h * (std::sin(heading(a)) + 4.0 * std::sin(heading(m)) + std::sin(heading(b))) / 6.0;
      }
    }
    return p;
  }
  // pose (world <- lidar) at path parameter tau, local frame
  void pose_tau(double tau, M3& R, V3& t) const {
    V3 p = xy_at(tau);
    const double roll = 0.03 * std::sin(0.9 * tau + 0.3), pitch = 0.03 * std::sin(0.6 * tau + 1.1);
    if (scene == VMP_SCENE_ROOM) p.z = 1.2 + 0.05 * std::sin(0.7 * tau);
    else p.z = ground(p.x, p.y) + 2.0;
    R = rpy(roll, pitch, heading(tau));
    t = p;
  }
  void pose_clock(double t, M3& R, V3& p) const {
    double tau, sp, sd;
    tau_at(t, tau, sp, sd);
    pose_tau(tau, R, p);
  }
  // world acceleration at clock time t (second derivative of the position)
  V3 accel_clock(double t) const {
    double tau, sp, sd;
    tau_at(t, tau, sp, sd);
    const double h = heading(tau), hr = heading_rate(tau);
    const double c = std::cos(h), sn = std::sin(h);
    V3 a{sd * c - sp * sp * hr * sn, sd * sn + sp * sp * hr * c, 0.0};
    if (scene == VMP_SCENE_ROOM) {
      const double dz = 0.035 * std::cos(0.7 * tau), ddz = -0.0245 * std::sin(0.7 * tau);
      a.z = sd * dz + sp * sp * ddz;
    } else {
      a.z = 0.02 * a.x + 0.005 * a.y;
    }
    return a;
  }

  // pose at time t = kScanDt * (k + 1), local (un-offset) frame
  void pose_local(int k, M3& R, V3& t) const {
    double tm = kScanDt * (k + 1);
    V3 p = xy_table[size_t(k) + 1];
    double roll = 0.03 * std::sin(0.9 * tm + 0.3), pitch = 0.03 * std::sin(0.6 * tm + 1.1);
    if (scene == VMP_SCENE_ROOM) p.z = 1.2 + 0.05 * std::sin(0.7 * tm);
    else p.z = ground(p.x, p.y) + 2.0;
    R = rpy(roll, pitch, heading(tm));
    t = p;
  }

  void build_scene() {
    Rng rng{mix64(seed ^ 0xA11CEull)};
    if (scene == VMP_SCENE_ROOM) {
      // path samples for keep-out
      std::vector<V3> path;
      for (size_t k = 0; k < xy_table.size(); k += 5) path.push_back(xy_table[k]);
      while (boxes.size() < 40) {
        OBox b;
        b.h = {0.5 * rng.uni(0.5, 3.0), 0.5 * rng.uni(0.5, 3.0), 0.5 * rng.uni(0.5, 3.0)};
        b.c = {rng.uni(-13.0, 13.0), rng.uni(-13.0, 13.0), b.h.z};
        double yaw = rng.uni(0.0, kPi);
        b.cy = std::cos(yaw);
        b.sy = std::sin(yaw);
        double rad = std::sqrt(b.h.x * b.h.x + b.h.y * b.h.y) + 1.5;
        bool ok = true;
        for (const V3& q : path) {
          double dx = q.x - b.c.x, dy = q.y - b.c.y;
          if (dx * dx + dy * dy < rad * rad) { ok = false; break; }
        }
        if (ok) boxes.push_back(b);
      }
    } else {
      for (int side = -1; side <= 1; side += 2) {
        double x = -110.0;
        while (x < 110.0) {
          double len = rng.uni(10.0, 25.0);
          bool gap = rng.uni() < 0.2;
          double setback = rng.uni(8.0, 12.0), height = rng.uni(6.0, 25.0);
          if (!gap) {
            OBox f;
            double y0 = side * setback, y1 = side * (setback + 12.0);
            f.c = {x + 0.5 * len, 0.5 * (y0 + y1), 0.5 * (-5.0 + height)};
            f.h = {0.5 * len, 6.0, 0.5 * (height + 5.0)};
            f.cy = 1.0;
            f.sy = 0.0;
            facades.push_back(f);
          }
          x += len;
        }
        for (double px = -108.0; px < 110.0; px += 12.0) {
          Pole p;
          p.x = px + rng.uni(-1.0, 1.0);
          p.y = side * 6.5;
          p.r = 0.3;
          p.z0 = -5.0;
          p.z1 = ground(p.x, p.y) + 6.0;
          poles.push_back(p);
        }
      }
    }
  }

  // nearest hit distance along unit ray (o, d) in the local frame; <0 = miss
  double cast(V3 o, V3 d) const {
    double best = 1e30;
    if (scene == VMP_SCENE_ROOM) {
      double t;
      if (slab(o, d, {-15, -15, 0}, {15, 15, 6}, t)) best = t;  // interior exit
      for (const OBox& b : boxes) {
        V3 lo = o - b.c;
        V3 ol{b.cy * lo.x + b.sy * lo.y, -b.sy * lo.x + b.cy * lo.y, lo.z};
        V3 dl{b.cy * d.x + b.sy * d.y, -b.sy * d.x + b.cy * d.y, d.z};
        if (slab(ol, dl, {-b.h.x, -b.h.y, -b.h.z}, {b.h.x, b.h.y, b.h.z}, t) && t < best) best = t;
      }
    } else {
      // ground plane z = 0.02 x + 0.005 y
      double nx = -0.02, ny = -0.005, nz = 1.0;
      double den = nx * d.x + ny * d.y + nz * d.z;
      if (std::fabs(den) > 1e-12) {
        double t = -(nx * o.x + ny * o.y + nz * o.z) / den;
        if (t > 1e-6) {
          V3 q = o + t * d;
          if (std::fabs(q.x) <= 110.0 && std::fabs(q.y) <= 40.0 && t < best) best = t;
        }
      }
      double t;
      for (const OBox& f : facades) {
        V3 lo{f.c.x - f.h.x, f.c.y - f.h.y, f.c.z - f.h.z}, hi{f.c.x + f.h.x, f.c.y + f.h.y, f.c.z + f.h.z};
        if (slab(o, d, lo, hi, t) && t < best) best = t;
      }
      for (const Pole& p : poles) {
        if (std::fabs(p.x - o.x) > kMaxRange + 1.0) continue;
        double ox = o.x - p.x, oy = o.y - p.y;
        double a = d.x * d.x + d.y * d.y;
        if (a < 1e-14) continue;
        double b = ox * d.x + oy * d.y, c = ox * ox + oy * oy - p.r * p.r;
        double disc = b * b - a * c;
        if (disc < 0) continue;
        double sq = std::sqrt(disc);
        double t0 = (-b - sq) / a;
        if (t0 < 1e-6) t0 = (-b + sq) / a;
        if (t0 < 1e-6) continue;
        double z = o.z + t0 * d.z;
        if (z < p.z0 || z > p.z1) continue;
        if (t0 < best) best = t0;
      }
    }
    return best < 1e29 ? best : -1.0;
  }
};

extern "C" int vmp_synth_create(int scene, uint64_t seed, int32_t pts_per_scan,
                                const double origin_offset[3], vmp_synth** out) {
  if (!out || pts_per_scan <= 0 || (scene != VMP_SCENE_ROOM && scene != VMP_SCENE_STREET)) return -1;
  vmp_synth* s = new (std::nothrow) vmp_synth();
  if (!s) return -3;
  s->scene = scene;
  s->seed = seed;
  s->n = pts_per_scan;
  s->offset = origin_offset ? V3{origin_offset[0], origin_offset[1], origin_offset[2]} : V3{0, 0, 0};
  s->build_trajectory(scene == VMP_SCENE_ROOM ? 6000 : 1800);
  s->build_scene();
  *out = s;
  return 0;
}

extern "C" void vmp_synth_destroy(vmp_synth* s) { delete s; }

static void store_cm(const M3& R, double out[9]) {
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) out[c * 3 + r] = R.m[r][c];
}

extern "C" int vmp_synth_pose(const vmp_synth* s, int32_t k, double R[9], double t[3]) {
  if (!s || k < 0 || size_t(k) + 1 >= s->xy_table.size()) return -1;
  M3 Rl;
  V3 tl;
  s->pose_local(k, Rl, tl);
  if (R) store_cm(Rl, R);
  if (t) {
    t[0] = tl.x + s->offset.x;
    t[1] = tl.y + s->offset.y;
    t[2] = tl.z + s->offset.z;
  }
  return 0;
}

extern "C" int vmp_synth_scan_n(const vmp_synth* s, int32_t k, int32_t n, float* x, float* y, float* z,
                                float* t_ms) {
  if (!s || k < 0 || n <= 0 || size_t(k) + 1 >= s->xy_table.size()) return -1;
  M3 R;
  V3 o;
  s->pose_local(k, R, o);
  const double a1 = 0.7548776662466927, a2 = 0.5698402909980532;
  const double b1 = 0.1370 * k, b2 = 0.4210 * k;
  const double el_lo = -7.0 * kPi / 180.0, el_span = 59.0 * kPi / 180.0;
  const uint64_t sseed = hash3(s->seed, 0x5CA17ull, uint64_t(k));
  const int chunk = n + n / 2 + 64;
  std::vector<float> px(chunk), py(chunk), pz(chunk);
  std::vector<unsigned char> ok(chunk);
  int filled = 0;
  long long ray0 = 0;
  while (filled < n) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < chunk; ++i) {
      long long j = ray0 + i;
      double fa = a1 * double(j) + b1, fe = a2 * double(j) + b2;
      double az = 2.0 * kPi * (fa - std::floor(fa));
      double el = el_lo + el_span * (fe - std::floor(fe));
      V3 dl{std::cos(el) * std::cos(az), std::cos(el) * std::sin(az), std::sin(el)};
      V3 dw = mul(R, dl);
      double r = s->cast(o, dw);
      uint64_t h0 = hash3(sseed, uint64_t(j), 1), h1 = hash3(sseed, uint64_t(j), 2);
      if (s->scene == VMP_SCENE_STREET && u01(hash3(sseed, uint64_t(j), 3)) < 0.15) {
        double rmax = (r > 0 && r < kMaxRange) ? r : kMaxRange;  // foliage scattered along the ray
        r = rmax * (0.05 + 0.95 * u01(hash3(sseed, uint64_t(j), 4)));
      }
      ok[i] = 0;
      if (r > 0 && r <= kMaxRange) {
        double rm = r + kRangeSigma * gauss(h0, h1);
        if (rm >= kMinRange && rm <= kMaxRange) {
          px[i] = float(rm * dl.x);
          py[i] = float(rm * dl.y);
          pz[i] = float(rm * dl.z);
          ok[i] = 1;
        }
      }
    }
    for (int i = 0; i < chunk && filled < n; ++i) {
      if (!ok[i]) continue;
      if (x) x[filled] = px[i];
      if (y) y[filled] = py[i];
      if (z) z[filled] = pz[i];
      if (t_ms) t_ms[filled] = float(1e-3 * double(ray0 + i));
      ++filled;
    }
    ray0 += chunk;
    if (ray0 > 64LL * n + 100000) return -2;  // scene produced too few returns
  }
  return 0;
}

extern "C" int vmp_synth_scan(const vmp_synth* s, int32_t k, float* x, float* y, float* z, float* t_ms) {
  if (!s) return -1;
  return vmp_synth_scan_n(s, k, s->n, x, y, z, t_ms);
}

extern "C" int vmp_synth_motion(const vmp_synth* s, int32_t k, double dR[9], double dp[3]) {
  if (!s || k < 1 || size_t(k) + 1 >= s->xy_table.size()) return -1;
  M3 R0, R1;
  V3 t0, t1;
  s->pose_local(k - 1, R0, t0);
  s->pose_local(k, R1, t1);
  M3 R0t = transpose(R0);
  M3 d = mulm(R0t, R1);
  V3 p = mul(R0t, t1 - t0);
  const uint64_t ms = hash3(s->seed, 0x3071ull, uint64_t(k));
  const double sr = 0.2 * kPi / 180.0, sp = 0.02;
  V3 er{sr * gauss(hash3(ms, 1, 0), hash3(ms, 1, 1)), sr * gauss(hash3(ms, 2, 0), hash3(ms, 2, 1)),
        sr * gauss(hash3(ms, 3, 0), hash3(ms, 3, 1))};
  V3 ep{sp * gauss(hash3(ms, 4, 0), hash3(ms, 4, 1)), sp * gauss(hash3(ms, 5, 0), hash3(ms, 5, 1)),
        sp * gauss(hash3(ms, 6, 0), hash3(ms, 6, 1))};
  M3 dn = mulm(d, expso3(er));
  if (dR) store_cm(dn, dR);
  if (dp) {
    dp[0] = p.x + ep.x;
    dp[1] = p.y + ep.y;
    dp[2] = p.z + ep.z;
  }
  return 0;
}

extern "C" int vmp_synth_create_imu(int scene, uint64_t seed, int32_t pts_per_scan, const double origin_offset[3],
                                    vmp_synth** out) {
  const int rc = vmp_synth_create(scene, seed, pts_per_scan, origin_offset, out);
  if (rc != 0) return rc;
  vmp_synth* s = *out;
  s->imu_mode = 1;
  // bias random walks per IMU sample over the trajectory's length
  const size_t m = size_t(double(s->xy_table.size()) * kScanDt / kImuDt) + 64;
  s->bg.resize(m);
  s->ba.resize(m);
  const uint64_t bs = hash3(seed, 0xB1A5ull, 0);
  auto g = [&](uint64_t a, uint64_t b) { return gauss(hash3(bs, a, 2 * b), hash3(bs, a, 2 * b + 1)); };
  V3 bgc{2e-3 * g(0, 0), 2e-3 * g(0, 1), 2e-3 * g(0, 2)};
  V3 bac{1e-2 * g(1, 0), 1e-2 * g(1, 1), 1e-2 * g(1, 2)};
  for (size_t j = 0; j < m; ++j) {
    s->bg[j] = bgc;
    s->ba[j] = bac;
    bgc = bgc + V3{1e-5 * g(10 + j, 0), 1e-5 * g(10 + j, 1), 1e-5 * g(10 + j, 2)};
    bac = bac + V3{1e-4 * g(10 + m + j, 0), 1e-4 * g(10 + m + j, 1), 1e-4 * g(10 + m + j, 2)};
  }
  return 0;
}

extern "C" int vmp_synth_pose_at(const vmp_synth* s, double t, double R[9], double p[3]) {
  if (!s || t < 0.0 || t > double(s->xy_table.size() - 1) * kScanDt) return -1;
  M3 Rl;
  V3 tl;
  s->pose_clock(t, Rl, tl);
  if (R) store_cm(Rl, R);
  if (p) {
    p[0] = tl.x + s->offset.x;
    p[1] = tl.y + s->offset.y;
    p[2] = tl.z + s->offset.z;
  }
  return 0;
}

extern "C" int vmp_synth_imu(const vmp_synth* s, int32_t j0, int32_t count, double* t, double* acc, double* gyro) {
  if (!s || !s->imu_mode || j0 < 0 || count < 0 || size_t(j0) + size_t(count) > s->bg.size()) return -1;
  const double h = 1e-4;
  for (int32_t q = 0; q < count; ++q) {
    const int32_t j = j0 + q;
    const double tj = kImuDt * j;
    M3 R, Rm, Rp;
    V3 p, pm, pp;
    s->pose_clock(tj, R, p);
    s->pose_clock(tj > h ? tj - h : 0.0, Rm, pm);
    s->pose_clock(tj + h, Rp, pp);
    const double span = tj > h ? 2.0 * h : tj + h;
    const V3 w = (1.0 / span) * logso3(mulm(transpose(Rm), Rp));  // body rate
    V3 f = s->accel_clock(tj);
    f.z += 9.81;  // specific force: a - g, g = (0, 0, -9.81)
    const V3 fb = mul(transpose(R), f);
    const uint64_t ns = hash3(s->seed, 0x1A7Aull, uint64_t(j));
    auto g = [&](uint64_t a) { return gauss(hash3(ns, a, 0), hash3(ns, a, 1)); };
    if (t) t[q] = tj;
    if (gyro) {
      gyro[3 * q + 0] = w.x + s->bg[size_t(j)].x + 1e-3 * g(1);
      gyro[3 * q + 1] = w.y + s->bg[size_t(j)].y + 1e-3 * g(2);
      gyro[3 * q + 2] = w.z + s->bg[size_t(j)].z + 1e-3 * g(3);
    }
    if (acc) {
      acc[3 * q + 0] = fb.x + s->ba[size_t(j)].x + 1e-2 * g(4);
      acc[3 * q + 1] = fb.y + s->ba[size_t(j)].y + 1e-2 * g(5);
      acc[3 * q + 2] = fb.z + s->ba[size_t(j)].z + 1e-2 * g(6);
    }
  }
  return 0;
}

// Raw scan k: ray j leaves at clock time 0.1 k + 0.1 j / J from the pose at
// that time; a return is expressed in the LiDAR frame of its own time (what
// undistortCloud receives).  J rays span the 0.1 s sweep and are sized from
// the return rate of a first sweep so that the n-th return comes late in it;
// a sweep with fewer than n returns is redone with more rays.
extern "C" int vmp_synth_scan_raw(const vmp_synth* s, int32_t k, int32_t n, float* x, float* y, float* z,
                                  float* t_ms) {
  if (!s || k < 0 || n <= 0 || size_t(k) + 1 >= s->xy_table.size()) return -1;
  const double a1 = 0.7548776662466927, a2 = 0.5698402909980532;
  const double b1 = 0.1370 * k, b2 = 0.4210 * k;
  const double el_lo = -7.0 * kPi / 180.0, el_span = 59.0 * kPi / 180.0;
  const uint64_t sseed = hash3(s->seed, 0x5CA17ull, uint64_t(k));
  const double t0 = kScanDt * k;
  // poses along the sweep at kTab + 1 knots, linearly interpolated per ray
  // (interpolation error ~1e-9 m and rad at these rates)
  constexpr int kTab = 4096;
  std::vector<M3> tabR(kTab + 1);
  std::vector<V3> tabP(kTab + 1);
#pragma omp parallel for schedule(static)
  for (int q = 0; q <= kTab; ++q) s->pose_clock(t0 + kScanDt * double(q) / kTab, tabR[size_t(q)], tabP[size_t(q)]);
  long long J = 2LL * n + 64;
  for (int attempt = 0; attempt < 8; ++attempt) {
    std::vector<float> px(static_cast<size_t>(J)), py(static_cast<size_t>(J)), pz(static_cast<size_t>(J));
    std::vector<unsigned char> ok(static_cast<size_t>(J));
    // the sizing sweep casts every 16th ray only
    const long long stride = attempt == 0 ? 16 : 1;
#pragma omp parallel for schedule(static)
    for (long long j = 0; j < J; j += stride) {
      const double u = double(kTab) * double(j) / double(J);
      const int q = std::min(int(u), kTab - 1);
      const double w = u - q;
      M3 R;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) R.m[a][b] = (1.0 - w) * tabR[size_t(q)].m[a][b] + w * tabR[size_t(q) + 1].m[a][b];
      const V3 o = (1.0 - w) * tabP[size_t(q)] + w * tabP[size_t(q) + 1];
      double fa = a1 * double(j) + b1, fe = a2 * double(j) + b2;
      double az = 2.0 * kPi * (fa - std::floor(fa));
      double el = el_lo + el_span * (fe - std::floor(fe));
      V3 dl{std::cos(el) * std::cos(az), std::cos(el) * std::sin(az), std::sin(el)};
      V3 dw = mul(R, dl);
      double r = s->cast(o, dw);
      uint64_t h0 = hash3(sseed, uint64_t(j), 1), h1 = hash3(sseed, uint64_t(j), 2);
      if (s->scene == VMP_SCENE_STREET && u01(hash3(sseed, uint64_t(j), 3)) < 0.15) {
        double rmax = (r > 0 && r < kMaxRange) ? r : kMaxRange;
        r = rmax * (0.05 + 0.95 * u01(hash3(sseed, uint64_t(j), 4)));
      }
      ok[size_t(j)] = 0;
      if (r > 0 && r <= kMaxRange) {
        double rm = r + kRangeSigma * gauss(h0, h1);
        if (rm >= kMinRange && rm <= kMaxRange) {
          px[size_t(j)] = float(rm * dl.x);
          py[size_t(j)] = float(rm * dl.y);
          pz[size_t(j)] = float(rm * dl.z);
          ok[size_t(j)] = 1;
        }
      }
    }
    long long returns = 0;
    for (long long j = 0; j < J; j += stride) returns += ok[size_t(j)];
    if (attempt == 0) {  // size the sweep: n returns at ~99 % of it
      const double rate = double(std::max(returns, 1LL)) / double((J + stride - 1) / stride);
      J = (long long)(double(n) / rate * 1.01) + 32;
      continue;
    }
    if (returns < n) {  // short of returns: a longer sweep
      J += J / 20 + 64;
      continue;
    }
    int filled = 0;
    for (long long j = 0; j < J && filled < n; ++j) {
      if (!ok[size_t(j)]) continue;
      if (x) x[filled] = px[size_t(j)];
      if (y) y[filled] = py[size_t(j)];
      if (z) z[filled] = pz[size_t(j)];
      if (t_ms) t_ms[filled] = float(1e3 * kScanDt * double(j) / double(J));
      ++filled;
    }
    return 0;
  }
  return -2;
}
