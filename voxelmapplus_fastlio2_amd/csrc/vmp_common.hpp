// Shared definitions of the MI355X registration path: voxel keys, the device
// hash map layout, voxel records, and FP64 state algebra usable on host and
// device.  The library is compiled with -ffp-contract=off so every FP64
// expression here rounds exactly like the reference's (no fused multiply-add);
// FP32 per-point kernels opt back into contraction locally.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstddef>

#include "vmp.h"

namespace vmp {

// ---------------------------------------------------------------- keys ----
// VoxelKey{int64 x,y,z} (voxel_map.h:17-34) packed into 63 bits: 21 bits per
// axis, biased by 2^20 (+-524 km at 0.5 m voxels).  Two bit patterns with the
// top bit set mark empty and deleted hash slots.
constexpr uint64_t kKeyEmpty = ~0ull;
constexpr uint64_t kKeyTomb = ~0ull - 1;
constexpr int64_t kKeyBias = int64_t(1) << 20;
constexpr uint64_t kKeyMask = (uint64_t(1) << 21) - 1;

__host__ __device__ inline bool key_ok(int64_t k) { return k >= -kKeyBias && k < kKeyBias; }
__host__ __device__ inline uint64_t pack_key(int64_t x, int64_t y, int64_t z) {
  return uint64_t(x + kKeyBias) | (uint64_t(y + kKeyBias) << 21) | (uint64_t(z + kKeyBias) << 42);
}
__host__ __device__ inline void unpack_key(uint64_t k, int64_t& x, int64_t& y, int64_t& z) {
  x = int64_t(k & kKeyMask) - kKeyBias;
  y = int64_t((k >> 21) & kKeyMask) - kKeyBias;
  z = int64_t((k >> 42) & kKeyMask) - kKeyBias;
}
// neighbour in the fixed order of voxel_map.cpp:140-146: -x,-y,-z,+x,+y,+z
__host__ __device__ inline uint64_t key_neighbour(uint64_t k, int j) {
  const uint64_t unit = j % 3 == 0 ? 1ull : (j % 3 == 1 ? (1ull << 21) : (1ull << 42));
  return j < 3 ? k - unit : k + unit;
}
__host__ __device__ inline bool key_neighbour_ok(uint64_t k, int j) {
  const int sh = (j % 3) * 21;
  const uint64_t c = (k >> sh) & kKeyMask;
  return j < 3 ? c > 0 : c < kKeyMask;
}
// Home slot of a key: the 2x2x2 brick of the voxel is hashed and its eight
// voxels take eight consecutive slots (one 128-byte line of 16-byte entries),
// so spatially coherent lookups of a scan share cache lines.
__host__ __device__ inline uint32_t hash_slot(uint64_t k, uint32_t mask) {
  const uint64_t sub = (k & 1ull) | ((k >> 20) & 2ull) | ((k >> 40) & 4ull);
  uint64_t h = (k >> 1) & ~((1ull << 20) | (1ull << 41));  // drop the low bit of every axis
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return uint32_t((h << 3) | sub) & mask;
}

struct __align__(16) HashEntry {
  uint64_t key;  // packed key, kKeyEmpty or kKeyTomb
  int32_t vid;   // voxel id, -1 while a new key waits for its voxel
  int32_t pad;
};

// ------------------------------------------------------------ records -----
// Voxel flags (VoxelGrid members, voxel_map.h:84-95)
enum : uint32_t {
  kInit = 1u,
  kPlane = 2u,
  kEnable = 4u,   // update_enable
  kMerged = 8u,
  kDirty = 16u,   // a merge() call on this voxel may change state
  kAlive = 32u,
};

// 32 B record read by the observe kernel.  mean is relative to the voxel's
// lower corner (key * voxel_size) so fp32 keeps ~1e-8 m resolution.
struct __align__(16) VoxHot {
  float mx, my, mz;
  uint32_t flags;
  float nx, ny, nz;
  float pad;
};

// Per-voxel bookkeeping for the map update.  Fields named *_t are point
// indices within the current scan and are valid only when tag == scan tag.
// The per-scan fields fill bytes 40-79 (one 8-byte and two 16-byte groups),
// so a new scan tag is written with three vector stores.
struct __align__(16) VoxState {
  uint64_t key;
  uint64_t group_id;
  uint64_t stamp;     // LRU: global index of the last point that touched it
  int32_t n;          // Plane::n
  int32_t newly_add;  // newly_add_point
  uint32_t flags;
  int32_t pad;
  int32_t merge_pos;  // position in the segment where merge() calls start
  int32_t dirty_t;    // latest point time at which a merge() call could act
  int32_t tag;        // scan tag of the per-scan fields
  int32_t create_t;   // created by point create_t of this scan (-1: older)
  int32_t evict_t;    // evicted at point evict_t of this scan (INT32_MAX: no)
  int32_t freeze_t;   // froze at point freeze_t of this scan (-1: no)
  int32_t prev_vid;   // incarnation of the same key evicted in this scan
  int32_t first_t;    // first point of this scan that touched it
  int32_t seg;        // its segment in this scan's key-sorted points
  int32_t taint_t;    // first point time a merge changed its neighbourhood
};
static_assert(sizeof(VoxState) == 80, "VoxState is 80 bytes");
static_assert(offsetof(VoxState, merge_pos) == 40 && offsetof(VoxState, tag) == 48 &&
                  offsetof(VoxState, prev_vid) == 64,
              "per-scan VoxState fields are grouped for vector stores");

// Pointers loaded from memory compile to flat accesses, which also count
// against the LDS wait counter; this names the global address space.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* as_global(T* p) {
  return (__attribute__((address_space(1))) T*)(p);
}
typedef int32_t vmp_i2 __attribute__((ext_vector_type(2)));
typedef int32_t vmp_i4 __attribute__((ext_vector_type(4)));
// 8- and 16-byte stores of int32 groups (p aligned to 8 / 16 bytes)
__device__ __forceinline__ void st_global2(int32_t* p, int32_t a, int32_t b) {
  *as_global(reinterpret_cast<vmp_i2*>(p)) = vmp_i2{a, b};
}
__device__ __forceinline__ void st_global4(int32_t* p, int32_t a, int32_t b, int32_t c, int32_t d) {
  *as_global(reinterpret_cast<vmp_i4*>(p)) = vmp_i4{a, b, c, d};
}

// Plane record in FP64 (what merge() reads and writes, voxel_map.cpp:138-186):
// mean xyz, normal xyz, trace of the normal block and of the centre block of
// Plane::cov (only these traces are ever consumed, voxel_map.cpp:161-164).
constexpr int kPlaneD = 8;
// Refit moments, centre-local: S1q(3) S2q(6) S0C(6) S1C(18) S2C(36) over
// every point except the pending ones (below); then the running S1q over all
// points (the plane mean), the number of pending points, one pad.  Points
// added by a scan with no refit wait in the voxel's pending list (kPendF
// floats each: the point relative to the voxel centre, its world cov) until
// the next refit folds them in.
constexpr int kMomD = 74;
constexpr int kMomRun = 69;
constexpr int kMomPend = 72;
constexpr int kPendF = 9;

// ------------------------------------------------------------- params -----
// Per-pass constants derived from the filter state in FP64.
struct ObsParams {
  double r_wl[9];      // row-major rot*rot_ext
  double p_wl[3];
  float Rh[9], Rl[9];  // r_wl split into fp32 hi + lo (row-major)
  float th[3], tl[3];  // p_wl - anchor, hi + lo
  int64_t anchor[3];   // anchor voxel (key of p_wl)
  float rot[9];        // state.rot (row-major) for the Jacobian
  float rwl32[9];      // r_wl for the extrinsic Jacobian
  float ext_R[9];      // rot_ext
  float ext_t[3];      // pos_ext
  float Prr[6], Ppp[6];  // symmetric blocks of the prior P (xx xy xz yy yz zz)
  float vs, inv_vs;
  float s2;            // sin^2(angle_cov)
  float rv;            // ranging_cov^2
  double vs64;
  int32_t estimate_ext;
  int32_t active;      // 0 once the IESKF converged: the pass is a no-op
};

struct UpdParams {
  float m[9];          // fp32 insertion transform (PCL Matrix4f), row-major
  float t[3];
  float rwl32[9];      // r_wl for the insertion covariance
  float Prr[6], Ppp[6];  // posterior P blocks
  float s2, rv;
  double vs64;
};

// -------------------------------------------------- FP64 state algebra -----
// A small column-vector/row-major matrix toolkit mirroring the summation
// order of the reference's Eigen expressions (inner index left to right).
struct D3 {
  double v[3];
};
struct D33 {
  double m[9];  // row-major
};

__host__ __device__ inline D33 d33_mul(const D33& A, const D33& B) {
  D33 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = A.m[i * 3] * B.m[j];
      s += A.m[i * 3 + 1] * B.m[3 + j];
      s += A.m[i * 3 + 2] * B.m[6 + j];
      C.m[i * 3 + j] = s;
    }
  return C;
}
__host__ __device__ inline D3 d33_mulv(const D33& A, const D3& x) {
  D3 y;
  for (int i = 0; i < 3; ++i) {
    double s = A.m[i * 3] * x.v[0];
    s += A.m[i * 3 + 1] * x.v[1];
    s += A.m[i * 3 + 2] * x.v[2];
    y.v[i] = s;
  }
  return y;
}
__host__ __device__ inline D33 d33_t(const D33& A) {
  D33 T;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T.m[i * 3 + j] = A.m[j * 3 + i];
  return T;
}
__host__ __device__ inline D33 d33_from_cm(const double* cm) {
  D33 A;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) A.m[r * 3 + c] = cm[c * 3 + r];
  return A;
}
__host__ __device__ inline void d33_to_cm(const D33& A, double* cm) {
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) cm[c * 3 + r] = A.m[r * 3 + c];
}

// r_wl = rot * rot_ext, p_wl = rot * pos_ext + pos (lio_builder.cpp:252-253)
__host__ __device__ inline void state_lidar_pose(const vmp_state& x, D33& r_wl, D3& p_wl) {
  D33 rot = d33_from_cm(x.rot), ext = d33_from_cm(x.rot_ext);
  r_wl = d33_mul(rot, ext);
  D3 pe{{x.pos_ext[0], x.pos_ext[1], x.pos_ext[2]}};
  D3 a = d33_mulv(rot, pe);
  for (int i = 0; i < 3; ++i) p_wl.v[i] = a.v[i] + x.pos[i];
}

__host__ __device__ inline void split_f32(double v, float& hi, float& lo) {
  hi = float(v);
  lo = float(v - double(hi));
}

__host__ __device__ inline void sym6_from_P(const double* P, int o, float out[6]) {
  // P column-major 23x23, block (o..o+2, o..o+2)
  auto at = [&](int r, int c) { return P[(o + c) * 23 + (o + r)]; };
  out[0] = float(at(0, 0));
  out[1] = float(at(0, 1));
  out[2] = float(at(0, 2));
  out[3] = float(at(1, 1));
  out[4] = float(at(1, 2));
  out[5] = float(at(2, 2));
}

// sin^2 of the angular resolution (commons.cpp:26-28)
__host__ __device__ inline float angle_s2(const vmp_map_cfg& cfg) {
  const double sa = sin(cfg.angle_cov * 3.14159265358979323846 / 180.0);
  return float(sa * sa);
}

__host__ __device__ inline void make_obs_params(const vmp_state& x, const double* P, const vmp_map_cfg& cfg,
                                                ObsParams& o) {
  D33 r;
  D3 p;
  state_lidar_pose(x, r, p);
  for (int i = 0; i < 9; ++i) o.r_wl[i] = r.m[i];
  for (int i = 0; i < 3; ++i) o.p_wl[i] = p.v[i];
  for (int i = 0; i < 9; ++i) split_f32(r.m[i], o.Rh[i], o.Rl[i]);
  for (int i = 0; i < 3; ++i) {
    o.anchor[i] = int64_t(floor(p.v[i] / cfg.voxel_size));
    double loc = p.v[i] - double(o.anchor[i]) * cfg.voxel_size;
    split_f32(loc, o.th[i], o.tl[i]);
  }
  D33 rot = d33_from_cm(x.rot), ext = d33_from_cm(x.rot_ext);
  for (int i = 0; i < 9; ++i) {
    o.rot[i] = float(rot.m[i]);
    o.rwl32[i] = float(r.m[i]);
    o.ext_R[i] = float(ext.m[i]);
  }
  for (int i = 0; i < 3; ++i) o.ext_t[i] = float(x.pos_ext[i]);
  sym6_from_P(P, 3, o.Prr);
  sym6_from_P(P, 0, o.Ppp);
  o.vs = float(cfg.voxel_size);
  o.inv_vs = float(1.0 / cfg.voxel_size);
  o.vs64 = cfg.voxel_size;
  o.s2 = angle_s2(cfg);
  o.rv = float(cfg.ranging_cov * cfg.ranging_cov);
  o.estimate_ext = cfg.estimate_ext;
  o.active = 1;
}

__host__ __device__ inline void make_upd_params(const vmp_state& x, const double* P, const vmp_map_cfg& cfg,
                                                UpdParams& u) {
  D33 r;
  D3 p;
  state_lidar_pose(x, r, p);
  for (int i = 0; i < 9; ++i) {
    u.m[i] = float(r.m[i]);  // (rot*rot_ext).cast<float>()
    u.rwl32[i] = float(r.m[i]);
  }
  for (int i = 0; i < 3; ++i) u.t[i] = float(p.v[i]);
  sym6_from_P(P, 3, u.Prr);
  sym6_from_P(P, 0, u.Ppp);
  u.s2 = angle_s2(cfg);
  u.rv = float(cfg.ranging_cov * cfg.ranging_cov);
  u.vs64 = cfg.voxel_size;
}

// -------------------------------------------------- FP32 point helpers -----
// Insertion point: PCL transformPointCloud with a Matrix4f, evaluated as
// m0*x + (m1*y + (m2*z + t)) in fp32 without contraction.
__host__ __device__ inline void ins_transform(const float* m, const float* t, float x, float y, float z,
                                              float out[3]) {
  for (int r = 0; r < 3; ++r) {
    float a = m[r * 3 + 2] * z + t[r];
    float b = m[r * 3 + 1] * y + a;
    out[r] = m[r * 3 + 0] * x + b;
  }
}

// Per-point world covariance C = R C_l R^T + [p]x Prr [p]x^T + Ppp in fp32,
// with calcBodyCov in closed form: C_l = a I + (b/|p|^2) p p^T,
// a = |p|^2 sin^2(angle), b = ranging^2 - a (commons.cpp:18-45: the basis N
// spans the plane orthogonal to the ray, so A A^T = |p|^2 (I - d d^T)).
__host__ __device__ inline void point_world_cov(const float* R, const float* Prr, const float* Ppp, float s2,
                                                float rv, float px, float py, float pz, float C[6]) {
  float r2 = px * px + py * py + pz * pz;
  float a = r2 * s2;
  float bb = (rv - a) / r2;
  float w0 = R[0] * px + R[1] * py + R[2] * pz;
  float w1 = R[3] * px + R[4] * py + R[5] * pz;
  float w2 = R[6] * px + R[7] * py + R[8] * pz;
  // K = [p]x ; M = K Prr K^T
  const float K[9] = {0.f, -pz, py, pz, 0.f, -px, -py, px, 0.f};
  const float P[9] = {Prr[0], Prr[1], Prr[2], Prr[1], Prr[3], Prr[4], Prr[2], Prr[4], Prr[5]};
  float KP[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) KP[i * 3 + j] = K[i * 3] * P[j] + K[i * 3 + 1] * P[3 + j] + K[i * 3 + 2] * P[6 + j];
  auto M = [&](int i, int j) { return KP[i * 3] * K[j * 3] + KP[i * 3 + 1] * K[j * 3 + 1] + KP[i * 3 + 2] * K[j * 3 + 2]; };
  C[0] = a + bb * w0 * w0 + M(0, 0) + Ppp[0];
  C[1] = bb * w0 * w1 + M(0, 1) + Ppp[1];
  C[2] = bb * w0 * w2 + M(0, 2) + Ppp[2];
  C[3] = a + bb * w1 * w1 + M(1, 1) + Ppp[3];
  C[4] = bb * w1 * w2 + M(1, 2) + Ppp[4];
  C[5] = a + bb * w2 * w2 + M(2, 2) + Ppp[5];
}

}  // namespace vmp
