/*
 * vmp_synth.h — seeded synthetic LiDAR sequences for the voxel_plus
 * registration path (C ABI, host-only, deterministic on every x86-64 box).
 *
 * The reference has no test data of its own (its demo rosbag is external), so
 * every workload here is generated:
 *   scene 0 "room":   30 x 30 x 6 m room (walls, floor, ceiling) + 40 seeded
 *                     yawed boxes with sides U[0.5, 3] m.
 *   scene 1 "street": 200 m street along x with a sloped ground plane,
 *                     building facades on both sides, 0.3 m-radius poles and a
 *                     15 % share of foliage returns scattered uniformly along
 *                     the ray.
 * Rays follow a non-repeating rosette-like pattern over 360 deg azimuth and
 * -7..52 deg elevation (Livox MID-360 field of view), range noise N(0, 0.02 m),
 * kept ranges 0.1..40 m.  Every scan has exactly `pts_per_scan` points, in
 * ascending, unique per-point time order (the order lio_builder.cpp:75-76 sorts
 * by; utils.cpp:20 time convention).  Points are emitted already motion-
 * compensated: each is expressed in the LiDAR frame at the scan end pose, which
 * is what undistortCloud (lio_builder.cpp:65-153) hands to the registration
 * path.  Trajectory: 1 m/s, |yaw rate| <= 0.5 rad/s, small roll/pitch sway;
 * scans at 10 Hz.
 *
 * Matrices are 3x3 column-major (Eigen's default layout).
 */
#ifndef VMP_SYNTH_H
#define VMP_SYNTH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vmp_synth vmp_synth;

enum { VMP_SCENE_ROOM = 0, VMP_SCENE_STREET = 1 };

/* origin_offset: added to every world coordinate (precision study: 10 km). */
int vmp_synth_create(int scene, uint64_t seed, int32_t pts_per_scan,
                     const double origin_offset[3], vmp_synth** out);
void vmp_synth_destroy(vmp_synth* s);

/* Ground-truth LiDAR pose (world <- lidar) at the end of scan k (10 Hz). */
int vmp_synth_pose(const vmp_synth* s, int32_t k, double R[9], double t[3]);

/* Scan k: xyz are LiDAR-frame fp32 points (n = pts_per_scan each), t_ms the
 * per-point time offset in ms (strictly ascending).  Any output may be NULL. */
int vmp_synth_scan(const vmp_synth* s, int32_t k, float* x, float* y, float* z,
                   float* t_ms);
/* Same scan with n points instead of pts_per_scan: the first n returns of
 * the same ray sequence (a partial sweep). */
int vmp_synth_scan_n(const vmp_synth* s, int32_t k, int32_t n, float* x, float* y,
                     float* z, float* t_ms);

/* Noisy odometry increment between scans k-1 and k, expressed in the LiDAR
 * frame of scan k-1: pose_k ~= pose_{k-1} * [dR | dp].  Seeded noise of
 * 2 cm / 0.2 deg per axis stands in for IMU propagation error. */
int vmp_synth_motion(const vmp_synth* s, int32_t k, double dR[9], double dp[3]);

/* IMU mode: the same scene and path, traversed from rest (0.5 s still, then
 * a 1 s smooth speed ramp to the trajectory's 1 m/s), with a 200 Hz IMU
 * (gyro / accelerometer white noise 1e-3 rad/s / 1e-2 m/s^2 and random-walk
 * biases) and raw scans whose points are taken along the 0.1 s sweep. */
int vmp_synth_create_imu(int scene, uint64_t seed, int32_t pts_per_scan,
                         const double origin_offset[3], vmp_synth** out);
/* Ground-truth pose (world <- lidar = imu) at clock time t seconds. */
int vmp_synth_pose_at(const vmp_synth* s, double t, double R[9], double p[3]);
/* IMU samples j0 .. j0+count-1, taken at t_j = 0.005 j: acc (specific
 * force, m/s^2) and gyro (rad/s) in the body frame, 3 per sample. */
int vmp_synth_imu(const vmp_synth* s, int32_t j0, int32_t count, double* t, double* acc, double* gyro);
/* Raw scan k over [0.1 k, 0.1 k + 0.1): n returns, each in the LiDAR frame
 * at its own time; t_ms = time since 0.1 k in ms (strictly ascending). */
int vmp_synth_scan_raw(const vmp_synth* s, int32_t k, int32_t n, float* x, float* y, float* z, float* t_ms);

#ifdef __cplusplus
}
#endif
#endif
