// kf_synth.cpp - seeded synthetic depth stream (box room, spheres, boxes).
//
// The reference reads recorded Kinect v1 frames from data/dep%04d.dat
// (system_utility.cpp:58-68) and has no ground truth.  This generator renders
// the same kind of uint16 millimetre frames analytically, with ground-truth
// camera-to-world poses, so that every stage and the trajectory can be
// checked.  Everything is FP64 on the host and deterministic for a seed.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/kinfu_amd.h"

namespace {

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// uniform in (0, 1]
double u01(uint64_t& s) { return ((splitmix64(s) >> 11) + 1) * (1.0 / 9007199254740992.0); }

double lerp(double a, double b, double t) { return a + (b - a) * t; }

const double kRoomMin[3] = {-1950.0, -1950.0, -1150.0};
const double kRoomMax[3] = {1550.0, 1550.0, 1750.0};

template <class Fn>
void parallel_rows(int H, int n_threads, Fn fn) {
  if (n_threads <= 1) {
    for (int y = 0; y < H; ++y) fn(y);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < n_threads; ++t)
    th.emplace_back([=] {
      for (int y = t; y < H; y += n_threads) fn(y);
    });
  for (auto& x : th) x.join();
}

bool inv3(const double* a, double* r) {
  const double c00 = a[4] * a[8] - a[5] * a[7];
  const double c01 = a[5] * a[6] - a[3] * a[8];
  const double c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  if (det == 0.0) return false;
  const double id = 1.0 / det;
  r[0] = c00 * id;
  r[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  r[3] = c01 * id;
  r[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  r[6] = c02 * id;
  r[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  return true;
}

// Nearest positive hit along o + t*d (d not normalised).  Returns +inf when
// nothing is hit.
double trace(const kf_synth_scene& s, const double* o, const double* d) {
  double best = INFINITY;
  // room, seen from inside: the exit distance
  {
    double t = INFINITY;
    for (int i = 0; i < 3; ++i) {
      if (d[i] > 0) t = std::min(t, ((double)s.room_max[i] - o[i]) / d[i]);
      if (d[i] < 0) t = std::min(t, ((double)s.room_min[i] - o[i]) / d[i]);
    }
    if (t > 0) best = t;
  }
  const double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  for (int k = 0; k < s.n_spheres; ++k) {
    const double oc[3] = {o[0] - s.spheres[k][0], o[1] - s.spheres[k][1], o[2] - s.spheres[k][2]};
    const double r = s.spheres[k][3];
    const double b = 2.0 * (d[0] * oc[0] + d[1] * oc[1] + d[2] * oc[2]);
    const double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - r * r;
    const double disc = b * b - 4.0 * a * c;
    if (disc < 0) continue;
    const double sq = sqrt(disc);
    double t = (-b - sq) / (2.0 * a);
    if (t <= 1e-9) t = (-b + sq) / (2.0 * a);
    if (t > 1e-9 && t < best) best = t;
  }
  for (int k = 0; k < s.n_boxes; ++k) {
    double tn = -INFINITY, tf = INFINITY;
    bool miss = false;
    for (int i = 0; i < 3; ++i) {
      const double lo = s.boxes[k][i], hi = s.boxes[k][3 + i];
      if (d[i] == 0) {
        if (o[i] < lo || o[i] > hi) miss = true;
        continue;
      }
      double t0 = (lo - o[i]) / d[i], t1 = (hi - o[i]) / d[i];
      if (t0 > t1) std::swap(t0, t1);
      tn = std::max(tn, t0);
      tf = std::min(tf, t1);
    }
    if (miss || tn > tf) continue;
    const double t = tn > 1e-9 ? tn : tf;
    if (t > 1e-9 && t < best) best = t;
  }
  return best;
}

}  // namespace

extern "C" {

int kf_synth_build_scene(int32_t n_spheres, int32_t n_boxes, uint64_t seed,
                         const float room_min[3], const float room_max[3], kf_synth_scene* out) {
  if (!out || n_spheres < 0 || n_spheres > KF_SYNTH_MAX_SPHERES || n_boxes < 0 ||
      n_boxes > KF_SYNTH_MAX_BOXES)
    return KF_ERR_INVALID;
  memset(out, 0, sizeof *out);
  for (int i = 0; i < 3; ++i) {
    out->room_min[i] = room_min ? room_min[i] : (float)kRoomMin[i];
    out->room_max[i] = room_max ? room_max[i] : (float)kRoomMax[i];
  }
  uint64_t s = seed * 0x2545f4914f6cdd1dull + 0x1234567ull;
  // Objects sit in front of the first camera (z forward, y down): x in
  // [-1000, 900], y in [-800, 1000], z from 600 mm to 100 mm before the far
  // wall, so the orbit (z <= 200 mm) never enters one and every object is
  // farther than the 300 mm depth gate.
  const double zfar = out->room_max[2] - 100.0;
  out->n_spheres = n_spheres;
  for (int k = 0; k < n_spheres; ++k) {
    const double r = lerp(200.0, 500.0, u01(s));
    const double zlo = 600.0 + r, zhi = std::max(zlo, zfar - r);
    out->spheres[k][0] = (float)lerp(-1000.0 + r, 900.0 - r, u01(s));
    out->spheres[k][1] = (float)lerp(-800.0 + r, 1000.0 - r, u01(s));
    out->spheres[k][2] = (float)lerp(zlo, zhi, u01(s));
    out->spheres[k][3] = (float)r;
  }
  out->n_boxes = n_boxes;
  for (int k = 0; k < n_boxes; ++k) {
    const double sx = lerp(200.0, 600.0, u01(s)), sy = lerp(200.0, 600.0, u01(s)),
                 sz = lerp(200.0, 600.0, u01(s));
    const double x0 = lerp(-1000.0, 900.0 - sx, u01(s));
    const double y0 = lerp(-800.0, 1000.0 - sy, u01(s));
    const double z0 = lerp(600.0, std::max(600.0, zfar - sz), u01(s));
    const double b[6] = {x0, y0, z0, x0 + sx, y0 + sy, z0 + sz};
    for (int i = 0; i < 6; ++i) out->boxes[k][i] = (float)b[i];
  }
  return KF_OK;
}

int kf_synth_preset_scene(int32_t which, kf_synth_scene* out) {
  int ns = 6, nb = 4;
  if (which == KF_PRESET_ROOM_HD_F16) ns = 12, nb = 8;
  else if (which == KF_PRESET_SPHERES_128) ns = 3, nb = 0;
  else if (which != KF_PRESET_ROOM_CPU && which != KF_PRESET_ROOM &&
           which != KF_PRESET_ROOM_5MM_F16 && which != KF_PRESET_REFERENCE)
    return KF_ERR_INVALID;
  return kf_synth_build_scene(ns, nb, 0, nullptr, nullptr, out);
}

// phi = 2 pi f / 300; position (250 sin phi, 50 sin 2phi, 100 (1 - cos phi)) mm;
// R = Ry(12 deg sin phi) Rx(4 deg sin 2phi).  Per frame: <= 6.1 mm, <= 0.31 deg.
int kf_synth_orbit_pose(int32_t frame, float T[16]) {
  if (!T) return KF_ERR_INVALID;
  const double pi = 3.14159265358979323846;
  const double phi = 2.0 * pi * frame / 300.0;
  const double yaw = 12.0 * pi / 180.0 * sin(phi);
  const double pitch = 4.0 * pi / 180.0 * sin(2.0 * phi);
  const double cy = cos(yaw), sy = sin(yaw), cp = cos(pitch), sp = sin(pitch);
  // Ry * Rx
  const double R[9] = {cy, sy * sp, sy * cp, 0.0, cp, -sp, -sy, cy * sp, cy * cp};
  const double p[3] = {250.0 * sin(phi), 50.0 * sin(2.0 * phi), 100.0 * (1.0 - cos(phi))};
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T[r * 4 + c] = (float)R[r * 3 + c];
    T[r * 4 + 3] = (float)p[r];
  }
  T[12] = T[13] = T[14] = 0.0f;
  T[15] = 1.0f;
  return KF_OK;
}

int kf_synth_render(const kf_synth_scene* scene, const float K[9], const float T[16],
                    int32_t H, int32_t W, float noise_scale, float hole_frac, float max_depth_mm,
                    uint64_t noise_seed, uint16_t* depth, float* clean_mm, int32_t n_threads) {
  if (!scene || !K || !T || !depth || H <= 0 || W <= 0) return KF_ERR_INVALID;
  double Kd[9], Ki[9];
  for (int i = 0; i < 9; ++i) Kd[i] = K[i];
  if (!inv3(Kd, Ki)) return KF_ERR_INVALID;
  double R[9], o[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) R[r * 3 + c] = T[r * 4 + c];
    o[r] = T[r * 4 + 3];
  }
  const kf_synth_scene s = *scene;
  if (n_threads <= 0) n_threads = (int)std::max(1u, std::thread::hardware_concurrency());
  parallel_rows(H, n_threads, [&](int y) {
    for (int x = 0; x < W; ++x) {
      // pixel (x, y) looks along K^-1 (x+1, y+1, 1): the reference's 1-indexed
      // convention, so that back-projection recovers the rendered points.
      const double u = x + 1.0, v = y + 1.0;
      const double dc[3] = {Ki[0] * u + Ki[1] * v + Ki[2], Ki[3] * u + Ki[4] * v + Ki[5],
                            Ki[6] * u + Ki[7] * v + Ki[8]};
      const double dcn[3] = {dc[0] / dc[2], dc[1] / dc[2], 1.0};  // unit z
      const double dw[3] = {R[0] * dcn[0] + R[1] * dcn[1] + R[2] * dcn[2],
                            R[3] * dcn[0] + R[4] * dcn[1] + R[5] * dcn[2],
                            R[6] * dcn[0] + R[7] * dcn[1] + R[8] * dcn[2]};
      const double z = trace(s, o, dw);  // t along a unit-z camera ray = depth
      const int64_t p = (int64_t)y * W + x;
      if (clean_mm) clean_mm[p] = std::isfinite(z) ? (float)z : 0.0f;
      double zn = std::isfinite(z) ? z : 0.0;
      uint64_t st = noise_seed * 0x9e3779b97f4a7c15ull ^ ((uint64_t)p * 0xd1b54a32d192ed03ull);
      const double u1 = u01(st), u2 = u01(st), u3 = u01(st);
      if (noise_scale > 0 && zn > 0) {
        const double zm = zn / 1000.0;
        const double sigma_mm = 1000.0 * (0.0012 + 0.0019 * (zm - 0.4) * (zm - 0.4));
        const double g = sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
        zn += noise_scale * sigma_mm * g;
      }
      if (u3 <= hole_frac) zn = 0.0;
      double q = floor(zn + 0.5);
      if (q < 0) q = 0;
      if (q > 65535) q = 65535;
      if (q > max_depth_mm) q = 0;
      depth[p] = (uint16_t)q;
    }
  });
  return KF_OK;
}

int kf_synth_stream(const kf_synth_scene* scene, const float K[9], int32_t H, int32_t W,
                    int32_t first, int32_t n, float noise_scale, float hole_frac,
                    float max_depth_mm, uint16_t* frames, float* gt_poses, int32_t n_threads) {
  if (!scene || !K || !frames || n <= 0) return KF_ERR_INVALID;
  for (int32_t i = 0; i < n; ++i) {
    float T[16];
    kf_synth_orbit_pose(first + i, T);
    if (gt_poses) memcpy(gt_poses + (int64_t)i * 16, T, sizeof T);
    const int rc = kf_synth_render(scene, K, T, H, W, noise_scale, hole_frac, max_depth_mm,
                                   1000ull + (uint64_t)(first + i),
                                   frames + (int64_t)i * H * W, nullptr, n_threads);
    if (rc != KF_OK) return rc;
  }
  return KF_OK;
}

}  // extern "C"
