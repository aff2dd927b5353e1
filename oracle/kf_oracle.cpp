// kf_oracle.cpp - CPU oracle for the KinectFusion hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is loaded by tests/, by the smoke
// check and by bench.py's cpu_baseline leg, as the checker or as the timed
// CPU baseline.  The product library never links, loads or calls it.
//
// Each function restates one reference kernel (PoChang007/Kinect_Fusion_SLAM,
// src/parallel_processing/*.cu) in plain C++: same operations in the same
// order, float arithmetic with -ffp-contract=off, IEEE division and sqrt.
// Templates on the scalar type S give the FP32 restatement (S = float) and an
// FP64 twin (S = double) used as the accuracy reference for the FP16 volume.
//
// Where the reference's behaviour is undefined or depends on an unpinned
// library, the oracle follows the conventions documented with each function;
// the HIP kernels follow the same ones.  Arithmetic shared with the product
// (FP64 4x4/3x3 inverse and product, LDL^T solve, exp, float<->half) comes
// from kf_hd_math.h and is pinned independently by tests against numpy.
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../kinect_fusion_slam_amd/csrc/kf_hd_math.h"

namespace {

// 180/pi as the reference computes it: 180.0f / CUDART_PI_F (3.141592654f).
const float kDegF = 180.0f / 3.141592654f;

inline int round_i(float v) { return kf_round_to_int(v); }
inline int round_i(double v) {
  if (!(v > -2147483648.0 && v < 2147483648.0)) return (int)0x80000000u;
  return (int)std::round(v);
}
inline int floor_i(float v) { return kf_floor_to_int(v); }
inline int floor_i(double v) {
  if (!(v > -2147483648.0 && v < 2147483648.0)) return (int)0x80000000u;
  return (int)std::floor(v);
}

// cuda_functions.cuh:2-10  v *= 1 / sqrt(v.v)
template <class S>
inline void normalize3(S& x, S& y, S& z) {
  const S t = S(1) / std::sqrt(x * x + y * y + z * z);
  x = x * t;
  y = y * t;
  z = z * t;
}

// Bilateral samples are read by linear index y*W + x, as the reference does
// (bilateral_filtering.cu:26); its inclusive bound y <= H - r reads one row
// past the image.  Convention: indices outside [0, H*W) read 0.
inline float linear_sample(const float* img, int64_t n, int64_t idx) {
  return (idx >= 0 && idx < n) ? img[idx] : 0.0f;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// A1  system_utility.cpp:58-68 (uint16 mm -> float) and :77-86 (cv::inRange,
// inclusive at both ends, then multiply by the 0/1 mask).
void oracle_range_gate(const uint16_t* in, float* out, int64_t n, float dmin, float dmax) {
  for (int64_t i = 0; i < n; ++i) {
    const float d = (float)in[i];
    out[i] = (d >= dmin && d <= dmax) ? d : 0.0f;
  }
}

// system_utility.cpp:88-113: X holds column offsets, Y row offsets in
// [-r, r]; w = exp(-(X.*X + Y.*Y) / (2 sigma^2)).  OpenCV's MatExpr scaling
// and cv::exp are not pinned by the reference; convention: float arithmetic
// in the written order and kf_expf.
void oracle_spatial_kernel(int r, float sigma, float* w) {
  const int L = 2 * r + 1;
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < L; ++j) {
      const float X = (float)(j - r), Y = (float)(i - r);
      const float t = -(X * X + Y * Y) / (2 * sigma * sigma);
      w[i * L + j] = kf_expf(t);
    }
}

// ---------------------------------------------------------------------------
// A4  bilateral_filtering.cu:3-47 with the wrapper's in-place semantics
// (:77, :94): pixels outside the filtered window keep the zero the output was
// initialised with.
void oracle_bilateral(const float* depth, float* out, int H, int W, const float* ws, int r,
                      float sigma_r) {
  const int64_t n = (int64_t)H * W;
  const int L = 2 * r + 1;
#pragma omp parallel
  {
    std::vector<float> nb((size_t)L * L);
#pragma omp for schedule(static)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const int64_t p = (int64_t)y * W + x;
        if (!(y >= r && y <= H - r && x >= r && x <= W - r)) {
          out[p] = 0.0f;
          continue;
        }
        const float dc = depth[p];
        float sum = 0.f;
        for (int yy = y - r; yy <= y + r; ++yy)
          for (int xx = x - r; xx <= x + r; ++xx) {
            float t = linear_sample(depth, n, (int64_t)yy * W + xx) - dc;
            t = kf_expf(-(t * t) / (2 * sigma_r * sigma_r));
            const int k = ((yy - y) + r) * L + ((xx - x) + r);
            const float w = ws[k] * t + 0.000000001f;
            nb[k] = w;
            sum = sum + w;
          }
        float fin = 0.f;
        for (int yy = y - r; yy <= y + r; ++yy)
          for (int xx = x - r; xx <= x + r; ++xx) {
            const int k = ((yy - y) + r) * L + ((xx - x) + r);
            const float wr = nb[k] / sum;
            fin = fin + wr * linear_sample(depth, n, (int64_t)yy * W + xx);
          }
        out[p] = fin;
      }
  }
}

// ---------------------------------------------------------------------------
// A3  raw_vertices.cu:3-51.  Outputs start at zero (CleanDepthData,
// kinfu_pipeline.cpp:206-216); the mask is set before the neighbour test.
void oracle_vertex_normals(const float* d, const float* Ki, int H, int W, float* vx, float* vy,
                           float* nx, float* ny, float* nz, uint8_t* mask) {
  const int64_t n = (int64_t)H * W;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    vx[i] = vy[i] = nx[i] = ny[i] = nz[i] = 0.0f;
    mask[i] = 0;
    const int y = (int)(i / W), x = (int)(i % W);
    if (d[i] != 0 && y != (H - 1) && x != (W - 1)) {
      mask[i] = 1;
      const int64_t nr = (int64_t)(y + 1) * W + x;
      const int64_t nc = (int64_t)y * W + (x + 1);
      if (nr > 0 && nc > 0 && d[nr] != 0 && d[nc] != 0) {
        const float X1 = (float)(x + 1), Y1 = (float)(y + 1);
        const float X2 = (float)(x + 2), Y2 = (float)(y + 2);
        const float rvx = d[i] * (Ki[0] * X1 + Ki[1] * Y1 + Ki[2]);
        const float rvy = d[i] * (Ki[3] * X1 + Ki[4] * Y1 + Ki[5]);
        const float rvx_r = d[nr] * (Ki[0] * X1 + Ki[1] * Y2 + Ki[2]);
        const float rvx_c = d[nc] * (Ki[0] * X2 + Ki[1] * Y1 + Ki[2]);
        const float rvy_r = d[nr] * (Ki[3] * X1 + Ki[4] * Y2 + Ki[5]);
        const float rvy_c = d[nc] * (Ki[3] * X2 + Ki[4] * Y1 + Ki[5]);
        const float nI = (rvy_r - rvy) * (d[nc] - d[i]) - (d[nr] - d[i]) * (rvy_c - rvy);
        const float nJ = (d[nr] - d[i]) * (rvx_c - rvx) - (rvx_r - rvx) * (d[nc] - d[i]);
        const float nK = (rvx_r - rvx) * (rvy_c - rvy) - (rvy_r - rvy) * (rvx_c - rvx);
        vx[i] = rvx;
        vy[i] = rvy;
        nx[i] = nI / sqrtf(nI * nI + nJ * nJ + nK * nK);
        ny[i] = nJ / sqrtf(nI * nI + nJ * nJ + nK * nK);
        nz[i] = nK / sqrtf(nI * nI + nJ * nJ + nK * nK);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Coarse-to-fine ICP pyramid.  No reference counterpart: the reference runs
// every ICP iteration at full resolution (sensor_pose_estimation.cu:339).
// Level l+1 has floor(H/2) x floor(W/2) pixels; its pixel (x, y) covers
// pixels (2x + i, 2y + j), i, j in {0, 1}, of level l.  Vertex and normal
// maps of a level are oracle_vertex_normals on its depth with its K^-1.

// Depth: the block mean ((d00 + d01) + (d10 + d11)) * 0.25 where all four
// depths are > 0 and max - min <= max_step, else 0.
void oracle_pyramid_depth(const float* d, int H, int W, float max_step, float* out) {
  const int Ho = H / 2, Wo = W / 2;
  for (int y = 0; y < Ho; ++y)
    for (int x = 0; x < Wo; ++x) {
      const float a = d[(int64_t)(2 * y) * W + 2 * x], b = d[(int64_t)(2 * y) * W + 2 * x + 1];
      const float c = d[(int64_t)(2 * y + 1) * W + 2 * x];
      const float e = d[(int64_t)(2 * y + 1) * W + 2 * x + 1];
      const float lo = std::min(std::min(a, b), std::min(c, e));
      const float hi = std::max(std::max(a, b), std::max(c, e));
      const bool ok = a > 0.0f && b > 0.0f && c > 0.0f && e > 0.0f && hi - lo <= max_step;
      out[(int64_t)y * Wo + x] = ok ? ((a + b) + (c + e)) * 0.25f : 0.0f;
    }
}

// Model maps: where all four pixels hold a surface (ok_in), the block mean
// of the vertices and the sum of the normals scaled by 1 / sqrtf(|sum|^2)
// (the normalisation of cuda_functions.cuh:2-10); elsewhere zeros, ok 0.
void oracle_pyramid_model(const float* const in[6], const uint8_t* ok_in, int H, int W,
                          float* const out[6], uint8_t* ok_out) {
  const int Ho = H / 2, Wo = W / 2;
  for (int y = 0; y < Ho; ++y)
    for (int x = 0; x < Wo; ++x) {
      const int64_t o = (int64_t)y * Wo + x;
      const int64_t p[4] = {(int64_t)(2 * y) * W + 2 * x, (int64_t)(2 * y) * W + 2 * x + 1,
                            (int64_t)(2 * y + 1) * W + 2 * x, (int64_t)(2 * y + 1) * W + 2 * x + 1};
      const bool ok = ok_in[p[0]] && ok_in[p[1]] && ok_in[p[2]] && ok_in[p[3]];
      ok_out[o] = ok ? 1 : 0;
      for (int k = 0; k < 6; ++k) out[k][o] = 0.0f;
      if (!ok) continue;
      for (int k = 0; k < 3; ++k)
        out[k][o] = ((in[k][p[0]] + in[k][p[1]]) + (in[k][p[2]] + in[k][p[3]])) * 0.25f;
      float n[3];
      for (int k = 0; k < 3; ++k)
        n[k] = (in[3 + k][p[0]] + in[3 + k][p[1]]) + (in[3 + k][p[2]] + in[3 + k][p[3]]);
      const float t = 1 / sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int k = 0; k < 3; ++k) out[3 + k][o] = n[k] * t;
    }
}

// Intrinsics of the next coarser level: a coarse pixel u (1-indexed) covers
// the fine pixels 2u - 1 and 2u, centred at 2u - 0.5, so u_coarse =
// (u_fine + 0.5) / 2: fx and the skew halve, cx -> (cx + 0.5) / 2 (and y).
void oracle_pyramid_intrinsics(const float* K, float* Kc) {
  for (int i = 0; i < 9; ++i) Kc[i] = K[i];
  Kc[0] = K[0] * 0.5f;
  Kc[1] = K[1] * 0.5f;
  Kc[2] = (K[2] + 0.5f) * 0.5f;
  Kc[4] = K[4] * 0.5f;
  Kc[5] = (K[5] + 0.5f) * 0.5f;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// A8  projective_tsdf.cu:5-140.  One voxel at a time, exactly as written:
// the voxel centre is (i + start) * d + d / 2 (kinfu_pipeline.cpp:53-55),
// K^-1 and T^-1 come from kf_m3_inverse / kf_m4_inverse (FP64, rounded).
namespace {

template <class S>
struct VolumeF {  // separate TSDF and weight arrays, the reference's layout
  S* tsdf;
  S* weight;
  inline bool load(int64_t i, S& F, S& Wt) const {
    F = tsdf[i];
    Wt = weight[i];
    return true;
  }
};

template <class S>
inline bool integrate_voxel(int ix, int iy, int iz, const int* start, S d, const S* K,
                            const S* Ki, const S* T, const S* Ti, const float* depth, int H,
                            int W, S& frk_out) {
  const S cx = (S)(ix + start[0]) * d + d / S(2);
  const S cy = (S)(iy + start[1]) * d + d / S(2);
  const S cz = (S)(iz + start[2]) * d + d / S(2);
  const S lx = Ti[0] * cx + Ti[1] * cy + Ti[2] * cz + Ti[3];
  const S ly = Ti[4] * cx + Ti[5] * cy + Ti[6] * cz + Ti[7];
  const S lz = Ti[8] * cx + Ti[9] * cy + Ti[10] * cz + Ti[11];
  const S px = K[0] * lx + K[1] * ly + K[2] * lz;
  const S py = K[3] * lx + K[4] * ly + K[5] * lz;
  const S pz = K[6] * lx + K[7] * ly + K[8] * lz;
  const int u = round_i(px / pz);
  const int v = round_i(py / pz);
  if (!(u > 0 && u <= W)) return false;
  if (!(v > 0 && v <= H)) return false;
  const S dv = (S)depth[(int64_t)(v - 1) * W + (u - 1)];
  if (!(dv != 0)) return false;
  const S U = (S)u, V = (S)v;
  const S lamx = Ki[0] * U + Ki[1] * V + Ki[2] * S(1);
  const S lamy = Ki[3] * U + Ki[4] * V + Ki[5] * S(1);
  const S lamz = Ki[6] * U + Ki[7] * V + Ki[8] * S(1);
  const S rsx = T[0] * lamx + T[1] * lamy + T[2] * lamz + T[3];
  const S rsy = T[4] * lamx + T[5] * lamy + T[6] * lamz + T[7];
  const S rsz = T[8] * lamx + T[9] * lamy + T[10] * lamz + T[11];
  const S rdx = T[0] * 2 * lamx + T[1] * 2 * lamy + T[2] * 2 * lamz + T[3];
  const S rdy = T[4] * 2 * lamx + T[5] * 2 * lamy + T[6] * 2 * lamz + T[7];
  const S rdz = T[8] * 2 * lamx + T[9] * 2 * lamy + T[10] * 2 * lamz + T[11];
  const S dx = rdx - rsx, dy = rdy - rsy, dz = rdz - rsz;
  const S rm = std::sqrt(dx * dx + dy * dy + dz * dz);
  const S tdx = cx - T[3], tdy = cy - T[7], tdz = cz - T[11];
  // front test with the reference's swapped principal point (K[5], K[2])
  const S cax = Ki[0] * K[5] + Ki[1] * K[2] + Ki[2] * S(1);
  const S cay = Ki[3] * K[5] + Ki[4] * K[2] + Ki[5] * S(1);
  const S caz = Ki[6] * K[5] + Ki[7] * K[2] + Ki[8] * S(1);
  const S gax = T[0] * cax + T[1] * cay + T[2] * caz + T[3];
  const S gay = T[4] * cax + T[5] * cay + T[6] * caz + T[7];
  const S gaz = T[8] * cax + T[9] * cay + T[10] * caz + T[11];
  const S cbx = 2 * (Ki[0] * K[5] + Ki[1] * K[2] + Ki[2] * S(1));
  const S cby = 2 * (Ki[3] * K[5] + Ki[4] * K[2] + Ki[5] * S(1));
  const S cbz = 2 * (Ki[6] * K[5] + Ki[7] * K[2] + Ki[8] * S(1));
  const S gbx = T[0] * cbx + T[1] * cby + T[2] * cbz + T[3];
  const S gby = T[4] * cbx + T[5] * cby + T[6] * cbz + T[7];
  const S gbz = T[8] * cbx + T[9] * cby + T[10] * cbz + T[11];
  const S fx = gbx - gax, fy = gby - gay, fz = gbz - gaz;
  const S mag = fx * tdx + fy * tdy + fz * tdz;
  if (!(mag > 0)) return false;
  const S tm = std::sqrt(tdx * tdx + tdy * tdy + tdz * tdz);
  const S p3x = dv * lamx, p3y = dv * lamy, p3z = dv * lamz;
  const S pgx = T[0] * p3x + T[1] * p3y + T[2] * p3z + T[3];
  const S pgy = T[4] * p3x + T[5] * p3y + T[6] * p3z + T[7];
  const S pgz = T[8] * p3x + T[9] * p3y + T[10] * p3z + T[11];
  const S rx = pgx - T[3], ry = pgy - T[7], rz = pgz - T[11];
  const S rpd = std::sqrt(rx * rx + ry * ry + rz * rz);
  const S pd = rpd / rm;
  frk_out = pd - ((S(1) / rm) * tm);
  return true;
}

template <class S>
void integrate_impl(const float* depth, int H, int W, const float* Kf, const float* Tf,
                    const int* dims, const int* start, float voxel_mm, float trunc, float smin,
                    float smax, int first, S* tsdf, S* weight) {
  float Kif[9], Tif[16];
  kf_m3_inverse(Kf, Kif);
  kf_m4_inverse(Tf, Tif);
  S K[9], Ki[9], T[16], Ti[16];
  for (int i = 0; i < 9; ++i) K[i] = (S)Kf[i], Ki[i] = (S)Kif[i];
  for (int i = 0; i < 16; ++i) T[i] = (S)Tf[i], Ti[i] = (S)Tif[i];
  const S d = (S)voxel_mm, tr = (S)trunc, lo = (S)smin, hi = (S)smax;
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  const S WRK = S(1);
#pragma omp parallel for collapse(2) schedule(static)
  for (int ix = 0; ix < Nx; ++ix)
    for (int iy = 0; iy < Ny; ++iy)
      for (int iz = 0; iz < Nz; ++iz) {
        S frk;
        if (!integrate_voxel<S>(ix, iy, iz, start, d, K, Ki, T, Ti, depth, H, W, frk)) continue;
        const int64_t x = ((int64_t)ix * Ny + iy) * Nz + iz;
        if (first) {
          if (frk > 0 && frk <= hi) {
            tsdf[x] = std::fmin(S(1), (frk / tr));
            weight[x] = WRK;
          } else if (frk <= 0 && frk >= lo) {
            tsdf[x] = std::fmax(S(-1), frk / tr);
            weight[x] = WRK;
          }
        } else {
          if (frk > 0 && frk <= hi) {
            if (std::isnan(tsdf[x])) tsdf[x] = S(0);
            if (std::isnan(weight[x])) weight[x] = S(0);
            frk = std::fmin(S(1), (frk / tr));
            tsdf[x] = (weight[x] * tsdf[x] + WRK * frk) / (weight[x] + WRK);
            weight[x] = weight[x] + WRK;
          } else if (frk <= 0 && frk >= lo) {
            if (std::isnan(tsdf[x])) tsdf[x] = S(0);
            if (std::isnan(weight[x])) weight[x] = S(0);
            frk = std::fmax(S(-1), frk / tr);
            tsdf[x] = (weight[x] * tsdf[x] + WRK * frk) / (weight[x] + WRK);
            weight[x] = weight[x] + WRK;
          }
        }
      }
}

}  // namespace

extern "C" {

void oracle_integrate_f32(const float* depth, int H, int W, const float* K, const float* T,
                          const int* dims, const int* start, float voxel_mm, float trunc,
                          float smin, float smax, int first, float* tsdf, float* weight) {
  integrate_impl<float>(depth, H, W, K, T, dims, start, voxel_mm, trunc, smin, smax, first, tsdf,
                        weight);
}

// FP64 twin: the same restatement evaluated in double on a double volume.
void oracle_integrate_f64(const float* depth, int H, int W, const float* K, const float* T,
                          const int* dims, const int* start, float voxel_mm, float trunc,
                          float smin, float smax, int first, double* tsdf, double* weight) {
  integrate_impl<double>(depth, H, W, K, T, dims, start, voxel_mm, trunc, smin, smax, first, tsdf,
                         weight);
}

// FP16 storage: the FP32 restatement with the TSDF held as IEEE half (NaN =
// never observed) and the weight as uint16 capped at 65535.  Each update is
// computed in float from the widened values and rounded to half once.
void oracle_integrate_f16(const float* depth, int H, int W, const float* Kf, const float* Tf,
                          const int* dims, const int* start, float voxel_mm, float trunc,
                          float smin, float smax, int first, uint16_t* tsdf_h, uint16_t* w16) {
  float Ki[9], Ti[16];
  kf_m3_inverse(Kf, Ki);
  kf_m4_inverse(Tf, Ti);
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
#pragma omp parallel for collapse(2) schedule(static)
  for (int ix = 0; ix < Nx; ++ix)
    for (int iy = 0; iy < Ny; ++iy)
      for (int iz = 0; iz < Nz; ++iz) {
        float frk;
        if (!integrate_voxel<float>(ix, iy, iz, start, voxel_mm, Kf, Ki, Tf, Ti, depth, H, W,
                                    frk))
          continue;
        const int64_t x = ((int64_t)ix * Ny + iy) * Nz + iz;
        const bool pos = frk > 0 && frk <= smax;
        const bool neg = !pos && frk <= 0 && frk >= smin;
        if (!pos && !neg) continue;
        const float f = pos ? fminf(1.f, frk / trunc) : fmaxf(-1.f, frk / trunc);
        if (first) {
          tsdf_h[x] = kf_f32_to_f16_bits(f);
          w16[x] = 1;
        } else {
          float F = kf_f16_bits_to_f32(tsdf_h[x]);
          if (std::isnan(F)) F = 0.0f;
          const float Wt = (float)w16[x];
          const float Fn = (Wt * F + 1.0f * f) / (Wt + 1.0f);
          tsdf_h[x] = kf_f32_to_f16_bits(Fn);
          w16[x] = (uint16_t)(w16[x] == 65535 ? 65535 : w16[x] + 1);
        }
      }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// A9/A10  voxel_traversal.cu:4-424 and the wrapper's bounds (:439-446).
//
// Conventions for what the reference leaves undefined:
//  * XY, XZ, YX, YZ, ZX, ZY (read before any step on that axis,
//    voxel_traversal.cu:136) start at the clamped start voxel's indices.
//  * The normal flip acos((double)c)*(180/pi_f) > 90 uses libm acos.
//  * fix_bounds = 0 keeps the reference's bound checks (:233-281), which test
//    XY/XZ against voxel_length, YX/YZ against voxel_width and ZX/ZY against
//    voxel_height; fix_bounds = 1 tests each index against its own axis.
namespace {

template <class S>
inline bool ray_box(S& tmin_out, S dx, S dy, S dz, S mnx, S mny, S mnz, S mxx, S mxy, S mxz,
                    S sx, S sy, S sz) {
  S tmin, tmax, tymin, tymax, tzmin, tzmax;
  if (dx >= 0) {
    tmin = (mnx - sx) / dx;
    tmax = (mxx - sx) / dx;
  } else {
    tmin = (mxx - sx) / dx;
    tmax = (mnx - sx) / dx;
  }
  if (dy >= 0) {
    tymin = (mny - sy) / dy;
    tymax = (mxy - sy) / dy;
  } else {
    tymin = (mxy - sy) / dy;
    tymax = (mny - sy) / dy;
  }
  if ((tmin > tymax) || (tymin > tmax)) {
    tmin_out = -1;
    return false;
  }
  if (tymin > tmin) tmin = tymin;
  if (tymax < tmax) tmax = tymax;
  if (dz >= 0) {
    tzmin = (mnz - sz) / dz;
    tzmax = (mxz - sz) / dz;
  } else {
    tzmin = (mxz - sz) / dz;
    tzmax = (mnz - sz) / dz;
  }
  if ((tmin > tzmax) || (tzmin > tmax)) {
    tmin_out = -1;
    return false;
  }
  if (tzmin > tmin) tmin = tzmin;
  if (tzmax < tmax) tmax = tzmax;
  tmin_out = tmin;
  return true;
}

struct RayStats {
  int64_t steps;
  int64_t hits;
};

// Trilinear zero-crossing refinement (the differentiable raycast mode; no
// reference counterpart).  After the walk's crossing voxel, entered at ray
// parameter te, the TSDF is sampled trilinearly at p_k = sg + (k d) dir for
// k = floor(te/d) - 2 ... + 5; the first valid pair with F_k > 0 >= F_k+1
// gives t* = t_k + d F_k / (F_k - F_k+1), p* = sg + t* dir, the camera depth
// sc + t*/|lambda| and the normal grad F(p*) / |grad F(p*)|.  Samples are in
// voxel-centre coordinates u = (p - min)/d - 1/2, valid when all 8 corners
// are inside the volume and observed.
template <class S>
struct TriS {
  S v, g[3], w[3];
  int i, j, k;
  S c[8];  // corner (dx<<2)|(dy<<1)|dz
};

template <class S>
struct TriCtx {
  S* out7;          // x, y, z, nx, ny, nz, depth of this pixel
  int32_t* disc;    // discrete state of the hit: sample k and 3 cells (nullable)
  // adjoint (gF non-null): dL/d(depth), dL/d(normal)
  S gD;
  const S* gN;      // 3 values (nullable)
  S* gF;            // dense dL/dF, reference layout
  S* gxi;           // 6 values
};

template <class S, class Vol>
inline bool tri_sample(const Vol& F, const int* dims, const int* start, S d, S px, S py, S pz,
                       TriS<S>& s) {
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  const S ux = (px - (S)(float)start[0] * d) / d - S(0.5);
  const S uy = (py - (S)(float)start[1] * d) / d - S(0.5);
  const S uz = (pz - (S)(float)start[2] * d) / d - S(0.5);
  const int i = floor_i(ux), j = floor_i(uy), k = floor_i(uz);
  if (!(i >= 0 && i < Nx - 1 && j >= 0 && j < Ny - 1 && k >= 0 && k < Nz - 1)) return false;
  for (int q = 0; q < 8; ++q) {
    const int64_t lin = ((int64_t)(i + (q >> 2)) * Ny + (j + ((q >> 1) & 1))) * Nz + (k + (q & 1));
    s.c[q] = F(lin);
    if (std::isnan(s.c[q])) return false;
  }
  const S fx = ux - (S)i, fy = uy - (S)j, fz = uz - (S)k;
  const S ox = S(1) - fx, oy = S(1) - fy, oz = S(1) - fz;
  const S c00 = s.c[0] * ox + s.c[4] * fx;
  const S c01 = s.c[1] * ox + s.c[5] * fx;
  const S c10 = s.c[2] * ox + s.c[6] * fx;
  const S c11 = s.c[3] * ox + s.c[7] * fx;
  const S c0 = c00 * oy + c10 * fy;
  const S c1 = c01 * oy + c11 * fy;
  s.v = c0 * oz + c1 * fz;
  const S e00 = s.c[4] - s.c[0], e01 = s.c[5] - s.c[1];
  const S e10 = s.c[6] - s.c[2], e11 = s.c[7] - s.c[3];
  s.g[0] = ((e00 * oy + e10 * fy) * oz + (e01 * oy + e11 * fy) * fz) / d;
  s.g[1] = ((c10 - c00) * oz + (c11 - c01) * fz) / d;
  s.g[2] = (c1 - c0) / d;
  s.w[0] = fx;
  s.w[1] = fy;
  s.w[2] = fz;
  s.i = i;
  s.j = j;
  s.k = k;
  return true;
}

// dL/dF of the corners of s: a * weight + b . grad(weight)
template <class S>
inline void tri_scatter(S* gF, const int* dims, S d, const TriS<S>& s, S a, S bx, S by, S bz) {
  const int Ny = dims[1], Nz = dims[2];
  const S fx = s.w[0], fy = s.w[1], fz = s.w[2];
  const S ox = S(1) - fx, oy = S(1) - fy, oz = S(1) - fz;
  for (int q = 0; q < 8; ++q) {
    const int qx = q >> 2, qy = (q >> 1) & 1, qz = q & 1;
    const S wx = qx ? fx : ox, wy = qy ? fy : oy, wz = qz ? fz : oz;
    const S sx = qx ? S(1) : S(-1), sy = qy ? S(1) : S(-1), sz = qz ? S(1) : S(-1);
    const S val = a * (wx * wy * wz) +
                  (bx * (sx * wy * wz) + by * (wx * sy * wz) + bz * (wx * wy * sz)) / d;
    const int64_t lin = ((int64_t)(s.i + qx) * Ny + (s.j + qy)) * Nz + (s.k + qz);
#pragma omp atomic
    gF[lin] += val;
  }
}

template <class S, class Vol>
inline bool tri_cross(const Vol& F, const int* dims, const int* start, S d, S sgx, S sgy, S sgz,
                      S dx, S dy, S dz, S te, S sc, S lam, const TriCtx<S>& tc) {
  if (!(te >= S(0) && te < S(1.0e9))) return false;
  const int kb = floor_i(te / d) - 2;
  TriS<S> a{}, b{};
  bool va = false, found = false;
  int ka = 0;
  for (int k = kb; k < kb + 6; ++k) {
    const S t = (S)(float)k * d;
    const bool ok = tri_sample<S>(F, dims, start, d, sgx + t * dx, sgy + t * dy, sgz + t * dz, b);
    if (ok && va && a.v > S(0) && b.v <= S(0)) {
      ka = k - 1;
      found = true;
      break;
    }
    va = ok;
    if (ok) a = b;
  }
  if (!found) return false;
  const S ta = (S)(float)ka * d;
  const S den = a.v - b.v;
  const S ts = ta + d * (a.v / den);
  const S psx = sgx + ts * dx, psy = sgy + ts * dy, psz = sgz + ts * dz;
  TriS<S> s;
  if (!tri_sample<S>(F, dims, start, d, psx, psy, psz, s)) return false;
  const S gn = std::sqrt(s.g[0] * s.g[0] + s.g[1] * s.g[1] + s.g[2] * s.g[2]);
  if (!(gn > S(0))) return false;
  const S nx = s.g[0] / gn, ny = s.g[1] / gn, nz = s.g[2] / gn;
  S* o = tc.out7;
  o[0] = psx;
  o[1] = psy;
  o[2] = psz;
  o[3] = nx;
  o[4] = ny;
  o[5] = nz;
  o[6] = sc + ts / lam;
  if (tc.disc) {
    const int Ny = dims[1], Nz = dims[2];
    tc.disc[0] = ka;
    tc.disc[1] = (a.i * Ny + a.j) * Nz + a.k;
    tc.disc[2] = (b.i * Ny + b.j) * Nz + b.k;
    tc.disc[3] = (s.i * Ny + s.j) * Nz + s.k;
  }
  if (!tc.gF) return true;
  // adjoint
  S dts = tc.gD / lam;
  const S gx = tc.gN ? tc.gN[0] : S(0), gy = tc.gN ? tc.gN[1] : S(0), gz = tc.gN ? tc.gN[2] : S(0);
  const S ndot = nx * gx + ny * gy + nz * gz;
  const S lgx = (gx - nx * ndot) / gn, lgy = (gy - ny * ndot) / gn, lgz = (gz - nz * ndot) / gn;
  tri_scatter<S>(tc.gF, dims, d, s, S(0), lgx, lgy, lgz);
  const S fx = s.w[0], fy = s.w[1], fz = s.w[2];
  const S ox = S(1) - fx, oy = S(1) - fy, oz = S(1) - fz;
  const S e00 = s.c[4] - s.c[0], e01 = s.c[5] - s.c[1];
  const S e10 = s.c[6] - s.c[2], e11 = s.c[7] - s.c[3];
  const S c00 = s.c[0] * ox + s.c[4] * fx, c01 = s.c[1] * ox + s.c[5] * fx;
  const S c10 = s.c[2] * ox + s.c[6] * fx, c11 = s.c[3] * ox + s.c[7] * fx;
  const S dd = d * d;
  const S hxy = ((e10 - e00) * oz + (e11 - e01) * fz) / dd;
  const S hxz = ((e01 * oy + e11 * fy) - (e00 * oy + e10 * fy)) / dd;
  const S hyz = ((c11 - c01) - (c10 - c00)) / dd;
  const S lpx = hxy * lgy + hxz * lgz;
  const S lpy = hxy * lgx + hyz * lgz;
  const S lpz = hxz * lgx + hyz * lgy;
  dts += lpx * dx + lpy * dy + lpz * dz;
  S xi[6];
  xi[0] = psy * lpz - psz * lpy;
  xi[1] = psz * lpx - psx * lpz;
  xi[2] = psx * lpy - psy * lpx;
  xi[3] = lpx;
  xi[4] = lpy;
  xi[5] = lpz;
  const S den2 = den * den;
  const S dFa = dts * d * (-b.v) / den2;
  const S dFb = dts * d * a.v / den2;
  tri_scatter<S>(tc.gF, dims, d, a, dFa, S(0), S(0), S(0));
  tri_scatter<S>(tc.gF, dims, d, b, dFb, S(0), S(0), S(0));
  const S tb = ta + d;
  const S pax = sgx + ta * dx, pay = sgy + ta * dy, paz = sgz + ta * dz;
  const S pbx = sgx + tb * dx, pby = sgy + tb * dy, pbz = sgz + tb * dz;
  xi[0] += dFa * (pay * a.g[2] - paz * a.g[1]) + dFb * (pby * b.g[2] - pbz * b.g[1]);
  xi[1] += dFa * (paz * a.g[0] - pax * a.g[2]) + dFb * (pbz * b.g[0] - pbx * b.g[2]);
  xi[2] += dFa * (pax * a.g[1] - pay * a.g[0]) + dFb * (pbx * b.g[1] - pby * b.g[0]);
  xi[3] += dFa * a.g[0] + dFb * b.g[0];
  xi[4] += dFa * a.g[1] + dFb * b.g[1];
  xi[5] += dFa * a.g[2] + dFb * b.g[2];
  for (int q = 0; q < 6; ++q) {
#pragma omp atomic
    tc.gxi[q] += xi[q];
  }
  return true;
}

template <class S, class Vol>
inline int raycast_pixel(int x, int y, int W, float dval_f, const S* Ki, const S* T,
                         const Vol& F, const int* dims, const int* start, S d, S trunc,
                         int fix_bounds, S* out6, int64_t* steps,
                         const TriCtx<S>* tri = nullptr) {
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  const S dval = (S)dval_f;
  const S X = (S)(float)x + S(1), Y = (S)(float)y + S(1);
  const S lamx = Ki[0] * X + Ki[1] * Y + Ki[2] * S(1);
  const S lamy = Ki[3] * X + Ki[4] * Y + Ki[5] * S(1);
  const S lamz = Ki[6] * X + Ki[7] * Y + Ki[8] * S(1);
  const S lam = std::sqrt(lamx * lamx + lamy * lamy + lamz * lamz);
  const S slx = ((lam * dval - trunc) / lam) * lamx;
  const S sly = ((lam * dval - trunc) / lam) * lamy;
  const S slz = ((lam * dval - trunc) / lam) * lamz;
  const S sgx = T[0] * slx + T[1] * sly + T[2] * slz + T[3] * S(1);
  const S sgy = T[4] * slx + T[5] * sly + T[6] * slz + T[7] * S(1);
  const S sgz = T[8] * slx + T[9] * sly + T[10] * slz + T[11] * S(1);
  const S camx = T[3], camy = T[7], camz = T[11];
  const S rgx = T[0] * lamx + T[1] * lamy + T[2] * lamz + T[3] * S(1);
  const S rgy = T[4] * lamx + T[5] * lamy + T[6] * lamz + T[7] * S(1);
  const S rgz = T[8] * lamx + T[9] * lamy + T[10] * lamz + T[11] * S(1);
  S dx = rgx - camx, dy = rgy - camy, dz = rgz - camz;
  normalize3(dx, dy, dz);
  // wrapper bounds (voxel_traversal.cu:439-443)
  const S mnx = (S)(float)start[0] * d, mny = (S)(float)start[1] * d, mnz = (S)(float)start[2] * d;
  const S mxx = (S)(float)(start[0] + Nx) * d, mxy = (S)(float)(start[1] + Ny) * d,
          mxz = (S)(float)(start[2] + Nz) * d;
  S tmin;
  if (!ray_box<S>(tmin, dx, dy, dz, mnx, mny, mnz, mxx, mxy, mxz, sgx, sgy, sgz)) return -1;
  if (tmin < 0) tmin = 0;
  const S stx = sgx + tmin * dx, sty = sgy + tmin * dy, stz = sgz + tmin * dz;
  const S bx = mxx - mnx, by = mxy - mny, bz = mxz - mnz;
  int ix = floor_i(((stx - mnx) / bx) * (S)Nx) + 1;
  int iy = floor_i(((sty - mny) / by) * (S)Ny) + 1;
  int iz = floor_i(((stz - mnz) / bz) * (S)Nz) + 1;
  if (ix == Nx + 1) ix = ix - 1;
  if (iy == Ny + 1) iy = iy - 1;
  if (iz == Nz + 1) iz = iz - 1;
  int XY = iy, XZ = iz, YX = ix, YZ = iz, ZX = ix, ZY = iy;
  S tvx, tvy, tvz;
  int sx_, sy_, sz_;
  if (dx >= 0) {
    tvx = (S)ix / (S)Nx;
    sx_ = 1;
  } else {
    tvx = (S)(ix - 1) / (S)Nx;
    sx_ = -1;
  }
  if (dy >= 0) {
    tvy = (S)iy / (S)Ny;
    sy_ = 1;
  } else {
    tvy = (S)(iy - 1) / (S)Ny;
    sy_ = -1;
  }
  if (dz >= 0) {
    tvz = (S)iz / (S)Nz;
    sz_ = 1;
  } else {
    tvz = (S)(iz - 1) / (S)Nz;
    sz_ = -1;
  }
  const S vmx = mnx + tvx * bx, vmy = mny + tvy * by, vmz = mnz + tvz * bz;
  S tmx = tmin + ((vmx - stx) / dx);
  S tmy = tmin + ((vmy - sty) / dy);
  S tmz = tmin + ((vmz - stz) / dz);
  const S tdx = d / std::fabs(dx), tdy = d / std::fabs(dy), tdz = d / std::fabs(dz);
  S prev = 0;
  S te = tmin;  // ray parameter where the walk entered the current voxel
  const int64_t NyNz = (int64_t)Ny * Nz;
  int64_t vi = (int64_t)(ix - 1) * NyNz + (int64_t)(iy - 1) * Nz + (iz - 1);
  const int bXY = fix_bounds ? Ny : Nx, bXZ = fix_bounds ? Nz : Nx;
  const int bYX = fix_bounds ? Nx : Ny, bYZ = fix_bounds ? Nz : Ny;
  const int bZX = fix_bounds ? Nx : Nz, bZY = fix_bounds ? Ny : Nz;
  while (ix < Nx - 1 && ix > 0 && iy < Ny - 1 && iy > 0 && iz < Nz - 1 && iz > 0) {
    ++*steps;
    const S f = F(vi);
    if (std::isnan(f) && prev < 0) break;
    if (f < 0 && std::isnan(prev)) break;
    if (f > 0 && prev < 0) break;
    if (f < 0 && prev > 0 && tri) {
      const S sc = (lam * dval - trunc) / lam;
      return tri_cross<S>(F, dims, start, d, sgx, sgy, sgz, dx, dy, dz, te, sc, lam, *tri)
                 ? (int)vi
                 : -1;
    }
    if (f < 0 && prev > 0) {
      const int xn = (ix - 1) - 1, xp = (ix - 1) + 1;
      const int yn = (iy - 1) - 1, yp = (iy - 1) + 1;
      const int zn = (iz - 1) - 1, zp = (iz - 1) + 1;
      if (xn >= Nx || xn < 0) break;
      if (xp >= Nx || xp < 0) break;
      if ((XY - 1) >= bXY || (XY - 1) < 0) break;
      if ((XZ - 1) >= bXZ || (XZ - 1) < 0) break;
      if (yn >= Ny || yn < 0) break;
      if (yp >= Ny || yp < 0) break;
      if ((YX - 1) >= bYX || (YX - 1) < 0) break;
      if ((YZ - 1) >= bYZ || (YZ - 1) < 0) break;
      if (zn >= Nz || zn < 0) break;
      if (zp >= Nz || zp < 0) break;
      if ((ZX - 1) >= bZX || (ZX - 1) < 0) break;
      if ((ZY - 1) >= bZY || (ZY - 1) < 0) break;
      const int64_t x_np = xn * NyNz + (int64_t)(XY - 1) * Nz + XZ - 1;
      const int64_t x_pp = xp * NyNz + (int64_t)(XY - 1) * Nz + XZ - 1;
      const int64_t y_np = (int64_t)(YX - 1) * NyNz + (int64_t)yn * Nz + YZ - 1;
      const int64_t y_pp = (int64_t)(YX - 1) * NyNz + (int64_t)yp * Nz + YZ - 1;
      const int64_t z_np = (int64_t)(ZX - 1) * NyNz + (int64_t)(ZY - 1) * Nz + zn;
      const int64_t z_pp = (int64_t)(ZX - 1) * NyNz + (int64_t)(ZY - 1) * Nz + zp;
      const int64_t x_nn = (int64_t)(ix - 2) * NyNz + (int64_t)(iy - 1) * Nz + (iz - 1);
      const int64_t x_pn = (int64_t)(ix)*NyNz + (int64_t)(iy - 1) * Nz + (iz - 1);
      const int64_t y_nn = (int64_t)(ix - 1) * NyNz + (int64_t)(iy - 2) * Nz + (iz - 1);
      const int64_t y_pn = (int64_t)(ix - 1) * NyNz + (int64_t)(iy)*Nz + (iz - 1);
      const int64_t z_nn = (int64_t)(ix - 1) * NyNz + (int64_t)(iy - 1) * Nz + (iz - 2);
      const int64_t z_pn = (int64_t)(ix - 1) * NyNz + (int64_t)(iy - 1) * Nz + (iz);
      if (std::isnan(F(x_np)) || std::isnan(F(x_pp)) || std::isnan(F(y_np)) ||
          std::isnan(F(y_pp)) || std::isnan(F(z_np)) || std::isnan(F(z_pp)) || std::isnan(F(vi)))
        break;
      if (std::isnan(F(x_nn)) || std::isnan(F(x_pn)) || std::isnan(F(y_nn)) ||
          std::isnan(F(y_pn)) || std::isnan(F(z_nn)) || std::isnan(F(z_pn)))
        break;
      const S fc = F(vi);
      S spx, spy, spz;
      if (dx > 0) {
        const S g = (S)(xn + start[0]) * d + d / S(2);
        spx = g + d * (std::fabs(F(x_np)) / (std::fabs(F(x_np) - fc)));
      } else {
        const S g = (S)(xp + start[0]) * d + d / S(2);
        spx = g - d * (std::fabs(F(x_pp)) / (std::fabs(F(x_pp) - fc)));
      }
      if (dy > 0) {
        const S g = (S)(yn + start[1]) * d + d / S(2);
        spy = g + d * (std::fabs(F(y_np)) / (std::fabs(F(y_np) - fc)));
      } else {
        const S g = (S)(yp + start[1]) * d + d / S(2);
        spy = g - d * (std::fabs(F(y_pp)) / (std::fabs(F(y_pp) - fc)));
      }
      if (dz > 0) {
        const S g = (S)(zn + start[2]) * d + d / S(2);
        spz = g + d * (std::fabs(F(z_np)) / (std::fabs(F(z_np) - fc)));
      } else {
        const S g = (S)(zp + start[2]) * d + d / S(2);
        spz = g - d * (std::fabs(F(z_pp)) / (std::fabs(F(z_pp) - fc)));
      }
      const S snx = (F(x_pn) - F(x_nn)) / S(2);
      const S sny = (F(y_pn) - F(y_nn)) / S(2);
      const S snz = (F(z_pn) - F(z_nn)) / S(2);
      S ax = camx - spx, ay = camy - spy, az = camz - spz;
      normalize3(ax, ay, az);
      S bxn = snx, byn = sny, bzn = snz;
      normalize3(bxn, byn, bzn);
      const double cos_theta = (double)(ax * bxn + ay * byn + az * bzn);
      const double angle = std::acos(cos_theta) * (double)kDegF;
      if (angle > 90) {
        bxn = -bxn;
        byn = -byn;
        bzn = -bzn;
      }
      out6[0] = spx;
      out6[1] = spy;
      out6[2] = spz;
      out6[3] = bxn;
      out6[4] = byn;
      out6[5] = bzn;
      return (int)vi;
    }
    prev = f;
    if (tmx < tmy) {
      if (tmx < tmz) {
        te = tmx;
        ix = ix + sx_;
        XY = iy;
        XZ = iz;
        tmx = tmx + tdx;
      } else {
        te = tmz;
        iz = iz + sz_;
        ZX = ix;
        ZY = iy;
        tmz = tmz + tdz;
      }
    } else {
      if (tmy < tmz) {
        te = tmy;
        iy = iy + sy_;
        YX = ix;
        YZ = iz;
        tmy = tmy + tdy;
      } else {
        te = tmz;
        iz = iz + sz_;
        ZX = ix;
        ZY = iy;
        tmz = tmz + tdz;
      }
    }
    vi = (int64_t)(ix - 1) * NyNz + (int64_t)(iy - 1) * Nz + (iz - 1);
  }
  return -1;
}

template <class S, class Vol>
void raycast_impl(const float* depth, int H, int W, const float* Kf, const float* Tf,
                  const Vol& F, const int* dims, const int* start, float voxel_mm, float trunc,
                  int fix_bounds, S* vx, S* vy, S* vz, S* nx, S* ny, S* nz, int32_t* hit,
                  int64_t* stats) {
  float Kif[9];
  kf_m3_inverse(Kf, Kif);
  S Ki[9], T[16];
  for (int i = 0; i < 9; ++i) Ki[i] = (S)Kif[i];
  for (int i = 0; i < 16; ++i) T[i] = (S)Tf[i];
  const int64_t n = (int64_t)H * W;
  int64_t steps = 0, hits = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : steps, hits)
  for (int64_t p = 0; p < n; ++p) {
    S o[6] = {0, 0, 0, 0, 0, 0};
    int h = -1;
    if (depth[p] != 0) {
      h = raycast_pixel<S>((int)(p % W), (int)(p / W), W, depth[p], Ki, T, F, dims, start,
                           (S)voxel_mm, (S)trunc, fix_bounds, o, &steps);
      if (h >= 0) ++hits;
      else
        for (int k = 0; k < 6; ++k) o[k] = 0;
    }
    vx[p] = o[0];
    vy[p] = o[1];
    vz[p] = o[2];
    nx[p] = o[3];
    ny[p] = o[4];
    nz[p] = o[5];
    if (hit) hit[p] = h;
  }
  if (stats) {
    stats[0] = steps;
    stats[1] = hits;
  }
}

// Trilinear raycast over an image; T in the scalar type S (FP64 for the
// finite-difference checks).  out: x, y, z, nx, ny, nz, depth planes; disc
// (nullable): 4 ints per pixel.  With gF non-null, also the adjoint for
// dL/d(depth) gD (and dL/d(normal) gN, 3 per pixel, nullable).
template <class S, class Vol>
void raycast_tri_impl(const float* depth, int H, int W, const float* Kf, const S* T,
                      const Vol& F, const int* dims, const int* start, float voxel_mm,
                      float trunc, S* const* out, int32_t* hit, int32_t* disc, const S* gD,
                      const S* gN, S* gF, S* gxi) {
  float Kif[9];
  kf_m3_inverse(Kf, Kif);
  S Ki[9];
  for (int i = 0; i < 9; ++i) Ki[i] = (S)Kif[i];
  const int64_t n = (int64_t)H * W;
  if (gxi)
    for (int q = 0; q < 6; ++q) gxi[q] = 0;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p = 0; p < n; ++p) {
    S o[7] = {0, 0, 0, 0, 0, 0, 0};
    int h = -1;
    int32_t dsc[4] = {-1, -1, -1, -1};
    if (depth[p] != 0) {
      TriCtx<S> tc{o, dsc, gD ? gD[p] : S(0), gN ? gN + 3 * p : nullptr, gF, gxi};
      S o6[6];
      int64_t steps = 0;
      h = raycast_pixel<S>((int)(p % W), (int)(p / W), W, depth[p], Ki, T, F, dims, start,
                           (S)voxel_mm, (S)trunc, 0, o6, &steps, &tc);
      if (h < 0)
        for (int k = 0; k < 7; ++k) o[k] = 0;
    }
    for (int k = 0; k < 7; ++k)
      if (out[k]) out[k][p] = o[k];
    if (hit) hit[p] = h;
    if (disc)
      for (int k = 0; k < 4; ++k) disc[4 * p + k] = h >= 0 ? dsc[k] : -1;
  }
}

struct F32Reader {
  const float* t;
  inline float operator()(int64_t i) const { return t[i]; }
};
struct F64Reader {
  const double* t;
  inline double operator()(int64_t i) const { return t[i]; }
};
struct F16Reader {
  const uint16_t* t;
  inline float operator()(int64_t i) const { return kf_f16_bits_to_f32(t[i]); }
};

}  // namespace

extern "C" {

// stats (nullable): [0] total traversal steps, [1] rays that hit.
void oracle_raycast_f32(const float* depth, int H, int W, const float* K, const float* T,
                        const float* tsdf, const int* dims, const int* start, float voxel_mm,
                        float trunc, int fix_bounds, float* vx, float* vy, float* vz, float* nx,
                        float* ny, float* nz, int32_t* hit, int64_t* stats) {
  raycast_impl<float>(depth, H, W, K, T, F32Reader{tsdf}, dims, start, voxel_mm, trunc,
                      fix_bounds, vx, vy, vz, nx, ny, nz, hit, stats);
}

void oracle_raycast_f16(const float* depth, int H, int W, const float* K, const float* T,
                        const uint16_t* tsdf_h, const int* dims, const int* start,
                        float voxel_mm, float trunc, int fix_bounds, float* vx, float* vy,
                        float* vz, float* nx, float* ny, float* nz, int32_t* hit,
                        int64_t* stats) {
  raycast_impl<float>(depth, H, W, K, T, F16Reader{tsdf_h}, dims, start, voxel_mm, trunc,
                      fix_bounds, vx, vy, vz, nx, ny, nz, hit, stats);
}

void oracle_raycast_f64(const float* depth, int H, int W, const float* K, const float* T,
                        const double* tsdf, const int* dims, const int* start, float voxel_mm,
                        float trunc, int fix_bounds, double* vx, double* vy, double* vz,
                        double* nx, double* ny, double* nz, int32_t* hit, int64_t* stats) {
  raycast_impl<double>(depth, H, W, K, T, F64Reader{tsdf}, dims, start, voxel_mm, trunc,
                       fix_bounds, vx, vy, vz, nx, ny, nz, hit, stats);
}

// Trilinear raycast, FP32 (parity with the kernels).  out7: 7 planes.
void oracle_raycast_tri_f32(const float* depth, int H, int W, const float* K, const float* T,
                            const float* tsdf, const int* dims, const int* start,
                            float voxel_mm, float trunc, float** out7, int32_t* hit,
                            int32_t* disc) {
  raycast_tri_impl<float>(depth, H, W, K, T, F32Reader{tsdf}, dims, start, voxel_mm, trunc,
                          out7, hit, disc, nullptr, nullptr, nullptr, nullptr);
}

void oracle_raycast_tri_f16(const float* depth, int H, int W, const float* K, const float* T,
                            const uint16_t* tsdf_h, const int* dims, const int* start,
                            float voxel_mm, float trunc, float** out7, int32_t* hit,
                            int32_t* disc) {
  raycast_tri_impl<float>(depth, H, W, K, T, F16Reader{tsdf_h}, dims, start, voxel_mm, trunc,
                          out7, hit, disc, nullptr, nullptr, nullptr, nullptr);
}

// FP64 forward with an FP64 pose, and its adjoint when gF is non-null
// (gF: dense dL/dF in the reference layout, zeroed by the caller).
void oracle_raycast_tri_f64(const float* depth, int H, int W, const float* K, const double* T,
                            const double* tsdf, const int* dims, const int* start,
                            float voxel_mm, float trunc, double** out7, int32_t* hit,
                            int32_t* disc, const double* gD, const double* gN, double* gF,
                            double* gxi) {
  raycast_tri_impl<double>(depth, H, W, K, T, F64Reader{tsdf}, dims, start, voxel_mm, trunc,
                           out7, hit, disc, gD, gN, gF, gxi);
}

// The literal gates, for tests of the cosine thresholds.
int oracle_icp_angle_gate(float c, float gate_deg) {
  const float ang = (float)std::acos((double)c) * kDegF;
  return (ang <= gate_deg && ang >= 0.0f) ? 1 : 0;
}
int oracle_normal_flip(float c) {
  const double angle = std::acos((double)c) * (double)kDegF;
  return angle > 90 ? 1 : 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// A5-A7  sensor_pose_estimation.cu:7-136 (per-pixel rows), :225-245 + :363-
// 387 (reductions), :255-259 + :339-429 (iteration loop, solve, update).
//
// reduce_mode 0: the reference's order - each of the 27 quantities summed by
//                a 64-wide shared-memory tree per block of 64 pixels, then a
//                sequential float sum over the block partials.
// reduce_mode 1: FP64 sequential sums (accuracy twin).
namespace {

struct IcpInputs {
  const float *vx, *vy, *vz, *nx, *ny, *nz;
  const uint8_t* mask;
  const float *sx, *sy, *sz, *snx, *sny, *snz;
  int H, W;
};

// One pixel's association; returns 1 and the 27 values (21 J^TJ upper
// triangle, 6 J^Tr) when the pixel passes every gate.
inline int icp_pixel(const IcpInputs& in, int64_t x, const float* K, const float* F,
                     const float* Te, float dist_gate, float angle_gate, float* row) {
  const int W = in.W, H = in.H;
  if (!(in.vz[x] != 0 && in.mask[x] != 0)) return 0;
  const float Vx = in.vx[x], Vy = in.vy[x], Vz = in.vz[x];
  const float cx = F[0] * Vx + F[1] * Vy + F[2] * Vz + F[3];
  const float cy = F[4] * Vx + F[5] * Vy + F[6] * Vz + F[7];
  const float cz = F[8] * Vx + F[9] * Vy + F[10] * Vz + F[11];
  const float px = K[0] * cx + K[1] * cy + K[2] * cz;
  const float py = K[3] * cx + K[4] * cy + K[5] * cz;
  const float pz = K[6] * cx + K[7] * cy + K[8] * cz;
  const int u = kf_round_to_int(px / pz);
  const int v = kf_round_to_int(py / pz);
  if (!(u > 0 && u <= W)) return 0;
  if (!(v > 0 && v <= H)) return 0;
  const float ux = Te[0] * Vx + Te[1] * Vy + Te[2] * Vz + Te[3];
  const float uy = Te[4] * Vx + Te[5] * Vy + Te[6] * Vz + Te[7];
  const float uz = Te[8] * Vx + Te[9] * Vy + Te[10] * Vz + Te[11];
  const int64_t j = (int64_t)(v - 1) * W + (u - 1);
  const float ddx = ux - in.sx[j], ddy = uy - in.sy[j], ddz = uz - in.sz[j];
  const float dist = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
  if (!(dist <= dist_gate && dist >= 0.0f)) return 0;
  float nnx = Te[0] * in.nx[x] + Te[1] * in.ny[x] + Te[2] * in.nz[x];
  float nny = Te[4] * in.nx[x] + Te[5] * in.ny[x] + Te[6] * in.nz[x];
  float nnz = Te[8] * in.nx[x] + Te[9] * in.ny[x] + Te[10] * in.nz[x];
  normalize3(nnx, nny, nnz);
  float ax = in.snx[j], ay = in.sny[j], az = in.snz[j];
  normalize3(ax, ay, az);
  const float c = (nnx * ax + nny * ay + nnz * az);
  const float ang = (float)std::acos((double)c) * kDegF;
  if (!(ang <= angle_gate && ang >= 0.0f)) return 0;
  const float Sx = in.snx[j], Sy = in.sny[j], Sz = in.snz[j];
  float a[6];
  a[0] = uz * Sy + (-uy * Sz);
  a[1] = -uz * Sx + (ux * Sz);
  a[2] = uy * Sx + (-ux * Sy);
  a[3] = Sx;
  a[4] = Sy;
  a[5] = Sz;
  const float b = Sx * (in.sx[j] - ux) + Sy * (in.sy[j] - uy) + Sz * (in.sz[j] - uz);
  int k = 0;
  for (int r = 0; r < 6; ++r)
    for (int s = r; s < 6; ++s) row[k++] = a[r] * a[s];
  for (int r = 0; r < 6; ++r) row[21 + r] = a[r] * b;
  return 1;
}

}  // namespace

extern "C" {

// T_inout: previous pose on entry (the pose the model maps were raycast at),
// refined pose on return.  T_start (nullable) replaces the starting estimate,
// which is otherwise T_inout itself as in the reference (:259); it lets a
// test replay one iteration from any estimate.
// Outputs (nullable): T_iters [iters][16], sums [iters][28] (as doubles; in
// reduce_mode 0 these are the float sums widened), status [iters],
// masks [iters][H*W].
void oracle_icp(const float* vx, const float* vy, const float* vz, const float* nx,
                const float* ny, const float* nz, const uint8_t* mask, const float* sx,
                const float* sy, const float* sz, const float* snx, const float* sny,
                const float* snz, int H, int W, const float* K, float* T_inout,
                const float* T_start, int iters, float dist_gate, float angle_gate,
                int reduce_mode, float* T_iters, double* sums_out, int32_t* status_out,
                uint8_t* masks, double min_inliers, float stop_rot, float stop_trans,
                int32_t* iters_done) {
  const IcpInputs in{vx, vy, vz, nx, ny, nz, mask, sx, sy, sz, snx, sny, snz, H, W};
  const int64_t n = (int64_t)H * W;
  const int64_t nblk = (n + 63) / 64;
  float Tprev_inv[16], Te[16];
  kf_m4_inverse(T_inout, Tprev_inv);
  for (int i = 0; i < 16; ++i) Te[i] = T_start ? T_start[i] : T_inout[i];
  std::vector<float> part((size_t)nblk * 28);
  int done = 0;
  for (int it = 0; it < iters; ++it) {
    float F[16];
    kf_mat4_mul(Tprev_inv, Te, F);
    double sums[28];
    if (reduce_mode == 0) {
#pragma omp parallel for schedule(static)
      for (int64_t bk = 0; bk < nblk; ++bk) {
        float sd[64][28];
        for (int t = 0; t < 64; ++t) {
          const int64_t x = bk * 64 + t;
          float row[27];
          int ok = 0;
          if (x < n) ok = icp_pixel(in, x, K, F, Te, dist_gate, angle_gate, row);
          for (int k = 0; k < 27; ++k) sd[t][k] = ok ? row[k] : 0.0f;
          sd[t][27] = ok ? 1.0f : 0.0f;
          if (masks && x < n) masks[(int64_t)it * n + x] = (uint8_t)ok;
        }
        for (int s = 32; s > 0; s >>= 1)
          for (int t = 0; t < s; ++t)
            for (int k = 0; k < 28; ++k) sd[t][k] += sd[t + s][k];
        for (int k = 0; k < 28; ++k) part[(size_t)bk * 28 + k] = sd[0][k];
      }
      for (int k = 0; k < 28; ++k) {
        float acc = 0.0f;
        for (int64_t bk = 0; bk < nblk; ++bk) acc += part[(size_t)bk * 28 + k];
        sums[k] = (double)acc;
      }
    } else {
      for (int k = 0; k < 28; ++k) sums[k] = 0.0;
      for (int64_t x = 0; x < n; ++x) {
        float row[27];
        const int ok = icp_pixel(in, x, K, F, Te, dist_gate, angle_gate, row);
        if (masks) masks[(int64_t)it * n + x] = (uint8_t)ok;
        if (!ok) continue;
        for (int k = 0; k < 27; ++k) sums[k] += (double)row[k];
        sums[27] += 1.0;
      }
    }
    // min_inliers < 0: the reference's update; else the robust one
    // (kf_icp_update_gated, written out to keep the increment)
    float inc[16];
    const int st = kf_icp_increment_gated(sums, inc, min_inliers);
    kf_mat4_mul(inc, Te, Te);
    if (T_iters)
      for (int i = 0; i < 16; ++i) T_iters[it * 16 + i] = Te[i];
    if (sums_out)
      for (int k = 0; k < 28; ++k) sums_out[it * 28 + k] = sums[k];
    if (status_out) status_out[it] = st;
    done = it + 1;
    // opt-in early stop (no reference counterpart; 0 thresholds never stop)
    if (kf_icp_converged(inc, stop_rot, stop_trans)) break;
  }
  for (int i = 0; i < 16; ++i) T_inout[i] = Te[i];
  if (iters_done) *iters_done = done;
}

// ---------------------------------------------------------------------------
// KinfuPipeline frame sequence (kinfu_pipeline.cpp:91-193) over n uint16
// frames, FP32 volume: frame 0 = InitialProcessing (integrate first, raycast),
// frames 1.. = IncomingFrameProcessing (vertex/normal, ICP, integrate,
// raycast).  The bilateral filter (r > 0) runs on every frame right after the
// range gate, replacing the depth as the reference's in-place wrapper does.
// poses_out [n][16]; tsdf/weight (N voxels each) are caller-owned and are
// reset to NaN here.  stage_ms (nullable, 6 doubles) accumulates wall time
// per stage in the KF_STAGE_* order.
void oracle_pipeline(const uint16_t* frames, int n_frames, int H, int W, const float* K,
                     const int* dims, const int* start, float voxel_mm, float trunc, float smin,
                     float smax, float dmin, float dmax, int bilateral_r, float sigma_s,
                     float sigma_r, int icp_iters, float dist_gate, float angle_gate,
                     int reduce_mode, int fix_bounds, float* tsdf, float* weight,
                     float* poses_out, double* stage_ms, float stop_rot, float stop_trans) {
  const int64_t n = (int64_t)H * W;
  const int64_t nv = (int64_t)dims[0] * dims[1] * dims[2];
  std::fill_n(tsdf, nv, NAN);
  std::fill_n(weight, nv, NAN);
  std::vector<float> depth(n), tmp(n), vx(n), vy(n), nx(n), ny(n), nz(n);
  std::vector<uint8_t> mask(n);
  std::vector<float> sx(n), sy(n), sz(n), snx(n), sny(n), snz(n);
  std::vector<float> ws;
  if (bilateral_r > 0) {
    ws.resize((size_t)(2 * bilateral_r + 1) * (2 * bilateral_r + 1));
    oracle_spatial_kernel(bilateral_r, sigma_s, ws.data());
  }
  float Ki[9];
  kf_m3_inverse(K, Ki);
  float T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  double acc[6] = {0, 0, 0, 0, 0, 0};
  auto now = [] { return omp_get_wtime() * 1e3; };
  for (int f = 0; f < n_frames; ++f) {
    double t0 = now();
    oracle_range_gate(frames + (int64_t)f * n, depth.data(), n, dmin, dmax);
    double t1 = now();
    acc[0] += t1 - t0;
    if (bilateral_r > 0) {
      oracle_bilateral(depth.data(), tmp.data(), H, W, ws.data(), bilateral_r, sigma_r);
      depth.swap(tmp);
    }
    double t2 = now();
    acc[1] += t2 - t1;
    if (f > 0) {
      oracle_vertex_normals(depth.data(), Ki, H, W, vx.data(), vy.data(), nx.data(), ny.data(),
                            nz.data(), mask.data());
      double t3 = now();
      acc[2] += t3 - t2;
      oracle_icp(vx.data(), vy.data(), depth.data(), nx.data(), ny.data(), nz.data(),
                 mask.data(), sx.data(), sy.data(), sz.data(), snx.data(), sny.data(),
                 snz.data(), H, W, K, T, nullptr, icp_iters, dist_gate, angle_gate, reduce_mode,
                 nullptr, nullptr, nullptr, nullptr, -1.0, stop_rot, stop_trans, nullptr);
      acc[3] += now() - t3;
    }
    double t4 = now();
    oracle_integrate_f32(depth.data(), H, W, K, T, dims, start, voxel_mm, trunc, smin, smax,
                         f == 0 ? 1 : 0, tsdf, weight);
    double t5 = now();
    acc[4] += t5 - t4;
    oracle_raycast_f32(depth.data(), H, W, K, T, tsdf, dims, start, voxel_mm, trunc, fix_bounds,
                       sx.data(), sy.data(), sz.data(), snx.data(), sny.data(), snz.data(),
                       nullptr, nullptr);
    acc[5] += now() - t5;
    for (int i = 0; i < 16; ++i) poses_out[(int64_t)f * 16 + i] = T[i];
  }
  if (stage_ms)
    for (int i = 0; i < 6; ++i) stage_ms[i] += acc[i];
}

int oracle_max_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"

extern "C" {
// IEEE half conversions used by the FP16 volume (exposed for tests).
void oracle_f32_to_f16(const float* in, uint16_t* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = kf_f32_to_f16_bits(in[i]);
}
void oracle_f16_to_f32(const uint16_t* in, float* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = kf_f16_bits_to_f32(in[i]);
}
}  // extern "C"

extern "C" {
// Per-pixel traversal step counts of the F32 raycast (for the cost model of
// the raycast and its tests).
void oracle_raycast_steps_f32(const float* depth, int H, int W, const float* Kf, const float* Tf,
                              const float* tsdf, const int* dims, const int* start,
                              float voxel_mm, float trunc, int fix_bounds, int32_t* steps) {
  float Kif[9];
  kf_m3_inverse(Kf, Kif);
  const int64_t n = (int64_t)H * W;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p = 0; p < n; ++p) {
    int64_t s = 0;
    float o[6];
    if (depth[p] != 0)
      raycast_pixel<float>((int)(p % W), (int)(p / W), W, depth[p], Kif, Tf, F32Reader{tsdf}, dims,
                           start, voxel_mm, trunc, fix_bounds, o, &s);
    steps[p] = (int32_t)s;
  }
}
}  // extern "C"
