// kf_hd_math.h - arithmetic shared by the gfx950 kernels, the host runtime
// and the CPU oracle.
//
// Everything here is built only from IEEE-754 add, sub, mul, div, sqrt,
// rounding-to-integer and bit manipulation, with no libm transcendental
// calls.  Compiled with -ffp-contract=off on both sides (g++ for the host,
// hipcc for gfx950), each function therefore returns bit-identical results
// on the CPU and on the GPU.  That is what lets the kernels be compared
// bit-exactly with the oracle.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef __HIPCC__
#define KF_HD __host__ __device__ inline
// full unrolling keeps the small fixed-size arrays of the solvers in registers
#define KF_UNROLL _Pragma("unroll")
#else
#define KF_HD inline
#define KF_UNROLL
#endif

// ---------------------------------------------------------------------------
// Float -> int conversions with defined results.
//
// The reference writes (int)roundf(p/pz) (projective_tsdf.cu:28-29,
// sensor_pose_estimation.cu:30-31) and (int)floor(..) (voxel_traversal.cu:133-
// 135).  C leaves the conversion undefined for NaN and values outside int32;
// x86 and AMD hardware also disagree there.  Every caller follows the
// conversion by a 1-based bounds test, so mapping those inputs to INT32_MIN
// gives the same accept/reject decision as any in-range value would, on both
// CPU and GPU.
KF_HD int kf_round_to_int(float v) {
  if (!(v > -2147483648.0f && v < 2147483648.0f)) return (int)0x80000000u;
  return (int)roundf(v);
}

KF_HD int kf_floor_to_int(float v) {
  if (!(v > -2147483648.0f && v < 2147483648.0f)) return (int)0x80000000u;
  return (int)floorf(v);
}

// ---------------------------------------------------------------------------
// exp for float arguments (bilateral range weight bilateral_filtering.cu:27 and
// the spatial kernel system_utility.cpp:112).
//
// Method: Cody & Waite (1980) additive range reduction x = k ln2 + r with ln2
// split into a 16-bit head and a float tail (so k*ln2_hi is exact for
// |k| <= 255), |r| <= ln2/2, then the degree-7 Taylor polynomial of e^r in
// Horner form (truncation error r^8/8! < 6e-9 relative) and an exact scaling
// by 2^k assembled from exponent bits.  Error is within 2 ulp of the true
// value over the whole float range; it is identical on host and device.
KF_HD float kf_pow2i(int k) {  // 2^k for -126 <= k <= 127
  uint32_t bits = (uint32_t)(k + 127) << 23;
  float f;
  memcpy(&f, &bits, 4);
  return f;
}

KF_HD float kf_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return INFINITY;
  if (x < -103.972084045410156f) return 0.0f;
  const float kf = rintf(x * 1.44269502162933349609375f);  // x / ln2
  const int k = (int)kf;
  float r = x - kf * 0.693145751953125f;             // ln2 head (16 bits)
  r = r - kf * 1.428606765330187045e-06f;            // ln2 tail
  float p = 1.98412698412698e-04f;                   // 1/7!
  p = p * r + 1.38888888888889e-03f;                 // 1/6!
  p = p * r + 8.33333333333333e-03f;                 // 1/5!
  p = p * r + 4.16666666666667e-02f;                 // 1/4!
  p = p * r + 1.66666666666667e-01f;                 // 1/3!
  p = p * r + 0.5f;                                  // 1/2!
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  if (k > 127) return (p * kf_pow2i(k - 1)) * 2.0f;
  if (k < -126) return (p * kf_pow2i(k + 64)) * kf_pow2i(-64);
  return p * kf_pow2i(k);
}

// ---------------------------------------------------------------------------
// Small matrix helpers.  Inputs and outputs are float (the reference keeps K,
// T and every pose as CV_32F); the arithmetic is FP64 and rounded once on
// store.  The reference uses OpenCV's cv::Mat::inv and operator* here
// (projective_tsdf.cu:148-149, sensor_pose_estimation.cu:255-256, 343, 428),
// whose last-bit behaviour is not pinned by anything in the reference; these
// FP64 versions are the convention on both sides of the comparison.

// The factor the reference's normalize (cuda_functions.cuh:2-10) scales a
// 3-vector by: 1 / |v| with the float sum x^2 + y^2 + z^2 left to right and
// a correctly rounded square root and division (+inf for the zero vector).
KF_HD float kf_inv_norm3(float x, float y, float z) { return 1 / sqrtf(x * x + y * y + z * z); }

// Element (i, j) of A * B for 4x4 row-major float matrices, FP64 products
// summed left to right and rounded once.
KF_HD float kf_mat4_mul_elem(const float* A, const float* B, int i, int j) {
  double s = (double)A[i * 4 + 0] * (double)B[0 * 4 + j];
  s = s + (double)A[i * 4 + 1] * (double)B[1 * 4 + j];
  s = s + (double)A[i * 4 + 2] * (double)B[2 * 4 + j];
  s = s + (double)A[i * 4 + 3] * (double)B[3 * 4 + j];
  return (float)s;
}

// C = A * B, 4x4 row-major; C may alias A or B.
KF_HD void kf_mat4_mul(const float* A, const float* B, float* C) {
  float out[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[i * 4 + j] = kf_mat4_mul_elem(A, B, i, j);
  for (int i = 0; i < 16; ++i) C[i] = out[i];
}

// General 4x4 inverse by Laplace expansion over 2x2 minors of the top and
// bottom row pairs.  Returns 1 (and leaves out untouched) when the
// determinant is zero or not finite.  out may alias in.
KF_HD int kf_m4_inverse(const float* in, float* out) {
  double a[16];
  for (int i = 0; i < 16; ++i) a[i] = (double)in[i];
  // minors of rows 0,1
  const double s0 = a[0] * a[5] - a[4] * a[1];
  const double s1 = a[0] * a[6] - a[4] * a[2];
  const double s2 = a[0] * a[7] - a[4] * a[3];
  const double s3 = a[1] * a[6] - a[5] * a[2];
  const double s4 = a[1] * a[7] - a[5] * a[3];
  const double s5 = a[2] * a[7] - a[6] * a[3];
  // minors of rows 2,3
  const double c5 = a[10] * a[15] - a[14] * a[11];
  const double c4 = a[9] * a[15] - a[13] * a[11];
  const double c3 = a[9] * a[14] - a[13] * a[10];
  const double c2 = a[8] * a[15] - a[12] * a[11];
  const double c1 = a[8] * a[14] - a[12] * a[10];
  const double c0 = a[8] * a[13] - a[12] * a[9];
  const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  if (!(det != 0.0) || !(det - det == 0.0)) return 1;
  const double id = 1.0 / det;
  double r[16];
  r[0] = (a[5] * c5 - a[6] * c4 + a[7] * c3) * id;
  r[1] = (-a[1] * c5 + a[2] * c4 - a[3] * c3) * id;
  r[2] = (a[13] * s5 - a[14] * s4 + a[15] * s3) * id;
  r[3] = (-a[9] * s5 + a[10] * s4 - a[11] * s3) * id;
  r[4] = (-a[4] * c5 + a[6] * c2 - a[7] * c1) * id;
  r[5] = (a[0] * c5 - a[2] * c2 + a[3] * c1) * id;
  r[6] = (-a[12] * s5 + a[14] * s2 - a[15] * s1) * id;
  r[7] = (a[8] * s5 - a[10] * s2 + a[11] * s1) * id;
  r[8] = (a[4] * c4 - a[5] * c2 + a[7] * c0) * id;
  r[9] = (-a[0] * c4 + a[1] * c2 - a[3] * c0) * id;
  r[10] = (a[12] * s4 - a[13] * s2 + a[15] * s0) * id;
  r[11] = (-a[8] * s4 + a[9] * s2 - a[11] * s0) * id;
  r[12] = (-a[4] * c3 + a[5] * c1 - a[6] * c0) * id;
  r[13] = (a[0] * c3 - a[1] * c1 + a[2] * c0) * id;
  r[14] = (-a[12] * s3 + a[13] * s1 - a[14] * s0) * id;
  r[15] = (a[8] * s3 - a[9] * s1 + a[10] * s0) * id;
  for (int i = 0; i < 16; ++i) out[i] = (float)r[i];
  return 0;
}

// 3x3 inverse by the adjugate.  Returns 1 when singular.
KF_HD int kf_m3_inverse(const float* in, float* out) {
  double a[9];
  for (int i = 0; i < 9; ++i) a[i] = (double)in[i];
  const double c00 = a[4] * a[8] - a[5] * a[7];
  const double c01 = a[5] * a[6] - a[3] * a[8];
  const double c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  if (!(det != 0.0) || !(det - det == 0.0)) return 1;
  const double id = 1.0 / det;
  double r[9];
  r[0] = c00 * id;
  r[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  r[3] = c01 * id;
  r[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  r[6] = c02 * id;
  r[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  for (int i = 0; i < 9; ++i) out[i] = (float)r[i];
  return 0;
}

// ---------------------------------------------------------------------------
// 6x6 normal equations (sensor_pose_estimation.cu:389-400).
//
// The reference solves with cv::Cholesky on float data and ignores its
// return value; when a pivot is not positive the right-hand side is left in
// place and used as the increment.  Here: LDL^T without pivoting in FP64;
// a pivot below FLT_EPSILON (or NaN) reports failure and the caller then uses
// b itself, which reproduces that behaviour.  Each update of a running sum
// is one fused multiply-add with a stored product LD[j][k] = L[j][k] d_k
// (the column's sum before its division by the pivot d_k), so that the dependent
// chain of the solve, which every ICP iteration waits for, is one operation
// per term.  fma is correctly rounded on the host and on the device, so
// both compute the same bits.
#define KF_LDLT_EPS 1.1920928955078125e-07

KF_HD int kf_ldlt6_solve(const double* A, const double* b, double* x) {
  double L[6][6], LD[6][6];
  double rd[6];  // reciprocals of the pivots (6 divisions in total)
  KF_UNROLL for (int i = 0; i < 6; ++i) {
    KF_UNROLL for (int j = 0; j < i; ++j) {
      double s = A[i * 6 + j];
      KF_UNROLL for (int k = 0; k < j; ++k) s = fma(-L[i][k], LD[j][k], s);
      LD[i][j] = s;
      L[i][j] = s * rd[j];
    }
    double s = A[i * 6 + i];
    KF_UNROLL for (int k = 0; k < i; ++k) s = fma(-L[i][k], LD[i][k], s);
    if (!(s >= KF_LDLT_EPS)) return 1;
    rd[i] = 1.0 / s;
  }
  double y[6];
  KF_UNROLL for (int i = 0; i < 6; ++i) {
    double s = b[i];
    KF_UNROLL for (int k = 0; k < i; ++k) s = fma(-L[i][k], y[k], s);
    y[i] = s;
  }
  KF_UNROLL for (int i = 0; i < 6; ++i) y[i] = y[i] * rd[i];
  KF_UNROLL for (int i = 5; i >= 0; --i) {
    double s = y[i];
    KF_UNROLL for (int k = i + 1; k < 6; ++k) s = fma(-L[k][i], x[k], s);
    x[i] = s;
  }
  return 0;
}

// Order of the 28 ICP sums: upper triangle of J^TJ (row-major, i <= j), then
// J^Tr, then the inlier count.
KF_HD int kf_sym_index(int i, int j) {  // i <= j
  return i * 6 - (i * (i - 1)) / 2 + (j - i);
}

// The ICP increment from FP64 sums (sensor_pose_estimation.cu:399-425):
// solves the symmetric system and returns
// x = (beta, gamma, alpha, tx, ty, tz) rounded to float as the small-angle
//   inc = [[1, x2, -x1, x3], [-x2, 1, x0, x4], [x1, -x0, 1, x5], [0, 0, 0, 1]].
// Returns the solver status (0 solved, 1 fallback to the right-hand side).
KF_HD int kf_icp_increment(const double* sums, float* inc) {
  double A[36], b[6], xd[6];
  KF_UNROLL for (int i = 0; i < 6; ++i)
    KF_UNROLL for (int j = i; j < 6; ++j) {
      const double v = sums[kf_sym_index(i, j)];
      A[i * 6 + j] = v;
      A[j * 6 + i] = v;
    }
  KF_UNROLL for (int i = 0; i < 6; ++i) b[i] = sums[21 + i];
  const int status = kf_ldlt6_solve(A, b, xd);
  float x[6];
  KF_UNROLL for (int i = 0; i < 6; ++i) x[i] = (float)(status ? b[i] : xd[i]);
  const float m[16] = {1.0f, x[2], -x[1], x[3],  //
                       -x[2], 1.0f, x[0], x[4],  //
                       x[1], -x[0], 1.0f, x[5],  //
                       0.0f, 0.0f, 0.0f, 1.0f};
  KF_UNROLL for (int i = 0; i < 16; ++i) inc[i] = m[i];
  return status;
}

// The tracking-robust increment (no reference counterpart): an iteration
// whose solve fails (status 1), or that solved with fewer than min_inliers
// inliers (sums[27], status 2), gets the identity increment instead.
KF_HD int kf_icp_increment_robust(const double* sums, float* inc, double min_inliers) {
  int status = kf_icp_increment(sums, inc);
  if (status == 0 && sums[27] < min_inliers) status = 2;
  if (status != 0)
    KF_UNROLL for (int i = 0; i < 16; ++i) inc[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  return status;
}

// The increment of either update: the robust one when min_inliers >= 0,
// else the reference's.
KF_HD int kf_icp_increment_gated(const double* sums, float* inc, double min_inliers) {
  return min_inliers >= 0.0 ? kf_icp_increment_robust(sums, inc, min_inliers)
                            : kf_icp_increment(sums, inc);
}

// Opt-in early stop (no reference counterpart; the reference always runs its
// 10 iterations, sensor_pose_estimation.cu:339): an increment whose rotation
// parameters (x0, x1, x2, radians) are all below stop_rot in magnitude and
// whose translation (x3, x4, x5, mm) is below stop_trans is the last one
// applied.  inc is kf_icp_increment's matrix, so x0 = inc[6], x1 = inc[8],
// x2 = inc[1], x3 = inc[3], x4 = inc[7], x5 = inc[11].  A threshold of 0
// never stops, and neither does a NaN parameter.
KF_HD bool kf_icp_converged(const float* inc, float stop_rot, float stop_trans) {
  return fabsf(inc[6]) < stop_rot && fabsf(inc[8]) < stop_rot && fabsf(inc[1]) < stop_rot &&
         fabsf(inc[3]) < stop_trans && fabsf(inc[7]) < stop_trans && fabsf(inc[11]) < stop_trans;
}

// One ICP update: T <- inc * T (sensor_pose_estimation.cu:428).  Returns the
// solver status.
KF_HD int kf_icp_update(const double* sums, float* T) {
  float inc[16];
  const int status = kf_icp_increment(sums, inc);
  kf_mat4_mul(inc, T, T);
  return status;
}

// The same with the robust increment when min_inliers >= 0.
KF_HD int kf_icp_update_gated(const double* sums, float* T, double min_inliers) {
  if (!(min_inliers >= 0.0)) return kf_icp_update(sums, T);
  float inc[16];
  const int status = kf_icp_increment_robust(sums, inc, min_inliers);
  kf_mat4_mul(inc, T, T);
  return status;
}

// ---------------------------------------------------------------------------
// IEEE half <-> float for the FP16 volume (round to nearest even on store).
KF_HD uint16_t kf_f32_to_f16_bits(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7fffffffu;
  if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);           // NaN
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);          // overflow
  if (ax >= 0x38800000u) {                                           // normal
    uint32_t m = ax + 0xc8000fffu + ((ax >> 13) & 1u);  // rebias, RNE
    return (uint16_t)(sign | (m >> 13));
  }
  if (ax < 0x33000001u) return (uint16_t)sign;                       // to zero
  // subnormal: shift the 24-bit significand, round to nearest even
  const uint32_t e = ax >> 23;
  const uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
  const uint32_t shift = 126u - e;  // 14..24
  uint32_t h = mant >> shift;
  const uint32_t rem = mant & ((1u << shift) - 1u);
  const uint32_t half = 1u << (shift - 1u);
  if (rem > half || (rem == half && (h & 1u))) h += 1u;
  return (uint16_t)(sign | h);
}

KF_HD float kf_f16_bits_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu;
  const uint32_t m = h & 0x3ffu;
  uint32_t bits;
  if (e == 0) {
    if (m == 0) {
      bits = sign;
    } else {  // subnormal half -> normal float
      int sh = 0;
      uint32_t mm = m;
      while (!(mm & 0x400u)) {
        mm <<= 1;
        ++sh;
      }
      bits = sign | ((uint32_t)(113 - sh) << 23) | ((mm & 0x3ffu) << 13);
    }
  } else if (e == 31) {
    bits = sign | 0x7f800000u | (m << 13);
  } else {
    bits = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &bits, 4);
  return f;
}
