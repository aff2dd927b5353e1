// kf_kernels.hip - gfx950 kernels of the KinectFusion frame sequence.
//
// Every kernel evaluates the reference's float expressions in the reference's
// order (built with -ffp-contract=off, correctly rounded divide and sqrt, no
// denormal flush), so given identical inputs its outputs are bit-identical to
// the CPU oracle.  What changes is where the work and the data live:
//   * the volume, the images and the ICP state stay resident in HBM;
//   * voxel centres are computed from the voxel index instead of being read
//     from three 155 MB arrays (kinfu_pipeline.cpp:38-58);
//   * per-pixel terms of the integration are computed once per pixel
//     (k_integ_table) instead of once per voxel, which changes no value;
//   * ICP reduces its 27 sums in registers (wave64 DPP) and LDS, and the
//     last block to finish solves the 6x6 system in FP64 and updates the pose
//     on the device: no 51.6 MB zero uploads and no 420 blocking copies per
//     frame (sensor_pose_estimation.cu:339-387).
#include <hip/hip_fp16.h>

#include <type_traits>

#include "kf_internal.h"

namespace {

// ---------------------------------------------------------------------------
// helpers

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Pointer qualifier for the global address space: loads through it are
// global_load (vmcnt only) rather than flat_load (vmcnt and lgkmcnt).  The
// host compilation pass of the kernels has no such address space.
#if defined(__HIP_DEVICE_COMPILE__)
#define KF_GLOBAL __attribute__((address_space(1)))
#else
#define KF_GLOBAL
#endif

// Uniform load of a small float array that the previous kernel wrote.
template <int N>
__device__ __forceinline__ void load_uniform(const float* __restrict__ src, float* dst) {
#pragma unroll
  for (int i = 0; i < N; ++i) dst[i] = src[i];
}


// cuda_functions.cuh:2-10
__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
  const float t = kf_inv_norm3(x, y, z);
  x = x * t;
  y = y * t;
  z = z * t;
}

// Compensated (TwoSum) combination of (s, c) pairs: s + c carries the sum
// with the rounding error of every addition folded into c.
__device__ __forceinline__ void two_sum_add(float& s, float& c, float s2, float c2) {
  const float t = s + s2;
  const float bp = t - s;
  const float e = (s - (t - bp)) + (s2 - bp);
  s = t;
  c = c + c2 + e;
}

// Cross-lane moves for float and double (a double moves as two dwords).
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, false);
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_mov(float v) {
  return __uint_as_float(dpp_u32<CTRL, ROW_MASK>(__float_as_uint(v)));
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_mov(double v) {
  const uint64_t u = __double_as_longlong(v);
  const uint32_t lo = dpp_u32<CTRL, ROW_MASK>((uint32_t)u);
  const uint32_t hi = dpp_u32<CTRL, ROW_MASK>((uint32_t)(u >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// lane ^ 4 within each 32 lanes
__device__ __forceinline__ float swz_xor4(float v) {
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x101F));
}
__device__ __forceinline__ double swz_xor4(double v) {
  const uint64_t u = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)u, 0x101F);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)(u >> 32), 0x101F);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// v_permlane32_swap: the upper 32 lanes of a trade places with the lower 32
// lanes of b.  v_permlane16_swap: rows 1 and 3 of a with rows 0 and 2 of b.
template <bool W32>
__device__ __forceinline__ void lane_swap_u32(uint32_t& a, uint32_t& b) {
  if constexpr (W32) {
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    a = r[0];
    b = r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
    a = r[0];
    b = r[1];
  }
}
template <bool W32>
__device__ __forceinline__ void lane_swap(float& a, float& b) {
  uint32_t x = __float_as_uint(a), y = __float_as_uint(b);
  lane_swap_u32<W32>(x, y);
  a = __uint_as_float(x);
  b = __uint_as_float(y);
}
template <bool W32>
__device__ __forceinline__ void lane_swap(double& a, double& b) {
  const uint64_t ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  uint32_t alo = (uint32_t)ua, ahi = (uint32_t)(ua >> 32);
  uint32_t blo = (uint32_t)ub, bhi = (uint32_t)(ub >> 32);
  lane_swap_u32<W32>(alo, blo);
  lane_swap_u32<W32>(ahi, bhi);
  a = __longlong_as_double((long long)(((uint64_t)ahi << 32) | alo));
  b = __longlong_as_double((long long)(((uint64_t)bhi << 32) | blo));
}

// Sum over the 64 lanes of a wave (DPP: quad_perm [1,0,3,2] and [2,3,0,1],
// row_half_mirror, row_mirror give every lane its row's sum; row_bcast:15 and
// row_bcast:31 fold rows 0..3 into row 3).  Lane 63 holds the total; the
// order of additions is fixed.
template <typename T>
__device__ __forceinline__ T wave_total(T v) {
  v = v + dpp_mov<0xB1>(v);
  v = v + dpp_mov<0x4E>(v);
  v = v + dpp_mov<0x141>(v);
  v = v + dpp_mov<0x140>(v);
  v = v + dpp_mov<0x142, 0xA>(v);
  v = v + dpp_mov<0x143, 0xC>(v);
  return v;
}

// Transposing wave reduction of 32 values per lane.  Each of the steps
// o = 32, 16, 8, 4, 2 halves the values a lane keeps (lanes with bit o set
// keep the upper half and take the partner's upper half, lane ^ o, the others
// the lower), so the wave moves 16 + 8 + 4 + 2 + 1 + 1 values instead of
// reducing 32 values separately.  Steps 32 and 16 are v_permlane32_swap /
// v_permlane16_swap of the two halves (after the swap each lane adds its own
// and its partner's element), steps 8 and 2 are DPP row_ror:8 and quad_perm,
// step 4 is ds_swizzle.  On return s[0] (with its compensation c[0] when
// KAHAN) in lane l holds the 64-lane total of value
//   ((l >> 5) & 1) * 16 + ((l >> 4) & 1) * 8 + ((l >> 3) & 1) * 4
//   + ((l >> 2) & 1) * 2 + ((l >> 1) & 1),
// the same in lanes l and l ^ 1.  The additions form a fixed tree.
__device__ __forceinline__ int wave_transpose_index(int l) {
  return ((l >> 5) & 1) * 16 + ((l >> 4) & 1) * 8 + ((l >> 3) & 1) * 4 + ((l >> 2) & 1) * 2 +
         ((l >> 1) & 1);
}

template <bool KAHAN, typename T>
__device__ __forceinline__ void transpose_add(T& s, T* c, T gs, T gc) {
  if constexpr (KAHAN)
    two_sum_add(s, *c, gs, gc);
  else
    s = s + gs;
}

template <bool KAHAN, typename T>
__device__ __forceinline__ void wave_transpose_sum(T (&s)[32], T* c) {
  const int ln = lane_id();
  T cdummy = T(0);
  // steps 32 and 16: swap the halves across lanes, then add
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    T a = s[i], b = s[i + 16];
    lane_swap<true>(a, b);
    T ca = T(0), cb = T(0);
    if constexpr (KAHAN) {
      ca = c[i];
      cb = c[i + 16];
      lane_swap<true>(ca, cb);
    }
    transpose_add<KAHAN>(a, KAHAN ? &ca : &cdummy, b, cb);
    s[i] = a;
    if constexpr (KAHAN) c[i] = ca;
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    T a = s[i], b = s[i + 8];
    lane_swap<false>(a, b);
    T ca = T(0), cb = T(0);
    if constexpr (KAHAN) {
      ca = c[i];
      cb = c[i + 8];
      lane_swap<false>(ca, cb);
    }
    transpose_add<KAHAN>(a, KAHAN ? &ca : &cdummy, b, cb);
    s[i] = a;
    if constexpr (KAHAN) c[i] = ca;
  }
  // steps 8, 4, 2: each lane keeps one half and takes the partner's copy of it
  auto pull = [&](auto move, int o, int h) {
    const bool up = (ln & o) != 0;
#pragma unroll
    for (int i = 0; i < h; ++i) {
      // the empty asm keeps the two halves as values, so that the selects do
      // not become a load from a selected address (which would put the array
      // in scratch memory)
      T lo = s[i], hi = s[i + h];
      asm volatile("" : "+v"(lo), "+v"(hi));
      T ks = up ? hi : lo;
      const T gs = move(up ? lo : hi);
      T kc = T(0), gc = T(0);
      if constexpr (KAHAN) {
        T clo = c[i], chi = c[i + h];
        asm volatile("" : "+v"(clo), "+v"(chi));
        kc = up ? chi : clo;
        gc = move(up ? clo : chi);
      }
      transpose_add<KAHAN>(ks, KAHAN ? &kc : &cdummy, gs, gc);
      s[i] = ks;
      if constexpr (KAHAN) c[i] = kc;
    }
  };
  pull([](T v) { return dpp_mov<0x128>(v); }, 8, 4);  // row_ror:8 = lane ^ 8
  pull([](T v) { return swz_xor4(v); }, 4, 2);
  pull([](T v) { return dpp_mov<0x4E>(v); }, 2, 1);    // quad_perm [2,3,0,1]
  // step 1: both lanes of a pair add in lane order, so they agree
  const T gs = dpp_mov<0xB1>(s[0]);                     // quad_perm [1,0,3,2]
  if constexpr (KAHAN) {
    const T gc = dpp_mov<0xB1>(c[0]);
    if (ln & 1) {
      T t = gs, tc = gc;
      two_sum_add(t, tc, s[0], c[0]);
      s[0] = t;
      c[0] = tc;
    } else {
      two_sum_add(s[0], c[0], gs, gc);
    }
  } else {
    s[0] = (ln & 1) ? gs + s[0] : s[0] + gs;
  }
}

// ---------------------------------------------------------------------------
// TSDF readers over the brick-tiled arrays (index = kf_store_index)

template <int PREC>
struct VolRead;
template <>
struct VolRead<KF_TSDF_F32> {
  const float* __restrict__ v;
  __device__ __forceinline__ float operator()(int i) const { return v[i]; }
};
template <>
struct VolRead<KF_TSDF_F16> {
  const uint16_t* __restrict__ v;
  __device__ __forceinline__ float operator()(int i) const {
    return __half2float(__ushort_as_half(v[i]));
  }
};

// ---------------------------------------------------------------------------
// A1  uint16 mm -> float mm with the inclusive range gate
// (system_utility.cpp:58-68, 77-86).
__global__ __launch_bounds__(256) void k_depth_u16(const uint16_t* __restrict__ src,
                                                   float* __restrict__ dst, int n, float dmin,
                                                   float dmax) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float d = (float)src[i];
  dst[i] = (d >= dmin && d <= dmax) ? d : 0.0f;
}

// ---------------------------------------------------------------------------
// A4  bilateral filter (bilateral_filtering.cu:3-47), LDS-tiled.
//
// A 16x16 block stages a (16+2R)^2 window of the depth in LDS.  The window is
// read by linear index y*W + x exactly as the reference reads it, so the
// column past the right edge wraps to the next row as it does there; indices
// outside the image read 0.  Pixels outside [R, H-R] x [R, W-R] are written
// as 0 (the reference's zero-initialised output replaces the depth).
// U16: the input is the raw uint16 frame and the range gate of k_depth_u16
// is applied as the tile is loaded (the gated frame itself is not stored).
// Its depths are whole millimetres, so a tap's range weight
// expf(-(t * t) / den) depends only on the integer |t|: with lut_n > 0 it is
// read from the table of those exact float results (staged in LDS), entries
// past lut_n being exactly 0.
// The filtered value of the pixel whose window starts at tile[ty][tx] (the
// window is tile[ty .. ty + 2R][tx .. tx + 2R], centre tile[ty + R][tx + R]),
// with the reference's arithmetic (bilateral_filtering.cu:26-44).  lut_n > 0:
// range weights from the table of whole-millimetre differences (s_lut).
// a / b, correctly rounded, from y = 1 / b (correctly rounded): q = a * y,
// then one fused correction q + (a - b q) y (Markstein's theorem: Markstein,
// IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of Floating-Point
// Arithmetic, ch. 4).  Exact whenever no intermediate overflows or
// underflows, which holds for the filter's weights (1e-9 <= a <= b <= the
// tap count) and for the trilinear samples' quotients by the voxel size and
// by the gradient norm (mm-scale or unit-scale values, or exact zeros, whose
// quotient may come out +0 where division gives -0).  The filter normalises
// every tap by the same sum, so one division per pixel replaces one per tap.
__device__ __forceinline__ float div_by_recip(float a, float b, float y) {
  const float q = a * y;
  const float r = fmaf(-q, b, a);
  return fmaf(r, y, q);
}

// Windows of up to this many taps keep their weights in registers; larger
// ones compute each weight twice, one window row at a time (with the range
// table: 5x5 9.3 us and 7x7 15.5 us in registers against 9.6 and 20.7 us
// recomputed; 9x9 21.7 us recomputed against 38.3 us in registers).
constexpr int kBilateralRegTaps = 49;

template <int R, int PITCH>
__device__ __forceinline__ float bilateral_value(const float (*tile)[PITCH], int ty, int tx,
                                                 const float* sw, const float* s_lut, int lut_n,
                                                 float sigma_r) {
  constexpr int L = 2 * R + 1;
  const float dc = tile[ty + R][tx + R];
  const float den = 2 * sigma_r * sigma_r;
  float sum = 0.f, fin = 0.f;
  if (lut_n > 0) {
    if constexpr (L * L <= kBilateralRegTaps) {
      float w[L * L];
#pragma unroll
      for (int k = 0; k < L * L; ++k) {
        const int j = (int)fabsf(tile[ty + k / L][tx + k % L] - dc);
        const float t = j < lut_n ? s_lut[j] : 0.0f;
        w[k] = sw[k] * t + 0.000000001f;
        sum = sum + w[k];
      }
      const float y = 1.0f / sum;
#pragma unroll
      for (int k = 0; k < L * L; ++k) {
        const float wr = div_by_recip(w[k], sum, y);  // = w[k] / sum
        fin = fin + wr * tile[ty + k / L][tx + k % L];
      }
    } else {
      // large windows: the (identical) weights again in the second pass,
      // one window row at a time, instead of L * L weights in registers
#pragma unroll 1
      for (int r = 0; r < L; ++r) {
#pragma unroll
        for (int cc = 0; cc < L; ++cc) {
          const int j = (int)fabsf(tile[ty + r][tx + cc] - dc);
          const float t = j < lut_n ? s_lut[j] : 0.0f;
          sum = sum + (sw[r * L + cc] * t + 0.000000001f);
        }
      }
      const float y = 1.0f / sum;
#pragma unroll 1
      for (int r = 0; r < L; ++r) {
#pragma unroll
        for (int cc = 0; cc < L; ++cc) {
          const float v = tile[ty + r][tx + cc];
          const int j = (int)fabsf(v - dc);
          const float t = j < lut_n ? s_lut[j] : 0.0f;
          const float wr = div_by_recip(sw[r * L + cc] * t + 0.000000001f, sum, y);
          fin = fin + wr * v;
        }
      }
    }
  } else if constexpr (L * L <= kBilateralRegTaps) {
    // small windows: keep the weights in registers (fully unrolled)
    float w[L * L];
#pragma unroll
    for (int k = 0; k < L * L; ++k) {
      float t = tile[ty + k / L][tx + k % L] - dc;
      t = kf_expf(-(t * t) / den);
      w[k] = sw[k] * t + 0.000000001f;
      sum = sum + w[k];
    }
    const float y = 1.0f / sum;
#pragma unroll
    for (int k = 0; k < L * L; ++k) {
      const float wr = div_by_recip(w[k], sum, y);  // = w[k] / sum
      fin = fin + wr * tile[ty + k / L][tx + k % L];
    }
  } else {
    // large windows: recompute the (identical) weights in the second pass
    for (int k = 0; k < L * L; ++k) {
      float t = tile[ty + k / L][tx + k % L] - dc;
      t = kf_expf(-(t * t) / den);
      sum = sum + (sw[k] * t + 0.000000001f);
    }
    const float y = 1.0f / sum;
    for (int k = 0; k < L * L; ++k) {
      const float v = tile[ty + k / L][tx + k % L];
      float t = v - dc;
      t = kf_expf(-(t * t) / den);
      const float wr = div_by_recip(sw[k] * t + 0.000000001f, sum, y);  // = w / sum
      fin = fin + wr * v;
    }
  }
  return fin;
}

// Stages the (SW)^2 window with top-left pixel (x0 - R, y0 - R) of the depth
// into tile, read by linear index y*W + x exactly as the reference reads it
// (a column past the right edge wraps to the next row; indices outside the
// image read 0).  U16: the raw frame with the range gate applied.
template <int R, int SW, bool U16>
__device__ __forceinline__ void bilateral_stage(const void* __restrict__ in_, float (*tile)[SW + 1],
                                                int x0, int y0, int H, int W, float dmin,
                                                float dmax) {
  const int n = H * W;
  for (int i = threadIdx.x; i < SW * SW; i += 256) {
    const int r = i / SW, cc = i - r * SW;
    const int idx = (y0 - R + r) * W + (x0 - R + cc);
    float v = 0.0f;
    if (idx >= 0 && idx < n) {
      if constexpr (U16) {
        const float d = (float)static_cast<const uint16_t*>(in_)[idx];
        v = (d >= dmin && d <= dmax) ? d : 0.0f;
      } else {
        v = static_cast<const float*>(in_)[idx];
      }
    }
    tile[r][cc] = v;
  }
}

template <int R, bool U16>
__global__ __launch_bounds__(256) void k_bilateral(const void* __restrict__ in_,
                                                   float* __restrict__ out,
                                                   const float* __restrict__ ws, int H, int W,
                                                   float sigma_r, float dmin, float dmax,
                                                   const float* __restrict__ lut, int lut_n) {
  constexpr int TS = 16, L = 2 * R + 1, SW = TS + 2 * R;
  __shared__ float tile[SW][SW + 1];
  __shared__ float sw[L * L];
  __shared__ float s_lut[U16 ? KF_RANGE_LUT_MAX : 1];
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * TS, y0 = blockIdx.y * TS;
  bilateral_stage<R, SW, U16>(in_, tile, x0, y0, H, W, dmin, dmax);
  for (int i = tid; i < L * L; i += 256) sw[i] = ws[i];
  if constexpr (U16)
    for (int i = tid; i < lut_n; i += 256) s_lut[i] = lut[i];
  __syncthreads();
  const int tx = tid & (TS - 1), ty = tid / TS;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= W || y >= H) return;
  float result = 0.0f;
  if (y >= R && y <= H - R && x >= R && x <= W - R)
    result = bilateral_value<R, SW + 1>(tile, ty, tx, sw, s_lut, U16 ? lut_n : 0, sigma_r);
  out[y * W + x] = result;
}

// ---------------------------------------------------------------------------
// A3  vertex and normal maps (raw_vertices.cu:3-51).  Writes every pixel, so
// the zero defaults of the reference's CleanDepthData need no separate fill.
// Vertex, normal and mask of pixel (x, y) from its depth di and the depths
// of the pixels below (dr) and to the right (dcn); dr and dcn are read only
// where the reference reads them (di != 0, not the last row or column).
// The stored Vz is the depth where the mask is 1 and 0 elsewhere: ICP uses a
// pixel where Vz != 0 and the mask is set (sensor_pose_estimation.cu:20),
// and a set mask implies di != 0, so "stored Vz != 0" is that test, with the
// same Vz wherever it passes.
__device__ __forceinline__ void vertex_at(int x, int y, int H, int W, const float* Ki, float di,
                                          float dr, float dcn, kf_p3& V, kf_p3& N,
                                          uint8_t& m) {
  V = kf_p3{0.f, 0.f, 0.f};
  N = kf_p3{0.f, 0.f, 0.f};
  m = 0;
  if (di != 0 && y != (H - 1) && x != (W - 1)) {
    m = 1;
    V.z = di;
    const int nr = (y + 1) * W + x;
    const int nc = y * W + (x + 1);
    if (nr > 0 && nc > 0 && dr != 0 && dcn != 0) {
      const float X1 = (float)(x + 1), Y1 = (float)(y + 1);
      const float X2 = (float)(x + 2), Y2 = (float)(y + 2);
      const float rvx = di * (Ki[0] * X1 + Ki[1] * Y1 + Ki[2]);
      const float rvy = di * (Ki[3] * X1 + Ki[4] * Y1 + Ki[5]);
      const float rvx_r = dr * (Ki[0] * X1 + Ki[1] * Y2 + Ki[2]);
      const float rvx_c = dcn * (Ki[0] * X2 + Ki[1] * Y1 + Ki[2]);
      const float rvy_r = dr * (Ki[3] * X1 + Ki[4] * Y2 + Ki[5]);
      const float rvy_c = dcn * (Ki[3] * X2 + Ki[4] * Y1 + Ki[5]);
      const float nI = (rvy_r - rvy) * (dcn - di) - (dr - di) * (rvy_c - rvy);
      const float nJ = (dr - di) * (rvx_c - rvx) - (rvx_r - rvx) * (dcn - di);
      const float nK = (rvx_r - rvx) * (rvy_c - rvy) - (rvy_r - rvy) * (rvx_c - rvx);
      const float nn = sqrtf(nI * nI + nJ * nJ + nK * nK);
      V.x = rvx;
      V.y = rvy;
      N.x = nI / nn;
      N.y = nJ / nn;
      N.z = nK / nn;
    }
  }
}

__global__ __launch_bounds__(256) void k_vertex(const float* __restrict__ d, kf_cam cam, int H,
                                                int W, kf_p3* __restrict__ vmap,
                                                kf_p3* __restrict__ nmap,
                                                uint8_t* __restrict__ mask) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= H * W) return;
  const int y = i / W, x = i - y * W;
  // the three loads are unconditional (clamped indices), so that they are in
  // flight together; vertex_at uses the neighbours only where they exist
  const int n = H * W;
  const float di = d[i];
  const float dr = d[min((y + 1) * W + x, n - 1)];
  const float dcn = d[min(y * W + (x + 1), n - 1)];
  kf_p3 V, N;
  uint8_t m;
  vertex_at(x, y, H, W, cam.Kinv, di, dr, dcn, V, N, m);
  vmap[i] = V;
  nmap[i] = N;
  mask[i] = m;
}

// ---------------------------------------------------------------------------
// Coarse-to-fine pyramid for ICP (opt-in; the reference runs every iteration
// at full resolution, sensor_pose_estimation.cu:339).  Level l+1 has
// floor(H/2) x floor(W/2) pixels; its pixel (x, y) covers pixels (2x + i,
// 2y + j), i, j in {0, 1}, of level l.

// Depth: the mean of the block, ((d00 + d01) + (d10 + d11)) * 0.25, where all
// four are > 0 and max - min <= max_step; 0 elsewhere (NaN fails both tests).
__global__ __launch_bounds__(256) void k_depth_down(const float* __restrict__ d_in, int Wi,
                                                    float* __restrict__ d_out, int Ho, int Wo,
                                                    float max_step) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= Ho * Wo) return;
  const int y = i / Wo, x = i - y * Wo;
  const float* r0 = d_in + (size_t)(2 * y) * Wi + 2 * x;
  const float* r1 = r0 + Wi;
  const float a = r0[0], b = r0[1], c = r1[0], d = r1[1];
  const float lo = fminf(fminf(a, b), fminf(c, d)), hi = fmaxf(fmaxf(a, b), fmaxf(c, d));
  const bool ok = a > 0.0f && b > 0.0f && c > 0.0f && d > 0.0f && hi - lo <= max_step;
  d_out[i] = ok ? ((a + b) + (c + d)) * 0.25f : 0.0f;
}

// Model maps: where all four pixels hold a surface (hit index >= 0 at level
// 0, ok set at coarser levels), the mean vertex ((v00 + v01) + (v10 + v11)) *
// 0.25 and the normalised sum of the normals (normalize3); elsewhere zeros
// and ok = 0, as a raycast pixel without a hit.  t_out: kf_inv_norm3 of the
// normal written.
__global__ __launch_bounds__(256) void k_model_down(
    const kf_p3* __restrict__ v_in, const kf_p3* __restrict__ n_in,
    const int32_t* __restrict__ hit_in, const uint8_t* __restrict__ ok_in, int Wi,
    kf_p3* __restrict__ v_out, kf_p3* __restrict__ n_out, float* __restrict__ t_out,
    uint8_t* __restrict__ ok_out, int Ho, int Wo) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= Ho * Wo) return;
  const int y = i / Wo, x = i - y * Wo;
  const size_t p[4] = {(size_t)(2 * y) * Wi + 2 * x, (size_t)(2 * y) * Wi + 2 * x + 1,
                       (size_t)(2 * y + 1) * Wi + 2 * x, (size_t)(2 * y + 1) * Wi + 2 * x + 1};
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) ok = ok && (hit_in ? hit_in[p[k]] >= 0 : ok_in[p[k]] != 0);
  kf_p3 V{0.f, 0.f, 0.f}, N{0.f, 0.f, 0.f};
  if (ok) {
    const kf_p3 v0 = v_in[p[0]], v1 = v_in[p[1]], v2 = v_in[p[2]], v3 = v_in[p[3]];
    const kf_p3 n0 = n_in[p[0]], n1 = n_in[p[1]], n2 = n_in[p[2]], n3 = n_in[p[3]];
    V = kf_p3{((v0.x + v1.x) + (v2.x + v3.x)) * 0.25f, ((v0.y + v1.y) + (v2.y + v3.y)) * 0.25f,
              ((v0.z + v1.z) + (v2.z + v3.z)) * 0.25f};
    float nx = (n0.x + n1.x) + (n2.x + n3.x), ny = (n0.y + n1.y) + (n2.y + n3.y),
          nz = (n0.z + n1.z) + (n2.z + n3.z);
    normalize3(nx, ny, nz);
    N = kf_p3{nx, ny, nz};
  }
  v_out[i] = V;
  n_out[i] = N;
  t_out[i] = kf_inv_norm3(N.x, N.y, N.z);
  ok_out[i] = ok ? 1 : 0;
}

// Consecutive block ids are dealt round-robin to the 8 XCDs, each with its
// own L2.  xcd_contiguous maps a block id to a work index so that each XCD
// takes a contiguous run of the n_blocks work items (neighbouring pixels
// gather neighbouring model pixels; neighbouring tiles read neighbouring
// bricks): XCD x receives hardware blocks x, x+8, ...: q+1 of them for
// x < r, q otherwise.  It is a bijection for any n_blocks, so results do not
// depend on how the hardware actually deals blocks.
__device__ __forceinline__ int xcd_contiguous(int bid, int n_blocks) {
  const int q = n_blocks >> 3, r = n_blocks & 7;
  const int xcd = bid & 7, loc = bid >> 3;
  const int base = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  return base + loc;
}

// ---------------------------------------------------------------------------
// A5-A7  ICP.

// One kernel per iteration, and no cross-block synchronisation inside a
// kernel.  The blocks of iteration i first finish iteration i-1: each block
// reads the block partials iteration i-1 wrote (ROWS rows of nb floats, rows
// 2k and 2k+1 = sum and compensation of quantity k in compensated mode, row k
// in plain mode), adds them in FP64 in a fixed order, solves the 6x6 system
// and composes the pose.  Every block computes bit-identical values; block 0
// records them.  The blocks of iteration 0 start the frame instead:
// T_prev := T, T_prev^-1 and F = T_prev^-1 T (sensor_pose_estimation.cu:
// 255-259, 343).  A one-block kernel finishes the last iteration.  Partials
// are double-buffered by iteration parity, and the pose an iteration starts
// from is read from the per-iteration record, so nothing a kernel reads is
// written by that kernel.

template <bool KAHAN>
using IcpPartial = typename std::conditional<KAHAN, float, double>::type;

template <int ROWS>
struct IcpShared {
  double rowsum[ROWS];
  double sums[KF_ICP_SUMS];
  float inc[16], T0[16], Tp[16], Tn[16], F[16];
  int status;
};

// Sums each of the ROWS rows of nb partials into sh.rowsum (all threads
// take part).  Wave w takes rows w, w + NW, ...; each lane loads its PPL
// entries of those rows (nb <= 64 PPL) with unconditional (clamped) loads,
// so that they are all in flight together (icp_load_partials), adds them in
// FP64, and the 64 lane sums of a row are added by a DPP tree, a fixed order
// (icp_sum_partials, which ends with a block barrier; icp_update_pose forms
// the sums).
template <int ROWS, int NT, int PPL, typename PT>
struct IcpPartialRegs {
  static constexpr int NW = NT / 64;
  static constexpr int RPW = (ROWS + NW - 1) / NW;  // rows per wave
  PT v[RPW][PPL];
};

template <int ROWS, int NT, int PPL, typename PT, typename PP>
__device__ __forceinline__ void icp_load_partials(PP P, int nb,
                                                  IcpPartialRegs<ROWS, NT, PPL, PT>& R) {
  using Regs = IcpPartialRegs<ROWS, NT, PPL, PT>;
  const int wave = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < Regs::RPW; ++r) {
    const int row = min(wave + r * Regs::NW, ROWS - 1);
#pragma unroll
    for (int j = 0; j < PPL; ++j)
      R.v[r][j] = P[(size_t)row * nb + min(ln + 64 * j, nb - 1)];
  }
}

template <int ROWS, int NT, int PPL, typename PT>
__device__ __forceinline__ void icp_sum_partials(const IcpPartialRegs<ROWS, NT, PPL, PT>& R,
                                                 int nb, IcpShared<ROWS>& sh) {
  using Regs = IcpPartialRegs<ROWS, NT, PPL, PT>;
  const int wave = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < Regs::RPW; ++r) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < PPL; ++j)
      t += ln + 64 * j < nb ? (double)R.v[r][j] : 0.0;
    t = wave_total(t);
    const int row = wave + r * Regs::NW;
    if (ln == 63 && row < ROWS) sh.rowsum[row] = t;
  }
  __syncthreads();
}

template <int ROWS, int NT, int PPL, typename PT>
__device__ __forceinline__ void icp_reduce_partials(const PT* __restrict__ P, int nb,
                                                    IcpShared<ROWS>& sh) {
  IcpPartialRegs<ROWS, NT, PPL, PT> R;
  icp_load_partials<ROWS, NT, PPL, PT>(P, nb, R);  // NOLINT
  icp_sum_partials<ROWS, NT, PPL, PT>(R, nb, sh);
}

// LDS visibility between the lanes of one wave.
__device__ __forceinline__ void icp_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Finishes an iteration in wave 0: the sums from the row sums (a sum and
// its compensation row in the compensated variant), the solve (one lane),
// T_new = inc(sums) * T0 and F = Tp * T_new (16 lanes each), then one block
// barrier.  min_inliers >= 0 selects the robust increment.
template <int ROWS>
__device__ __forceinline__ void icp_update_pose(IcpShared<ROWS>& sh, float min_inliers) {
  if (threadIdx.x < 64) {
    if (threadIdx.x < KF_ICP_SUMS) {
      const int k = threadIdx.x;
      sh.sums[k] = ROWS == 2 * KF_ICP_SUMS ? sh.rowsum[2 * k] + sh.rowsum[2 * k + 1] : sh.rowsum[k];
    }
    icp_wave_sync();
    if (threadIdx.x == 0)
      sh.status = min_inliers >= 0.0f
                      ? kf_icp_increment_robust(sh.sums, sh.inc, (double)min_inliers)
                      : kf_icp_increment(sh.sums, sh.inc);
    icp_wave_sync();
    if (threadIdx.x < 16)
      sh.Tn[threadIdx.x] = kf_mat4_mul_elem(sh.inc, sh.T0, threadIdx.x >> 2, threadIdx.x & 3);
    icp_wave_sync();
    if (threadIdx.x < 16)
      sh.F[threadIdx.x] = kf_mat4_mul_elem(sh.Tp, sh.Tn, threadIdx.x >> 2, threadIdx.x & 3);
  }
  __syncthreads();
}

// Records iteration `it`'s result (block 0 / the finishing kernel only).
template <int ROWS>
__device__ __forceinline__ void icp_record(kf_dev_state* st, int it, const IcpShared<ROWS>& sh) {
  if (threadIdx.x < 16) st->iter_T[it][threadIdx.x] = sh.Tn[threadIdx.x];
  if (threadIdx.x < KF_ICP_SUMS) st->iter_sums[it][threadIdx.x] = sh.sums[threadIdx.x];
  if (threadIdx.x == 0) {
    st->iter_status[it] = sh.status;
    st->icp_iters_done = it + 1;
  }
}

// Per-pixel association and gates (sensor_pose_estimation.cu:16-136) plus the
// block reduction, for iteration `iter`.  nb_in is the grid of the iteration
// before (the row stride of its partials): the levels of the opt-in
// coarse-to-fine pyramid have grids of their own.  FROM_DEPTH: the frame's
// vertex and normal maps are not read but computed from the depth (three
// loads per pixel, raw_vertices.cu:3-51 through vertex_at) and written to
// vmap_out / nmap_out / vmask_out for the later iterations, so that the
// pipeline needs no separate vertex kernel.
// STOP: the opt-in early stop is compiled in (kf_icp_converged, with the
// thresholds and the first iteration that may stop in the device state), so
// that the reference's fixed iteration count runs the kernel without it.
// NT: the normal gate normalises the model normal by mt[j] (kf_inv_norm3 of
// mn[j], kept with the model maps: it does not change between iterations)
// instead of normalize3, the same float operations.
// BLK: threads per block, KF_ICP_BLOCK, or 2 KF_ICP_BLOCK for images above
// one block of 4 KF_ICP_BLOCK pixels per compute unit (icp_grid): every
// block reads every block partial of the iteration before, so halving the
// block count quarters those reads.
template <bool KAHAN, int PX, int PPL, bool FROM_DEPTH, bool STOP, bool NT, int BLK>
__global__ __launch_bounds__(BLK) void k_icp_iter(
    const float* __restrict__ depth, kf_p3* __restrict__ vmap_out, kf_p3* __restrict__ nmap_out,
    uint8_t* __restrict__ vmask_out, const kf_p3* __restrict__ vmap, const kf_p3* __restrict__ nmap,
    const kf_p3* __restrict__ mv, const kf_p3* __restrict__ mn, const float* __restrict__ mt,
    kf_cam cam, int H, int W,
    float dist2_hi, float cos_lo, float min_inliers, kf_dev_state* __restrict__ st,
    const void* __restrict__ part_in_, void* __restrict__ part_out_,
    uint8_t* __restrict__ mask_out, int iter, int bpx, int nb_in) {
  constexpr int NS = KF_ICP_SUMS;
  constexpr int ROWS = KAHAN ? 2 * NS : NS;
  // block partials: FP64 sums, or FP32 (sum, compensation) pairs when KAHAN
  using PT = IcpPartial<KAHAN>;
  const PT* __restrict__ part_in = static_cast<const PT*>(part_in_);
  PT* __restrict__ part_out = static_cast<PT*>(part_out_);
  __shared__ IcpShared<ROWS> sh;
  __shared__ PT red_s[BLK / 64][NS];
  __shared__ PT red_c[KAHAN ? BLK / 64 : 1][NS];
  const int nb = gridDim.x;

  // chunk c takes pixels [c bpx, (c + 1) bpx) of the image, thread t the
  // pixels t, t + BLK, ... below bpx (bpx <= PX BLK); each
  // XCD takes a contiguous band of chunks, so the model rows its blocks
  // gather are shared in its L2 instead of fetched by two XCDs at every
  // chunk boundary.  The partials stay indexed by block.
  const int n = H * W;
  const int chunk = xcd_contiguous(blockIdx.x, nb);
  const int base = chunk * bpx + threadIdx.x;
  const int lim = min(bpx, n - chunk * bpx) - (int)threadIdx.x;  // k valid: k * BLK < lim
  // phase 1: the previous iteration's block partials and the two poses it
  // needs, then this thread's vertex and normal loads, all issued before
  // the previous iteration is finished (the maps do not depend on the pose).
  // Loads complete in issue order, so the prologue's loads go first: it
  // waits for them and not for the maps.  Every load is unconditional, with
  // in-bounds indices (iteration 0 loads the other parity's partials and
  // ignores them), so that no branch makes the compiler wait at its join.
  // The order is pinned for the compiler: the pointers are passed through
  // empty asm statements (which drops what their __restrict__ lets it
  // assume), and an asm memory clobber separates the two groups of loads.
  const PT* pin_ = part_in;
  const kf_dev_state* stp_ = st;
  const kf_p3* vm_ = vmap;
  const kf_p3* nm_ = nmap;
  const float* dp_ = depth;
  asm volatile("" : "+s"(pin_), "+s"(stp_), "+s"(vm_), "+s"(nm_), "+s"(dp_));
  // back in the global address space (global, not flat, loads)
  const KF_GLOBAL PT* pin = (const KF_GLOBAL PT*)pin_;
  const KF_GLOBAL kf_dev_state* stp = (const KF_GLOBAL kf_dev_state*)stp_;
  const KF_GLOBAL kf_p3* vm = (const KF_GLOBAL kf_p3*)vm_;
  const KF_GLOBAL kf_p3* nm = (const KF_GLOBAL kf_p3*)nm_;
  const KF_GLOBAL float* dp = (const KF_GLOBAL float*)dp_;
  IcpPartialRegs<ROWS, BLK, PPL, PT> R;
  icp_load_partials<ROWS, BLK, PPL, PT>(pin, nb_in, R);
  const int e16 = threadIdx.x & 15;
  const float T0e = iter <= 1 ? stp->Tprev[e16] : stp->iter_T[iter - 2][e16];
  const float Tpe = stp->Tprev_inv[e16];
  // opt-in early stop: the first iteration that may stop, and whether an
  // earlier iteration of this frame stopped the ICP
  const int stop_from = STOP ? stp->icp_stop_from : 0;
  const int stop_at = STOP ? stp->icp_stop_at : 0;
  asm volatile("" ::: "memory");
  kf_p3 V[PX], Nc[PX];
  // FROM_DEPTH: each pixel's depth and those below and to the right of it
  constexpr int PD = FROM_DEPTH ? PX : 1;
  float Di[PD], Dr[PD], Dc[PD];
#pragma unroll
  for (int k = 0; k < PX; ++k) {
    const int x = min(base + k * BLK, n - 1);
    if constexpr (FROM_DEPTH) {
      const int py = x / W, px = x - py * W;
      Di[k] = dp[x];
      Dr[k] = dp[min((py + 1) * W + px, n - 1)];
      Dc[k] = dp[min(py * W + (px + 1), n - 1)];
    } else {
      V[k] = vm[x];
      Nc[k] = nm[x];
    }
  }
  if (iter == 0) {
    // frame start
    if (threadIdx.x == 0) {
      float T[16], Ti[16];
      for (int i = 0; i < 16; ++i) T[i] = st->T[i];
      if (kf_m4_inverse(T, Ti) != 0)
        for (int i = 0; i < 16; ++i) Ti[i] = __builtin_nanf("");
      for (int i = 0; i < 16; ++i) sh.Tp[i] = Ti[i], sh.Tn[i] = T[i];
    }
    __syncthreads();
    if (threadIdx.x < 16)
      sh.F[threadIdx.x] = kf_mat4_mul_elem(sh.Tp, sh.Tn, threadIdx.x >> 2, threadIdx.x & 3);
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x < 16) {
      st->Tprev[threadIdx.x] = sh.Tn[threadIdx.x];
      st->Tprev_inv[threadIdx.x] = sh.Tp[threadIdx.x];
    }
    if (STOP && blockIdx.x == 0 && threadIdx.x == 0) st->icp_stop_at = 0;
  } else {
    if (STOP && iter > stop_from && stop_at) return;  // uniform: one record for every block
    // finish iteration iter-1, which started from the pose recorded for
    // iter-2 (or from T_prev for iter-1 == 0)
    if (threadIdx.x < 16) {
      sh.T0[threadIdx.x] = T0e;
      sh.Tp[threadIdx.x] = Tpe;
    }
    icp_sum_partials<ROWS, BLK, PPL, PT>(R, nb_in, sh);
    icp_update_pose<ROWS>(sh, min_inliers);
    if (blockIdx.x == 0) icp_record<ROWS>(st, iter - 1, sh);
    // opt-in early stop after a small increment (every block decides alike
    // from the same LDS values): iteration iter-1 was the last; the later
    // iterations and the finishing kernel read icp_stop_at
    if (STOP && iter >= stop_from &&
        kf_icp_converged(sh.inc, st->icp_stop_rot, st->icp_stop_trans)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) st->icp_stop_at = iter;
      return;
    }
  }
  if constexpr (FROM_DEPTH) {
    // the frame's vertex, normal and mask (raw_vertices.cu:3-51), kept for
    // the later iterations
#pragma unroll
    for (int k = 0; k < PX; ++k) {
      const int x = min(base + k * BLK, n - 1);
      const int py = x / W, px = x - py * W;
      uint8_t m;
      vertex_at(px, py, H, W, cam.Kinv, Di[k], Dr[k], Dc[k], V[k], Nc[k], m);
      if (k * BLK < lim) {
        vmap_out[x] = V[k];
        nmap_out[x] = Nc[k];
        vmask_out[x] = m;
      }
    }
  }
  float F[12], Te[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) F[i] = sh.F[i], Te[i] = sh.Tn[i];
  const float* K = cam.K;

  float acc[NS], comp[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) acc[k] = 0.f, comp[k] = 0.f;

#pragma unroll
  for (int k = 0; k < PX; ++k)
    if (k * BLK >= lim) V[k] = Nc[k] = kf_p3{0.f, 0.f, 0.f};
  // phase 2: projective association, then the model gathers, issued together.
  // Vz != 0 is the reference's "Vz != 0 and mask" (the vertex map stores 0
  // where the mask is 0).
  int jx[PX];
#pragma unroll
  for (int k = 0; k < PX; ++k) {
    jx[k] = -1;
    if (V[k].z != 0) {
      const float cx = F[0] * V[k].x + F[1] * V[k].y + F[2] * V[k].z + F[3];
      const float cy = F[4] * V[k].x + F[5] * V[k].y + F[6] * V[k].z + F[7];
      const float cz = F[8] * V[k].x + F[9] * V[k].y + F[10] * V[k].z + F[11];
      const float px = K[0] * cx + K[1] * cy + K[2] * cz;
      const float py = K[3] * cx + K[4] * cy + K[5] * cz;
      const float pz = K[6] * cx + K[7] * cy + K[8] * cz;
      const int u = kf_round_to_int(px / pz);
      const int v = kf_round_to_int(py / pz);
      if (u > 0 && u <= W && v > 0 && v <= H) jx[k] = (v - 1) * W + (u - 1);
    }
  }
  // the model gathers, unconditional (pixel 0 for unassociated pixels, whose
  // values are not used).  Each empty asm statement takes a pixel's six
  // gathered values as register inputs, so all 2 PX loads are issued before
  // the first of them is waited for; without it the compiler moves each load
  // into the branch that uses it, and the gathers of a thread's pixels
  // (vertex, then normal after the distance gate) become 2 PX dependent
  // memory round trips.
  kf_p3 S[PX], Sn[PX];
  float St[NT ? PX : 1];
#pragma unroll
  for (int k = 0; k < PX; ++k) {
    S[k] = mv[max(jx[k], 0)];
    Sn[k] = mn[max(jx[k], 0)];
    if constexpr (NT) St[k] = mt[max(jx[k], 0)];
  }
#pragma unroll
  for (int k = 0; k < PX; ++k) {
    asm volatile("" ::"v"(S[k].x), "v"(S[k].y), "v"(S[k].z), "v"(Sn[k].x), "v"(Sn[k].y),
                 "v"(Sn[k].z));
    if constexpr (NT) asm volatile("" ::"v"(St[k]));
  }
  // phase 3: gates, Jacobian rows, accumulation
#pragma unroll
  for (int k = 0; k < PX; ++k) {
    const int x = base + k * BLK;
    bool ok = false;
    float a[6], b = 0.f;
    if (jx[k] >= 0) {
      const kf_p3 Vk = V[k];
      const float ux = Te[0] * Vk.x + Te[1] * Vk.y + Te[2] * Vk.z + Te[3];
      const float uy = Te[4] * Vk.x + Te[5] * Vk.y + Te[6] * Vk.z + Te[7];
      const float uz = Te[8] * Vk.x + Te[9] * Vk.y + Te[10] * Vk.z + Te[11];
      const kf_p3 Sk = S[k];
      const float ddx = ux - Sk.x, ddy = uy - Sk.y, ddz = uz - Sk.z;
      // sqrtf(d2) <= gate (sensor_pose_estimation.cu:44-48) as d2 <= the
      // largest float whose correctly rounded square root passes; NaN fails
      // both forms
      const float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
      if (d2 <= dist2_hi) {
        const kf_p3 Nk = Nc[k];
        float nnx = Te[0] * Nk.x + Te[1] * Nk.y + Te[2] * Nk.z;
        float nny = Te[4] * Nk.x + Te[5] * Nk.y + Te[6] * Nk.z;
        float nnz = Te[8] * Nk.x + Te[9] * Nk.y + Te[10] * Nk.z;
        normalize3(nnx, nny, nnz);
        const kf_p3 Snk = Sn[k];
        float ax = Snk.x, ay = Snk.y, az = Snk.z;
        if constexpr (NT) {
          const float tn = St[k];
          ax = ax * tn;
          ay = ay * tn;
          az = az * tn;
        } else {
          normalize3(ax, ay, az);
        }
        const float c = (nnx * ax + nny * ay + nnz * az);
        if (c >= cos_lo && c <= 1.0f) {
          ok = true;
          a[0] = uz * Snk.y + (-uy * Snk.z);
          a[1] = -uz * Snk.x + (ux * Snk.z);
          a[2] = uy * Snk.x + (-ux * Snk.y);
          a[3] = Snk.x;
          a[4] = Snk.y;
          a[5] = Snk.z;
          b = Snk.x * (Sk.x - ux) + Snk.y * (Sk.y - uy) + Snk.z * (Sk.z - uz);
        }
      }
    }
    if (mask_out && k * BLK < lim) mask_out[x] = ok ? 1 : 0;
    if (ok) {
      int q = 0;
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int s = r; s < 6; ++s) {
          const float p = a[r] * a[s];
          if constexpr (KAHAN) {
            const float y = p - comp[q];
            const float t = acc[q] + y;
            comp[q] = (t - acc[q]) - y;
            acc[q] = t;
          } else {
            acc[q] = acc[q] + p;
          }
          ++q;
        }
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        const float p = a[r] * b;
        if constexpr (KAHAN) {
          const float y = p - comp[21 + r];
          const float t = acc[21 + r] + y;
          comp[21 + r] = (t - acc[21 + r]) - y;
          acc[21 + r] = t;
        } else {
          acc[21 + r] = acc[21 + r] + p;
        }
      }
      acc[27] = acc[27] + 1.0f;
    }
  }

  // wave reduction -> LDS.  Kahan keeps s - c as the value (c is the running
  // error with the opposite sign), so the pair (s, -c) is reduced in FP32
  // TwoSum arithmetic; otherwise the lane sums (FP32 over this thread's
  // pixels, as the per-pixel products) are added in FP64 from here on.
  const int wave = threadIdx.x >> 6, ln = lane_id();
  {
    PT rs[32], rc[KAHAN ? 32 : 1];
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      rs[k] = k < NS ? (PT)acc[k] : (PT)0;
      if constexpr (KAHAN) rc[k] = k < NS ? -comp[k] : 0.f;
    }
    wave_transpose_sum<KAHAN>(rs, rc);
    const int k = wave_transpose_index(ln);
    if ((ln & 1) == 0 && k < NS) {
      red_s[wave][k] = rs[0];
      if constexpr (KAHAN) red_c[wave][k] = rc[0];
    }
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    const int k = threadIdx.x;
    PT s = red_s[0][k], c = KAHAN ? red_c[0][k] : (PT)0;
#pragma unroll
    for (int w = 1; w < BLK / 64; ++w) {
      if constexpr (KAHAN)
        two_sum_add(s, c, red_s[w][k], red_c[w][k]);
      else
        s = s + red_s[w][k];
    }
    if constexpr (KAHAN) {
      part_out[(size_t)(2 * k) * nb + blockIdx.x] = s;
      part_out[(size_t)(2 * k + 1) * nb + blockIdx.x] = c;
    } else {
      part_out[(size_t)k * nb + blockIdx.x] = s;
    }
  }
}

// Finishes the last iteration: final pose, its inverse and the record.
constexpr int kIcpFinishThreads = 1024;

// STOP: the opt-in early stop may have ended the ICP before the last
// iteration (compiled only into the kernel of the runs that use it).
template <bool KAHAN, int PPL, bool STOP>
__global__ __launch_bounds__(kIcpFinishThreads) void k_icp_finish(kf_dev_state* __restrict__ st,
                                                             const void* __restrict__ part_in,
                                                             int nb, int last, float min_inliers) {
  constexpr int ROWS = KAHAN ? 2 * KF_ICP_SUMS : KF_ICP_SUMS;
  __shared__ IcpShared<ROWS> sh;
  const int stopped = STOP ? st->icp_stop_at : 0;
  if (STOP && stopped) {
    // the ICP stopped early: the final pose is the one recorded for the last
    // iteration that ran
    if (threadIdx.x < 16) {
      sh.Tn[threadIdx.x] = st->iter_T[stopped - 1][threadIdx.x];
      sh.Tp[threadIdx.x] = st->Tprev_inv[threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x < 16)
      sh.F[threadIdx.x] = kf_mat4_mul_elem(sh.Tp, sh.Tn, threadIdx.x >> 2, threadIdx.x & 3);
    __syncthreads();
  } else {
    if (threadIdx.x < 16) {
      sh.T0[threadIdx.x] = last == 0 ? st->Tprev[threadIdx.x] : st->iter_T[last - 1][threadIdx.x];
      sh.Tp[threadIdx.x] = st->Tprev_inv[threadIdx.x];
    }
    icp_reduce_partials<ROWS, kIcpFinishThreads, PPL>(
        static_cast<const IcpPartial<KAHAN>*>(part_in), nb, sh);
    icp_update_pose<ROWS>(sh, min_inliers);
    icp_record<ROWS>(st, last, sh);
  }
  if (threadIdx.x < 16) {
    st->T[threadIdx.x] = sh.Tn[threadIdx.x];
    st->F[threadIdx.x] = sh.F[threadIdx.x];
  }
  if (threadIdx.x == 0) {
    float Ti[16];
    if (kf_m4_inverse(sh.Tn, Ti) != 0)
      for (int i = 0; i < 16; ++i) Ti[i] = __builtin_nanf("");
    for (int i = 0; i < 16; ++i) st->Tinv[i] = Ti[i];
  }
}

// ---------------------------------------------------------------------------
// A8  projective TSDF integration (projective_tsdf.cu:5-140).

// Front-test direction of integration (projective_tsdf.cu:57-71): the ray
// through the pixel at the swapped principal point (K[5], K[2]), through T.
__device__ __forceinline__ void integ_center_dir(const float* K, const float* Ki,
                                                 const float* T, float* cd) {
  const float cax = Ki[0] * K[5] + Ki[1] * K[2] + Ki[2] * 1.f;
  const float cay = Ki[3] * K[5] + Ki[4] * K[2] + Ki[5] * 1.f;
  const float caz = Ki[6] * K[5] + Ki[7] * K[2] + Ki[8] * 1.f;
  const float gax = T[0] * cax + T[1] * cay + T[2] * caz + T[3];
  const float gay = T[4] * cax + T[5] * cay + T[6] * caz + T[7];
  const float gaz = T[8] * cax + T[9] * cay + T[10] * caz + T[11];
  const float cbx = 2 * (Ki[0] * K[5] + Ki[1] * K[2] + Ki[2] * 1.f);
  const float cby = 2 * (Ki[3] * K[5] + Ki[4] * K[2] + Ki[5] * 1.f);
  const float cbz = 2 * (Ki[6] * K[5] + Ki[7] * K[2] + Ki[8] * 1.f);
  const float gbx = T[0] * cbx + T[1] * cby + T[2] * cbz + T[3];
  const float gby = T[4] * cbx + T[5] * cby + T[6] * cbz + T[7];
  const float gbz = T[8] * cbx + T[9] * cby + T[10] * cbz + T[11];
  cd[0] = gbx - gax;
  cd[1] = gby - gay;
  cd[2] = gbz - gaz;
}

// Integration's box culls (see k_integrate): the FP64 affine quantities the
// per-voxel tests depend on, at the voxel centre (x, y, z) of a box corner,
// and which of the five rejections hold there.  Each rejection set is convex
// (a linear form against a convex margin), so one that holds at all 8
// corners of a box holds on the whole box.
//
// any_side: the four image-side tests hold without "pz > M".  A voxel with
// camera coordinates l and p = K l lies on the ray l = pz K^-1 (u, v, 1)
// (u = px / pz, v = py / pz), so the front test's dot is
// mag = lc . l = pz (lc . K^-1 (u, v, 1)), lc = K^-1 (K[5], K[2], 1).  When
// lc . K^-1 (u, v, 1) > 0 over the image (checked at its corners on the
// host), a voxel behind the camera plane that projects into the image has
// mag < 0 and fails the front test, so the side tests alone decide; and a
// voxel near the plane projects far outside the image (|px| or |py| is at
// least M while pz is near 0).
struct BoxCorner {
  double cx, cy, cz;  // voxel centre
  double px, py, pz;  // K T^-1 c
  double tx, ty, tz;  // c - t
  bool front, r1, r2, r3, r4, r5;
};

__device__ __forceinline__ BoxCorner integ_corner(int x, int y, int z, const kf_vol& vg,
                                                  const float* Ti, const float* t,
                                                  const float* cd, const float* K, int H,
                                                  int W, bool any_side) {
  BoxCorner o;
  const float d = vg.d, hd = d / 2.f;
  o.cx = (double)((float)(x + vg.sx) * d + hd);
  o.cy = (double)((float)(y + vg.sy) * d + hd);
  o.cz = (double)((float)(z + vg.sz) * d + hd);
  const double lx = (double)Ti[0] * o.cx + (double)Ti[1] * o.cy + (double)Ti[2] * o.cz + (double)Ti[3];
  const double ly = (double)Ti[4] * o.cx + (double)Ti[5] * o.cy + (double)Ti[6] * o.cz + (double)Ti[7];
  const double lz = (double)Ti[8] * o.cx + (double)Ti[9] * o.cy + (double)Ti[10] * o.cz + (double)Ti[11];
  o.px = (double)K[0] * lx + (double)K[1] * ly + (double)K[2] * lz;
  o.py = (double)K[3] * lx + (double)K[4] * ly + (double)K[5] * lz;
  o.pz = (double)K[6] * lx + (double)K[7] * ly + (double)K[8] * lz;
  double kabs = 0.0;
  for (int i = 0; i < 9; ++i) kabs = fmax(kabs, fabs((double)K[i]));
  const double M = 16.0 + 3e-5 * kabs * (fabs(lx) + fabs(ly) + fabs(lz) + 1.0);
  o.tx = o.cx - (double)t[0], o.ty = o.cy - (double)t[1], o.tz = o.cz - (double)t[2];
  const double mag = (double)cd[0] * o.tx + (double)cd[1] * o.ty + (double)cd[2] * o.tz;
  const double Mm = 1.0 + 1e-5 * (fabs((double)cd[0]) + fabs((double)cd[1]) + fabs((double)cd[2])) *
                              (fabs(o.tx) + fabs(o.ty) + fabs(o.tz));
  o.front = o.pz > M;
  const bool side = any_side || o.front;
  o.r1 = mag < -Mm;
  o.r2 = side && o.px < 0.45 * o.pz - M;
  o.r3 = side && o.px > ((double)W + 0.55) * o.pz + M;
  o.r4 = side && o.py < 0.45 * o.pz - M;
  o.r5 = side && o.py > ((double)H + 0.55) * o.pz + M;
  return o;
}

// The 8 boxes of a wave (box j = lanes 8j .. 8j+7, one corner each): bit j
// is set when no rejection holds at all 8 corners of box j.
__device__ __forceinline__ uint32_t wave_box_keep(bool r1, bool r2, bool r3, bool r4, bool r5,
                                                  bool r6) {
  const uint64_t m1 = __ballot(r1), m2 = __ballot(r2), m3 = __ballot(r3), m4 = __ballot(r4),
                 m5 = __ballot(r5), m6 = __ballot(r6);
  uint32_t keep = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int sh = 8 * j;
    const bool rej = ((m1 >> sh) & 0xff) == 0xff || ((m2 >> sh) & 0xff) == 0xff ||
                     ((m3 >> sh) & 0xff) == 0xff || ((m4 >> sh) & 0xff) == 0xff ||
                     ((m5 >> sh) & 0xff) == 0xff || ((m6 >> sh) & 0xff) == 0xff;
    keep |= rej ? 0u : 1u << j;
  }
  return keep;
}

// Integration culls in two levels.  Groups of 2x2x2 bricks are tested first
// (by extra blocks of k_integ_table, with the frustum and front tests), and
// the surviving groups are appended to a list; k_integrate then tests the
// bricks of the listed groups only.  The list has KF_RAY_PARTS partitions,
// each with its own counter line (two counter slots used by alternate
// integrations, the integration clearing the other slot); the table block
// of group-block index gb appends to partition gb % KF_RAY_PARTS.
struct GroupList {
  int* glist;         // surviving groups: KF_RAY_PARTS x gcap ids
  uint32_t* gcount;   // their counters (this integration's slot)
  uint32_t* gnext;    // the next integration's slot
  int gcap;           // groups per partition
  int* blist;         // surviving bricks: KF_RAY_PARTS x bcap ids
  uint32_t* bcount;   // their counters (this integration's slot)
  uint32_t* bnext;    // the next integration's slot
  int bcap;           // bricks per partition
  int gx, gy, gz;     // groups per axis
  int any_side;       // side tests without the front requirement (see integ_corner)
};

// Consumers of a partitioned list: wave w of the grid serves partition
// w % KF_RAY_PARTS and takes its entries w / KF_RAY_PARTS + k * (waves /
// KF_RAY_PARTS), reading one counter.

// Per-pixel part: for a pixel (u, v) with depth d != 0, the ray magnitude
// |T(2 lambda) - T(lambda)|, the predicted depth |T(d lambda) - t| / |.| and
// its reciprocal depend only on the pixel, so they are computed here once
// with the reference's expressions (:39-52, :78-94).  Depth 0 stores NaN.
// Thread 0 also stores the front-test direction (:57-71), which depends only
// on K and T.
//
// A 256-thread block is a 16x16 pixel tile.  It also stores the tile's band
// bounds: a voxel at camera distance tm projecting to pixel p is updated only
// when sdf_min <= frk = tab.x - tab.y tm <= sdf_max, i.e. (tab.y > 0) when
// tm lies in [(tab.x - sdf_max) / tab.y, (tab.x - sdf_min) / tab.y].  The
// tile keeps the smallest lower and the largest upper end over its pixels
// with depth (FP64, then rounded outwards to float); a tile without depth
// stores {+inf, -inf}.  Integration culls bricks against them.
__global__ __launch_bounds__(256) void k_integ_table(const float* __restrict__ depth, kf_cam cam,
                                                     int H, int W, kf_dev_state* __restrict__ st,
                                                     float2* __restrict__ tab,
                                                     float2* __restrict__ tiles, float smin,
                                                     float smax, int tiles_y, kf_vol vg,
                                                     GroupList gl) {
  float T[12];
  load_uniform<12>(st->T, T);
  const float* Ki = cam.Kinv;
  const float* K = cam.K;
  if (blockIdx.y >= (unsigned)tiles_y) {
    // group cull: lane = (group of the block, corner), 32 groups per block;
    // the surviving groups are appended to the group list
    const int gb = (blockIdx.y - tiles_y) * gridDim.x + blockIdx.x;
    const int lane = threadIdx.x & 63, corner = lane & 7;
    const int g0 = gb * 32 + (int)(threadIdx.x >> 6) * 8;  // this wave's first group
    const int g = g0 + (lane >> 3);
    bool r1 = true, r2 = true, r3 = true, r4 = true, r5 = true;
    if (g < gl.gx * gl.gy * gl.gz) {
      float Ti[12], t[3], cd[3];
      load_uniform<12>(st->Tinv, Ti);
      t[0] = T[3], t[1] = T[7], t[2] = T[11];
      integ_center_dir(K, Ki, T, cd);
      const int gz = g % gl.gz, q = g / gl.gz;
      const int gy = q % gl.gy, gx = q / gl.gy;
      const int x = min(gx * 16 + ((corner & 4) ? 15 : 0), vg.nx - 1);
      const int y = min(gy * 16 + ((corner & 2) ? 15 : 0), vg.ny - 1);
      const int z = min(gz * 16 + ((corner & 1) ? 15 : 0), vg.nz - 1);
      const BoxCorner c = integ_corner(x, y, z, vg, Ti, t, cd, K, H, W, gl.any_side != 0);
      r1 = c.r1, r2 = c.r2, r3 = c.r3, r4 = c.r4, r5 = c.r5;
    }
    const uint32_t keep = wave_box_keep(r1, r2, r3, r4, r5, false);
    if (keep) {
      const int part = gb & (KF_RAY_PARTS - 1);
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&gl.gcount[part * KF_RAY_COUNT_STRIDE], (uint32_t)__popc(keep));
      base = __shfl(base, 0);
      if (lane < 8 && ((keep >> lane) & 1u))
        gl.glist[part * gl.gcap + base + __popc(keep & ((1u << lane) - 1u))] = g0 + lane;
    }
    return;
  }
  const int px = blockIdx.x * 16 + (threadIdx.x & 15), py = blockIdx.y * 16 + (threadIdx.x >> 4);
  const bool inside = px < W && py < H;
  const int i = inside ? py * W + px : -1;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    float cd[3];
    integ_center_dir(K, Ki, T, cd);
    st->center_dir[0] = cd[0];
    st->center_dir[1] = cd[1];
    st->center_dir[2] = cd[2];
    st->center_dir[3] = 0.f;
  }
  double lo = __builtin_inf(), hi = -__builtin_inf();
  if (inside) {
    const float dv = depth[i];
    float2 out = make_float2(__builtin_nanf(""), __builtin_nanf(""));
    if (dv != 0) {
      const int y = i / W, x = i - y * W;
      const float U = (float)(x + 1), V = (float)(y + 1);
      const float lamx = Ki[0] * U + Ki[1] * V + Ki[2] * 1.f;
      const float lamy = Ki[3] * U + Ki[4] * V + Ki[5] * 1.f;
      const float lamz = Ki[6] * U + Ki[7] * V + Ki[8] * 1.f;
      const float rsx = T[0] * lamx + T[1] * lamy + T[2] * lamz + T[3];
      const float rsy = T[4] * lamx + T[5] * lamy + T[6] * lamz + T[7];
      const float rsz = T[8] * lamx + T[9] * lamy + T[10] * lamz + T[11];
      const float rdx = T[0] * 2 * lamx + T[1] * 2 * lamy + T[2] * 2 * lamz + T[3];
      const float rdy = T[4] * 2 * lamx + T[5] * 2 * lamy + T[6] * 2 * lamz + T[7];
      const float rdz = T[8] * 2 * lamx + T[9] * 2 * lamy + T[10] * 2 * lamz + T[11];
      const float dx = rdx - rsx, dy = rdy - rsy, dz = rdz - rsz;
      const float rm = sqrtf(dx * dx + dy * dy + dz * dz);
      const float p3x = dv * lamx, p3y = dv * lamy, p3z = dv * lamz;
      const float pgx = T[0] * p3x + T[1] * p3y + T[2] * p3z + T[3];
      const float pgy = T[4] * p3x + T[5] * p3y + T[6] * p3z + T[7];
      const float pgz = T[8] * p3x + T[9] * p3y + T[10] * p3z + T[11];
      const float rx = pgx - T[3], ry = pgy - T[7], rz = pgz - T[11];
      const float rpd = sqrtf(rx * rx + ry * ry + rz * rz);
      out.x = rpd / rm;
      out.y = 1.f / rm;
      lo = ((double)out.x - (double)smax) / (double)out.y;
      hi = ((double)out.x - (double)smin) / (double)out.y;
    }
    tab[i] = out;
  }
  // tile bounds: wave min / max, then over the 4 waves
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o));
    hi = fmax(hi, __shfl_xor(hi, o));
  }
  __shared__ double s_lo[4], s_hi[4];
  if ((threadIdx.x & 63) == 0) {
    s_lo[threadIdx.x >> 6] = lo;
    s_hi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double l = fmin(fmin(s_lo[0], s_lo[1]), fmin(s_lo[2], s_lo[3]));
    const double h = fmax(fmax(s_hi[0], s_hi[1]), fmax(s_hi[2], s_hi[3]));
    tiles[blockIdx.y * gridDim.x + blockIdx.x] =
        make_float2(__double2float_rd(l), __double2float_ru(h));
  }
}

// Per-voxel part, on the bricks (8x8x8 voxels) that survive three culls.
//
// * Box culls.  Lanes evaluate, in FP64, the affine quantities the
//   per-voxel test depends on at the 8 corner voxel centres of a box: the
//   projection p = K T^-1 c and the front-test dot mag = dir . (c - t).  If
//   one of "mag < -M", "px < 0.45 pz - M", "px > (W + 0.55) pz + M" or the
//   same two for py/H holds at all 8 corners, it holds on the whole box
//   (integ_corner), and the margin M (16 + 3e-5 |K| |l|, many times the
//   rounding error of the float evaluation) guarantees that the reference's
//   float test rejects every voxel of the box.  k_integ_table applies these
//   to groups of 2x2x2 bricks, then to the bricks of the surviving groups,
//   and lists the surviving bricks; no decision changes.
// * Band cull (here, per listed brick): the brick is skipped when its
//   distance range to the camera misses the tile intervals of updatable
//   distances under its footprint (k_integ_table).
// * One listed brick per wave, so that no wave updates several bricks in a
//   row.  Lane (lx, ly) walks its z column of 8 voxels.  The reference sums
//   T^-1 c left to right, so Ti0*cx + Ti1*cy (and the x/y parts of the front
//   test and of |c - t|^2) are the same for the column and are computed once,
//   bit-identically.
// * Division pre-check.  When pz > 0 and px < 0.25 pz (or > (W+1) pz, and
//   the same for y, mirrored for pz < 0), the rounded quotient cannot land in
//   [1, W]; the voxel is rejected before the two correctly rounded divisions.
// * A brick costs one memory round trip: each lane loads its column's 8
//   TSDF and weight values up front (vector loads; bricks are padded), then
//   computes the 8 projections and issues the 8 table gathers together.
// A voxel is written only when its signed distance falls in [sdf_min,
// sdf_max], as in the reference.

// Updates one brick with the calling wave: lane (lx, ly) walks its z column.
// Returns whether this lane stored a negative TSDF value (NEG; else false).
template <int PREC, bool FIRST, bool NEG>
__device__ __forceinline__ bool integrate_brick(void* __restrict__ tsdf_, void* __restrict__ weight_,
                                                const float2* __restrict__ tab, const float* K,
                                                const kf_vol& vg, int H, int W, float trunc,
                                                float smin, float smax, const float* Ti,
                                                const float* t, const float* cd, int brick) {
  const int lane = threadIdx.x & 63;
  const float d = vg.d;
  const float hd = d / 2.f;
  const float Wp1 = (float)(W + 1), Hp1 = (float)(H + 1);
  bool wneg = false;
  do {
    const int bz = brick % vg.bz;
    const int q = brick / vg.bz;
    const int by = q % vg.by;
    const int bx = q / vg.by;
    const int ix = bx * 8 + (lane >> 3), iy = by * 8 + (lane & 7);
    if (ix >= vg.nx || iy >= vg.ny) continue;
    const float cx = (float)(ix + vg.sx) * d + hd;
    const float cy = (float)(iy + vg.sy) * d + hd;
    const float Px = Ti[0] * cx + Ti[1] * cy;
    const float Py = Ti[4] * cx + Ti[5] * cy;
    const float Pz = Ti[8] * cx + Ti[9] * cy;
    const float tdx = cx - t[0], tdy = cy - t[1];
    const float Pm = cd[0] * tdx + cd[1] * tdy;
    const float Pt = tdx * tdx + tdy * tdy;
    const int base = (brick << 9) | (lane << 3);
    const int zn = min(8, vg.nz - bz * 8);
    // The column's 8 stored values (bricks are padded, so all 8 exist), loaded
    // before anything else so that the brick costs one memory round trip:
    // projections first, then the 8 table gathers together.
    float F0[8], W0[8];  // F32 weight
    uint32_t Wc[8];      // u16 observation count (F16 volume)
    if constexpr (!FIRST) {
      if constexpr (PREC == KF_TSDF_F32) {
        const float4* T4 = reinterpret_cast<const float4*>(reinterpret_cast<const float*>(tsdf_) + base);
        const float4* W4 = reinterpret_cast<const float4*>(reinterpret_cast<const float*>(weight_) + base);
        const float4 a = T4[0], b = T4[1], c = W4[0], e = W4[1];
        F0[0] = a.x, F0[1] = a.y, F0[2] = a.z, F0[3] = a.w;
        F0[4] = b.x, F0[5] = b.y, F0[6] = b.z, F0[7] = b.w;
        W0[0] = c.x, W0[1] = c.y, W0[2] = c.z, W0[3] = c.w;
        W0[4] = e.x, W0[5] = e.y, W0[6] = e.z, W0[7] = e.w;
      } else {
        const uint4 a = *reinterpret_cast<const uint4*>(reinterpret_cast<const uint16_t*>(tsdf_) + base);
        const uint4 c = *reinterpret_cast<const uint4*>(reinterpret_cast<const uint16_t*>(weight_) + base);
        const uint32_t ta[4] = {a.x, a.y, a.z, a.w}, wa[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const uint32_t h = (ta[k >> 1] >> (16 * (k & 1))) & 0xffffu;
          F0[k] = __half2float(__ushort_as_half((uint16_t)h));
          Wc[k] = (wa[k >> 1] >> (16 * (k & 1))) & 0xffffu;
        }
      }
    }
    int tix[8];
#pragma unroll
    for (int lz = 0; lz < 8; ++lz) {
      tix[lz] = -1;
      if (lz >= zn) continue;
      const float cz = (float)(bz * 8 + lz + vg.sz) * d + hd;
      const float lx = (Px + Ti[2] * cz) + Ti[3];
      const float ly = (Py + Ti[6] * cz) + Ti[7];
      const float lzz = (Pz + Ti[10] * cz) + Ti[11];
      const float px = K[0] * lx + K[1] * ly + K[2] * lzz;
      const float py = K[3] * lx + K[4] * ly + K[5] * lzz;
      const float pz = K[6] * lx + K[7] * ly + K[8] * lzz;
      if (pz > 0.f) {
        if (px < 0.25f * pz || px > Wp1 * pz || py < 0.25f * pz || py > Hp1 * pz) continue;
      } else if (pz < 0.f) {
        if (px > 0.25f * pz || px < Wp1 * pz || py > 0.25f * pz || py < Hp1 * pz) continue;
      }
      const int u = kf_round_to_int(px / pz);
      const int pv = kf_round_to_int(py / pz);
      if (!(u > 0 && u <= W && pv > 0 && pv <= H)) continue;
      tix[lz] = (pv - 1) * W + (u - 1);
    }
    float2 tb[8];
#pragma unroll
    for (int lz = 0; lz < 8; ++lz) tb[lz] = tab[max(tix[lz], 0)];
#pragma unroll
    for (int lz = 0; lz < 8; ++lz) {
      if (tix[lz] < 0) continue;
      if (tb[lz].x != tb[lz].x) continue;  // depth 0 at that pixel
      const float cz = (float)(bz * 8 + lz + vg.sz) * d + hd;
      const float tdz = cz - t[2];
      const float mag = Pm + cd[2] * tdz;
      if (!(mag > 0)) continue;
      const float tm = sqrtf(Pt + tdz * tdz);
      const float frk = tb[lz].x - (tb[lz].y * tm);
      const bool pos = frk > 0 && frk <= smax;
      const bool neg = frk <= 0 && frk >= smin;
      if (!pos && !neg) continue;
      const float f = pos ? fminf(1.f, (frk / trunc)) : fmaxf(-1.f, frk / trunc);
      const int v = base + lz;
      if constexpr (PREC == KF_TSDF_F32) {
        float* T = reinterpret_cast<float*>(tsdf_);
        float* Wt = reinterpret_cast<float*>(weight_);
        if constexpr (FIRST) {
          T[v] = f;
          Wt[v] = 1.0f;
          if constexpr (NEG) wneg = wneg || f < 0.f;
        } else {
          float Fo = F0[lz], Wo = W0[lz];
          if (Fo != Fo) Fo = 0.0f;
          if (Wo != Wo) Wo = 0.0f;
          const float Fn = (Wo * Fo + 1.0f * f) / (Wo + 1.0f);
          T[v] = Fn;
          Wt[v] = Wo + 1.0f;
          if constexpr (NEG) wneg = wneg || Fn < 0.f;
        }
      } else {
        uint16_t* T = reinterpret_cast<uint16_t*>(tsdf_);
        uint16_t* Wt = reinterpret_cast<uint16_t*>(weight_);
        if constexpr (FIRST) {
          const __half h = __float2half_rn(f);
          T[v] = __half_as_ushort(h);
          Wt[v] = 1;
          if constexpr (NEG) wneg = wneg || __half2float(h) < 0.f;
        } else {
          float Fo = F0[lz];
          if (Fo != Fo) Fo = 0.0f;
          const uint32_t w = Wc[lz];
          const float Wo = (float)w;
          const float Fn = (Wo * Fo + 1.0f * f) / (Wo + 1.0f);
          const __half h = __float2half_rn(Fn);
          T[v] = __half_as_ushort(h);
          Wt[v] = (uint16_t)(w == 65535u ? 65535u : w + 1u);
          if constexpr (NEG) wneg = wneg || __half2float(h) < 0.f;
        }
      }
    }
  } while (0);
  return wneg;
}

// Skip flags of the second raycast pass, one byte per brick: a brick's flag
// is set when it or one of its 26 neighbours holds (or once held) a negative
// TSDF value.  Flags are only ever set between volume resets, so they stay a
// superset of that property as integration changes the volume.  Lane q < 27
// of the calling wave sets the flag of neighbour (q / 9 - 1, q / 3 % 3 - 1,
// q % 3 - 1) of brick (bx, by, bz).
__device__ __forceinline__ void set_skip_flags(uint8_t* __restrict__ flags, const kf_vol& vg,
                                               int bx, int by, int bz) {
  const int q = threadIdx.x & 63;
  const int x = bx + q / 9 - 1, y = by + (q / 3) % 3 - 1, z = bz + q % 3 - 1;
  if (q < 27 && x >= 0 && x < vg.bx && y >= 0 && y < vg.by && z >= 0 && z < vg.bz)
    flags[(x * vg.by + y) * vg.bz + z] = 1;
}

// Rebuilds the skip flags from the whole volume (after a volume is uploaded
// or written through its device address): one brick per wave, lane l reads
// voxels 8l .. 8l + 7 of the brick.  The flags must be all zero before.
template <int PREC>
__global__ __launch_bounds__(256) void k_skip_flags(const void* __restrict__ tsdf,
                                                    uint8_t* __restrict__ flags, kf_vol vg,
                                                    int n_bricks) {
  const int brick = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (brick >= n_bricks) return;
  const int lane = threadIdx.x & 63;
  const int base = (brick << 9) | (lane << 3);
  bool neg = false;
  if constexpr (PREC == KF_TSDF_F32) {
    const float4* T4 = reinterpret_cast<const float4*>(reinterpret_cast<const float*>(tsdf) + base);
    const float4 a = T4[0], b = T4[1];
    neg = a.x < 0.f || a.y < 0.f || a.z < 0.f || a.w < 0.f || b.x < 0.f || b.y < 0.f ||
          b.z < 0.f || b.w < 0.f;
  } else {
    const uint4 a = *reinterpret_cast<const uint4*>(reinterpret_cast<const uint16_t*>(tsdf) + base);
    const uint32_t ta[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t h = (ta[k >> 1] >> (16 * (k & 1))) & 0xffffu;
      neg = neg || __half2float(__ushort_as_half((uint16_t)h)) < 0.f;
    }
  }
  if (__ballot(neg)) {
    const int bz = brick % vg.bz, q = brick / vg.bz;
    set_skip_flags(flags, vg, q / vg.by, q % vg.by, bz);
  }
}

__global__ __launch_bounds__(256) void k_fill_u8(uint8_t* __restrict__ p, int n, uint8_t v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// Brick cull: one listed group per wave, lane = (brick of the group,
// corner): the frustum tests and the band cull; the surviving bricks are
// appended to the brick list (block b to partition b % KF_RAY_PARTS).
__global__ __launch_bounds__(256) void k_integ_cull(kf_cam cam, kf_vol vg, int H, int W,
                                                    const kf_dev_state* __restrict__ st,
                                                    GroupList gl,
                                                    const float2* __restrict__ tiles,
                                                    int tiles_x) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float Ti[12], t[3], cd[3];
  load_uniform<12>(st->Tinv, Ti);
  t[0] = st->T[3];
  t[1] = st->T[7];
  t[2] = st->T[11];
  load_uniform<3>(st->center_dir, cd);
  const float* K = cam.K;
  const int wid = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave);
  const int gpart = wid & (KF_RAY_PARTS - 1);
  const uint32_t n_groups = __builtin_amdgcn_readfirstlane(gl.gcount[gpart * KF_RAY_COUNT_STRIDE]);
  // the next integration appends to the other slot, which nothing reads now
  if (blockIdx.x == 0 && threadIdx.x < KF_RAY_PARTS) gl.gnext[lane * KF_RAY_COUNT_STRIDE] = 0u;
  const uint32_t stride = gridDim.x * 4 / KF_RAY_PARTS;
  const int corner = lane & 7, b = lane >> 3;
  const int part = blockIdx.x & (KF_RAY_PARTS - 1);
  for (uint32_t e = (uint32_t)wid / KF_RAY_PARTS; e < n_groups; e += stride) {
    const int g = __builtin_amdgcn_readfirstlane(gl.glist[gpart * gl.gcap + e]);
    const int gz = g % gl.gz, q = g / gl.gz;
    const int gy = q % gl.gy, gx = q / gl.gy;
    const int bx = gx * 2 + (b >> 2), by = gy * 2 + ((b >> 1) & 1), bz = gz * 2 + (b & 1);
    const bool inside = bx < vg.bx && by < vg.by && bz < vg.bz;
    bool r1 = true, r2 = true, r3 = true, r4 = true, r5 = true, r6 = false;
    if (inside) {
      const int x = min(bx * 8 + ((corner & 4) ? 7 : 0), vg.nx - 1);
      const int y = min(by * 8 + ((corner & 2) ? 7 : 0), vg.ny - 1);
      const int z = min(bz * 8 + ((corner & 1) ? 7 : 0), vg.nz - 1);
      const BoxCorner c = integ_corner(x, y, z, vg, Ti, t, cd, K, H, W, gl.any_side != 0);
      r1 = c.r1, r2 = c.r2, r3 = c.r3, r4 = c.r4, r5 = c.r5;
      const bool front = c.front;
      const double px = c.px, py = c.py, pz = c.pz, tx = c.tx, ty = c.ty, tz = c.tz;
      const double cx = c.cx, cy = c.cy, cz = c.cz;
      // band cull: every voxel of the brick is nearer than the updatable
      // distances of every pixel it can project to, or farther
      double u = front ? px / pz : 0.0, v = front ? py / pz : 0.0;
      double umin = u, umax = u, vmin = v, vmax = v;
      double dmax = sqrt(tx * tx + ty * ty + tz * tz);
      double xmin = cx, xmax = cx, ymin = cy, ymax = cy, zmin = cz, zmax = cz;
      int all_front = front ? 1 : 0;
#pragma unroll
      for (int o = 1; o < 8; o <<= 1) {
        umin = fmin(umin, __shfl_xor(umin, o, 8));
        umax = fmax(umax, __shfl_xor(umax, o, 8));
        vmin = fmin(vmin, __shfl_xor(vmin, o, 8));
        vmax = fmax(vmax, __shfl_xor(vmax, o, 8));
        dmax = fmax(dmax, __shfl_xor(dmax, o, 8));
        xmin = fmin(xmin, __shfl_xor(xmin, o, 8));
        xmax = fmax(xmax, __shfl_xor(xmax, o, 8));
        ymin = fmin(ymin, __shfl_xor(ymin, o, 8));
        ymax = fmax(ymax, __shfl_xor(ymax, o, 8));
        zmin = fmin(zmin, __shfl_xor(zmin, o, 8));
        zmax = fmax(zmax, __shfl_xor(zmax, o, 8));
        all_front = all_front & __shfl_xor(all_front, o, 8);
      }
      if (all_front) {
        // footprint: the voxels project inside the hull of the corner
        // projections; one pixel more on each side covers roundf and the
        // float evaluation
        const int u0 = (int)fmax(floor(umin) - 1.0, 1.0), u1 = (int)fmin(ceil(umax) + 1.0, (double)W);
        const int v0 = (int)fmax(floor(vmin) - 1.0, 1.0), v1 = (int)fmin(ceil(vmax) + 1.0, (double)H);
        if (u0 <= u1 && v0 <= v1) {
          const int tx0 = (u0 - 1) >> 4, tx1 = (u1 - 1) >> 4;
          const int ty0 = (v0 - 1) >> 4, ty1 = (v1 - 1) >> 4;
          const int ntx = tx1 - tx0 + 1, nt = ntx * (ty1 - ty0 + 1);
          if (nt <= 64) {
            float lo = __builtin_inff(), hi = -__builtin_inff();
            for (int k = corner; k < nt; k += 8) {
              const float2 bb = tiles[(ty0 + k / ntx) * tiles_x + tx0 + k % ntx];
              lo = fminf(lo, bb.x);
              hi = fmaxf(hi, bb.y);
            }
#pragma unroll
            for (int o = 1; o < 8; o <<= 1) {
              lo = fminf(lo, __shfl_xor(lo, o, 8));
              hi = fmaxf(hi, __shfl_xor(hi, o, 8));
            }
            // nearest voxel centre of the brick
            const double ex = fmax(fmax(xmin - (double)t[0], (double)t[0] - xmax), 0.0);
            const double ey = fmax(fmax(ymin - (double)t[1], (double)t[1] - ymax), 0.0);
            const double ez = fmax(fmax(zmin - (double)t[2], (double)t[2] - zmax), 0.0);
            const double dmin = sqrt(ex * ex + ey * ey + ez * ez);
            const double margin = 1.0 + 1e-5 * dmax;
            r6 = dmax + margin < (double)lo || dmin - margin > (double)hi;
          }
        }
      }
    }
    const uint32_t keep = wave_box_keep(r1, r2, r3, r4, r5, r6);
    if (keep) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&gl.bcount[part * KF_RAY_COUNT_STRIDE], (uint32_t)__popc(keep));
      base = __shfl(base, 0);
      if (lane < 8 && ((keep >> lane) & 1u)) {
        const int kx = gx * 2 + (lane >> 2), ky = gy * 2 + ((lane >> 1) & 1), kz = gz * 2 + (lane & 1);
        gl.blist[part * gl.bcap + base + __popc(keep & ((1u << lane) - 1u))] =
            (kx * vg.by + ky) * vg.bz + kz;
      }
    }
  }
}

// Brick update: one listed brick per wave; with FLAGS, a brick that received
// a negative value sets the skip flags around it.
template <int PREC, bool FIRST, bool FLAGS>
__global__ __launch_bounds__(256) void k_integrate(
    void* __restrict__ tsdf_, void* __restrict__ weight_, const float2* __restrict__ tab,
    kf_cam cam, kf_vol vg, int H, int W, float trunc, float smin, float smax,
    const kf_dev_state* __restrict__ st, GroupList gl, uint8_t* __restrict__ skip_flags) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float Ti[12], t[3], cd[3];
  load_uniform<12>(st->Tinv, Ti);
  t[0] = st->T[3];
  t[1] = st->T[7];
  t[2] = st->T[11];
  load_uniform<3>(st->center_dir, cd);
  const int wid = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave);
  const int part = wid & (KF_RAY_PARTS - 1);
  const uint32_t n_bricks = __builtin_amdgcn_readfirstlane(gl.bcount[part * KF_RAY_COUNT_STRIDE]);
  if (blockIdx.x == 0 && threadIdx.x < KF_RAY_PARTS) gl.bnext[lane * KF_RAY_COUNT_STRIDE] = 0u;
  const uint32_t stride = gridDim.x * 4 / KF_RAY_PARTS;
  for (uint32_t e = (uint32_t)wid / KF_RAY_PARTS; e < n_bricks; e += stride) {
    const int brick = __builtin_amdgcn_readfirstlane(gl.blist[part * gl.bcap + e]);
    const bool neg = integrate_brick<PREC, FIRST, FLAGS>(tsdf_, weight_, tab, cam.K, vg, H, W,
                                                         trunc, smin, smax, Ti, t, cd, brick);
    if (FLAGS && __ballot(neg)) {
      const int bz = brick % vg.bz, q = brick / vg.bz;
      set_skip_flags(skip_flags, vg, q / vg.by, q % vg.by, bz);
    }
  }
}

// Never-observed fill: F32 TSDF and weight NaN; F16 TSDF half NaN, weight 0.
__global__ __launch_bounds__(256) void k_volume_fill(void* tsdf, void* weight, int prec,
                                                     int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if (prec == KF_TSDF_F32) {
    reinterpret_cast<float*>(tsdf)[i] = __builtin_nanf("");
    reinterpret_cast<float*>(weight)[i] = __builtin_nanf("");
  } else {
    reinterpret_cast<uint16_t*>(tsdf)[i] = 0x7e00u;
    reinterpret_cast<uint16_t*>(weight)[i] = 0u;
  }
}

// ---------------------------------------------------------------------------
// A9  voxel-traversal raycast (voxel_traversal.cu:75-424), one ray per lane.
//
// A 256-thread block covers 16x16 pixels as four 8x8 wave tiles, so a wave's
// rays walk neighbouring voxels.  Consecutive block ids are dealt round-robin
// to the 8 XCDs; the id is remapped so that each XCD gets a contiguous band
// of the image and its L2 sees the band's voxels.
// Traversal state of voxel_traversal.cu:133-205: the 1-based voxel, the
// indices recorded at the last step on each axis (XY, XZ: y and z at the
// last x step, ...) and the parametric distances to the next boundaries.
struct DdaState {
  int ix, iy, iz, XY, XZ, YX, YZ, ZX, ZY;
  float tmx, tmy, tmz;
};
struct DdaStep {
  int sx, sy, sz;
  float tdx, tdy, tdz;
};

// One Amanatides-Woo step with the reference's tie order (:384-417):
//   tmx < tmy ? (tmx < tmz ? x : z) : (tmy < tmz ? y : z),
// written with selects so that a wave steps its 64 rays without branching.
// NaN comparisons are false in both forms, so the choice is the same.
__device__ __forceinline__ void dda_step(DdaState& s, const DdaStep& p) {
  const bool xy = s.tmx < s.tmy;
  const bool xm = xy && s.tmx < s.tmz;
  const bool ym = !xy && s.tmy < s.tmz;
  const bool zm = !xm && !ym;
  const int ix = s.ix, iy = s.iy, iz = s.iz;
  const float nx = s.tmx + p.tdx, ny = s.tmy + p.tdy, nz = s.tmz + p.tdz;
  s.ix = xm ? ix + p.sx : ix;
  s.XY = xm ? iy : s.XY;
  s.XZ = xm ? iz : s.XZ;
  s.tmx = xm ? nx : s.tmx;
  s.iy = ym ? iy + p.sy : iy;
  s.YX = ym ? ix : s.YX;
  s.YZ = ym ? iz : s.YZ;
  s.tmy = ym ? ny : s.tmy;
  s.iz = zm ? iz + p.sz : iz;
  s.ZX = zm ? ix : s.ZX;
  s.ZY = zm ? iy : s.ZY;
  s.tmz = zm ? nz : s.tmz;
}

// Zero crossing found at voxel h (voxel_traversal.cu:222-377): bound checks,
// 13-tap NaN checks, per-axis interpolation, central-difference normal and
// the flip towards the camera.  Returns the 0-based voxel index, or -1 when a
// check rejects the crossing.
template <int PREC>
__device__ __forceinline__ int crossing(const VolRead<PREC>& F, const DdaState& h, float f,
                                        const kf_vol& vg, int fix_bounds, float camx,
                                        float camy, float camz, float dx, float dy, float dz,
                                        float flip_hi, float4& ov, float4& on) {
  const int Nx = vg.nx, Ny = vg.ny, Nz = vg.nz;
  const int NyNz = Ny * Nz;
  const float d = vg.d;
  const int bXY = fix_bounds ? Ny : Nx, bXZ = fix_bounds ? Nz : Nx;
  const int bYX = fix_bounds ? Nx : Ny, bYZ = fix_bounds ? Nz : Ny;
  const int bZX = fix_bounds ? Nx : Nz, bZY = fix_bounds ? Ny : Nz;
  const int ix = h.ix, iy = h.iy, iz = h.iz;
  const int XY = h.XY, XZ = h.XZ, YX = h.YX, YZ = h.YZ, ZX = h.ZX, ZY = h.ZY;
  const int vi = (ix - 1) * NyNz + (iy - 1) * Nz + (iz - 1);
  const int xn = (ix - 1) - 1, xp = (ix - 1) + 1;
  const int yn = (iy - 1) - 1, yp = (iy - 1) + 1;
  const int zn = (iz - 1) - 1, zp = (iz - 1) + 1;
  if (xn >= Nx || xn < 0 || xp >= Nx || xp < 0) return -1;
  if ((XY - 1) >= bXY || (XY - 1) < 0 || (XZ - 1) >= bXZ || (XZ - 1) < 0) return -1;
  if (yn >= Ny || yn < 0 || yp >= Ny || yp < 0) return -1;
  if ((YX - 1) >= bYX || (YX - 1) < 0 || (YZ - 1) >= bYZ || (YZ - 1) < 0) return -1;
  if (zn >= Nz || zn < 0 || zp >= Nz || zp < 0) return -1;
  if ((ZX - 1) >= bZX || (ZX - 1) < 0 || (ZY - 1) >= bZY || (ZY - 1) < 0) return -1;
  const float f_xnp = F(kf_store_index(vg, xn, XY - 1, XZ - 1));
  const float f_xpp = F(kf_store_index(vg, xp, XY - 1, XZ - 1));
  const float f_ynp = F(kf_store_index(vg, YX - 1, yn, YZ - 1));
  const float f_ypp = F(kf_store_index(vg, YX - 1, yp, YZ - 1));
  const float f_znp = F(kf_store_index(vg, ZX - 1, ZY - 1, zn));
  const float f_zpp = F(kf_store_index(vg, ZX - 1, ZY - 1, zp));
  const float f_xnn = F(kf_store_index(vg, ix - 2, iy - 1, iz - 1));
  const float f_xpn = F(kf_store_index(vg, ix, iy - 1, iz - 1));
  const float f_ynn = F(kf_store_index(vg, ix - 1, iy - 2, iz - 1));
  const float f_ypn = F(kf_store_index(vg, ix - 1, iy, iz - 1));
  const float f_znn = F(kf_store_index(vg, ix - 1, iy - 1, iz - 2));
  const float f_zpn = F(kf_store_index(vg, ix - 1, iy - 1, iz));
  if (f_xnp != f_xnp || f_xpp != f_xpp || f_ynp != f_ynp || f_ypp != f_ypp || f_znp != f_znp ||
      f_zpp != f_zpp || f != f)
    return -1;
  if (f_xnn != f_xnn || f_xpn != f_xpn || f_ynn != f_ynn || f_ypn != f_ypn || f_znn != f_znn ||
      f_zpn != f_zpn)
    return -1;
  const float hd = d / 2.f;
  float spx, spy, spz;
  if (dx > 0) {
    const float g = (float)(xn + vg.sx) * d + hd;
    spx = g + d * (fabsf(f_xnp) / (fabsf(f_xnp - f)));
  } else {
    const float g = (float)(xp + vg.sx) * d + hd;
    spx = g - d * (fabsf(f_xpp) / (fabsf(f_xpp - f)));
  }
  if (dy > 0) {
    const float g = (float)(yn + vg.sy) * d + hd;
    spy = g + d * (fabsf(f_ynp) / (fabsf(f_ynp - f)));
  } else {
    const float g = (float)(yp + vg.sy) * d + hd;
    spy = g - d * (fabsf(f_ypp) / (fabsf(f_ypp - f)));
  }
  if (dz > 0) {
    const float g = (float)(zn + vg.sz) * d + hd;
    spz = g + d * (fabsf(f_znp) / (fabsf(f_znp - f)));
  } else {
    const float g = (float)(zp + vg.sz) * d + hd;
    spz = g - d * (fabsf(f_zpp) / (fabsf(f_zpp - f)));
  }
  float bxn = (f_xpn - f_xnn) / 2.0f;
  float byn = (f_ypn - f_ynn) / 2.0f;
  float bzn = (f_zpn - f_znn) / 2.0f;
  float ax = camx - spx, ay = camy - spy, az = camz - spz;
  normalize3(ax, ay, az);
  normalize3(bxn, byn, bzn);
  const float c = ax * bxn + ay * byn + az * bzn;
  if (c < flip_hi && c >= -1.0f) {
    bxn = -bxn;
    byn = -byn;
    bzn = -bzn;
  }
  ov = make_float4(spx, spy, spz, 0.f);
  on = make_float4(bxn, byn, bzn, 0.f);
  return vi;
}

// A ray's traversal state between the two raycast passes.
struct RayState {
  DdaState cur;
  float prev;
  float dx, dy, dz;
  DdaStep stp;
  int pix;
  float tent;  // ray parameter where the walk entered cur (trilinear modes)
};

// Launch constants of the raycast kernels.
struct RayIO {
  kf_vol vg;
  int fix_bounds;
  float flip_hi;
  float trunc;
  int W;
  float Ki[9];
  const float* depth;
  // adjoint inputs/outputs (MODE 2)
  const float* gD;     // dL/d(depth), H*W
  const float4* gN;    // dL/d(normal), H*W (nullable)
  float* gF;           // dL/d(TSDF), native store order (null: not wanted)
  uint32_t* touched;   // bricks that receive dL/dF (one bit each)
  double* xi_part;     // dL/dxi partials, 6 per block
  int part_base;       // first partial slot of this launch
  float* gn_part;      // refinement partials, kGnSums per block
  float max_res;       // refinement: residuals beyond this (mm) are ignored
  // Refinement: a walk whose voxel-entry parameter (mm from the skip point)
  // exceeds this can only find crossings whose residual exceeds max_res, so
  // it stops there (+inf in the other modes).
  float walk_tmax;
  int refine_batches;  // refinement: first-pass batches (enough to reach walk_tmax)
  int pass1_batches;   // first-pass batches of the other modes (multi-batch kernels)
  uint32_t* qcount;    // this raycast's partition append counters
  uint32_t* qnext;     // the next raycast's counters (cleared by the second pass)
  int qcap;            // rays per queue partition
  const uint8_t* skip_flags;  // per-brick skip flags of the second pass (null: no skipping)
};

// Raycast modes: 0 = the reference's hit handling (voxel_traversal.cu:222-
// 377); 1 = trilinear zero-crossing refinement (vertex, normal, depth);
// 2 = adjoint of mode 1.
constexpr int kModeReference = 0;
constexpr int kModeTrilinear = 1;
constexpr int kModeAdjoint = 2;
// 3 = Gauss-Newton pose refinement on the depth residual of mode 1: per
// pixel J = d(depth)/d(xi) and r = depth - observed depth.
constexpr int kModeRefine = 3;
constexpr int kGnSums = 29;  // 21 J^TJ (upper, row-major), 6 J^Tr, count, r^2

// World-frame unit direction of the ray through the camera-frame point
// (lamx, lamy, lamz) = K^-1 (x + 1, y + 1, 1) (voxel_traversal.cu:90-200).
__device__ __forceinline__ void ray_direction(float lamx, float lamy, float lamz, const float* T,
                                              float& dx, float& dy, float& dz) {
  const float camx = T[3], camy = T[7], camz = T[11];
  const float rgx = T[0] * lamx + T[1] * lamy + T[2] * lamz + T[3] * 1.f;
  const float rgy = T[4] * lamx + T[5] * lamy + T[6] * lamz + T[7] * 1.f;
  const float rgz = T[8] * lamx + T[9] * lamy + T[10] * lamz + T[11] * 1.f;
  dx = rgx - camx, dy = rgy - camy, dz = rgz - camz;
  normalize3(dx, dy, dz);
}

// Ray setup (voxel_traversal.cu:90-200): ray skipping to (depth - trunc),
// slab test, 1-based start voxel, steps and boundary distances.  Returns
// false when the ray never enters the traversal.
__device__ __forceinline__ bool ray_setup(int x, int y, float dval, const float* Ki,
                                          const float* T, const kf_vol& vg, float trunc,
                                          RayState& r) {
  const int Nx = vg.nx, Ny = vg.ny, Nz = vg.nz;
  const float d = vg.d;
  const float X = (float)x + 1.0f, Y = (float)y + 1.0f;
  const float lamx = Ki[0] * X + Ki[1] * Y + Ki[2] * 1.f;
  const float lamy = Ki[3] * X + Ki[4] * Y + Ki[5] * 1.f;
  const float lamz = Ki[6] * X + Ki[7] * Y + Ki[8] * 1.f;
  const float lam = sqrtf(lamx * lamx + lamy * lamy + lamz * lamz);
  const float sc = (lam * dval - trunc) / lam;
  const float slx = sc * lamx, sly = sc * lamy, slz = sc * lamz;
  const float sgx = T[0] * slx + T[1] * sly + T[2] * slz + T[3] * 1.f;
  const float sgy = T[4] * slx + T[5] * sly + T[6] * slz + T[7] * 1.f;
  const float sgz = T[8] * slx + T[9] * sly + T[10] * slz + T[11] * 1.f;
  float dx, dy, dz;
  ray_direction(lamx, lamy, lamz, T, dx, dy, dz);
  const float mnx = (float)vg.sx * d, mny = (float)vg.sy * d, mnz = (float)vg.sz * d;
  const float mxx = (float)(vg.sx + Nx) * d, mxy = (float)(vg.sy + Ny) * d,
              mxz = (float)(vg.sz + Nz) * d;
  // Ray_Box_Intersection (voxel_traversal.cu:4-73)
  float tmin, tmax, tymin, tymax, tzmin, tzmax;
  if (dx >= 0) {
    tmin = (mnx - sgx) / dx;
    tmax = (mxx - sgx) / dx;
  } else {
    tmin = (mxx - sgx) / dx;
    tmax = (mnx - sgx) / dx;
  }
  if (dy >= 0) {
    tymin = (mny - sgy) / dy;
    tymax = (mxy - sgy) / dy;
  } else {
    tymin = (mxy - sgy) / dy;
    tymax = (mny - sgy) / dy;
  }
  if ((tmin > tymax) || (tymin > tmax)) return false;
  if (tymin > tmin) tmin = tymin;
  if (tymax < tmax) tmax = tymax;
  if (dz >= 0) {
    tzmin = (mnz - sgz) / dz;
    tzmax = (mxz - sgz) / dz;
  } else {
    tzmin = (mxz - sgz) / dz;
    tzmax = (mnz - sgz) / dz;
  }
  if ((tmin > tzmax) || (tzmin > tmax)) return false;
  if (tzmin > tmin) tmin = tzmin;
  if (tmin < 0) tmin = 0;
  const float stx = sgx + tmin * dx, sty = sgy + tmin * dy, stz = sgz + tmin * dz;
  const float bxs = mxx - mnx, bys = mxy - mny, bzs = mxz - mnz;
  int ix = kf_floor_to_int(((stx - mnx) / bxs) * (float)Nx) + 1;
  int iy = kf_floor_to_int(((sty - mny) / bys) * (float)Ny) + 1;
  int iz = kf_floor_to_int(((stz - mnz) / bzs) * (float)Nz) + 1;
  if (ix == Nx + 1) ix = ix - 1;
  if (iy == Ny + 1) iy = iy - 1;
  if (iz == Nz + 1) iz = iz - 1;
  float tvx, tvy, tvz;
  int stx_, sty_, stz_;
  if (dx >= 0) {
    tvx = (float)(ix) / (float)Nx;
    stx_ = 1;
  } else {
    tvx = (float)(ix - 1) / (float)Nx;
    stx_ = -1;
  }
  if (dy >= 0) {
    tvy = (float)(iy) / (float)Ny;
    sty_ = 1;
  } else {
    tvy = (float)(iy - 1) / (float)Ny;
    sty_ = -1;
  }
  if (dz >= 0) {
    tvz = (float)(iz) / (float)Nz;
    stz_ = 1;
  } else {
    tvz = (float)(iz - 1) / (float)Nz;
    stz_ = -1;
  }
  const float vmx = mnx + tvx * bxs, vmy = mny + tvy * bys, vmz = mnz + tvz * bzs;
  r.cur.ix = ix;
  r.cur.iy = iy;
  r.cur.iz = iz;
  // XY..ZY start at the start voxel (the reference reads them before the
  // first step on that axis, voxel_traversal.cu:136)
  r.cur.XY = iy;
  r.cur.XZ = iz;
  r.cur.YX = ix;
  r.cur.YZ = iz;
  r.cur.ZX = ix;
  r.cur.ZY = iy;
  r.cur.tmx = tmin + ((vmx - stx) / dx);
  r.cur.tmy = tmin + ((vmy - sty) / dy);
  r.cur.tmz = tmin + ((vmz - stz) / dz);
  r.stp = DdaStep{stx_, sty_, stz_, d / fabsf(dx), d / fabsf(dy), d / fabsf(dz)};
  r.prev = 0.0f;
  r.dx = dx;
  r.dy = dy;
  r.dz = dz;
  r.tent = tmin;
  return true;
}

// Ray parameter at which the next dda_step enters its new voxel (the
// boundary distance of the axis it steps).
__device__ __forceinline__ float dda_entry(const DdaState& s) {
  const bool xy = s.tmx < s.tmy;
  const bool xm = xy && s.tmx < s.tmz;
  const bool ym = !xy && s.tmy < s.tmz;
  return xm ? s.tmx : (ym ? s.tmy : s.tmz);
}

// ---------------------------------------------------------------------------
// Trilinear zero-crossing refinement: the differentiable raycast mode (no
// reference counterpart; the reference interpolates per axis from voxel
// values only, voxel_traversal.cu:318-349, which has no pose gradient).
//
// After the walk finds its crossing voxel, the TSDF is sampled trilinearly at
// p_k = sg + (k d) dir, a grid of one-voxel spacing anchored at the skip
// point sg, for k from floor(t_entry / d) - 2 on.  The first pair of valid
// samples with F(p_k) > 0 >= F(p_k+1) gives
//   t* = t_k + d F_k / (F_k - F_k+1),  p* = sg + t* dir,
//   depth = sc + t* / |lambda|  (sc = (|lambda| depth_in - trunc)/|lambda|),
//   normal = grad F(p*) / |grad F(p*)|.
// A sample is taken in voxel-centre coordinates u = (p - min)/d - 1/2 and is
// valid when its 8 corners lie inside the volume and are observed.
constexpr int kTriSamples = 6;

struct Tri {
  float v;       // value
  float g[3];    // gradient (per mm)
  float w[3];    // fractional position in the cell
  int i, j, k;   // lower corner, 0-based voxel indices
  float c[8];    // corner values, index (dx << 2) | (dy << 1) | dz
};

template <int PREC>
__device__ __forceinline__ bool tri_sample(const VolRead<PREC>& F, const kf_vol& vg, float px,
                                           float py, float pz, Tri& s) {
  const float d = vg.d, rd = 1.0f / d;  // quotients by d: div_by_recip
  const float ux = div_by_recip(px - (float)vg.sx * d, d, rd) - 0.5f;
  const float uy = div_by_recip(py - (float)vg.sy * d, d, rd) - 0.5f;
  const float uz = div_by_recip(pz - (float)vg.sz * d, d, rd) - 0.5f;
  const int i = kf_floor_to_int(ux), j = kf_floor_to_int(uy), k = kf_floor_to_int(uz);
  if (!(i >= 0 && i < vg.nx - 1 && j >= 0 && j < vg.ny - 1 && k >= 0 && k < vg.nz - 1))
    return false;
  bool ok = true;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    s.c[q] = F(kf_store_index(vg, i + (q >> 2), j + ((q >> 1) & 1), k + (q & 1)));
    ok = ok && s.c[q] == s.c[q];
  }
  if (!ok) return false;
  const float fx = ux - (float)i, fy = uy - (float)j, fz = uz - (float)k;
  const float ox = 1.0f - fx, oy = 1.0f - fy, oz = 1.0f - fz;
  const float c00 = s.c[0] * ox + s.c[4] * fx;  // y0 z0
  const float c01 = s.c[1] * ox + s.c[5] * fx;  // y0 z1
  const float c10 = s.c[2] * ox + s.c[6] * fx;  // y1 z0
  const float c11 = s.c[3] * ox + s.c[7] * fx;  // y1 z1
  const float c0 = c00 * oy + c10 * fy;
  const float c1 = c01 * oy + c11 * fy;
  s.v = c0 * oz + c1 * fz;
  const float e00 = s.c[4] - s.c[0], e01 = s.c[5] - s.c[1];
  const float e10 = s.c[6] - s.c[2], e11 = s.c[7] - s.c[3];
  s.g[0] = div_by_recip((e00 * oy + e10 * fy) * oz + (e01 * oy + e11 * fy) * fz, d, rd);
  s.g[1] = div_by_recip((c10 - c00) * oz + (c11 - c01) * fz, d, rd);
  s.g[2] = div_by_recip(c1 - c0, d, rd);
  s.w[0] = fx;
  s.w[1] = fy;
  s.w[2] = fz;
  s.i = i;
  s.j = j;
  s.k = k;
  return true;
}

// Where the adjoint's dL/dF contributions go.  The walk's first pass and the
// adjoint from saved crossings sum them in a per-block LDS table keyed by
// trilinear cell (the lower corner's linear voxel index): a key and the 8
// corners' sums per entry, so that a sample costs one lookup, and lanes
// sharing a cell (neighbouring rays) add to its 8 slots in lane-rotated
// order; the block then adds each touched corner to global memory once.  A
// sample that finds no entry within kCellProbes slots, and the walk's second
// pass, use global atomics.
constexpr int kCellBits = 9;
constexpr int kCells = 1 << kCellBits;
constexpr int kCellProbes = 16;

// LDS-qualified pointers, so that the table's atomics are LDS instructions
// (through a generic pointer the float adds become flat atomics).
typedef __attribute__((address_space(3))) int lds_int;
typedef __attribute__((address_space(3))) float lds_float;

struct Sink {
  float* gF;          // global dL/dF (native order); null: no TSDF gradient
  bool table;         // the LDS cell table below is in use (else global atomics only)
  lds_int* key;       // cell keys (-1 = empty), kCells
  lds_float* val;     // corner sums, kCells x 8
  uint32_t* touched;  // one bit per brick that received dL/dF (for zeroing)
};

// Marks brick b as holding dL/dF (a bit already seen set is not set again).
__device__ __forceinline__ void mark_brick(uint32_t* touched, int b) {
  const uint32_t m = 1u << (b & 31);
  if (!(touched[b >> 5] & m)) atomicOr(&touched[b >> 5], m);
}

// dL/dF into voxel idx through a global atomic.  Its brick is marked (test,
// then set if clear) only when it differs from *last, the brick this lane
// marked last: the corners of a cell nearly always share a brick, a test per
// corner puts a dependent load before every add, and setting without the
// test sends every ray's lanes to the same few words.
__device__ __forceinline__ void sink_global(const Sink& k, int idx, float v, int* last) {
  const int b = idx >> 9;
  if (b != *last) {
    mark_brick(k.touched, b);
    *last = b;
  }
  atomicAdd(&k.gF[idx], v);
}

// dL/dF of the 8 corners q (index (dx << 2) | (dy << 1) | dz) of sample s
// through its value: a * weight_q.
__device__ __forceinline__ void tri_corner_value(const Tri& s, float a, float (&val)[8]) {
  const float fx = s.w[0], fy = s.w[1], fz = s.w[2];
  const float ox = 1.0f - fx, oy = 1.0f - fy, oz = 1.0f - fz;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int qx = q >> 2, qy = (q >> 1) & 1, qz = q & 1;
    const float wx = qx ? fx : ox, wy = qy ? fy : oy, wz = qz ? fz : oz;
    val[q] = a * (wx * wy * wz);
  }
}

// ... and through its gradient: b . grad(weight_q).
__device__ __forceinline__ void tri_corner_grad(const kf_vol& vg, const Tri& s, float bx,
                                                float by, float bz, float (&val)[8]) {
  const float d = vg.d;
  const float fx = s.w[0], fy = s.w[1], fz = s.w[2];
  const float ox = 1.0f - fx, oy = 1.0f - fy, oz = 1.0f - fz;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int qx = q >> 2, qy = (q >> 1) & 1, qz = q & 1;
    const float wx = qx ? fx : ox, wy = qy ? fy : oy, wz = qz ? fz : oz;
    const float sx = qx ? 1.0f : -1.0f, sy = qy ? 1.0f : -1.0f, sz = qz ? 1.0f : -1.0f;
    val[q] = (bx * (sx * wy * wz) + by * (wx * sy * wz) + bz * (wx * wy * sz)) / d;
  }
}

// Adds val[q] to the 8 corners of the cell with lower corner (i, j, k).
// With the cell table: keys are never removed during a pass, so an entry that
// already holds the cell (a plain read, which lanes reading one address
// share) needs no compare-and-swap; only a cell's first insert swaps.  The 8
// adds go in lane-rotated corner order and return nothing.  A cell that finds
// no entry within kCellProbes slots takes global atomics.
__device__ __forceinline__ void sink_add_cell(const Sink& k, const kf_vol& vg, int i, int j,
                                              int kk, const float (&val)[8]) {
  int slot = -1;
  if (k.table) {
    const int cell = (i * vg.ny + j) * vg.nz + kk;
    uint32_t h = ((uint32_t)cell * 2654435761u) >> (32 - kCellBits);
    if (k.key[h] == cell) {
      slot = (int)h;
    } else {
#pragma unroll 1
      for (int probe = 0; probe < kCellProbes; ++probe) {
        int expected = -1;
        __hip_atomic_compare_exchange_strong(&k.key[h], &expected, cell, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (expected == -1 || expected == cell) {
          slot = (int)h;
          break;
        }
        h = (h + 1) & (kCells - 1);
      }
    }
  }
  if (slot >= 0) {
    const int r = lane_id() & 7;
    lds_float* e = k.val + slot * 8;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int q = (t + r) & 7;
      float v = val[0];
#pragma unroll
      for (int u = 1; u < 8; ++u) v = q == u ? val[u] : v;
      if (v != 0.0f) __hip_atomic_fetch_add(e + q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  } else {
    int last = -1;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (val[q] != 0.0f)
        sink_global(k, kf_store_index(vg, i + (q >> 2), j + ((q >> 1) & 1), kk + (q & 1)), val[q],
                    &last);
  }
}

__device__ __forceinline__ bool same_cell(const Tri& a, const Tri& b) {
  return a.i == b.i && a.j == b.j && a.k == b.k;
}

// Trilinear hit of a ray whose walk crossed at a voxel entered at parameter
// te.  MODE 1 writes vertex (w = depth) and normal (w = te); MODE 2 adds the
// adjoint contributions of this pixel (dL/dF scattered, dL/dxi into xi).
// Nothing else of the walk enters, so an adjoint at the inputs of a MODE 1
// raycast can start from its saved te (k_adjoint_saved).
template <int PREC, int MODE>
__device__ __forceinline__ bool tri_hit(const VolRead<PREC>& F, const RayIO& io, const float* T,
                                     int pix, float te, float dx, float dy, float dz,
                                     float4& ov, float4& on, double* xi, const Sink& sink,
                                     float* acc) {
  const kf_vol& vg = io.vg;
  const float* Ki = io.Ki;
  const int x = pix % io.W, y = pix / io.W;
  const float dval = io.depth[pix];
  // skip point, as in ray_setup
  const float X = (float)x + 1.0f, Y = (float)y + 1.0f;
  const float lamx = Ki[0] * X + Ki[1] * Y + Ki[2] * 1.f;
  const float lamy = Ki[3] * X + Ki[4] * Y + Ki[5] * 1.f;
  const float lamz = Ki[6] * X + Ki[7] * Y + Ki[8] * 1.f;
  const float lam = sqrtf(lamx * lamx + lamy * lamy + lamz * lamz);
  const float sc = (lam * dval - io.trunc) / lam;
  const float slx = sc * lamx, sly = sc * lamy, slz = sc * lamz;
  const float sgx = T[0] * slx + T[1] * sly + T[2] * slz + T[3] * 1.f;
  const float sgy = T[4] * slx + T[5] * sly + T[6] * slz + T[7] * 1.f;
  const float sgz = T[8] * slx + T[9] * sly + T[10] * slz + T[11] * 1.f;
  const float d = vg.d;
  if (!(te >= 0.0f && te < 1.0e9f)) return false;
  const int kb = kf_floor_to_int(te / d) - 2;
  Tri a, b;
  bool va = false, found = false;
  int ka = 0;
  for (int k = kb; k < kb + kTriSamples; ++k) {
    const float t = (float)k * d;
    const bool ok = tri_sample<PREC>(F, vg, sgx + t * dx, sgy + t * dy, sgz + t * dz, b);
    if (ok && va && a.v > 0.0f && b.v <= 0.0f) {
      ka = k - 1;
      found = true;
      break;
    }
    va = ok;
    if (ok) a = b;
  }
  if (!found) return false;
  const float ta = (float)ka * d;
  const float den = a.v - b.v;
  const float ts = ta + d * (a.v / den);
  const float psx = sgx + ts * dx, psy = sgy + ts * dy, psz = sgz + ts * dz;
  Tri s;
  if (!tri_sample<PREC>(F, vg, psx, psy, psz, s)) return false;
  const float gn = sqrtf(s.g[0] * s.g[0] + s.g[1] * s.g[1] + s.g[2] * s.g[2]);
  if (!(gn > 0.0f)) return false;
  const float rgn = 1.0f / gn;
  const float nx = div_by_recip(s.g[0], gn, rgn), ny = div_by_recip(s.g[1], gn, rgn),
              nz = div_by_recip(s.g[2], gn, rgn);
  if constexpr (MODE == kModeTrilinear) {
    ov = make_float4(psx, psy, psz, sc + ts / lam);
    on = make_float4(nx, ny, nz, te);  // te: what the adjoint needs of the walk
  } else if constexpr (MODE == kModeRefine) {
    const float r = (sc + ts / lam) - dval;
    if (!(fabsf(r) <= io.max_res)) return true;
    const float den2 = den * den;
    const float dts = 1.0f / lam;
    const float dFa = dts * d * (-b.v) / den2;
    const float dFb = dts * d * a.v / den2;
    const float tb = ta + d;
    const float pax = sgx + ta * dx, pay = sgy + ta * dy, paz = sgz + ta * dz;
    const float pbx = sgx + tb * dx, pby = sgy + tb * dy, pbz = sgz + tb * dz;
    float J[6];
    J[0] = dFa * (pay * a.g[2] - paz * a.g[1]) + dFb * (pby * b.g[2] - pbz * b.g[1]);
    J[1] = dFa * (paz * a.g[0] - pax * a.g[2]) + dFb * (pbz * b.g[0] - pbx * b.g[2]);
    J[2] = dFa * (pax * a.g[1] - pay * a.g[0]) + dFb * (pbx * b.g[1] - pby * b.g[0]);
    J[3] = dFa * a.g[0] + dFb * b.g[0];
    J[4] = dFa * a.g[1] + dFb * b.g[1];
    J[5] = dFa * a.g[2] + dFb * b.g[2];
    // the pixel's record (a ray hits at most once); gn_add_pixel forms the
    // sums' terms
    for (int i = 0; i < 6; ++i) acc[i] = J[i];
    acc[6] = r;
    acc[7] = 1.0f;
  } else {
    const float gD = io.gD[pix];
    const float4 gNv = io.gN ? io.gN[pix] : make_float4(0.f, 0.f, 0.f, 0.f);
    float dts = gD / lam;  // dL/dt* through the depth
    // normal: n = g/|g|, g = grad F(p*) (linear in the corners of p*)
    const float ndot = nx * gNv.x + ny * gNv.y + nz * gNv.z;
    const float lgx = div_by_recip(gNv.x - nx * ndot, gn, rgn);
    const float lgy = div_by_recip(gNv.y - ny * ndot, gn, rgn);
    const float lgz = div_by_recip(gNv.z - nz * ndot, gn, rgn);
    // dL/dF through the normal (zero without a normal gradient)
    const bool has_n = io.gN != nullptr;
    float vs[8] = {};
    if (sink.gF && has_n) tri_corner_grad(vg, s, lgx, lgy, lgz, vs);
    // dL/dp* = H lg, H the trilinear Hessian at p* (zero diagonal)
    const float fx = s.w[0], fy = s.w[1], fz = s.w[2];
    const float ox = 1.0f - fx, oy = 1.0f - fy, oz = 1.0f - fz;
    const float e00 = s.c[4] - s.c[0], e01 = s.c[5] - s.c[1];
    const float e10 = s.c[6] - s.c[2], e11 = s.c[7] - s.c[3];
    const float c00 = s.c[0] * ox + s.c[4] * fx, c01 = s.c[1] * ox + s.c[5] * fx;
    const float c10 = s.c[2] * ox + s.c[6] * fx, c11 = s.c[3] * ox + s.c[7] * fx;
    const float dd = d * d;
    const float hxy = ((e10 - e00) * oz + (e11 - e01) * fz) / dd;
    const float hxz = ((e01 * oy + e11 * fy) - (e00 * oy + e10 * fy)) / dd;
    const float hyz = ((c11 - c01) - (c10 - c00)) / dd;
    const float lpx = hxy * lgy + hxz * lgz;
    const float lpy = hxy * lgx + hyz * lgz;
    const float lpz = hxz * lgx + hyz * lgy;
    dts += lpx * dx + lpy * dy + lpz * dz;
    xi[0] += (double)(psy * lpz - psz * lpy);
    xi[1] += (double)(psz * lpx - psx * lpz);
    xi[2] += (double)(psx * lpy - psy * lpx);
    xi[3] += (double)lpx;
    xi[4] += (double)lpy;
    xi[5] += (double)lpz;
    // secant: t* = t_a + d F_a / (F_a - F_b)
    const float den2 = den * den;
    const float dFa = dts * d * (-b.v) / den2;
    const float dFb = dts * d * a.v / den2;
    if (sink.gF) {
      // dL/dF of the three samples' corners; samples in the same cell are
      // added in registers first (p* lies between the two depth samples)
      float va[8], vb[8];
      tri_corner_value(a, dFa, va);
      tri_corner_value(b, dFb, vb);
      const bool s_in_a = has_n && same_cell(s, a), s_in_b = has_n && !s_in_a && same_cell(s, b);
      const bool b_in_a = same_cell(b, a);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        va[q] = s_in_a ? va[q] + vs[q] : va[q];
        vb[q] = s_in_b ? vb[q] + vs[q] : vb[q];
        va[q] = b_in_a ? va[q] + vb[q] : va[q];
      }
      if (has_n && !s_in_a && !s_in_b) sink_add_cell(sink, vg, s.i, s.j, s.k, vs);
      sink_add_cell(sink, vg, a.i, a.j, a.k, va);
      if (!b_in_a) sink_add_cell(sink, vg, b.i, b.j, b.k, vb);
    }
    const float tb = ta + d;
    const float pax = sgx + ta * dx, pay = sgy + ta * dy, paz = sgz + ta * dz;
    const float pbx = sgx + tb * dx, pby = sgy + tb * dy, pbz = sgz + tb * dz;
    xi[0] += (double)(dFa * (pay * a.g[2] - paz * a.g[1]) + dFb * (pby * b.g[2] - pbz * b.g[1]));
    xi[1] += (double)(dFa * (paz * a.g[0] - pax * a.g[2]) + dFb * (pbz * b.g[0] - pbx * b.g[2]));
    xi[2] += (double)(dFa * (pax * a.g[1] - pay * a.g[0]) + dFb * (pbx * b.g[1] - pby * b.g[0]));
    xi[3] += (double)(dFa * a.g[0] + dFb * b.g[0]);
    xi[4] += (double)(dFa * a.g[1] + dFb * b.g[1]);
    xi[5] += (double)(dFa * a.g[2] + dFb * b.g[2]);
  }
  return true;
}

// Walks up to max_batches batches of the traversal (voxel_traversal.cu:203-
// 421).  The voxel sequence does not depend on the TSDF values, only where
// the walk stops does, so each batch generates kRayBatch steps ahead, issues
// their loads together and then applies the reference's per-voxel tests
// (:207-222) in order: the same voxels are visited and the same decisions
// taken.  Returns true when the ray is finished (hit or not); on false the
// state is ready to continue.
template <int PREC, int kRayBatch, bool COUNT, int MODE>
__device__ __forceinline__ bool ray_walk(const VolRead<PREC>& F, RayState& r, const RayIO& io,
                                         const float* T, int max_batches, float4& ov,
                                         float4& on, int& hit, int& visited, double* xi,
                                         const Sink& sink, float* gn) {
  const kf_vol& vg = io.vg;
  const int Nx = vg.nx, Ny = vg.ny, Nz = vg.nz;
  for (int bt = 0; bt < max_batches; ++bt) {
    // generate the next kRayBatch voxels of the walk (branch-free: the walk
    // keeps stepping past its exit, but those voxels are never read)
    int vis[kRayBatch];
    int nv = 0;
    bool alive = true;
    DdaState s = r.cur;
    float te_s = r.tent;
#pragma unroll
    for (int k = 0; k < kRayBatch; ++k) {
      const bool in =
          s.ix < Nx - 1 && s.ix > 0 && s.iy < Ny - 1 && s.iy > 0 && s.iz < Nz - 1 && s.iz > 0;
      alive = alive && in;
      vis[k] = alive ? kf_store_index(vg, s.ix - 1, s.iy - 1, s.iz - 1) : 0;
      nv += alive ? 1 : 0;
      if constexpr (MODE != kModeReference) te_s = dda_entry(s);
      dda_step(s, r.stp);
    }
    float fs[kRayBatch];
#pragma unroll
    for (int k = 0; k < kRayBatch; ++k) fs[k] = F(vis[k]);
    // the reference's per-voxel tests, in order
    bool done = false;
    int kc = -1;
    float fc = 0.0f;
    float prev = r.prev;
#pragma unroll
    for (int k = 0; k < kRayBatch; ++k) {
      const float f = fs[k];
      const bool valid = k < nv && !done;
      if constexpr (COUNT) visited += valid ? 1 : 0;
      const bool brk = (f != f && prev < 0.f) || (f < 0.f && prev != prev) || (f > 0.f && prev < 0.f);
      const bool cross = !brk && f < 0.f && prev > 0.f;
      kc = (valid && cross) ? k : kc;
      fc = (valid && cross) ? f : fc;
      prev = (valid && !brk && !cross) ? f : prev;
      done = done || (valid && (brk || cross));
    }
    r.prev = prev;
    if (kc >= 0) {
      DdaState h = r.cur;
      float te = r.tent;
      for (int j = 0; j < kc; ++j) {
        if constexpr (MODE != kModeReference) te = dda_entry(h);
        dda_step(h, r.stp);
      }
      if constexpr (MODE == kModeReference) {
        hit = crossing<PREC>(F, h, fc, vg, io.fix_bounds, T[3], T[7], T[11], r.dx, r.dy, r.dz,
                             io.flip_hi, ov, on);
      } else {
        const bool ok =
            tri_hit<PREC, MODE>(F, io, T, r.pix, te, r.dx, r.dy, r.dz, ov, on, xi, sink, gn);
        hit = ok ? (h.ix - 1) * (Ny * Nz) + (h.iy - 1) * Nz + (h.iz - 1) : -1;
      }
    }
    if (done || nv < kRayBatch) return true;
    r.cur = s;
    r.tent = te_s;
    if constexpr (MODE == kModeRefine) {
      if (r.tent > io.walk_tmax) return true;  // no crossing ahead can be used
    }
  }
  return false;
}

// ---------------------------------------------------------------------------
// Cooperative walk: one ray per wave, kWaveSteps steps of the traversal per
// round (the second pass, where the few long rays would otherwise run one
// lane each through a serial chain of several hundred steps).
//
// The walk chooses the axis with the smallest boundary distance, ties going
// to z, then y, then x (voxel_traversal.cu:384-417).  With finite distances
// that is the merge of the three per-axis sequences t_a(j + 1) = t_a(j) +
// td_a by the key (t, rank), rank z 0 < y 1 < x 2: each sequence is
// non-decreasing and a step always takes the smallest key among the heads.
// A round:
//   1. lanes 0-2 produce t_x, t_y, t_z (j = 0 .. kWaveSteps) with the float
//      additions the serial walk makes, into LDS;
//   2. lane j ranks the j-th element of each sequence in the merge (its index
//      plus how many elements of the two other sequences come before it);
//      the element of rank r < kWaveSteps is step r: its slot records the
//      per-axis step counts after it and its key, and each element records
//      the other two indices at its step (the reference's XY .. ZY);
//   3. lane k visits voxel V_k (the voxel after k steps), loads its TSDF and
//      applies the reference's tests with prev = the TSDF of V_{k-1} (lane
//      k - 1; the ray's prev for k = 0): prev always equals the previous
//      voxel's value because every visit that does not stop the walk sets it;
//   4. the first visit that breaks or crosses ends the walk (the crossing is
//      handled by that lane with the state of V_k rebuilt from the records),
//      as does the first voxel outside the traversal; otherwise the state
//      advances to V_kWaveSteps.
// Visits, decisions and the crossing state are those of the serial walk.
constexpr int kWaveSteps = 64;

struct alignas(16) WaveWalkLds {
  float seq[4][kWaveSteps + 4];  // t_a(j), j <= kWaveSteps (rows padded to 16 bytes);
                                 // row 3 takes the writes of lanes >= 3
  int2 rec[3][kWaveSteps];       // other two indices at the j-th step along a
  uint32_t cnt[kWaveSteps];      // step counts after step r: cx | cy << 8 | cz << 16
  float key[kWaveSteps];         // boundary distance of step r
};

// #{j < kWaveSteps : s[j] <= v} (LE) or #{j : s[j] < v}, for a non-decreasing
// s with s[j] close to s0 + j td: estimated from that, then corrected against
// the stored values.  An infinite sequence (s0 = +inf, inv_td = 0) makes the
// estimate NaN; it is clamped to 0 like any estimate that is not positive,
// and corrected.
template <bool LE>
__device__ __forceinline__ int seq_guess(float s0, float inv_td, float v) {
  const float q = floorf((v - s0) * inv_td) + 1.0f;
  return !(q > 0.f) ? 0 : (q >= (float)kWaveSteps ? kWaveSteps : (int)q);
}
// Whether guess c needs no correction, given lo = s[c - 1] and hi = s[c].
template <bool LE>
__device__ __forceinline__ bool seq_exact(int c, float lo, float hi, float v) {
  return (c == 0 || !(LE ? lo > v : lo >= v)) && (c == kWaveSteps || !(LE ? hi <= v : hi < v));
}
template <bool LE>
__device__ __forceinline__ int seq_fix(const float* s, int c, float v) {
  while (c > 0 && (LE ? s[c - 1] > v : s[c - 1] >= v)) --c;
  while (c < kWaveSteps && (LE ? s[c] <= v : s[c] < v)) ++c;
  return c;
}

// Wave-uniform (scalar) copies of values that every lane holds alike, so
// that the branches they decide are scalar branches.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Empty-space skipping of the second pass.  No step can stop the walk
// before it reaches a voxel with a negative TSDF value: a break or a
// crossing needs f < 0 or prev < 0, and while neither holds a visit only
// sets prev = f.  A query reads the skip flags at kSkipWindows points
// kSkipSpacing voxels apart along the ray: sample k is the continuous voxel
// coordinate (the walk's voxel at the query) + 0.5 + 8 k (dx, dy, dz), the
// ray's position at parameter tq + 8 k d (tq: the query's tent).  Window k
// holds the steps whose boundary distance lies within 4 d of tq + 8 k d.
// The voxel such a step enters lies within 4 voxels per axis of the sample,
// plus one for the number of boundaries the walk crosses on that axis in
// between (they are td apart), so its index is within 5 of the sample's
// rounded-down coordinate, inside the 8 voxels a set flag covers around its
// brick.  A window whose sample reads a clear flag therefore holds no
// negative voxel, and a skip round takes every step whose boundary distance
// lies below the first flagged window ahead: with the sequences of one
// round, it sets the state the merge would reach after those steps, and
// loads no TSDF value (the next full round loads the last one as its prev).
constexpr int kSkipSpacing = 8;
constexpr int kSkipWindows = 64;

// This lane's query sample: the skip flag of the brick that holds it
// (clamped into the volume, which moves it towards every voxel inside).
__device__ __forceinline__ uint8_t skip_flag_at(const uint8_t* __restrict__ flags, const kf_vol& vg,
                                                const DdaState& c, float dx, float dy, float dz) {
  const float o = (float)(kSkipSpacing * lane_id());
  const float fx = floorf(((float)(c.ix - 1) + 0.5f) + o * dx);
  const float fy = floorf(((float)(c.iy - 1) + 0.5f) + o * dy);
  const float fz = floorf(((float)(c.iz - 1) + 0.5f) + o * dz);
  const int x = (int)fminf(fmaxf(fx, 0.0f), (float)(vg.nx - 1));
  const int y = (int)fminf(fmaxf(fy, 0.0f), (float)(vg.ny - 1));
  const int z = (int)fminf(fmaxf(fz, 0.0f), (float)(vg.nz - 1));
  return flags[((x >> 3) * vg.by + (y >> 3)) * vg.bz + (z >> 3)];
}

// Walks ray r (the same state in every lane) to its end.  Returns the lane
// that holds the result (ov, on, hit and, for the adjoint and refinement
// modes, the gradient terms); the other lanes keep ov = on = 0, hit = -1.
template <int PREC, bool COUNT, int MODE, bool SKIP>
__device__ __forceinline__ int wave_walk(const VolRead<PREC>& F, RayState& r, const RayIO& io,
                                         const float* T, WaveWalkLds& L, float4& ov,
                                         float4& on, int& hit, int& visited, int& skipped,
                                         double* xi, const Sink& sink, float* gn) {
  const int lane = lane_id();
  const kf_vol& vg = io.vg;
  const int Nx = vg.nx, Ny = vg.ny, Nz = vg.nz;
  const DdaStep p = r.stp;
  // A direction component of +0 (rg - cam of equal floats is +0) gives that
  // axis td = +inf and, its first boundary being ahead, tm = +inf: the axis
  // never steps, in the serial walk (no comparison with +inf is true) and in
  // the merge (its elements are +inf, ranked after every finite element), so
  // the merge walks such rays too.  With finite td > 0 the heads stay finite
  // or reach +inf, with the same effect.  A NaN distance (the start point on
  // the boundary of an axis with a 0 component) changes the serial walk's
  // comparisons, and with every axis infinite nothing steps: those rays take
  // the serial walk, in one lane.
  const bool no_nan = !(isnan(p.tdx) || isnan(p.tdy) || isnan(p.tdz) || isnan(r.cur.tmx) ||
                        isnan(r.cur.tmy) || isnan(r.cur.tmz));
  const bool one_finite = (isfinite(p.tdx) && isfinite(r.cur.tmx)) ||
                          (isfinite(p.tdy) && isfinite(r.cur.tmy)) ||
                          (isfinite(p.tdz) && isfinite(r.cur.tmz));
  if (!(no_nan && one_finite && p.tdx > 0.f && p.tdy > 0.f && p.tdz > 0.f)) {
    if (lane == 0)
      ray_walk<PREC, 16, COUNT, MODE>(F, r, io, T, 0x7fffffff, ov, on, hit, visited, xi, sink,
                                      gn);
    return 0;
  }
  const float itx = 1.0f / p.tdx, ity = 1.0f / p.tdy, itz = 1.0f / p.tdz;
  // Skip rounds (SKIP, with io.skip_flags set): the windows of the last flag
  // query (bit k: window k flagged), its ray parameter tq, the parameter from
  // which the walk queries again, and the store index of the voxel whose TSDF
  // is the next visit's prev after a skip round (-1: r.prev holds it).
  const bool skip_on = SKIP && io.skip_flags != nullptr;
  const float s_mm = (float)kSkipSpacing * vg.d;
  uint64_t qmask = ~0ull;
  float tq = 0.0f, qagain = -INFINITY;
  int pidx = -1;
  // The parameter at which the walk entered its current voxel, which the
  // query's windows are aligned to.  r.tent holds it in the trilinear modes;
  // the reference mode's first pass does not keep it, so there it is the
  // last boundary crossed on any axis, max_a (t_a - td_a) (within an ulp;
  // an axis that has not stepped gives a value before the walk's start).
  // Rounds keep r.tent exact from then on.
  float tpos = r.tent;
  if constexpr (MODE == kModeReference)
    tpos = fmaxf(tpos, fmaxf(r.cur.tmx - p.tdx, fmaxf(r.cur.tmy - p.tdy, r.cur.tmz - p.tdz)));
  // Each round either ends the walk or advances it by kWaveSteps steps (a
  // skip round: at least one step) along the fixed step directions, so the
  // walk leaves the volume after at most Nx + Ny + Nz rounds.
  for (;;) {
    const DdaState c0 = r.cur;
    // the query's flag loads go out before the sequences and are used after
    const bool query = skip_on && !(tpos < qagain);
    uint8_t qflag = 1;
    if (query) qflag = skip_flag_at(io.skip_flags, vg, c0, r.dx, r.dy, r.dz);
    // 1. per-axis boundary distances (lanes 0-2; the others write row 3, so
    //    that no lane-dependent branch precedes the cross-lane steps below)
    {
      float t = lane == 0 ? c0.tmx : (lane == 1 ? c0.tmy : c0.tmz);
      const float td = lane == 0 ? p.tdx : (lane == 1 ? p.tdy : p.tdz);
      float4* sq = reinterpret_cast<float4*>(L.seq[min(lane, 3)]);
#pragma unroll
      for (int j = 0; j < kWaveSteps / 4; ++j) {
        float4 q;
        q.x = t;
        t = t + td;
        q.y = t;
        t = t + td;
        q.z = t;
        t = t + td;
        q.w = t;
        t = t + td;
        sq[j] = q;
      }
      L.seq[min(lane, 3)][kWaveSteps] = t;
    }
    wave_lds_sync();
    const float ex = L.seq[0][lane], ey = L.seq[1][lane], ez = L.seq[2][lane];
    if (query) {
      qmask = __ballot(qflag != 0);
      tq = tpos;
      qagain = uni(tq + (float)(kSkipWindows - 16) * s_mm);
    }
    // skip round: every step whose boundary distance lies below tr, where tr
    // is the first flagged window ahead, the round's sequence ends (t_a(64))
    // and the distance of each axis's first step out of the traversal.  It
    // visits the current voxel and those its steps enter but the last (the
    // next current voxel), all in clear windows; the current voxel must lie
    // in the traversal (a full round ends a walk whose current voxel does
    // not), and prev, the last visited voxel's TSDF, must not be negative
    // (after a skip round it is not)
    if (skip_on && (c0.ix < Nx - 1 && c0.ix > 0 && c0.iy < Ny - 1 && c0.iy > 0 &&
                    c0.iz < Nz - 1 && c0.iz > 0) &&
        (pidx >= 0 || !(r.prev < 0.f))) {
      // windows that end before tpos are ignored (rounded towards keeping
      // them); k0 <= kSkipWindows - 16 since the walk queries again there
      const int k0 = max(0, (int)floorf((tpos - tq) / s_mm) - 1);
      const uint64_t ahead = qmask >> k0;
      const float tneg = ahead ? tq + ((float)(k0 + __ffsll((unsigned long long)ahead) - 1) - 0.5f) * s_mm
                               : tq + ((float)kSkipWindows - 0.5f) * s_mm;
      if (fminf(c0.tmx, fminf(c0.tmy, c0.tmz)) < tneg) {
        float tr = fminf(tneg, fminf(L.seq[0][kWaveSteps],
                                     fminf(L.seq[1][kWaveSteps], L.seq[2][kWaveSteps])));
        // steps along an axis that keep the walk inside (0, N - 1) (>= 0: the
        // current voxel is inside)
        const int jx = p.sx > 0 ? Nx - 2 - c0.ix : c0.ix - 1;
        const int jy = p.sy > 0 ? Ny - 2 - c0.iy : c0.iy - 1;
        const int jz = p.sz > 0 ? Nz - 2 - c0.iz : c0.iz - 1;
        if (jx <= kWaveSteps) tr = fminf(tr, L.seq[0][jx]);
        if (jy <= kWaveSteps) tr = fminf(tr, L.seq[1][jy]);
        if (jz <= kWaveSteps) tr = fminf(tr, L.seq[2][jz]);
        const int nx = __popcll(__ballot(ex < tr));
        const int ny = __popcll(__ballot(ey < tr));
        const int nz = __popcll(__ballot(ez < tr));
        if (nx + ny + nz > 0) {
          // the state after those steps: the other indices at each axis's
          // last step (ranks with the merge's ties), the next boundary
          // distances and the last step's distance
          // distances of each axis's last step (-inf: none); the walk's last
          // step is the largest of them, ties going to x, then y
          const float vx = nx > 0 ? uni(L.seq[0][nx - 1]) : -INFINITY;
          const float vy = ny > 0 ? uni(L.seq[1][ny - 1]) : -INFINITY;
          const float vz = nz > 0 ? uni(L.seq[2][nz - 1]) : -INFINITY;
          DdaState& c = r.cur;
          if (nx > 0) {
            c.XY = c0.iy + p.sy * __popcll(__ballot(ey <= vx));
            c.XZ = c0.iz + p.sz * __popcll(__ballot(ez <= vx));
          }
          if (ny > 0) {
            c.YX = c0.ix + p.sx * __popcll(__ballot(ex < vy));
            c.YZ = c0.iz + p.sz * __popcll(__ballot(ez <= vy));
          }
          if (nz > 0) {
            c.ZX = c0.ix + p.sx * __popcll(__ballot(ex < vz));
            c.ZY = c0.iy + p.sy * __popcll(__ballot(ey < vz));
          }
          const bool lx = vx >= vy && vx >= vz;
          const bool ly = !lx && vy >= vz;
          const bool lz = !lx && !ly;
          c.ix = c0.ix + p.sx * nx;
          c.iy = c0.iy + p.sy * ny;
          c.iz = c0.iz + p.sz * nz;
          c.tmx = uni(L.seq[0][nx]);
          c.tmy = uni(L.seq[1][ny]);
          c.tmz = uni(L.seq[2][nz]);
          r.tent = fmaxf(vx, fmaxf(vy, vz));
          tpos = r.tent;
          // the last voxel the skipped steps visited: one step back along the
          // axis of the last step (its TSDF is the next visit's prev)
          pidx = kf_store_index(vg, c.ix - (lx ? p.sx : 0) - 1, c.iy - (ly ? p.sy : 0) - 1,
                                c.iz - (lz ? p.sz : 0) - 1);
          if constexpr (COUNT) {
            visited += lane == 0 ? nx + ny + nz : 0;
            skipped += lane == 0 ? nx + ny + nz : 0;
          }
          wave_lds_sync();  // every lane has read this round's sequences
          if constexpr (MODE == kModeRefine) {
            if (r.tent > io.walk_tmax) return 0;  // no crossing ahead can be used
          }
          continue;
        }
      }
    }
    // 2. ranks in the merge; ties: y and z before x, z before y.  Each rank
    //    is #{j : s[j] <= v} (or < v) over one sequence: a guess from s0 + j
    //    td, checked against the two stored values around it (all 12 loaded
    //    together), and corrected by a walk where the check fails.
    int y_x = seq_guess<true>(c0.tmy, ity, ex);
    int z_x = seq_guess<true>(c0.tmz, itz, ex);
    int x_y = seq_guess<false>(c0.tmx, itx, ey);
    int z_y = seq_guess<true>(c0.tmz, itz, ey);
    int x_z = seq_guess<false>(c0.tmx, itx, ez);
    int y_z = seq_guess<false>(c0.tmy, ity, ez);
    {
      const float* s0 = L.seq[0];
      const float* s1 = L.seq[1];
      const float* s2 = L.seq[2];
      const bool e0 = seq_exact<true>(y_x, s1[max(y_x - 1, 0)], s1[y_x], ex);
      const bool e1 = seq_exact<true>(z_x, s2[max(z_x - 1, 0)], s2[z_x], ex);
      const bool e2 = seq_exact<false>(x_y, s0[max(x_y - 1, 0)], s0[x_y], ey);
      const bool e3 = seq_exact<true>(z_y, s2[max(z_y - 1, 0)], s2[z_y], ey);
      const bool e4 = seq_exact<false>(x_z, s0[max(x_z - 1, 0)], s0[x_z], ez);
      const bool e5 = seq_exact<false>(y_z, s1[max(y_z - 1, 0)], s1[y_z], ez);
      if (!(e0 & e1 & e2 & e3 & e4 & e5)) {
        y_x = seq_fix<true>(s1, y_x, ex);
        z_x = seq_fix<true>(s2, z_x, ex);
        x_y = seq_fix<false>(s0, x_y, ey);
        z_y = seq_fix<true>(s2, z_y, ey);
        x_z = seq_fix<false>(s0, x_z, ez);
        y_z = seq_fix<false>(s1, y_z, ez);
      }
    }
    L.rec[0][lane] = make_int2(c0.iy + p.sy * y_x, c0.iz + p.sz * z_x);
    L.rec[1][lane] = make_int2(c0.ix + p.sx * x_y, c0.iz + p.sz * z_y);
    L.rec[2][lane] = make_int2(c0.ix + p.sx * x_z, c0.iy + p.sy * y_z);
    const int rx = lane + y_x + z_x, ry = lane + x_y + z_y, rz = lane + x_z + y_z;
    if (rx < kWaveSteps) {
      L.cnt[rx] = (uint32_t)(lane + 1) | ((uint32_t)y_x << 8) | ((uint32_t)z_x << 16);
      L.key[rx] = ex;
    }
    if (ry < kWaveSteps) {
      L.cnt[ry] = (uint32_t)x_y | ((uint32_t)(lane + 1) << 8) | ((uint32_t)z_y << 16);
      L.key[ry] = ey;
    }
    if (rz < kWaveSteps) {
      L.cnt[rz] = (uint32_t)x_z | ((uint32_t)y_z << 8) | ((uint32_t)(lane + 1) << 16);
      L.key[rz] = ez;
    }
    wave_lds_sync();
    // 3. lane k visits V_k (and after a skip round, the TSDF of V_0's
    //    predecessor is loaded with the visits)
    const float p0 = SKIP && pidx >= 0 ? F(pidx) : r.prev;
    const uint32_t ck = lane == 0 ? 0u : L.cnt[lane - 1];
    const int cx = ck & 0xff, cy = (ck >> 8) & 0xff, cz = ck >> 16;
    const int vx = c0.ix + p.sx * cx, vy = c0.iy + p.sy * cy, vz = c0.iz + p.sz * cz;
    const bool alive = vx < Nx - 1 && vx > 0 && vy < Ny - 1 && vy > 0 && vz < Nz - 1 && vz > 0;
    const float f = alive ? F(kf_store_index(vg, vx - 1, vy - 1, vz - 1)) : 0.0f;
    float prev = __shfl_up(f, 1);
    if (lane == 0) prev = p0;
    const bool brk =
        (f != f && prev < 0.f) || (f < 0.f && prev != prev) || (f > 0.f && prev < 0.f);
    const bool cross = !brk && f < 0.f && prev > 0.f;
    const uint64_t ev = __ballot(alive && (brk || cross));
    const uint64_t dead = __ballot(!alive);
    const int kc = ev ? __ffsll((unsigned long long)ev) - 1 : kWaveSteps;
    const int ke = dead ? __ffsll((unsigned long long)dead) - 1 : kWaveSteps;
    if constexpr (COUNT) visited += lane < min(kc + 1, ke) ? 1 : 0;
    if (kc < kWaveSteps) {
      // 4. the walk stops at V_kc (kc < ke: only live visits raise events)
      if (lane == kc && cross) {
        DdaState h;
        h.ix = vx;
        h.iy = vy;
        h.iz = vz;
        const int2 rx_ = cx > 0 ? L.rec[0][cx - 1] : make_int2(c0.XY, c0.XZ);
        const int2 ry_ = cy > 0 ? L.rec[1][cy - 1] : make_int2(c0.YX, c0.YZ);
        const int2 rz_ = cz > 0 ? L.rec[2][cz - 1] : make_int2(c0.ZX, c0.ZY);
        h.XY = rx_.x;
        h.XZ = rx_.y;
        h.YX = ry_.x;
        h.YZ = ry_.y;
        h.ZX = rz_.x;
        h.ZY = rz_.y;
        if constexpr (MODE == kModeReference) {
          hit = crossing<PREC>(F, h, f, vg, io.fix_bounds, T[3], T[7], T[11], r.dx, r.dy, r.dz,
                               io.flip_hi, ov, on);
        } else {
          const float te = lane == 0 ? r.tent : L.key[lane - 1];
          const bool ok =
              tri_hit<PREC, MODE>(F, io, T, r.pix, te, r.dx, r.dy, r.dz, ov, on, xi, sink, gn);
          hit = ok ? (vx - 1) * (Ny * Nz) + (vy - 1) * Nz + (vz - 1) : -1;
        }
      }
      return kc;
    }
    if (ke < kWaveSteps) return 0;  // left the traversal without a crossing
    // advance to V_kWaveSteps
    const uint32_t cn = __builtin_amdgcn_readfirstlane(L.cnt[kWaveSteps - 1]);
    const int nx = cn & 0xff, ny = (cn >> 8) & 0xff, nz = cn >> 16;
    // the ranks are a permutation, so the counts after the last step add up
    // to the round's steps; anything else would stall the walk: stop instead
    if (nx + ny + nz != kWaveSteps) return 0;
    DdaState& c = r.cur;
    c.ix = c0.ix + p.sx * nx;
    c.iy = c0.iy + p.sy * ny;
    c.iz = c0.iz + p.sz * nz;
    if (nx > 0) {
      const int2 q = L.rec[0][nx - 1];
      c.XY = uni(q.x);
      c.XZ = uni(q.y);
    }
    if (ny > 0) {
      const int2 q = L.rec[1][ny - 1];
      c.YX = uni(q.x);
      c.YZ = uni(q.y);
    }
    if (nz > 0) {
      const int2 q = L.rec[2][nz - 1];
      c.ZX = uni(q.x);
      c.ZY = uni(q.y);
    }
    c.tmx = uni(L.seq[0][nx]);
    c.tmy = uni(L.seq[1][ny]);
    c.tmz = uni(L.seq[2][nz]);
    r.tent = uni(L.key[kWaveSteps - 1]);
    r.prev = uni(__shfl(f, kWaveSteps - 1));
    pidx = -1;
    tpos = r.tent;
    wave_lds_sync();  // every lane has read this round's records
    if constexpr (MODE == kModeRefine) {
      if (r.tent > io.walk_tmax) return 0;  // no crossing ahead can be used
    }
  }
}

// Optional work counters (COUNT): voxels visited by the walks and rays that
// hit, summed into the device state for the algorithmic-traffic figure.
__device__ __forceinline__ void add_counts(kf_dev_state* st, int visited, bool hit,
                                           int skipped = 0) {
  uint64_t v = (uint64_t)visited, sk = (uint64_t)skipped;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o), sk += __shfl_xor(sk, o);
  const uint64_t h = __popcll(__ballot(hit));
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&st->ray_steps, (unsigned long long)v);
    atomicAdd(&st->ray_hits, (unsigned long long)h);
    if (sk) atomicAdd(&st->ray_skipped, (unsigned long long)sk);
  }
}

// Adjoint: the wave's dL/dxi partials into the device state.
// Adjoint: the block's dL/dxi partial sum, written to its slot of part
// (6 doubles per block; summed by k_adjoint_finish in a fixed order).
__device__ __forceinline__ void block_xi(const double* xi, double* __restrict__ part) {
  __shared__ double s_xi[4][6];
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double v = xi[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) s_xi[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6)
    part[threadIdx.x] = ((s_xi[0][threadIdx.x] + s_xi[1][threadIdx.x]) + s_xi[2][threadIdx.x]) +
                        s_xi[3][threadIdx.x];
}

// Refinement: a pixel record {J[6], r, used} into the kGnSums running sums
// (21 J^TJ upper, 6 J^T r, count, r^2).
constexpr int kGnPix = 8;
__device__ __forceinline__ void gn_add_pixel(const float* px, float* acc) {
  if (px[7] == 0.0f) return;
  int q = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j) acc[q++] += px[i] * px[j];
#pragma unroll
  for (int i = 0; i < 6; ++i) acc[21 + i] += px[i] * px[6];
  acc[27] += 1.0f;
  acc[28] += px[6] * px[6];
}

// Refinement: the block's kGnSums partial sums (FP32) into its slot of part.
__device__ __forceinline__ void block_gn(const float* gn, float* __restrict__ part) {
  __shared__ float s_gn[4][kGnSums];
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kGnSums; ++q) {
    float v = gn[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) s_gn[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < kGnSums)
    part[threadIdx.x] = ((s_gn[0][threadIdx.x] + s_gn[1][threadIdx.x]) + s_gn[2][threadIdx.x]) +
                        s_gn[3][threadIdx.x];
}

// Pass 1: every pixel with depth != 0 casts its ray (voxel_traversal.cu:
// 75-424) for up to kPass1Batches batches.  Rays still walking after that
// (a few per cent: long walks through unobserved or free space, such as the
// rays of bilateral-filtered holes, whose depth of a few 1e-5 mm puts their
// start behind the camera) are appended to a queue and finished by
// k_raycast_rest, a whole wave per ray, instead of keeping a wave of short
// rays waiting on one long one.
//
// A 256-thread block covers 16x16 pixels as four 8x8 wave tiles, so a wave's
// rays walk neighbouring voxels.  Consecutive block ids are dealt round-robin
// to the 8 XCDs; the id is remapped so that each XCD gets a contiguous band
// of the image and its L2 sees the band's voxels.
#ifndef KF_PASS1_BATCHES
#define KF_PASS1_BATCHES 1
#endif
#ifndef KF_PASS1_STEPS
#define KF_PASS1_STEPS 8
#endif
constexpr int kPass1Batches = KF_PASS1_BATCHES;
constexpr int kPass1Steps = KF_PASS1_STEPS;  // steps per batch

// Second-pass grid: 8192 waves, one queued ray per wave at the few thousand
// rays a 640x480 frame queues.  The adjoint and refinement modes keep one
// partial-sum slot per block.
#ifndef KF_REST_BLOCKS
#define KF_REST_BLOCKS 2048
#endif
constexpr int kRestBlocks = KF_REST_BLOCKS;

// Per-block dL/dF table of the adjoint: kCells cell keys (-1 free) with 8
// corner sums each, and the bricks the block has marked in io.touched (a
// small open-addressing set, -1 free).
constexpr int kBrickSet = 64;

__device__ __forceinline__ void table_clear(int* key, float* val, int* brick) {
  for (int e = threadIdx.x; e < kCells; e += 256) key[e] = -1;
  for (int e = threadIdx.x; e < kCells * 8; e += 256) val[e] = 0.0f;
  if (threadIdx.x < kBrickSet) brick[threadIdx.x] = -1;
  __syncthreads();
}

// Adds the table to io.gF once every wave of the block is done with it:
// thread t of the 256 takes the (entry, corner) pairs t, t + 256, ...; the first
// thread of the block to meet a brick marks it in io.touched.
__device__ __forceinline__ void table_flush(const int* key, const float* val, int* brick,
                                            const RayIO& io) {
  __syncthreads();
  const kf_vol& vg = io.vg;
  for (int f = threadIdx.x; f < kCells * 8; f += 256) {
    const int e = f >> 3, q = f & 7;
    const int cell = key[e];
    const float v = val[f];
    if (cell < 0 || v == 0.0f) continue;
    const int kk = cell % vg.nz, c2 = cell / vg.nz;
    const int j = c2 % vg.ny, i = c2 / vg.ny;
    const int idx = kf_store_index(vg, i + (q >> 2), j + ((q >> 1) & 1), kk + (q & 1));
    const int b = idx >> 9;
    uint32_t h = (((uint32_t)b * 2654435761u) >> 26) & (kBrickSet - 1);
    bool seen = false, placed = false;
    for (int probe = 0; probe < kBrickSet && !seen && !placed; ++probe) {
      const int old = atomicCAS(&brick[h], -1, b);
      seen = old == b;
      placed = old == -1;
      h = (h + 1) & (kBrickSet - 1);
    }
    if (!seen) mark_brick(io.touched, b);
    atomicAdd(&io.gF[idx], v);
  }
}

// One pixel of the model maps: vertex and normal as packed 3-vectors, the
// hit index, and in trilinear mode the camera depth and crossing parameter
// the walk keeps in the w lanes (the adjoint starts from that te).
template <int MODE>
__device__ __forceinline__ void store_maps(kf_p3* __restrict__ out_v, kf_p3* __restrict__ out_n,
                                           float2* __restrict__ out_aux,
                                           int32_t* __restrict__ out_hit,
                                           float* __restrict__ out_nt, int p, const float4& ov,
                                           const float4& on, int hit) {
  out_v[p] = kf_p3{ov.x, ov.y, ov.z};
  out_n[p] = kf_p3{on.x, on.y, on.z};
  out_hit[p] = hit;
  if (out_nt) out_nt[p] = kf_inv_norm3(on.x, on.y, on.z);
  if constexpr (MODE == kModeTrilinear) out_aux[p] = make_float2(ov.w, on.w);
}

template <int PREC, bool COUNT, int MODE, bool MULTI>
__global__ __launch_bounds__(256) void k_raycast(
    const void* __restrict__ volume, RayIO io, int H, kf_dev_state* __restrict__ st,
    kf_p3* __restrict__ out_v, kf_p3* __restrict__ out_n, float2* __restrict__ out_aux,
    int32_t* __restrict__ out_hit, float* __restrict__ out_nt, RayState* __restrict__ queue,
    int tiles_x, int n_blocks) {
  const int bid = xcd_contiguous(blockIdx.x, n_blocks);
  const int W = io.W;
  const int tile = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int bx = bid % tiles_x, by = bid / tiles_x;
  const int x = bx * 16 + (tile & 1) * 8 + (ln & 7);
  const int y = by * 16 + (tile >> 1) * 8 + (ln >> 3);
  const bool active = x < W && y < H;
  const int p = y * W + x;

  static_assert(MODE != kModeAdjoint, "the adjoint runs in k_adjoint_saved");
  VolRead<PREC> F;
  F.v = reinterpret_cast<decltype(F.v)>(volume);
  const Sink sink{};  // adjoint only
  float4 ov = make_float4(0.f, 0.f, 0.f, 0.f), on = make_float4(0.f, 0.f, 0.f, 0.f);
  int hit = -1;
  bool pending = false;
  int visited = 0;
  double xi[6] = {0, 0, 0, 0, 0, 0};
  float gn[MODE == kModeRefine ? kGnPix : 1] = {};  // refinement: the pixel record
  RayState r;
  float T[12];
  load_uniform<12>(st->T, T);
  const float dval = active ? io.depth[p] : 0.0f;
  if (dval != 0) {
    if (ray_setup(x, y, dval, io.Ki, T, io.vg, io.trunc, r)) {
      r.pix = p;
      pending = !ray_walk<PREC, kPass1Steps, COUNT, MODE>(
          F, r, io, T,
          MODE == kModeRefine ? io.refine_batches : (MULTI ? io.pass1_batches : kPass1Batches),
          ov, on, hit, visited, xi, sink, gn);
    }
  }
  if constexpr (COUNT) add_counts(st, visited, hit >= 0);
  if constexpr (MODE == kModeRefine) {
    float acc[kGnSums] = {};
    gn_add_pixel(gn, acc);
    block_gn(acc, io.gn_part + kGnSums * blockIdx.x);
  }
  // wave-aggregated append of the unfinished rays to this block's partition
  const uint64_t mask = __ballot(pending);
  if (mask) {
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const int part = blockIdx.x & (KF_RAY_PARTS - 1);
    uint32_t base = 0;
    if (ln == leader)
      base = atomicAdd(&io.qcount[part * KF_RAY_COUNT_STRIDE], (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (pending)
      queue[(size_t)part * io.qcap + base + __popcll(mask & ((1ull << ln) - 1ull))] = r;
  }
  if (!active || MODE == kModeRefine) return;
  store_maps<MODE>(out_v, out_n, out_aux, out_hit, out_nt, p, ov, on, hit);
}

__device__ __forceinline__ void uniform_ray(RayState& r) {
  DdaState& c = r.cur;
  c.ix = uni(c.ix), c.iy = uni(c.iy), c.iz = uni(c.iz);
  c.XY = uni(c.XY), c.XZ = uni(c.XZ), c.YX = uni(c.YX);
  c.YZ = uni(c.YZ), c.ZX = uni(c.ZX), c.ZY = uni(c.ZY);
  c.tmx = uni(c.tmx), c.tmy = uni(c.tmy), c.tmz = uni(c.tmz);
  r.prev = uni(r.prev);
  r.dx = uni(r.dx), r.dy = uni(r.dy), r.dz = uni(r.dz);
  r.stp.sx = uni(r.stp.sx), r.stp.sy = uni(r.stp.sy), r.stp.sz = uni(r.stp.sz);
  r.stp.tdx = uni(r.stp.tdx), r.stp.tdy = uni(r.stp.tdy), r.stp.tdz = uni(r.stp.tdz);
  r.pix = uni(r.pix);
  r.tent = uni(r.tent);
}

// Pass 2: finishes the queued rays, one ray per wave (wave_walk).  Rays are
// numbered over the concatenated partitions and dealt to waves in a fixed
// order (wave w takes rays w, w + waves, ...), so that the ray loop has no
// lane-dependent branch ahead of its cross-lane steps and the long rays,
// which cluster in some partitions, spread over all waves.  Lane l holds
// partition l's count and its inclusive prefix sum; ray q lies in the
// partition p = #{l : incl_l <= q}, at offset q - (incl_p - count_p).
template <int PREC, bool COUNT, int MODE, bool SKIP>
__global__ __launch_bounds__(256) void k_raycast_rest(
    const void* __restrict__ volume, RayIO io, kf_dev_state* __restrict__ st,
    kf_p3* __restrict__ out_v, kf_p3* __restrict__ out_n, float2* __restrict__ out_aux,
    int32_t* __restrict__ out_hit, float* __restrict__ out_nt,
    const RayState* __restrict__ queue) {
  __shared__ WaveWalkLds lds[4];
  static_assert(KF_RAY_PARTS == 64, "one queue partition per lane");
  const int ln = threadIdx.x & 63;
  const uint32_t pc = io.qcount[ln * KF_RAY_COUNT_STRIDE];
  uint32_t incl = pc;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    incl += ln >= o ? t : 0u;
  }
  const uint32_t excl = incl - pc;
  const uint32_t n = __builtin_amdgcn_readlane(incl, 63);
  // the next raycast appends to the other slot, which nothing reads now
  if (blockIdx.x == 0 && threadIdx.x < 64) io.qnext[ln * KF_RAY_COUNT_STRIDE] = 0u;
  static_assert(MODE != kModeAdjoint, "the adjoint runs in k_adjoint_saved");
  // a block none of whose waves receives a ray (n is the same in every wave)
  // only writes zero partial sums
  if constexpr (MODE == kModeRefine) {
    if (blockIdx.x * 4u >= n) {
      if (threadIdx.x < kGnSums)
        io.gn_part[kGnSums * (io.part_base + blockIdx.x) + threadIdx.x] = 0.0f;
      return;
    }
  }
  VolRead<PREC> F;
  F.v = reinterpret_cast<decltype(F.v)>(volume);
  float T[12];
  load_uniform<12>(st->T, T);
  const int wave = uni((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const uint32_t n_waves = gridDim.x * 4;
  double xi[6] = {0, 0, 0, 0, 0, 0};
  float gn[MODE == kModeRefine ? kGnSums : 1] = {};
  for (uint32_t q = blockIdx.x * 4 + wave; q < n; q += n_waves) {
    const int part = uni((int)__popcll(__ballot(incl <= q)));
    const uint32_t off = q - (uint32_t)__builtin_amdgcn_readlane((int)excl, part);
    RayState r = queue[(size_t)part * io.qcap + off];
    uniform_ray(r);
    float4 ov = make_float4(0.f, 0.f, 0.f, 0.f), on = make_float4(0.f, 0.f, 0.f, 0.f);
    int hit = -1, visited = 0, skipped = 0;
    const Sink sink{io.gF, false, nullptr, nullptr, io.touched};
    float px[MODE == kModeRefine ? kGnPix : 1] = {};  // refinement: the ray's pixel record
    const int owner = wave_walk<PREC, COUNT, MODE, SKIP>(F, r, io, T, lds[wave], ov, on, hit,
                                                         visited, skipped, xi, sink, px);
    if constexpr (MODE == kModeRefine) gn_add_pixel(px, gn);
    if constexpr (MODE != kModeRefine) {
      if (lane == owner)
        store_maps<MODE>(out_v, out_n, out_aux, out_hit, out_nt, r.pix, ov, on, hit);
    }
    if constexpr (COUNT) add_counts(st, visited, hit >= 0, skipped);
  }
  if constexpr (MODE == kModeRefine)
    block_gn(gn, io.gn_part + kGnSums * (io.part_base + blockIdx.x));
}

// The adjoint, from the crossings of a trilinear raycast at the same pose,
// camera, depth and volume: the walk is deterministic in those inputs, so
// each pixel whose raycast hit (hit index >= 0) runs tri_hit at the te that
// raycast saved in its normal map's w, with no walk.  Same tiles and XCD
// order as the raycast's first pass; dL/dF through the block's cell table,
// dL/dxi into the block's slot.
template <int PREC>
__global__ __launch_bounds__(256) void k_adjoint_saved(
    const void* __restrict__ volume, RayIO io, int H, kf_dev_state* __restrict__ st,
    const float2* __restrict__ model_aux, const int32_t* __restrict__ hit_map, int tiles_x,
    int n_blocks) {
  const int bid = xcd_contiguous(blockIdx.x, n_blocks);
  const int tile = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int bx = bid % tiles_x, by = bid / tiles_x;
  const int x = bx * 16 + (tile & 1) * 8 + (ln & 7);
  const int y = by * 16 + (tile >> 1) * 8 + (ln >> 3);
  const bool active = x < io.W && y < H;
  const int p = y * io.W + x;
  VolRead<PREC> F;
  F.v = reinterpret_cast<decltype(F.v)>(volume);
  __shared__ int s_key[kCells];
  __shared__ float s_val[kCells * 8];
  __shared__ int s_brick[kBrickSet];
  const bool use_table = io.gF != nullptr;
  if (use_table) table_clear(s_key, s_val, s_brick);
  const Sink sink{io.gF, use_table, (lds_int*)s_key, (lds_float*)s_val, io.touched};
  double xi[6] = {0, 0, 0, 0, 0, 0};
  float T[12];
  load_uniform<12>(st->T, T);
  const int h = active ? hit_map[p] : -1;
  const float te = active ? model_aux[p].y : 0.0f;
  if (h >= 0) {
    const float* Ki = io.Ki;
    const float X = (float)x + 1.0f, Y = (float)y + 1.0f;
    const float lamx = Ki[0] * X + Ki[1] * Y + Ki[2] * 1.f;
    const float lamy = Ki[3] * X + Ki[4] * Y + Ki[5] * 1.f;
    const float lamz = Ki[6] * X + Ki[7] * Y + Ki[8] * 1.f;
    float dx, dy, dz;
    ray_direction(lamx, lamy, lamz, T, dx, dy, dz);
    float4 ov, on;
    float gn[1] = {};
    tri_hit<PREC, kModeAdjoint>(F, io, T, p, te, dx, dy, dz, ov, on, xi, sink, gn);
  }
  block_xi(xi, io.xi_part + 6 * blockIdx.x);
  if (use_table) table_flush(s_key, s_val, s_brick, io);
}

// Zeroes dL/dF in the bricks the previous adjoint wrote and clears their
// bits: one wave per 32-brick word, 2 float4 per lane per brick.
__global__ __launch_bounds__(256) void k_zero_touched(float4* __restrict__ gF,
                                                      uint32_t* __restrict__ touched,
                                                      int n_words) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= n_words) return;
  uint32_t bits = touched[w];
  if (!bits) return;
  while (bits) {
    const int b = w * 32 + __ffs(bits) - 1;
    bits &= bits - 1;
    float4* p = gF + (size_t)b * 128;
    p[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    p[lane + 64] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (lane == 0) touched[w] = 0u;
}

// dL/dxi: the FP64 sums of the block partials in a fixed order (thread t adds
// blocks t, t + 256, ..., then a tree over the 256 threads).
// Sums the block partials in a fixed order: thread t adds partials t,
// t + 1024, ... (loaded 4 at a time, clamped, so that they are in flight
// together), then a DPP tree per wave and the 16 wave sums in order.
constexpr int kAdjFinishThreads = 1024;

__global__ __launch_bounds__(kAdjFinishThreads) void k_adjoint_finish(
    const double* __restrict__ part, int n_part, kf_dev_state* st) {
  constexpr int NT = kAdjFinishThreads, NW = NT / 64;
  __shared__ double red[NW][6];
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int b0 = 0; b0 < n_part; b0 += 4 * NT) {
    double t[4][6];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = min(b0 + u * NT + (int)threadIdx.x, n_part - 1);
#pragma unroll
      for (int q = 0; q < 6; ++q) t[u][q] = part[6 * b + q];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (b0 + u * NT + (int)threadIdx.x < n_part)
#pragma unroll
        for (int q = 0; q < 6; ++q) v[q] += t[u][q];
  }
  const int wave = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double w = wave_total(v[q]);
    if (ln == 63) red[wave][q] = w;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double sum = red[0][threadIdx.x];
    for (int w = 1; w < NW; ++w) sum += red[w][threadIdx.x];
    st->xi_grad[threadIdx.x] = sum;
  }
}


// One Gauss-Newton step of the pose refinement: the FP64 sums of the block
// partials, the 6x6 solve (LDLT, FP64) and T <- exp(-x^) T in FP64, with
// x = (omega, v) the solution of (J^T J) x = J^T r.  Records the mean squared
// residual before the step, the pixel count and the solver status.
constexpr int kRefineFinishThreads = 512;

__global__ __launch_bounds__(kRefineFinishThreads) void k_refine_finish(
    const float* __restrict__ part, int n_first, const uint32_t* __restrict__ qcount,
    kf_dev_state* st, int step) {
  // The first pass's n_first block slots, then the second pass's: only its
  // first ceil(queued / 4) blocks received rays (one per wave, four waves a
  // block); the others wrote zeros, which are not read.
  __shared__ double red[kGnSums][kRefineFinishThreads];
  __shared__ double s[kGnSums];
  __shared__ uint32_t s_queued;
  if (threadIdx.x < 64) {
    uint32_t n = qcount[threadIdx.x * KF_RAY_COUNT_STRIDE];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if (threadIdx.x == 0) s_queued = n;
  }
  __syncthreads();
  const int n_part = n_first + (int)min((s_queued + 3u) / 4u, (uint32_t)kRestBlocks);
  // FP64 sums of the partials in a fixed order: thread t adds blocks t,
  // t + kRefineFinishThreads, ..., then a tree over the threads
  {
    double v[kGnSums];
#pragma unroll
    for (int q = 0; q < kGnSums; ++q) v[q] = 0.0;
    for (int b = threadIdx.x; b < n_part; b += kRefineFinishThreads)
#pragma unroll
      for (int q = 0; q < kGnSums; ++q) v[q] += (double)part[kGnSums * b + q];
#pragma unroll
    for (int q = 0; q < kGnSums; ++q) red[q][threadIdx.x] = v[q];
  }
  __syncthreads();
  for (int w = kRefineFinishThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < kGnSums; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kGnSums) s[threadIdx.x] = red[threadIdx.x][0];
  __syncthreads();
  if (threadIdx.x != 0) return;
  double A[36], rhs[6], x[6] = {0, 0, 0, 0, 0, 0};
  int k = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j, ++k) A[i * 6 + j] = A[j * 6 + i] = s[k];
  for (int i = 0; i < 6; ++i) rhs[i] = s[21 + i];
  int status = s[27] >= 6.0 ? kf_ldlt6_solve(A, rhs, x) : 2;
  if (status != 0)
    for (int i = 0; i < 6; ++i) x[i] = 0.0;
  // exp of the twist (-x): Rodrigues rotation and the left Jacobian V
  const double w0 = -x[0], w1 = -x[1], w2 = -x[2];
  const double th2 = w0 * w0 + w1 * w1 + w2 * w2, th = sqrt(th2);
  double a, b, c;
  if (th < 1e-9) {
    a = 1.0 - th2 / 6.0;
    b = 0.5 - th2 / 24.0;
    c = 1.0 / 6.0 - th2 / 120.0;
  } else {
    a = sin(th) / th;
    b = (1.0 - cos(th)) / th2;
    c = (th - sin(th)) / (th2 * th);
  }
  const double W[9] = {0.0, -w2, w1, w2, 0.0, -w0, -w1, w0, 0.0};
  double W2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      W2[i * 3 + j] = W[i * 3] * W[j] + W[i * 3 + 1] * W[3 + j] + W[i * 3 + 2] * W[6 + j];
  double E[16] = {0};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) E[i * 4 + j] = (i == j ? 1.0 : 0.0) + a * W[i * 3 + j] + b * W2[i * 3 + j];
    double t = 0.0;
    for (int j = 0; j < 3; ++j)
      t += ((i == j ? 1.0 : 0.0) + b * W[i * 3 + j] + c * W2[i * 3 + j]) * (-x[3 + j]);
    E[i * 4 + 3] = t;
  }
  E[15] = 1.0;
  float Tn[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double v = 0.0;
      for (int m = 0; m < 4; ++m) v += E[i * 4 + m] * (double)st->T[m * 4 + j];
      Tn[i * 4 + j] = (float)v;
    }
  float Ti[16];
  if (kf_m4_inverse(Tn, Ti) == 0) {
    for (int i = 0; i < 16; ++i) {
      st->T[i] = Tn[i];
      st->Tinv[i] = Ti[i];
    }
  } else {
    status = 3;
  }
  if (step < KF_MAX_REFINE_STEPS) {
    st->refine_msr[step] = s[28] / (s[27] > 0.0 ? s[27] : 1.0);
    st->refine_count[step] = (int32_t)s[27];
    st->refine_status[step] = status;
  }
}
}  // namespace

// ===========================================================================
// launchers

hipError_t kf_launch_depth_u16(kf_context* c, const uint16_t* src) {
  const int n = c->H * c->W;
  k_depth_u16<<<(n + 255) / 256, 256, 0, c->stream>>>(src, c->depth, n, c->prm.depth_min_mm,
                                                      c->prm.depth_max_mm);
  return hipGetLastError();
}

// src: nullptr to filter the working depth, else a uint16 frame to gate and
// filter in one pass.
static hipError_t launch_bilateral(kf_context* c, const uint16_t* src) {
  const dim3 grid((c->W + 15) / 16, (c->H + 15) / 16);
  const float sr = c->prm.sigma_r_mm, lo = c->prm.depth_min_mm, hi = c->prm.depth_max_mm;
  const void* in = src ? static_cast<const void*>(src) : static_cast<const void*>(c->depth);
  switch (c->prm.bilateral_radius) {
#define KF_BIL(R)                                                                              \
  case R:                                                                                      \
    if (src)                                                                                   \
      k_bilateral<R, true><<<grid, 256, 0, c->stream>>>(in, c->depth_tmp, c->spatial, c->H,    \
                                                        c->W, sr, lo, hi, c->range_lut,        \
                                                        c->range_lut_n);                       \
    else                                                                                       \
      k_bilateral<R, false><<<grid, 256, 0, c->stream>>>(in, c->depth_tmp, c->spatial, c->H,   \
                                                         c->W, sr, lo, hi, nullptr, 0);        \
    break;
    KF_BIL(1)
    KF_BIL(2)
    KF_BIL(3)
    KF_BIL(4)
    KF_BIL(5)
    KF_BIL(6)
    KF_BIL(7)
    KF_BIL(8)
#undef KF_BIL
    default:
      return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  float* t = c->depth;  // the filtered image becomes the working depth
  c->depth = c->depth_tmp;
  c->depth_tmp = t;
  return hipSuccess;
}

hipError_t kf_launch_bilateral(kf_context* c) { return launch_bilateral(c, nullptr); }

hipError_t kf_launch_depth_bilateral(kf_context* c, const uint16_t* src) {
  return launch_bilateral(c, src);
}

hipError_t kf_launch_vertex(kf_context* c, const kf_cam& cam) {
  const int n = c->H * c->W;
  k_vertex<<<(n + 255) / 256, 256, 0, c->stream>>>(c->depth, cam, c->H, c->W, c->vmap, c->nmap,
                                                   c->mask);
  return hipGetLastError();
}

// One level of the ICP: its image, grid, camera and maps, and iterations.
// depth set: the first iteration computes the frame's maps from it and
// writes them to vmap_w / nmap_w / mask_w (the pipeline's level 0).
struct IcpLevel {
  int H, W, blocks, bpx, px, threads, iters;
  kf_cam cam;
  const kf_p3 *vmap, *nmap, *mv, *mn;
  const float* mt;  // kf_inv_norm3 of mn (null: the kernel normalises)
  const float* depth;
  kf_p3 *vmap_w, *nmap_w;
  uint8_t* mask_w;
};

template <bool KAHAN, int PX, int PPL>
static hipError_t launch_icp_iter_px(kf_context* c, const IcpLevel& L, const void* pin,
                                     void* pout, uint8_t* m, int it, int nb_in, float min_inl,
                                     bool from_depth, bool may_stop) {
  const float* d = from_depth ? L.depth : nullptr;
  kf_p3* vw = from_depth ? L.vmap_w : nullptr;
  kf_p3* nw = from_depth ? L.nmap_w : nullptr;
  uint8_t* mw = from_depth ? L.mask_w : nullptr;
#define KF_ICP_ITER_B(FD, ST, NT, B)                                                          \
  k_icp_iter<KAHAN, PX, PPL, FD, ST, NT, B><<<L.blocks, B, 0, c->stream>>>(                    \
      d, vw, nw, mw, L.vmap, L.nmap, L.mv, L.mn, L.mt, L.cam, L.H, L.W, c->icp_dist2_hi,       \
      c->icp_cos_lo, min_inl, c->state, pin, pout, m, it, L.bpx, nb_in)
#define KF_ICP_ITER(FD, ST, NT) KF_ICP_ITER_B(FD, ST, NT, KF_ICP_BLOCK)
  if (L.threads == 2 * KF_ICP_BLOCK) {
    // the large-image grid (icp_grid): 3 or 4 pixels per thread, at most one
    // block per compute unit (so 4 partials per lane cover the grid)
    if constexpr ((PX == 3 || PX == 4) && PPL == 4) {
      if (!L.mt) {
        if (may_stop)
          KF_ICP_ITER_B(false, true, false, 2 * KF_ICP_BLOCK);
        else
          KF_ICP_ITER_B(false, false, false, 2 * KF_ICP_BLOCK);
      } else if (may_stop) {
        if (from_depth)
          KF_ICP_ITER_B(true, true, true, 2 * KF_ICP_BLOCK);
        else
          KF_ICP_ITER_B(false, true, true, 2 * KF_ICP_BLOCK);
      } else {
        if (from_depth)
          KF_ICP_ITER_B(true, false, true, 2 * KF_ICP_BLOCK);
        else
          KF_ICP_ITER_B(false, false, true, 2 * KF_ICP_BLOCK);
      }
      return hipGetLastError();
    }
    return hipErrorInvalidConfiguration;  // icp_grid gives no other such grid
  }
  if (!L.mt) {
    // caller's model maps (the allocation-free entry point): no depth maps
    if (may_stop)
      KF_ICP_ITER(false, true, false);
    else
      KF_ICP_ITER(false, false, false);
  } else if (may_stop) {
    if (from_depth)
      KF_ICP_ITER(true, true, true);
    else
      KF_ICP_ITER(false, true, true);
  } else {
    if (from_depth)
      KF_ICP_ITER(true, false, true);
    else
      KF_ICP_ITER(false, false, true);
  }
#undef KF_ICP_ITER
#undef KF_ICP_ITER_B
  return hipGetLastError();
}

// The iterations of every level, coarsest first (lv[n_levels - 1] ... lv[0]),
// numbered 0, 1, ... across levels (the per-iteration record), then the
// finishing kernel.  Each iteration finishes the one before, whatever its
// level, from that one's partials (grid nb_in).
template <bool KAHAN, int PPL>
static hipError_t launch_icp_levels(kf_context* c, const IcpLevel* lv, int n_levels) {
  const int n = c->H * c->W;
  // one parity of partials: KF_ICP_SUMS doubles, or 2 * KF_ICP_SUMS floats
  // (compensated), per block of the largest grid (level 0)
  const size_t half = (size_t)KF_ICP_SUMS * 8 * c->icp_blocks;
  char* const pbase = reinterpret_cast<char*>(c->icp_partials);
  // robust update: the inlier threshold (>= 0); the reference's update: -1
  const float min_inl = c->prm.icp_robust ? (float)c->prm.icp_min_inliers : -1.0f;
  int total = 0;
  for (int l = 0; l < n_levels; ++l) total += lv[l].iters;
  // opt-in early stop (at full resolution only; its thresholds and first
  // iteration are in the device state, kf_icp_stop_from)
  const bool may_stop = kf_icp_may_stop(c->prm);
  int it = 0, nb_prev = lv[n_levels - 1].blocks;
  for (int l = n_levels - 1; l >= 0; --l) {
    const IcpLevel& L = lv[l];
    for (int k = 0; k < L.iters; ++k, ++it) {
      const bool fd = k == 0 && L.depth != nullptr;  // maps from the depth
      uint8_t* m =
          (c->prm.record_icp && c->icp_masks) ? c->icp_masks + (size_t)it * n : nullptr;
      const void* pin = pbase + (size_t)((it + 1) & 1) * half;
      void* pout = pbase + (size_t)(it & 1) * half;
      if (it == 0) kf_stage_begin(c, KF_STAGE_ICP_ITER);
      hipError_t e;
      switch (L.px) {
        case 1:
          e = launch_icp_iter_px<KAHAN, 1, PPL>(c, L, pin, pout, m, it, nb_prev, min_inl, fd,
                                                may_stop);
          break;
        case 2:
          e = launch_icp_iter_px<KAHAN, 2, PPL>(c, L, pin, pout, m, it, nb_prev, min_inl, fd,
                                                may_stop);
          break;
        case 3:
          e = launch_icp_iter_px<KAHAN, 3, PPL>(c, L, pin, pout, m, it, nb_prev, min_inl, fd,
                                                may_stop);
          break;
        default:
          e = launch_icp_iter_px<KAHAN, 4, PPL>(c, L, pin, pout, m, it, nb_prev, min_inl, fd,
                                                may_stop);
          break;
      }
      if (e != hipSuccess) return e;
      if (it == total - 1) kf_stage_end(c, KF_STAGE_ICP_ITER);
      nb_prev = L.blocks;
    }
  }
  if (total > 0) {
    const void* pin = pbase + (size_t)((total - 1) & 1) * half;
    if (may_stop)
      k_icp_finish<KAHAN, PPL, true><<<1, kIcpFinishThreads, 0, c->stream>>>(
          c->state, pin, nb_prev, total - 1, min_inl);
    else
      k_icp_finish<KAHAN, PPL, false><<<1, kIcpFinishThreads, 0, c->stream>>>(
          c->state, pin, nb_prev, total - 1, min_inl);
  }
  return hipGetLastError();
}

// Intrinsics of the next coarser level.  A pixel u (1-indexed) of the coarse
// image covers the fine pixels 2u - 1 and 2u, whose centre is at fine
// coordinate 2u - 0.5: u_coarse = (u_fine + 0.5) / 2, so fx and the skew
// halve and cx becomes (cx + 0.5) / 2 (likewise in y).
void kf_pyramid_intrinsics(const float* K, float* Kc) {
  for (int i = 0; i < 9; ++i) Kc[i] = K[i];
  Kc[0] = K[0] * 0.5f;
  Kc[1] = K[1] * 0.5f;
  Kc[2] = (K[2] + 0.5f) * 0.5f;
  Kc[4] = K[4] * 0.5f;
  Kc[5] = (K[5] + 0.5f) * 0.5f;
}

// Levels 1 .. icp_levels-1 from level 0: depth, vertex and normal maps of
// the frame, and the model maps (from the raycast's maps and hit index).
static hipError_t build_pyramid(kf_context* c, const kf_cam& cam0, IcpLevel* lv) {
  const float step = c->prm.pyramid_max_step_mm;
  float K[9];
  for (int i = 0; i < 9; ++i) K[i] = cam0.K[i];
  for (int l = 1; l < c->prm.icp_levels; ++l) {
    kf_pyr_level& P = c->pyr[l];
    const int Wi = l == 1 ? c->W : c->pyr[l - 1].W;
    const float* d_in = l == 1 ? c->depth : c->pyr[l - 1].depth;
    const kf_p3* v_in = l == 1 ? c->model_v : c->pyr[l - 1].mv;
    const kf_p3* n_in = l == 1 ? c->model_n : c->pyr[l - 1].mn;
    const int nl = P.H * P.W, grid = (nl + 255) / 256;
    k_depth_down<<<grid, 256, 0, c->stream>>>(d_in, Wi, P.depth, P.H, P.W, step);
    float Kl[9];
    kf_pyramid_intrinsics(K, Kl);
    for (int i = 0; i < 9; ++i) K[i] = Kl[i];
    kf_cam cam;
    kf_make_cam(Kl, &cam);
    k_vertex<<<grid, 256, 0, c->stream>>>(P.depth, cam, P.H, P.W, P.vmap, P.nmap, P.mask);
    k_model_down<<<grid, 256, 0, c->stream>>>(v_in, n_in, l == 1 ? c->hit : nullptr,
                                              l == 1 ? nullptr : c->pyr[l - 1].ok, Wi, P.mv,
                                              P.mn, P.mt, P.ok, P.H, P.W);
    lv[l] = IcpLevel{P.H,   P.W,    P.blocks, P.bpx, P.px, P.threads, c->prm.icp_coarse_iters[l - 1],
                     cam,   P.vmap, P.nmap,   P.mv,  P.mn,    P.mt,
                     nullptr, nullptr, nullptr, nullptr};
  }
  return hipGetLastError();
}

hipError_t kf_launch_icp(kf_context* c, const kf_cam& cam, bool maps_from_depth) {
  IcpLevel lv[KF_MAX_ICP_LEVELS];
  lv[0] = IcpLevel{c->H,    c->W,    c->icp_blocks, c->icp_bpx,
                   c->icp_px, c->icp_threads, c->prm.icp_iterations, cam, c->vmap,
                   c->nmap, c->model_v, c->model_n, c->model_nt,
                   maps_from_depth ? c->depth : nullptr, c->vmap, c->nmap, c->mask};
  const int n_levels = c->prm.icp_levels > 1 ? c->prm.icp_levels : 1;
  if (n_levels > 1) {
    if (!c->pyr[n_levels - 1].depth) return hipErrorInvalidValue;
    hipError_t e = build_pyramid(c, cam, lv);
    if (e != hipSuccess) return e;
  }
  // partials per lane of the iteration prologue: 4 covers up to 256 blocks
  // (one per compute unit on this GPU), 8 the maximum block count; level 0
  // has the most blocks
  const bool kahan = c->prm.icp_compensated != 0;
  if (c->icp_blocks <= 4 * 64)
    return kahan ? launch_icp_levels<true, 4>(c, lv, n_levels)
                 : launch_icp_levels<false, 4>(c, lv, n_levels);
  return kahan ? launch_icp_levels<true, KF_ICP_MAX_PART_PER_LANE>(c, lv, n_levels)
               : launch_icp_levels<false, KF_ICP_MAX_PART_PER_LANE>(c, lv, n_levels);
}

static int integ_groups(const kf_vol& vg, int* gx, int* gy, int* gz) {
  *gx = (vg.bx + 1) / 2;
  *gy = (vg.by + 1) / 2;
  *gz = (vg.bz + 1) / 2;
  return *gx * *gy * *gz;
}

// group-cull blocks of k_integ_table (32 groups each) and list entries per
// partition
static int integ_group_blocks(const kf_vol& vg) {
  int gx, gy, gz;
  return (integ_groups(vg, &gx, &gy, &gz) + 31) / 32;
}

// Grids of the brick cull (one listed group per wave) and of the brick
// update (one listed brick per wave); both loop over longer lists.
constexpr int kIntegCullBlocks = 512;
constexpr int kIntegBlocks = 2048;

// list entries per partition: the groups of the group-cull blocks that
// append to it, and the bricks of the brick-cull blocks that append to it
static int integ_group_cap(const kf_vol& vg) {
  return (integ_group_blocks(vg) + KF_RAY_PARTS - 1) / KF_RAY_PARTS * 32;
}

static_assert(kIntegCullBlocks * 4 % KF_RAY_PARTS == 0 && kIntegBlocks * 4 % KF_RAY_PARTS == 0,
              "list consumers need whole waves per partition");

static int integ_brick_cap(const kf_vol& vg) {
  // a brick-cull wave takes at most ceil(gcap / waves per partition) groups
  // of its partition, 8 bricks each; blocks b, b + 64, ... append to
  // partition b % 64
  const int per_part = kIntegCullBlocks * 4 / KF_RAY_PARTS;
  const int per_wave = (integ_group_cap(vg) + per_part - 1) / per_part;
  return (kIntegCullBlocks + KF_RAY_PARTS - 1) / KF_RAY_PARTS * 4 * per_wave * 8;
}

size_t kf_integ_group_entries(const kf_vol& vg) {
  return (size_t)KF_RAY_PARTS * (integ_group_cap(vg) + integ_brick_cap(vg));
}

// Whether lc . K^-1 (u, v, 1) > 0 over the image (at its four corners, with
// a margin), lc = K^-1 (K[5], K[2], 1): the condition under which the box
// culls may apply the image-side tests to boxes that reach behind the
// camera plane (see integ_corner).
static bool integ_side_cull_ok(const kf_cam& cam, int H, int W) {
  const float* Ki = cam.Kinv;
  const float* K = cam.K;
  double lc[3];
  for (int r = 0; r < 3; ++r)
    lc[r] = (double)Ki[3 * r] * K[5] + (double)Ki[3 * r + 1] * K[2] + (double)Ki[3 * r + 2];
  const double us[2] = {0.5, W + 0.5}, vs[2] = {0.5, H + 0.5};
  for (double u : us)
    for (double v : vs) {
      double dot = 0.0;
      for (int r = 0; r < 3; ++r)
        dot += lc[r] * ((double)Ki[3 * r] * u + (double)Ki[3 * r + 1] * v + (double)Ki[3 * r + 2]);
      const double nl = sqrt(lc[0] * lc[0] + lc[1] * lc[1] + lc[2] * lc[2]);
      double nr = 0.0;
      for (int r = 0; r < 3; ++r) {
        const double q = (double)Ki[3 * r] * u + (double)Ki[3 * r + 1] * v + (double)Ki[3 * r + 2];
        nr += q * q;
      }
      if (!(dot > 0.05 * nl * sqrt(nr))) return false;
    }
  return true;
}

// The lists and their counters: integ_groups holds the group list, then the
// brick list; integ_counts holds [group, brick] x [slot 0, slot 1] counter
// blocks of KF_RAY_PARTS lines.
static GroupList make_group_list(const kf_context* c, const kf_cam& cam) {
  GroupList gl;
  gl.any_side = integ_side_cull_ok(cam, c->H, c->W) ? 1 : 0;
  integ_groups(c->vg, &gl.gx, &gl.gy, &gl.gz);
  const size_t lines = (size_t)KF_RAY_PARTS * KF_RAY_COUNT_STRIDE;
  const int s = c->integ_slot;
  gl.gcap = integ_group_cap(c->vg);
  gl.bcap = integ_brick_cap(c->vg);
  gl.glist = c->integ_groups;
  gl.blist = c->integ_groups + (size_t)KF_RAY_PARTS * gl.gcap;
  gl.gcount = c->integ_counts + (size_t)s * lines;
  gl.gnext = c->integ_counts + (size_t)(s ^ 1) * lines;
  gl.bcount = c->integ_counts + (size_t)(2 + s) * lines;
  gl.bnext = c->integ_counts + (size_t)(2 + (s ^ 1)) * lines;
  return gl;
}

hipError_t kf_launch_integ_table(kf_context* c, const kf_cam& cam) {
  const int tiles_x = (c->W + 15) / 16, tiles_y = (c->H + 15) / 16;
  const int group_rows = (integ_group_blocks(c->vg) + tiles_x - 1) / tiles_x;
  const dim3 grid(tiles_x, tiles_y + group_rows);
  k_integ_table<<<grid, 256, 0, c->stream>>>(c->depth, cam, c->H, c->W, c->state, c->integ_tab,
                                             c->integ_tiles, c->prm.sdf_min_mm,
                                             c->prm.sdf_max_mm, tiles_y, c->vg,
                                             make_group_list(c, cam));
  return hipGetLastError();
}

hipError_t kf_launch_integrate(kf_context* c, const kf_cam& cam, int first) {
  const kf_params& p = c->prm;
  const GroupList gl = make_group_list(c, cam);
  // the skip flags are kept while raycast_skip is on; otherwise the first
  // raycast that skips rebuilds them
  uint8_t* flags = p.raycast_skip ? c->skip_flags : nullptr;
  if (!flags) c->skip_flags_stale = true;
  c->integ_slot ^= 1;
  k_integ_cull<<<kIntegCullBlocks, 256, 0, c->stream>>>(cam, c->vg, c->H, c->W, c->state, gl,
                                                         c->integ_tiles, (c->W + 15) / 16);
#define KF_INTEG(P, FI, FL)                                                               \
  k_integrate<P, FI, FL><<<kIntegBlocks, 256, 0, c->stream>>>(                              \
      c->tsdf, c->weight, c->integ_tab, cam, c->vg, c->H, c->W, p.truncation_mm,            \
      p.sdf_min_mm, p.sdf_max_mm, c->state, gl, flags)
#define KF_INTEG_P(P)              \
  do {                             \
    if (first && flags)            \
      KF_INTEG(P, true, true);     \
    else if (first)                \
      KF_INTEG(P, true, false);    \
    else if (flags)                \
      KF_INTEG(P, false, true);    \
    else                           \
      KF_INTEG(P, false, false);   \
  } while (0)
  if (c->vol.precision == KF_TSDF_F32)
    KF_INTEG_P(KF_TSDF_F32);
  else
    KF_INTEG_P(KF_TSDF_F16);
#undef KF_INTEG_P
#undef KF_INTEG
  return hipGetLastError();
}

// First-pass batches: enough 8-step batches for the walk from the skip point
// (depth - trunc) to the surface, about 1.3 trunc / d DDA steps on oblique
// rays plus the crossing, so that few rays are left for the second pass (one
// wave per ray).  Measured: 640x480 at trunc 30 mm / 10 mm voxels, one batch
// (two: -2.6 % frames/s); 1280x720 at 20 mm / 4 mm, two batches (+77 %
// frames/s of ICP-only tracking over one).  Either split walks the same
// voxels and gives the same results.
static int pass1_batches(const kf_context* c) {
  const double steps = 1.3 * c->prm.truncation_mm / c->vg.d + 2.0;
  const int b = (int)std::ceil(steps / kPass1Steps);
  return std::max(kPass1Batches, std::min(8, b));
}

static RayIO make_ray_io(const kf_context* c, const kf_cam& cam) {
  RayIO io;
  io.vg = c->vg;
  io.fix_bounds = c->prm.raycast_fix_bounds;
  io.flip_hi = c->flip_cos_hi;
  io.trunc = c->prm.truncation_mm;
  io.W = c->W;
  for (int i = 0; i < 9; ++i) io.Ki[i] = cam.Kinv[i];
  io.depth = c->depth;
  io.gD = c->adj_depth;
  io.gN = c->adj_normal_set ? c->adj_normal : nullptr;
  io.gF = c->adj_want_tsdf ? c->grad_tsdf : nullptr;
  io.touched = c->adj_touched;
  io.xi_part = c->adj_partials;
  io.part_base = 0;
  io.gn_part = c->gn_partials;
  io.max_res = c->prm.refine_max_residual_mm;
  io.walk_tmax = INFINITY;
  io.refine_batches = 1;
  io.pass1_batches = pass1_batches(c);
  io.skip_flags = nullptr;
  return io;
}

// Walk limit of the refinement.  A crossing found at voxel-entry parameter te
// (mm along the ray from the skip point) has t* >= te - 3d (the samples start
// at floor(te/d) - 2 voxels), and its residual is (t* - trunc) / |lambda|
// with |lambda| = |K^-1 (x+1, y+1, 1)|.  So beyond
//   te > trunc + max|lambda| * max_res + 3d + margin
// every residual exceeds max_res and the pixel is not used.  |lambda| is
// convex over the image, so its maximum is at a corner; the margin (2d +
// 1 mm) covers the FP32 rounding of the residual.
static float refine_walk_tmax(const kf_context* c, const kf_cam& cam) {
  const float* Ki = cam.Kinv;
  double lam = 0.0;
  for (int cy = 0; cy < 2; ++cy)
    for (int cx = 0; cx < 2; ++cx) {
      const double X = cx ? c->W : 1.0, Y = cy ? c->H : 1.0;
      const double lx = Ki[0] * X + Ki[1] * Y + Ki[2];
      const double ly = Ki[3] * X + Ki[4] * Y + Ki[5];
      const double lz = Ki[6] * X + Ki[7] * Y + Ki[8];
      lam = std::max(lam, std::sqrt(lx * lx + ly * ly + lz * lz));
    }
  const double d = c->vg.d, res = c->prm.refine_max_residual_mm;
  if (!(res >= 0.0) || !std::isfinite(res)) return INFINITY;
  return (float)(c->prm.truncation_mm + lam * res + 5.0 * d + 1.0);
}

// First-pass batches of the refinement: a walk stopped at walk_tmax has taken
// at most sqrt(3) tmax / d + 3 steps (a segment of length L crosses at most
// (|dx| + |dy| + |dz|) L / d + 3 voxel boundaries), so with that many steps
// every ray finishes in the first pass, one lane each, and the second pass's
// queue stays empty.  Capped at 32 batches (an unbounded walk keeps the
// default single batch and the second pass).
static int refine_pass1_batches(const kf_context* c, float tmax) {
  if (!std::isfinite(tmax)) return kPass1Batches;
  const double steps = std::sqrt(3.0) * std::max(0.0, (double)tmax) / c->vg.d + 4.0;
  const int b = (int)std::ceil(steps / kPass1Steps);
  return std::max(kPass1Batches, std::min(32, b));
}


// Model-map outputs of a raycast (the context's, or the adjoint's own).
struct RayMaps {
  kf_p3* v;
  kf_p3* n;
  float2* aux;
  int32_t* hit;
  float* nt;  // kf_inv_norm3 of each normal (null: not kept)
};

template <int PREC, bool COUNT, int MODE>
static void launch_raycast_t(kf_context* c, const RayIO& io, int tiles_x, int nb,
                             const RayMaps& m) {
  RayState* q = reinterpret_cast<RayState*>(c->ray_queue);
  if (MODE != kModeRefine && io.pass1_batches > kPass1Batches)
    k_raycast<PREC, COUNT, MODE, true><<<nb, 256, 0, c->stream>>>(
        c->tsdf, io, c->H, c->state, m.v, m.n, m.aux, m.hit, m.nt, q, tiles_x, nb);
  else
    k_raycast<PREC, COUNT, MODE, false><<<nb, 256, 0, c->stream>>>(
        c->tsdf, io, c->H, c->state, m.v, m.n, m.aux, m.hit, m.nt, q, tiles_x, nb);
  RayIO io2 = io;
  io2.part_base = nb;
  if (io.skip_flags)
    k_raycast_rest<PREC, COUNT, MODE, true><<<kRestBlocks, 256, 0, c->stream>>>(
        c->tsdf, io2, c->state, m.v, m.n, m.aux, m.hit, m.nt, q);
  else
    k_raycast_rest<PREC, COUNT, MODE, false><<<kRestBlocks, 256, 0, c->stream>>>(
        c->tsdf, io2, c->state, m.v, m.n, m.aux, m.hit, m.nt, q);
}

template <int PREC>
static void launch_raycast_p(kf_context* c, const RayIO& io, int tiles_x, int nb, bool count,
                             int mode, const RayMaps& m) {
  if (mode == kModeRefine)
    launch_raycast_t<PREC, false, kModeRefine>(c, io, tiles_x, nb, m);
  else if (mode == kModeTrilinear)
    count ? launch_raycast_t<PREC, true, kModeTrilinear>(c, io, tiles_x, nb, m)
          : launch_raycast_t<PREC, false, kModeTrilinear>(c, io, tiles_x, nb, m);
  else
    count ? launch_raycast_t<PREC, true, kModeReference>(c, io, tiles_x, nb, m)
          : launch_raycast_t<PREC, false, kModeReference>(c, io, tiles_x, nb, m);
}

// maps = null: the context's model maps; count = the context's work-counter
// switch (ignored by the modes that do not count).
static hipError_t launch_raycast_mode(kf_context* c, const kf_cam& cam, int mode,
                                      const RayMaps* maps = nullptr, bool count = true) {
  const RayMaps m =
      maps ? *maps : RayMaps{c->model_v, c->model_n, c->model_aux, c->hit, c->model_nt};
  if (!maps && mode != kModeRefine) c->maps_trilinear = mode == kModeTrilinear;
  const int tiles_x = (c->W + 15) / 16, tiles_y = (c->H + 15) / 16;
  const int nb = tiles_x * tiles_y;
  RayIO io = make_ray_io(c, cam);
  if (mode == kModeRefine) {
    io.walk_tmax = refine_walk_tmax(c, cam);
    io.refine_batches = refine_pass1_batches(c, io.walk_tmax);
  }
  io.qcount = c->ray_counts + (size_t)c->ray_slot * KF_RAY_PARTS * KF_RAY_COUNT_STRIDE;
  io.qnext = c->ray_counts + (size_t)(c->ray_slot ^ 1) * KF_RAY_PARTS * KF_RAY_COUNT_STRIDE;
  io.qcap = kf_ray_part_capacity(nb);
  c->ray_slot ^= 1;
  if (c->prm.raycast_skip && c->skip_flags) {
    if (c->skip_flags_stale) {
      const hipError_t e = kf_launch_skip_flags_rebuild(c);
      if (e != hipSuccess) return e;
    }
    io.skip_flags = c->skip_flags;
  }
  if (c->vol.precision == KF_TSDF_F32)
    launch_raycast_p<KF_TSDF_F32>(c, io, tiles_x, nb, count && c->count_rays, mode, m);
  else
    launch_raycast_p<KF_TSDF_F16>(c, io, tiles_x, nb, count && c->count_rays, mode, m);
  return hipGetLastError();
}

hipError_t kf_launch_raycast(kf_context* c, const kf_cam& cam) {
  return launch_raycast_mode(c, cam, c->prm.raycast_mode == 1 ? kModeTrilinear : kModeReference);
}

hipError_t kf_launch_refine(kf_context* c, const kf_cam& cam, int steps) {
  const int n_first = ((c->W + 15) / 16) * ((c->H + 15) / 16);
  for (int it = 0; it < steps; ++it) {
    // the raycast's append counters (launch_raycast_mode flips the slot)
    const uint32_t* qc =
        c->ray_counts + (size_t)c->ray_slot * KF_RAY_PARTS * KF_RAY_COUNT_STRIDE;
    hipError_t e = launch_raycast_mode(c, cam, kModeRefine);
    if (e != hipSuccess) return e;
    k_refine_finish<<<1, kRefineFinishThreads, 0, c->stream>>>(c->gn_partials, n_first, qc,
                                                                c->state, it);
  }
  return hipGetLastError();
}

size_t kf_refine_partials(const kf_context* c) {
  return (size_t)(((c->W + 15) / 16) * ((c->H + 15) / 16) + kRestBlocks) * kGnSums;
}

size_t kf_adjoint_partials(const kf_context* c) {
  return (size_t)(((c->W + 15) / 16) * ((c->H + 15) / 16)) * 6;  // one slot per tile
}

hipError_t kf_launch_raycast_adjoint(kf_context* c, const kf_cam& cam, bool saved) {
  if (c->adj_want_tsdf) {
    const int n_words = (c->n_bricks + 31) / 32;
    k_zero_touched<<<(unsigned)((n_words + 3) / 4), 256, 0, c->stream>>>(
        reinterpret_cast<float4*>(c->grad_tsdf), c->adj_touched, n_words);
  }
  // Without saved crossings for these inputs, the trilinear raycast first
  // walks every ray into the adjoint's own maps (the context's model maps
  // stay as they are); the adjoint then starts from those crossings.
  RayMaps m{c->model_v, c->model_n, c->model_aux, c->hit, c->model_nt};
  if (!saved) {
    m = RayMaps{c->adj_map_v, c->adj_map_n, c->adj_map_aux, c->adj_map_hit, nullptr};
    hipError_t e = launch_raycast_mode(c, cam, kModeTrilinear, &m, false);
    if (e != hipSuccess) return e;
  }
  const int tiles_x = (c->W + 15) / 16, nb = tiles_x * ((c->H + 15) / 16);
  const RayIO io = make_ray_io(c, cam);
  if (c->vol.precision == KF_TSDF_F32)
    k_adjoint_saved<KF_TSDF_F32><<<nb, 256, 0, c->stream>>>(c->tsdf, io, c->H, c->state,
                                                             m.aux, m.hit, tiles_x, nb);
  else
    k_adjoint_saved<KF_TSDF_F16><<<nb, 256, 0, c->stream>>>(c->tsdf, io, c->H, c->state,
                                                             m.aux, m.hit, tiles_x, nb);
  k_adjoint_finish<<<1, kAdjFinishThreads, 0, c->stream>>>(c->adj_partials, nb, c->state);
  return hipGetLastError();
}

size_t kf_ray_state_bytes() { return sizeof(RayState); }

size_t kf_ray_queue_entries(int H, int W) {
  return (size_t)KF_RAY_PARTS * kf_ray_part_capacity(((W + 15) / 16) * ((H + 15) / 16));
}

__global__ __launch_bounds__(256) void k_fill_f32(float* __restrict__ p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t kf_launch_fill_f32(float* p, int64_t n, float v, hipStream_t s) {
  k_fill_f32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t kf_launch_volume_fill(kf_context* c) {
  const int64_t n = c->n_store;
  k_volume_fill<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(c->tsdf, c->weight,
                                                                     c->vol.precision, n);
  return hipGetLastError();
}

hipError_t kf_launch_skip_flags_clear(kf_context* c) {
  if (c->skip_flags) {
    k_fill_u8<<<(unsigned)((c->n_bricks + 255) / 256), 256, 0, c->stream>>>(c->skip_flags,
                                                                             c->n_bricks, 0);
    c->skip_flags_stale = false;
  }
  return hipGetLastError();
}

hipError_t kf_launch_skip_flags_rebuild(kf_context* c) {
  if (!c->skip_flags) return hipSuccess;
  const hipError_t e = kf_launch_skip_flags_clear(c);
  if (e != hipSuccess) return e;
  const unsigned blocks = (unsigned)((c->n_bricks + 3) / 4);
  if (c->vol.precision == KF_TSDF_F32)
    k_skip_flags<KF_TSDF_F32><<<blocks, 256, 0, c->stream>>>(c->tsdf, c->skip_flags, c->vg,
                                                             c->n_bricks);
  else
    k_skip_flags<KF_TSDF_F16><<<blocks, 256, 0, c->stream>>>(c->tsdf, c->skip_flags, c->vg,
                                                             c->n_bricks);
  return hipGetLastError();
}
