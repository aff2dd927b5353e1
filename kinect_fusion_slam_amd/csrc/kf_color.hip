// kf_color.hip - the colour path on the device: the data-parallel half of a
// baseline JPEG decode (the reference's cv::imread, system_utility.cpp:70-75)
// and the coloured point cloud (ThreeDViewer::SetUpPointClouds,
// threeD_viewer.cpp:20-55).
//
// k_jpeg_idct: dequantisation and the IJG "ISLOW" inverse DCT, the
//   Loeffler-Ligtenberg-Moschytz factorisation libjpeg-turbo uses by
//   default (13-bit constants, 2 extra bits between the column and row
//   passes, the range-limit clamp).  Eight lanes per 8x8 block: lane r loads
//   row r of coefficients (16 bytes), the block goes through LDS for the
//   column pass, and lane r writes output row r (8 bytes).  Products are
//   64-bit, as in the C implementation, so out-of-range coefficients give
//   the same samples.
// k_jpeg_color: one pixel per lane: each component's sample at full
//   resolution by libjpeg's "fancy" upsampling (triangle filters for 2:1
//   factors, rows and columns at the image edge replicated, box replication
//   for other integral factors and for downsampled widths of 2 or less),
//   then YCbCr -> RGB with 16-bit fixed-point coefficients (or RGB / grey
//   as they are), written as B, G, R bytes as OpenCV stores them.
// k_point_cloud: the reference's per-pixel point record.
#include <hip/hip_runtime.h>

#include "kf_internal.h"
#include "kf_jpeg.h"

namespace {

constexpr int kIdctBlocksPerGroup = 32;  // 8 lanes each: 256 threads

// round(c * 2^13) of the ISLOW constants
constexpr int64_t C0_298631336 = 2446, C0_390180644 = 3196, C0_541196100 = 4433,
                  C0_765366865 = 6270, C0_899976223 = 7373, C1_175875602 = 9633,
                  C1_501321110 = 12299, C1_847759065 = 15137, C1_961570560 = 16069,
                  C2_053119869 = 16819, C2_562915447 = 20995, C3_072711026 = 25172;

__device__ __forceinline__ int32_t descale(int64_t x, int n) {
  return (int32_t)((x + ((int64_t)1 << (n - 1))) >> n);
}

// x masked to 10 bits, read as a signed value in [-512, 511], plus 128,
// clamped to [0, 255]
__device__ __forceinline__ uint32_t range_limit(int32_t x) {
  const int32_t m = ((x & 1023) ^ 512) - 512;
  return (uint32_t)min(max(m + 128, 0), 255);
}

// One 1-D pass over 8 values (the even / odd LLM butterflies); out[i] =
// undescaled result i.
__device__ __forceinline__ void llm8(const int64_t* in, int64_t* out) {
  int64_t z2 = in[2], z3 = in[6];
  int64_t z1 = (z2 + z3) * C0_541196100;
  const int64_t t2 = z1 - z3 * C1_847759065;
  const int64_t t3 = z1 + z2 * C0_765366865;
  const int64_t t0 = (in[0] + in[4]) * 8192;
  const int64_t t1 = (in[0] - in[4]) * 8192;
  const int64_t e10 = t0 + t3, e13 = t0 - t3, e11 = t1 + t2, e12 = t1 - t2;
  int64_t o0 = in[7], o1 = in[5], o2 = in[3], o3 = in[1];
  z1 = o0 + o3;
  z2 = o1 + o2;
  z3 = o0 + o2;
  int64_t z4 = o1 + o3;
  const int64_t z5 = (z3 + z4) * C1_175875602;
  o0 *= C0_298631336;
  o1 *= C2_053119869;
  o2 *= C3_072711026;
  o3 *= C1_501321110;
  z1 *= -C0_899976223;
  z2 *= -C2_562915447;
  z3 = z3 * -C1_961570560 + z5;
  z4 = z4 * -C0_390180644 + z5;
  o0 += z1 + z3;
  o1 += z2 + z4;
  o2 += z2 + z3;
  o3 += z1 + z4;
  out[0] = e10 + o3;
  out[7] = e10 - o3;
  out[1] = e11 + o2;
  out[6] = e11 - o2;
  out[2] = e12 + o1;
  out[5] = e12 - o1;
  out[3] = e13 + o0;
  out[4] = e13 - o0;
}

__global__ __launch_bounds__(256) void k_jpeg_idct(const int16_t* __restrict__ coef,
                                                   const uint16_t* __restrict__ qtab,
                                                   uint8_t* __restrict__ planes,
                                                   kf_jpeg_dev f) {
  __shared__ int32_t blk[kIdctBlocksPerGroup][8][9];  // [block][row][col], padded
  const int lb = threadIdx.x >> 3, r = threadIdx.x & 7;
  const int64_t b = (int64_t)blockIdx.x * kIdctBlocksPerGroup + lb;
  const bool live = b < f.n_blocks;
  int ci = 0;
  for (int k = 1; k < f.nc; ++k)
    if (b >= f.c[k].block0) ci = k;
  const kf_jpeg_dev_comp& c = f.c[ci];
  const uint16_t* q = qtab + c.tq * 64;
  // row r of coefficients, dequantised
  if (live) {
    const uint4 raw = *reinterpret_cast<const uint4*>(coef + b * 64 + r * 8);
    const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int16_t v = (int16_t)((w[k >> 1] >> ((k & 1) * 16)) & 0xFFFF);
      blk[lb][r][k] = (int32_t)v * (int32_t)q[r * 8 + k];
    }
  }
  __syncthreads();
  // column r
  int32_t colv[8];
  if (live) {
    int64_t in[8], out[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) in[k] = blk[lb][k][r];
    if ((in[1] | in[2] | in[3] | in[4] | in[5] | in[6] | in[7]) == 0) {
#pragma unroll
      for (int k = 0; k < 8; ++k) colv[k] = (int32_t)(in[0] * 4);
    } else {
      llm8(in, out);
#pragma unroll
      for (int k = 0; k < 8; ++k) colv[k] = descale(out[k], 11);
    }
  }
  __syncthreads();
  if (live)
#pragma unroll
    for (int k = 0; k < 8; ++k) blk[lb][k][r] = colv[k];
  __syncthreads();
  if (!live) return;
  // row r
  int64_t in[8], out[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) in[k] = blk[lb][r][k];
  uint32_t s[8];
  if ((in[1] | in[2] | in[3] | in[4] | in[5] | in[6] | in[7]) == 0) {
    const uint32_t v = range_limit(descale(in[0], 5));
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = v;
  } else {
    llm8(in, out);
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = range_limit(descale(out[k], 18));
  }
  const int64_t local = b - c.block0;
  const int by = (int)(local / c.bw), bx = (int)(local - (int64_t)by * c.bw);
  uint8_t* dst = planes + c.plane0 + (int64_t)(by * 8 + r) * c.stride + bx * 8;
  uint2 o;
  o.x = s[0] | s[1] << 8 | s[2] << 16 | s[3] << 24;
  o.y = s[4] | s[5] << 8 | s[6] << 16 | s[7] << 24;
  *reinterpret_cast<uint2*>(dst) = o;
}

// A component's sample at full-resolution pixel (x, y).
__device__ __forceinline__ int comp_sample(const uint8_t* __restrict__ planes,
                                           const kf_jpeg_dev_comp& c, int x, int y) {
  const uint8_t* p = planes + c.plane0;
  auto at = [&](int i, int j) -> int { return p[(int64_t)j * c.stride + i]; };
  if (c.fh == 1 && c.fv == 1) return at(x, y);
  if (c.fh == 2 && c.dw > 2 && c.fv == 1) {
    const int i = x >> 1;
    if ((x & 1) == 0) return i == 0 ? at(0, y) : (at(i, y) * 3 + at(i - 1, y) + 1) >> 2;
    return i == c.dw - 1 ? at(i, y) : (at(i, y) * 3 + at(i + 1, y) + 2) >> 2;
  }
  if (c.fh == 2 && c.dw > 2 && c.fv == 2) {
    const int i = x >> 1, j = y >> 1;
    const int j2 = (y & 1) ? min(j + 1, c.dh - 1) : max(j - 1, 0);
    const int cs = at(i, j) * 3 + at(i, j2);
    if ((x & 1) == 0) {
      if (i == 0) return (cs * 4 + 8) >> 4;
      return (cs * 3 + at(i - 1, j) * 3 + at(i - 1, j2) + 8) >> 4;
    }
    if (i == c.dw - 1) return (cs * 4 + 7) >> 4;
    return (cs * 3 + at(i + 1, j) * 3 + at(i + 1, j2) + 7) >> 4;
  }
  return at(x / c.fh, y / c.fv);
}

__global__ __launch_bounds__(256) void k_jpeg_color(const uint8_t* __restrict__ planes,
                                                    kf_jpeg_dev f, uint8_t* __restrict__ bgr) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)f.H * f.W) return;
  const int y = (int)(i / f.W), x = (int)(i - (int64_t)y * f.W);
  uint8_t* o = bgr + 3 * i;
  const int s0 = comp_sample(planes, f.c[0], x, y);
  if (f.nc == 1) {
    o[0] = o[1] = o[2] = (uint8_t)s0;
    return;
  }
  const int s1 = comp_sample(planes, f.c[1], x, y), s2 = comp_sample(planes, f.c[2], x, y);
  if (f.rgb) {
    o[0] = (uint8_t)s2;
    o[1] = (uint8_t)s1;
    o[2] = (uint8_t)s0;
    return;
  }
  // round(c * 2^16) of 1.40200, 1.77200, 0.34414, 0.71414
  const int cb = s1 - 128, cr = s2 - 128;
  const int r = s0 + ((91881 * cr + 32768) >> 16);
  const int b = s0 + ((116130 * cb + 32768) >> 16);
  const int g = s0 + ((-22554 * cb - 46802 * cr + 32768) >> 16);
  o[0] = (uint8_t)min(max(b, 0), 255);
  o[1] = (uint8_t)min(max(g, 0), 255);
  o[2] = (uint8_t)min(max(r, 0), 255);
}

// ThreeDViewer::SetUpPointClouds (threeD_viewer.cpp:20-55): where the raycast
// vertex's z != 0, {x, y, z, rgb} with rgb the bits of r << 16 | g << 8 | b
// of the same pixel; zeros elsewhere.
__global__ __launch_bounds__(256) void k_point_cloud(const kf_p3* __restrict__ mv,
                                                     const uint8_t* __restrict__ bgr, int n,
                                                     float4* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const kf_p3 v = mv[i];
  float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
  if (v.z != 0.f) {
    const uint32_t rgb = (uint32_t)bgr[3 * i + 2] << 16 | (uint32_t)bgr[3 * i + 1] << 8 |
                         (uint32_t)bgr[3 * i];
    o = make_float4(v.x, v.y, v.z, __uint_as_float(rgb));
  }
  out[i] = o;
}

}  // namespace

hipError_t kf_launch_jpeg_decode(kf_context* c, const kf_jpeg_dev& f) {
  const int64_t groups = (f.n_blocks + kIdctBlocksPerGroup - 1) / kIdctBlocksPerGroup;
  k_jpeg_idct<<<(unsigned)groups, 256, 0, c->stream>>>(c->jpeg_coef, c->jpeg_q, c->jpeg_planes, f);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t n = (int64_t)f.H * f.W;
  k_jpeg_color<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(c->jpeg_planes, f, c->color);
  return hipGetLastError();
}

hipError_t kf_launch_point_cloud(kf_context* c, float4* out) {
  const int n = c->H * c->W;
  k_point_cloud<<<(n + 255) / 256, 256, 0, c->stream>>>(c->model_v, c->color, n, out);
  return hipGetLastError();
}
