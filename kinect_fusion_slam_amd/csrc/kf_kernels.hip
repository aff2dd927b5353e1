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

#include "kf_internal.h"

namespace {

// ---------------------------------------------------------------------------
// helpers

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Uniform load of a small float array that the previous kernel wrote.
template <int N>
__device__ __forceinline__ void load_uniform(const float* __restrict__ src, float* dst) {
#pragma unroll
  for (int i = 0; i < N; ++i) dst[i] = src[i];
}

__device__ __forceinline__ uint32_t umulhi32(uint32_t a, uint32_t b) { return __umulhi(a, b); }

// floor(n / D) for n < 2^31 with the precomputed magic (m, s).
__device__ __forceinline__ uint32_t magic_div(uint32_t n, uint32_t m, int s) {
  return (umulhi32(n, m) + n) >> s;
}

// cuda_functions.cuh:2-10
__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
  const float t = 1 / sqrtf(x * x + y * y + z * z);
  x = x * t;
  y = y * t;
  z = z * t;
}

// wave64 DPP reduction: after the call lane 63 holds the sum of all lanes.
// quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror, row_mirror give every
// lane its row's sum; row_bcast:15 and row_bcast:31 then fold rows 0..3 into
// row 3.  The order of additions is fixed, so the result is deterministic.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_src(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xF, false));
}

__device__ __forceinline__ float wave_sum(float v) {
  v = v + dpp_src<0xB1, 0xF>(v);
  v = v + dpp_src<0x4E, 0xF>(v);
  v = v + dpp_src<0x141, 0xF>(v);
  v = v + dpp_src<0x140, 0xF>(v);
  v = v + dpp_src<0x142, 0xA>(v);
  v = v + dpp_src<0x143, 0xC>(v);
  return v;
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

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void dpp_pair_step(float& s, float& c) {
  const float s2 = dpp_src<CTRL, ROW_MASK>(s);
  const float c2 = dpp_src<CTRL, ROW_MASK>(c);
  two_sum_add(s, c, s2, c2);
}

__device__ __forceinline__ void wave_sum_comp(float& s, float& c) {
  dpp_pair_step<0xB1, 0xF>(s, c);
  dpp_pair_step<0x4E, 0xF>(s, c);
  dpp_pair_step<0x141, 0xF>(s, c);
  dpp_pair_step<0x140, 0xF>(s, c);
  dpp_pair_step<0x142, 0xA>(s, c);
  dpp_pair_step<0x143, 0xC>(s, c);
}

// ---------------------------------------------------------------------------
// volume accessors: F32 = float2 {tsdf, weight}; F16 = uint32 {half, u16}

__device__ __forceinline__ float f16_lo_to_f32(uint32_t p) {
  return __half2float(__ushort_as_half((unsigned short)(p & 0xffffu)));
}

template <int PREC>
struct VolRead;
template <>
struct VolRead<KF_TSDF_F32> {
  const float2* __restrict__ v;
  __device__ __forceinline__ float operator()(int i) const { return v[i].x; }
};
template <>
struct VolRead<KF_TSDF_F16> {
  const uint32_t* __restrict__ v;
  __device__ __forceinline__ float operator()(int i) const { return f16_lo_to_f32(v[i]); }
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
template <int R>
__global__ __launch_bounds__(256) void k_bilateral(const float* __restrict__ in,
                                                   float* __restrict__ out,
                                                   const float* __restrict__ ws, int H, int W,
                                                   float sigma_r) {
  constexpr int TS = 16, L = 2 * R + 1, SW = TS + 2 * R;
  __shared__ float tile[SW][SW + 1];
  __shared__ float sw[L * L];
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * TS, y0 = blockIdx.y * TS;
  const int n = H * W;
  for (int i = tid; i < SW * SW; i += 256) {
    const int r = i / SW, cc = i - r * SW;
    const int idx = (y0 - R + r) * W + (x0 - R + cc);
    tile[r][cc] = (idx >= 0 && idx < n) ? in[idx] : 0.0f;
  }
  for (int i = tid; i < L * L; i += 256) sw[i] = ws[i];
  __syncthreads();
  const int tx = tid & (TS - 1), ty = tid / TS;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= W || y >= H) return;
  float result = 0.0f;
  if (y >= R && y <= H - R && x >= R && x <= W - R) {
    const float dc = tile[ty + R][tx + R];
    const float den = 2 * sigma_r * sigma_r;
    float sum = 0.f, fin = 0.f;
    if constexpr (L * L <= 49) {
      // small windows: keep the weights in registers (fully unrolled)
      float w[L * L];
#pragma unroll
      for (int k = 0; k < L * L; ++k) {
        float t = tile[ty + k / L][tx + k % L] - dc;
        t = kf_expf(-(t * t) / den);
        w[k] = sw[k] * t + 0.000000001f;
        sum = sum + w[k];
      }
#pragma unroll
      for (int k = 0; k < L * L; ++k) {
        const float wr = w[k] / sum;
        fin = fin + wr * tile[ty + k / L][tx + k % L];
      }
    } else {
      // large windows: recompute the (identical) weights in the second pass
      for (int k = 0; k < L * L; ++k) {
        float t = tile[ty + k / L][tx + k % L] - dc;
        t = kf_expf(-(t * t) / den);
        sum = sum + (sw[k] * t + 0.000000001f);
      }
      for (int k = 0; k < L * L; ++k) {
        const float v = tile[ty + k / L][tx + k % L];
        float t = v - dc;
        t = kf_expf(-(t * t) / den);
        const float wr = (sw[k] * t + 0.000000001f) / sum;
        fin = fin + wr * v;
      }
    }
    result = fin;
  }
  out[y * W + x] = result;
}

// ---------------------------------------------------------------------------
// A3  vertex and normal maps (raw_vertices.cu:3-51).  Writes every pixel, so
// the zero defaults of the reference's CleanDepthData need no separate fill.
__global__ __launch_bounds__(256) void k_vertex(const float* __restrict__ d, kf_cam cam, int H,
                                                int W, float4* __restrict__ vmap,
                                                float4* __restrict__ nmap,
                                                uint8_t* __restrict__ mask) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= H * W) return;
  const int y = i / W, x = i - y * W;
  const float* Ki = cam.Kinv;
  const float di = d[i];
  float4 V = make_float4(0.f, 0.f, di, 0.f);
  float4 N = make_float4(0.f, 0.f, 0.f, 0.f);
  uint8_t m = 0;
  if (di != 0 && y != (H - 1) && x != (W - 1)) {
    m = 1;
    V.w = 1.0f;
    const int nr = (y + 1) * W + x;
    const int nc = y * W + (x + 1);
    const float dr = d[nr], dcn = d[nc];
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
  vmap[i] = V;
  nmap[i] = N;
  mask[i] = m;
}

// ---------------------------------------------------------------------------
// A5-A7  ICP.

// Frame start: T_prev := T, T_prev^-1, F = T_prev^-1 T (sensor_pose_estimation.cu:
// 255-259, 343 at i = 0).
__global__ void k_icp_begin(kf_dev_state* st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float T[16], Ti[16], F[16];
  for (int i = 0; i < 16; ++i) T[i] = st->T[i];
  if (kf_m4_inverse(T, Ti) != 0)
    for (int i = 0; i < 16; ++i) Ti[i] = __builtin_nanf("");
  kf_mat4_mul(Ti, T, F);
  for (int i = 0; i < 16; ++i) {
    st->Tprev[i] = T[i];
    st->Tprev_inv[i] = Ti[i];
    st->Tinv[i] = Ti[i];
    st->F[i] = F[i];
  }
  st->ticket = 0u;
  st->icp_iters_done = 0;
}

// One ICP iteration: association and gates of every pixel
// (sensor_pose_estimation.cu:16-136), reduction of the 27 unique products
// plus the inlier count, and - in the last block to finish - the FP64 solve
// and pose update (:389-429).
template <bool KAHAN>
__global__ __launch_bounds__(KF_ICP_BLOCK) void k_icp_iter(
    const float4* __restrict__ vmap, const float4* __restrict__ nmap,
    const float4* __restrict__ mv, const float4* __restrict__ mn, kf_cam cam, int H, int W,
    float dist_gate, float cos_lo, kf_dev_state* __restrict__ st,
    float* __restrict__ partials, uint8_t* __restrict__ mask_out, int iter, int last) {
  constexpr int NS = KF_ICP_SUMS;
  __shared__ float red_s[KF_ICP_BLOCK / 64][NS];
  __shared__ float red_c[KF_ICP_BLOCK / 64][NS];
  __shared__ double fin[NS][8];
  __shared__ double sums_d[NS];
  __shared__ int is_last;

  float F[12], Te[12];
  load_uniform<12>(st->F, F);
  load_uniform<12>(st->T, Te);
  const float* K = cam.K;

  float acc[NS], comp[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) acc[k] = 0.f, comp[k] = 0.f;

  const int n = H * W;
  const int base = blockIdx.x * (KF_ICP_BLOCK * KF_ICP_PIX_PER_THREAD) + threadIdx.x;
#pragma unroll 1
  for (int k = 0; k < KF_ICP_PIX_PER_THREAD; ++k) {
    const int x = base + k * KF_ICP_BLOCK;
    if (x >= n) break;
    const float4 V = vmap[x];
    bool ok = false;
    float a[6], b = 0.f;
    if (V.z != 0 && V.w != 0) {
      const float cx = F[0] * V.x + F[1] * V.y + F[2] * V.z + F[3];
      const float cy = F[4] * V.x + F[5] * V.y + F[6] * V.z + F[7];
      const float cz = F[8] * V.x + F[9] * V.y + F[10] * V.z + F[11];
      const float px = K[0] * cx + K[1] * cy + K[2] * cz;
      const float py = K[3] * cx + K[4] * cy + K[5] * cz;
      const float pz = K[6] * cx + K[7] * cy + K[8] * cz;
      const int u = kf_round_to_int(px / pz);
      const int v = kf_round_to_int(py / pz);
      if (u > 0 && u <= W && v > 0 && v <= H) {
        const float ux = Te[0] * V.x + Te[1] * V.y + Te[2] * V.z + Te[3];
        const float uy = Te[4] * V.x + Te[5] * V.y + Te[6] * V.z + Te[7];
        const float uz = Te[8] * V.x + Te[9] * V.y + Te[10] * V.z + Te[11];
        const int j = (v - 1) * W + (u - 1);
        const float4 S = mv[j];
        const float ddx = ux - S.x, ddy = uy - S.y, ddz = uz - S.z;
        const float dist = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
        if (dist <= dist_gate && dist >= 0.0f) {
          const float4 Nc = nmap[x];
          float nnx = Te[0] * Nc.x + Te[1] * Nc.y + Te[2] * Nc.z;
          float nny = Te[4] * Nc.x + Te[5] * Nc.y + Te[6] * Nc.z;
          float nnz = Te[8] * Nc.x + Te[9] * Nc.y + Te[10] * Nc.z;
          normalize3(nnx, nny, nnz);
          const float4 Sn = mn[j];
          float ax = Sn.x, ay = Sn.y, az = Sn.z;
          normalize3(ax, ay, az);
          const float c = (nnx * ax + nny * ay + nnz * az);
          if (c >= cos_lo && c <= 1.0f) {
            ok = true;
            a[0] = uz * Sn.y + (-uy * Sn.z);
            a[1] = -uz * Sn.x + (ux * Sn.z);
            a[2] = uy * Sn.x + (-ux * Sn.y);
            a[3] = Sn.x;
            a[4] = Sn.y;
            a[5] = Sn.z;
            b = Sn.x * (S.x - ux) + Sn.y * (S.y - uy) + Sn.z * (S.z - uz);
          }
        }
      }
    }
    if (mask_out) mask_out[x] = ok ? 1 : 0;
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
  // error with the opposite sign), so the pair (s, -c) is reduced.
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    if constexpr (KAHAN) {
      float s = acc[k], c = -comp[k];
      wave_sum_comp(s, c);
      if (lane_id() == 63) {
        red_s[wave][k] = s;
        red_c[wave][k] = c;
      }
    } else {
      const float s = wave_sum(acc[k]);
      if (lane_id() == 63) red_s[wave][k] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    const int k = threadIdx.x;
    float s = red_s[0][k], c = KAHAN ? red_c[0][k] : 0.f;
#pragma unroll
    for (int w = 1; w < KF_ICP_BLOCK / 64; ++w) {
      if constexpr (KAHAN)
        two_sum_add(s, c, red_s[w][k], red_c[w][k]);
      else
        s = s + red_s[w][k];
    }
    partials[((size_t)blockIdx.x * NS + k) * 2 + 0] = s;
    partials[((size_t)blockIdx.x * NS + k) * 2 + 1] = c;
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned t = atomicAdd(&st->ticket, 1u);
    is_last = (t == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();

  // Last block: FP64 sum of the block partials in block order.
  const int nb = gridDim.x;
  if (threadIdx.x < NS * 8) {
    const int k = threadIdx.x >> 3, ch = threadIdx.x & 7;
    double s = 0.0;
    for (int bk = ch; bk < nb; bk += 8) {
      const float* p = partials + ((size_t)bk * NS + k) * 2;
      s += (double)p[0] + (double)p[1];
    }
    fin[k][ch] = s;
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    double s = 0.0;
#pragma unroll
    for (int ch = 0; ch < 8; ++ch) s += fin[threadIdx.x][ch];
    sums_d[threadIdx.x] = s;
    st->iter_sums[iter][threadIdx.x] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float T[16], Tp[16], Fn[16];
    for (int i = 0; i < 16; ++i) T[i] = st->T[i], Tp[i] = st->Tprev_inv[i];
    const int status = kf_icp_update(sums_d, T, nullptr);
    kf_mat4_mul(Tp, T, Fn);
    for (int i = 0; i < 16; ++i) {
      st->T[i] = T[i];
      st->F[i] = Fn[i];
      st->iter_T[iter][i] = T[i];
    }
    st->iter_status[iter] = status;
    st->icp_iters_done = iter + 1;
    if (last) {
      float Ti[16];
      if (kf_m4_inverse(T, Ti) != 0)
        for (int i = 0; i < 16; ++i) Ti[i] = __builtin_nanf("");
      for (int i = 0; i < 16; ++i) st->Tinv[i] = Ti[i];
    }
    __threadfence();
    atomicExch(&st->ticket, 0u);
  }
}

// ---------------------------------------------------------------------------
// A8  projective TSDF integration (projective_tsdf.cu:5-140).

// Per-pixel part: for a pixel (u, v) with depth d != 0, the ray magnitude
// |T(2 lambda) - T(lambda)|, the predicted depth |T(d lambda) - t| / |.| and
// its reciprocal depend only on the pixel, so they are computed here once
// with the reference's expressions (:39-52, :78-94).  Depth 0 stores NaN.
// Thread 0 also stores the front-test direction (:57-71), which depends only
// on K and T.
__global__ __launch_bounds__(256) void k_integ_table(const float* __restrict__ depth, kf_cam cam,
                                                     int H, int W, kf_dev_state* __restrict__ st,
                                                     float2* __restrict__ tab) {
  float T[12];
  load_uniform<12>(st->T, T);
  const float* Ki = cam.Kinv;
  const float* K = cam.K;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i == 0) {
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
    st->center_dir[0] = gbx - gax;
    st->center_dir[1] = gby - gay;
    st->center_dir[2] = gbz - gaz;
    st->center_dir[3] = 0.f;
  }
  if (i >= H * W) return;
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
  }
  tab[i] = out;
}

// Per-voxel part.  One thread per voxel, 4 voxels per thread strided by the
// block so that a wave touches 64 consecutive z voxels (one coalesced 512 B
// F32 / 256 B F16 segment).  The voxel is read and written only when its
// signed distance falls in [sdf_min, sdf_max], as in the reference.
constexpr int kIntegVPT = 4;

template <int PREC, bool FIRST>
__global__ __launch_bounds__(256) void k_integrate(void* __restrict__ volume,
                                                   const float2* __restrict__ tab, kf_cam cam,
                                                   kf_vol vg, int H, int W, float trunc,
                                                   float smin, float smax,
                                                   const kf_dev_state* __restrict__ st, int nvox) {
  float Ti[12], t[3], cd[3];
  load_uniform<12>(st->Tinv, Ti);
  t[0] = st->T[3];
  t[1] = st->T[7];
  t[2] = st->T[11];
  load_uniform<3>(st->center_dir, cd);
  const float* K = cam.K;
  const float d = vg.d;
  const float hd = d / 2.f;
#pragma unroll
  for (int k = 0; k < kIntegVPT; ++k) {
    const int v = (blockIdx.x * kIntegVPT + k) * 256 + (int)threadIdx.x;
    if (v >= nvox) return;
    const uint32_t q1 = magic_div((uint32_t)v, vg.mz, vg.shz);
    const int iz = v - (int)q1 * vg.nz;
    const uint32_t q2 = magic_div(q1, vg.my, vg.shy);
    const int iy = (int)q1 - (int)q2 * vg.ny;
    const int ix = (int)q2;
    const float cx = (float)(ix + vg.sx) * d + hd;
    const float cy = (float)(iy + vg.sy) * d + hd;
    const float cz = (float)(iz + vg.sz) * d + hd;
    const float lx = Ti[0] * cx + Ti[1] * cy + Ti[2] * cz + Ti[3];
    const float ly = Ti[4] * cx + Ti[5] * cy + Ti[6] * cz + Ti[7];
    const float lz = Ti[8] * cx + Ti[9] * cy + Ti[10] * cz + Ti[11];
    const float px = K[0] * lx + K[1] * ly + K[2] * lz;
    const float py = K[3] * lx + K[4] * ly + K[5] * lz;
    const float pz = K[6] * lx + K[7] * ly + K[8] * lz;
    const int u = kf_round_to_int(px / pz);
    const int pv = kf_round_to_int(py / pz);
    if (!(u > 0 && u <= W && pv > 0 && pv <= H)) continue;
    const float2 tb = tab[(pv - 1) * W + (u - 1)];
    if (tb.x != tb.x) continue;  // depth 0 at that pixel
    const float tdx = cx - t[0], tdy = cy - t[1], tdz = cz - t[2];
    const float mag = cd[0] * tdx + cd[1] * tdy + cd[2] * tdz;
    if (!(mag > 0)) continue;
    const float tm = sqrtf(tdx * tdx + tdy * tdy + tdz * tdz);
    const float frk = tb.x - (tb.y * tm);
    const bool pos = frk > 0 && frk <= smax;
    const bool neg = frk <= 0 && frk >= smin;
    if (!pos && !neg) continue;
    const float f = pos ? fminf(1.f, (frk / trunc)) : fmaxf(-1.f, frk / trunc);
    if constexpr (PREC == KF_TSDF_F32) {
      float2* vol = reinterpret_cast<float2*>(volume);
      if constexpr (FIRST) {
        vol[v] = make_float2(f, 1.0f);
      } else {
        float2 fw = vol[v];
        if (fw.x != fw.x) fw.x = 0.0f;
        if (fw.y != fw.y) fw.y = 0.0f;
        fw.x = (fw.y * fw.x + 1.0f * f) / (fw.y + 1.0f);
        fw.y = fw.y + 1.0f;
        vol[v] = fw;
      }
    } else {
      uint32_t* vol = reinterpret_cast<uint32_t*>(volume);
      if constexpr (FIRST) {
        vol[v] = (1u << 16) | (uint32_t)__half_as_ushort(__float2half_rn(f));
      } else {
        const uint32_t p = vol[v];
        float F0 = f16_lo_to_f32(p);
        if (F0 != F0) F0 = 0.0f;
        const uint32_t w = p >> 16;
        const float Wt = (float)w;
        const float Fn = (Wt * F0 + 1.0f * f) / (Wt + 1.0f);
        const uint32_t w2 = w == 65535u ? 65535u : w + 1u;
        vol[v] = (w2 << 16) | (uint32_t)__half_as_ushort(__float2half_rn(Fn));
      }
    }
  }
}

// Never-observed fill: F32 {NaN, NaN}; F16 {half NaN, 0}.
__global__ __launch_bounds__(256) void k_volume_fill(void* volume, int prec, int64_t nvox) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nvox) return;
  if (prec == KF_TSDF_F32)
    reinterpret_cast<float2*>(volume)[i] = make_float2(__builtin_nanf(""), __builtin_nanf(""));
  else
    reinterpret_cast<uint32_t*>(volume)[i] = 0x7e00u;
}

// ---------------------------------------------------------------------------
// A9  voxel-traversal raycast (voxel_traversal.cu:75-424), one ray per lane.
//
// A 256-thread block covers 16x16 pixels as four 8x8 wave tiles, so a wave's
// rays walk neighbouring voxels.  Consecutive block ids are dealt round-robin
// to the 8 XCDs; the id is remapped so that each XCD gets a contiguous band
// of the image and its L2 sees the band's voxels.
template <int PREC>
__global__ __launch_bounds__(256) void k_raycast(
    const void* __restrict__ volume, const float* __restrict__ depth, kf_cam cam, kf_vol vg,
    int H, int W, float trunc, int fix_bounds, float flip_hi,
    const kf_dev_state* __restrict__ st, float4* __restrict__ out_v, float4* __restrict__ out_n,
    int32_t* __restrict__ out_hit, int tiles_x, int n_blocks) {
  int bid = blockIdx.x;
  {
    const int per = (n_blocks + 7) / 8;
    const int xcd = bid & 7, loc = bid >> 3;
    const int rem = n_blocks - (n_blocks / 8) * 8;  // XCDs with one block more
    const int base = xcd < rem ? xcd * per : rem * per + (xcd - rem) * (per - 1);
    bid = base + loc;
  }
  const int tile = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int bx = bid % tiles_x, by = bid / tiles_x;
  const int x = bx * 16 + (tile & 1) * 8 + (ln & 7);
  const int y = by * 16 + (tile >> 1) * 8 + (ln >> 3);
  if (x >= W || y >= H) return;
  const int p = y * W + x;

  VolRead<PREC> F;
  F.v = reinterpret_cast<decltype(F.v)>(volume);
  float4 ov = make_float4(0.f, 0.f, 0.f, 0.f), on = make_float4(0.f, 0.f, 0.f, 0.f);
  int hit = -1;
  const float dval = depth[p];
  if (dval != 0) {
    float T[12];
    load_uniform<12>(st->T, T);
    const float* Ki = cam.Kinv;
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
    const float camx = T[3], camy = T[7], camz = T[11];
    const float rgx = T[0] * lamx + T[1] * lamy + T[2] * lamz + T[3] * 1.f;
    const float rgy = T[4] * lamx + T[5] * lamy + T[6] * lamz + T[7] * 1.f;
    const float rgz = T[8] * lamx + T[9] * lamy + T[10] * lamz + T[11] * 1.f;
    float dx = rgx - camx, dy = rgy - camy, dz = rgz - camz;
    normalize3(dx, dy, dz);
    const float mnx = (float)vg.sx * d, mny = (float)vg.sy * d, mnz = (float)vg.sz * d;
    const float mxx = (float)(vg.sx + Nx) * d, mxy = (float)(vg.sy + Ny) * d,
                mxz = (float)(vg.sz + Nz) * d;
    // Ray_Box_Intersection (voxel_traversal.cu:4-73)
    bool flag = true;
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
    if ((tmin > tymax) || (tymin > tmax)) {
      flag = false;
    } else {
      if (tymin > tmin) tmin = tymin;
      if (tymax < tmax) tmax = tymax;
      if (dz >= 0) {
        tzmin = (mnz - sgz) / dz;
        tzmax = (mxz - sgz) / dz;
      } else {
        tzmin = (mxz - sgz) / dz;
        tzmax = (mnz - sgz) / dz;
      }
      if ((tmin > tzmax) || (tzmin > tmax)) {
        flag = false;
      } else {
        if (tzmin > tmin) tmin = tzmin;
      }
    }
    if (flag) {
      if (tmin < 0) tmin = 0;
      const float stx = sgx + tmin * dx, sty = sgy + tmin * dy, stz = sgz + tmin * dz;
      const float bxs = mxx - mnx, bys = mxy - mny, bzs = mxz - mnz;
      int ix = kf_floor_to_int(((stx - mnx) / bxs) * (float)Nx) + 1;
      int iy = kf_floor_to_int(((sty - mny) / bys) * (float)Ny) + 1;
      int iz = kf_floor_to_int(((stz - mnz) / bzs) * (float)Nz) + 1;
      if (ix == Nx + 1) ix = ix - 1;
      if (iy == Ny + 1) iy = iy - 1;
      if (iz == Nz + 1) iz = iz - 1;
      int XY = iy, XZ = iz, YX = ix, YZ = iz, ZX = ix, ZY = iy;
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
      float tmx = tmin + ((vmx - stx) / dx);
      float tmy = tmin + ((vmy - sty) / dy);
      float tmz = tmin + ((vmz - stz) / dz);
      const float tdx = d / fabsf(dx), tdy = d / fabsf(dy), tdz = d / fabsf(dz);
      float prev = 0.0f;
      const int NyNz = Ny * Nz;
      int vi = (ix - 1) * NyNz + (iy - 1) * Nz + (iz - 1);
      const int bXY = fix_bounds ? Ny : Nx, bXZ = fix_bounds ? Nz : Nx;
      const int bYX = fix_bounds ? Nx : Ny, bYZ = fix_bounds ? Nz : Ny;
      const int bZX = fix_bounds ? Nx : Nz, bZY = fix_bounds ? Ny : Nz;
      while (ix < Nx - 1 && ix > 0 && iy < Ny - 1 && iy > 0 && iz < Nz - 1 && iz > 0) {
        const float f = F(vi);
        if (f != f && prev < 0.f) break;
        if (f < 0.f && prev != prev) break;
        if (f > 0.f && prev < 0.f) break;
        if (f < 0.f && prev > 0.f) {
          const int xn = (ix - 1) - 1, xp = (ix - 1) + 1;
          const int yn = (iy - 1) - 1, yp = (iy - 1) + 1;
          const int zn = (iz - 1) - 1, zp = (iz - 1) + 1;
          if (xn >= Nx || xn < 0 || xp >= Nx || xp < 0) break;
          if ((XY - 1) >= bXY || (XY - 1) < 0 || (XZ - 1) >= bXZ || (XZ - 1) < 0) break;
          if (yn >= Ny || yn < 0 || yp >= Ny || yp < 0) break;
          if ((YX - 1) >= bYX || (YX - 1) < 0 || (YZ - 1) >= bYZ || (YZ - 1) < 0) break;
          if (zn >= Nz || zn < 0 || zp >= Nz || zp < 0) break;
          if ((ZX - 1) >= bZX || (ZX - 1) < 0 || (ZY - 1) >= bZY || (ZY - 1) < 0) break;
          const int x_np = xn * NyNz + (XY - 1) * Nz + XZ - 1;
          const int x_pp = xp * NyNz + (XY - 1) * Nz + XZ - 1;
          const int y_np = (YX - 1) * NyNz + yn * Nz + YZ - 1;
          const int y_pp = (YX - 1) * NyNz + yp * Nz + YZ - 1;
          const int z_np = (ZX - 1) * NyNz + (ZY - 1) * Nz + zn;
          const int z_pp = (ZX - 1) * NyNz + (ZY - 1) * Nz + zp;
          const int x_nn = vi - NyNz, x_pn = vi + NyNz;
          const int y_nn = vi - Nz, y_pn = vi + Nz;
          const int z_nn = vi - 1, z_pn = vi + 1;
          const float f_xnp = F(x_np), f_xpp = F(x_pp), f_ynp = F(y_np), f_ypp = F(y_pp);
          const float f_znp = F(z_np), f_zpp = F(z_pp);
          const float f_xnn = F(x_nn), f_xpn = F(x_pn), f_ynn = F(y_nn), f_ypn = F(y_pn);
          const float f_znn = F(z_nn), f_zpn = F(z_pn);
          if (f_xnp != f_xnp || f_xpp != f_xpp || f_ynp != f_ynp || f_ypp != f_ypp ||
              f_znp != f_znp || f_zpp != f_zpp || f != f)
            break;
          if (f_xnn != f_xnn || f_xpn != f_xpn || f_ynn != f_ynn || f_ypn != f_ypn ||
              f_znn != f_znn || f_zpn != f_zpn)
            break;
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
          hit = vi;
          break;
        }
        prev = f;
        if (tmx < tmy) {
          if (tmx < tmz) {
            ix = ix + stx_;
            XY = iy;
            XZ = iz;
            tmx = tmx + tdx;
          } else {
            iz = iz + stz_;
            ZX = ix;
            ZY = iy;
            tmz = tmz + tdz;
          }
        } else {
          if (tmy < tmz) {
            iy = iy + sty_;
            YX = ix;
            YZ = iz;
            tmy = tmy + tdy;
          } else {
            iz = iz + stz_;
            ZX = ix;
            ZY = iy;
            tmz = tmz + tdz;
          }
        }
        vi = (ix - 1) * NyNz + (iy - 1) * Nz + (iz - 1);
      }
    }
  }
  out_v[p] = ov;
  out_n[p] = on;
  out_hit[p] = hit;
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

hipError_t kf_launch_bilateral(kf_context* c) {
  const dim3 grid((c->W + 15) / 16, (c->H + 15) / 16);
  const float sr = c->prm.sigma_r_mm;
  switch (c->prm.bilateral_radius) {
#define KF_BIL(R)                                                                            \
  case R:                                                                                    \
    k_bilateral<R><<<grid, 256, 0, c->stream>>>(c->depth, c->depth_tmp, c->spatial, c->H,    \
                                                c->W, sr);                                   \
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

hipError_t kf_launch_vertex(kf_context* c, const kf_cam& cam) {
  const int n = c->H * c->W;
  k_vertex<<<(n + 255) / 256, 256, 0, c->stream>>>(c->depth, cam, c->H, c->W, c->vmap, c->nmap,
                                                   c->mask);
  return hipGetLastError();
}

hipError_t kf_launch_icp(kf_context* c, const kf_cam& cam) {
  k_icp_begin<<<1, 64, 0, c->stream>>>(c->state);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int iters = c->prm.icp_iterations;
  const int n = c->H * c->W;
  for (int it = 0; it < iters; ++it) {
    uint8_t* m = (c->prm.record_icp && c->icp_masks) ? c->icp_masks + (size_t)it * n : nullptr;
    const int last = it == iters - 1;
    if (c->prm.icp_compensated)
      k_icp_iter<true><<<c->icp_blocks, KF_ICP_BLOCK, 0, c->stream>>>(
          c->vmap, c->nmap, c->model_v, c->model_n, cam, c->H, c->W, c->prm.icp_dist_mm,
          c->icp_cos_lo, c->state, c->icp_partials, m, it, last);
    else
      k_icp_iter<false><<<c->icp_blocks, KF_ICP_BLOCK, 0, c->stream>>>(
          c->vmap, c->nmap, c->model_v, c->model_n, cam, c->H, c->W, c->prm.icp_dist_mm,
          c->icp_cos_lo, c->state, c->icp_partials, m, it, last);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t kf_launch_integ_table(kf_context* c, const kf_cam& cam) {
  const int n = c->H * c->W;
  k_integ_table<<<(n + 255) / 256, 256, 0, c->stream>>>(c->depth, cam, c->H, c->W, c->state,
                                                        c->integ_tab);
  return hipGetLastError();
}

hipError_t kf_launch_integrate(kf_context* c, const kf_cam& cam, int first) {
  const int nv = (int)c->n_voxels;
  const int blocks = (int)((c->n_voxels + 256 * kIntegVPT - 1) / (256 * kIntegVPT));
  const kf_params& p = c->prm;
#define KF_INTEG(P, FI)                                                                  \
  k_integrate<P, FI><<<blocks, 256, 0, c->stream>>>(c->volume, c->integ_tab, cam, c->vg, \
                                                    c->H, c->W, p.truncation_mm,         \
                                                    p.sdf_min_mm, p.sdf_max_mm, c->state, nv)
  if (c->vol.precision == KF_TSDF_F32) {
    if (first)
      KF_INTEG(KF_TSDF_F32, true);
    else
      KF_INTEG(KF_TSDF_F32, false);
  } else {
    if (first)
      KF_INTEG(KF_TSDF_F16, true);
    else
      KF_INTEG(KF_TSDF_F16, false);
  }
#undef KF_INTEG
  return hipGetLastError();
}

hipError_t kf_launch_raycast(kf_context* c, const kf_cam& cam) {
  const int tiles_x = (c->W + 15) / 16, tiles_y = (c->H + 15) / 16;
  const int nb = tiles_x * tiles_y;
  const kf_params& p = c->prm;
  if (c->vol.precision == KF_TSDF_F32)
    k_raycast<KF_TSDF_F32><<<nb, 256, 0, c->stream>>>(
        c->volume, c->depth, cam, c->vg, c->H, c->W, p.truncation_mm, p.raycast_fix_bounds,
        c->flip_cos_hi, c->state, c->model_v, c->model_n, c->hit, tiles_x, nb);
  else
    k_raycast<KF_TSDF_F16><<<nb, 256, 0, c->stream>>>(
        c->volume, c->depth, cam, c->vg, c->H, c->W, p.truncation_mm, p.raycast_fix_bounds,
        c->flip_cos_hi, c->state, c->model_v, c->model_n, c->hit, tiles_x, nb);
  return hipGetLastError();
}

hipError_t kf_launch_volume_fill(kf_context* c) {
  const int64_t nv = c->n_voxels;
  k_volume_fill<<<(unsigned)((nv + 255) / 256), 256, 0, c->stream>>>(c->volume,
                                                                      c->vol.precision, nv);
  return hipGetLastError();
}
