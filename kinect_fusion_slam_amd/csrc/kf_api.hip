// kf_api.hip - the C ABI: context lifetime, transfers, stage entry points and
// the KinfuPipeline frame sequence.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "kf_internal.h"
#include "kf_jpeg.h"

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_last_error;

void kf_set_error(const std::string& msg) { g_last_error = msg; }

int kf_fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define KF_REQUIRE(cond, code, msg) \
  do {                              \
    if (!(cond)) return kf_fail((code), (msg)); \
  } while (0)

#define KF_LAUNCH(expr)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return kf_fail(KF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

// ---------------------------------------------------------------------------
// small kernels owned by this file

namespace {

struct kf_pose_arg {
  float T[16];
  float Tinv[16];
};

__global__ void k_set_pose(kf_dev_state* st, kf_pose_arg p) {
  const int i = threadIdx.x;
  if (i < 16) {
    st->T[i] = p.T[i];
    st->Tinv[i] = p.Tinv[i];
  }
}

__global__ void k_set_icp_stop(kf_dev_state* st, int from, float rot, float trans) {
  if (threadIdx.x == 0) {
    st->icp_stop_at = 0;
    st->icp_stop_from = from;
    st->icp_stop_rot = rot;
    st->icp_stop_trans = trans;
  }
}

// Order-preserving map of floats to integers (for bisection over floats).
int64_t f2ord(float f) {
  int32_t b;
  memcpy(&b, &f, 4);
  return b >= 0 ? (int64_t)b : -(int64_t)(b & 0x7fffffff);
}
float ord2f(int64_t o) {
  int32_t b = o >= 0 ? (int32_t)o : (int32_t)((-o) | 0x80000000ll);
  float f;
  memcpy(&f, &b, 4);
  return f;
}

const float kDegF = 180.0f / 3.141592654f;  // 180.0f / CUDART_PI_F

// Smallest float c in [-1, 1] with (float)acos(c) * kDegF <= gate, i.e. the
// reference's ICP normal gate (sensor_pose_estimation.cu:63-69) for a
// correctly rounded acosf.  acos is monotone, so the gate holds exactly on
// [c, 1].
float icp_cos_threshold(float gate_deg) {
  auto pass = [&](float c) {
    const float ang = (float)acos((double)c) * kDegF;
    return ang <= gate_deg && ang >= 0.0f;
  };
  if (pass(-1.0f)) return -1.0f;
  if (!pass(1.0f)) return 2.0f;  // nothing passes
  int64_t lo = f2ord(-1.0f), hi = f2ord(1.0f);  // pass(lo) false, pass(hi) true
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (pass(ord2f(mid)))
      hi = mid;
    else
      lo = mid;
  }
  return ord2f(hi);
}

// Largest float s with the reference's distance gate passing for a squared
// distance s: sqrtf(s) <= gate && sqrtf(s) >= 0 (sensor_pose_estimation.cu:
// 47-48) for a correctly rounded sqrtf.  sqrtf is monotone, so the gate holds
// exactly on [0, s]; -1 when no distance passes.
float icp_dist2_threshold(float gate_mm) {
  auto pass = [&](float s) {
    const float d = sqrtf(s);
    return d <= gate_mm && d >= 0.0f;
  };
  const float inf = INFINITY;
  if (!pass(0.0f)) return -1.0f;
  if (pass(inf)) return inf;
  int64_t lo = f2ord(0.0f), hi = f2ord(inf);  // pass(lo) true, pass(hi) false
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (pass(ord2f(mid)))
      lo = mid;
    else
      hi = mid;
  }
  return ord2f(lo);
}

// Smallest float c with !(acos((double)c) * kDegF > 90): the normal is flipped
// (voxel_traversal.cu:361-369) exactly for -1 <= c < this value.
float flip_cos_threshold() {
  auto flip = [](float c) { return acos((double)c) * (double)kDegF > 90; };
  int64_t lo = f2ord(-1.0f), hi = f2ord(1.0f);  // flip(lo) true, flip(hi) false
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (flip(ord2f(mid)))
      lo = mid;
    else
      hi = mid;
  }
  return ord2f(hi);
}

bool params_valid(const kf_params* p, std::string* why) {
  if (p->bilateral_radius < 1 || p->bilateral_radius > KF_MAX_BILATERAL_RADIUS) {
    *why = "bilateral_radius must be in [1, 8]";
    return false;
  }
  if (p->icp_iterations < 0 || p->icp_iterations > KF_MAX_ICP_ITERS) {
    *why = "icp_iterations must be in [0, 64]";
    return false;
  }
  if (!(p->truncation_mm > 0) || !(p->sigma_s_px > 0) || !(p->sigma_r_mm > 0)) {
    *why = "truncation and sigmas must be positive";
    return false;
  }
  if (p->raycast_mode != KF_RAYCAST_REFERENCE && p->raycast_mode != KF_RAYCAST_TRILINEAR) {
    *why = "raycast_mode must be KF_RAYCAST_REFERENCE or KF_RAYCAST_TRILINEAR";
    return false;
  }
  if (p->refine_steps < 0 || p->refine_steps > KF_MAX_REFINE_STEPS ||
      !(p->refine_max_residual_mm > 0)) {
    *why = "refine_steps must be in [0, 64] and refine_max_residual_mm positive";
    return false;
  }
  if ((p->icp_robust != 0 && p->icp_robust != 1) || p->icp_min_inliers < 0) {
    *why = "icp_robust must be 0 or 1 and icp_min_inliers non-negative";
    return false;
  }
  if (p->icp_levels < 1 || p->icp_levels > KF_MAX_ICP_LEVELS || p->icp_coarse_iters[0] < 0 ||
      p->icp_coarse_iters[1] < 0 || !(p->pyramid_max_step_mm >= 0)) {
    *why = "icp_levels must be in [1, 3], icp_coarse_iters and pyramid_max_step_mm non-negative";
    return false;
  }
  if (!(p->icp_stop_rot >= 0 && p->icp_stop_rot < INFINITY) ||
      !(p->icp_stop_trans_mm >= 0 && p->icp_stop_trans_mm < INFINITY)) {
    *why = "icp_stop_rot and icp_stop_trans_mm must be finite and non-negative";
    return false;
  }
  if (p->raycast_skip != 0 && p->raycast_skip != 1) {
    *why = "raycast_skip must be 0 or 1";
    return false;
  }
  int total = p->icp_iterations;
  for (int l = 1; l < p->icp_levels; ++l) total += p->icp_coarse_iters[l - 1];
  if (total > KF_MAX_ICP_ITERS) {
    *why = "ICP iterations over all pyramid levels must be at most 64";
    return false;
  }
  return true;
}

// ICP grid: one block per compute unit, the image split evenly over all of
// them (a CU with two blocks would set the time of every iteration, and each
// block's two waves per SIMD hide each other's VALU latency).  Up to 4 pixels
// per thread of a KF_ICP_BLOCK-thread block.  Larger images take blocks of
// 2 KF_ICP_BLOCK threads (3 or 4 pixels each, still one block per compute
// unit): every block reads every block partial of the iteration before, so
// the block count squared sets those reads (at 1280x720, 450 blocks of 512
// threads read 45 MB per iteration, 256 of 1024 threads 15 MB).  Beyond
// that, 4 pixels per thread of KF_ICP_BLOCK threads and more blocks
// (KF_ICP_MAX_BLOCKS blocks cover 2^20 pixels).  The grid fixes the order of
// the ICP sums.
void icp_grid(size_t n, int n_cu, int* blocks, int* bpx, int* px, int* threads) {
  *threads = KF_ICP_BLOCK;
  *px = 4;
  *bpx = KF_ICP_BLOCK * 4;
  *blocks = (int)((n + *bpx - 1) / *bpx);
  const int nb = (int)std::min<size_t>((size_t)std::min(n_cu, KF_ICP_MAX_BLOCKS),
                                       (n + KF_ICP_BLOCK - 1) / KF_ICP_BLOCK);
  const int b = (int)((n + nb - 1) / nb);
  if (b <= 4 * KF_ICP_BLOCK) {
    *blocks = (int)((n + b - 1) / b);
    *bpx = b;
    *px = (b + KF_ICP_BLOCK - 1) / KF_ICP_BLOCK;
  } else if (b <= 8 * KF_ICP_BLOCK) {
    *threads = 2 * KF_ICP_BLOCK;
    *blocks = (int)((n + b - 1) / b);
    *bpx = b;
    *px = (b + 2 * KF_ICP_BLOCK - 1) / (2 * KF_ICP_BLOCK);  // 3 or 4
  }
}

int device_cus(int device) {
  int n_cu = 0;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) !=
          hipSuccess ||
      n_cu <= 0)
    n_cu = KF_ICP_MAX_BLOCKS;
  return n_cu;
}

void host_spatial_kernel(int r, float sigma, float* w) {
  const int L = 2 * r + 1;
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < L; ++j) {
      const float X = (float)(j - r), Y = (float)(i - r);
      const float t = -(X * X + Y * Y) / (2 * sigma * sigma);
      w[i * L + j] = kf_expf(t);
    }
}

}  // namespace

int kf_icp_stop_from(const kf_params& p) {
  int total = p.icp_iterations;
  for (int l = 1; l < p.icp_levels; ++l) total += p.icp_coarse_iters[l - 1];
  return total - p.icp_iterations + 1;
}

hipError_t kf_launch_set_icp_stop(kf_dev_state* st, const kf_params& p, hipStream_t s) {
  const bool on = kf_icp_may_stop(p);
  k_set_icp_stop<<<1, 64, 0, s>>>(st, on ? kf_icp_stop_from(p) : KF_MAX_ICP_ITERS + 1,
                                   on ? p.icp_stop_rot : 0.0f, on ? p.icp_stop_trans_mm : 0.0f);
  return hipGetLastError();
}


void kf_make_cam(const float* K, kf_cam* cam) {
  for (int i = 0; i < 9; ++i) cam->K[i] = K[i];
  if (kf_m3_inverse(K, cam->Kinv) != 0)
    for (int i = 0; i < 9; ++i) cam->Kinv[i] = NAN;
}

// ---------------------------------------------------------------------------
// stage timing: a pair of events per stage invocation, drained on demand

namespace {
struct PendingEv {
  int stage;
  hipEvent_t a, b;
};
}  // namespace

struct kf_timing {
  std::vector<PendingEv> pending;
  std::vector<hipEvent_t> pool;
  hipEvent_t open[KF_STAGE_COUNT] = {};
};

static kf_timing* timing_of(kf_context* c);

void kf_stage_begin(kf_context* c, int stage) {
  if (!c->prm.stage_timing) return;
  kf_timing* t = timing_of(c);
  hipEvent_t e;
  if (!t->pool.empty()) {
    e = t->pool.back();
    t->pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    return;
  }
  (void)hipEventRecord(e, c->stream);
  t->open[stage] = e;
}

void kf_stage_end(kf_context* c, int stage) {
  if (!c->prm.stage_timing) return;
  kf_timing* t = timing_of(c);
  if (!t->open[stage]) return;
  hipEvent_t e;
  if (!t->pool.empty()) {
    e = t->pool.back();
    t->pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    return;
  }
  (void)hipEventRecord(e, c->stream);
  t->pending.push_back({stage, t->open[stage], e});
  t->open[stage] = nullptr;
}

// The timing record hangs off the context through a side table so that
// kf_context stays a plain struct shared with the kernels' header.
#include <map>
#include <mutex>
static std::mutex g_timing_mu;
static std::map<const kf_context*, kf_timing*> g_timing;
static kf_timing* timing_of(kf_context* c) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  auto it = g_timing.find(c);
  if (it != g_timing.end()) return it->second;
  kf_timing* t = new kf_timing();
  g_timing[c] = t;
  return t;
}
static void timing_drop(kf_context* c) {
  kf_timing* t = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    auto it = g_timing.find(c);
    if (it == g_timing.end()) return;
    t = it->second;
    g_timing.erase(it);
  }
  for (auto& p : t->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : t->pool) (void)hipEventDestroy(e);
  for (auto e : t->open)
    if (e) (void)hipEventDestroy(e);
  delete t;
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char* kf_last_error(void) { return g_last_error.c_str(); }

int kf_struct_sizes(size_t sizes[3]) {
  if (!sizes) return kf_fail(KF_ERR_INVALID, "kf_struct_sizes: sizes is NULL");
  sizes[0] = sizeof(kf_params);
  sizes[1] = sizeof(kf_volume_desc);
  sizes[2] = sizeof(kf_synth_scene);
  return KF_OK;
}

int kf_preset(int32_t which, kf_volume_desc* vol, kf_params* prm, float K[9], int32_t* height,
              int32_t* width) {
  kf_volume_desc v{};
  kf_params p{};
  float k[9] = {525.0f, 0.0f, 319.5f, 0.0f, 525.0f, 239.5f, 0.0f, 0.0f, 1.0f};
  int H = 480, W = 640;
  // settings shared by the synthetic presets.  The SDF band keeps the
  // reference's shape [-trunc, 10 trunc] (kinfu_pipeline.h:52-55 has trunc
  // 300, band [-30, 300]): the traversal's interpolation taps reach back to
  // voxels several steps in front of the surface, which must be observed.
  v = {{360, 360, 300}, {-200, -200, -120}, 10.0f, KF_TSDF_F32};
  p.truncation_mm = 30.0f;
  p.sdf_min_mm = -30.0f;
  p.sdf_max_mm = 300.0f;
  p.depth_min_mm = 300.0f;
  p.depth_max_mm = 4000.0f;
  p.bilateral = 1;
  p.bilateral_radius = 2;
  p.sigma_s_px = 4.5f;
  p.sigma_r_mm = 30.0f;
  p.icp_iterations = 10;
  p.icp_dist_mm = 100.0f;
  p.icp_angle_deg = 20.0f;
  p.icp_compensated = 0;
  p.record_icp = 0;
  p.stage_timing = 0;
  p.raycast_fix_bounds = 0;
  p.raycast_mode = KF_RAYCAST_REFERENCE;
  p.refine_steps = 0;
  p.refine_max_residual_mm = 50.0f;
  p.icp_robust = 0;
  p.icp_min_inliers = 0;
  p.icp_levels = 1;  // the reference's single-resolution ICP
  p.icp_coarse_iters[0] = 5;
  p.icp_coarse_iters[1] = 4;
  p.pyramid_max_step_mm = 30.0f;
  p.icp_stop_rot = 0.0f;  // the reference's fixed iteration count
  p.icp_stop_trans_mm = 0.0f;
  p.raycast_skip = 0;  // skipping pays off on the long walks of large volumes (HD preset)
  switch (which) {
    case KF_PRESET_REFERENCE: {
      // kinfu_pipeline.cpp:14-24, kinfu_pipeline.h:42-62, system_interface.cpp:6-11,
      // sensor_pose_estimation.cu:48, 69, 339; bilateral disabled (kinfu_pipeline.cpp:141-145)
      const float kr[9] = {524.639f, 0.0f, 316.625f, 0.0f, 523.503f, 256.2318f, 0.0f, 0.0f, 1.0f};
      memcpy(k, kr, sizeof k);
      p.truncation_mm = 300.0f;
      p.sdf_min_mm = -30.0f;
      p.sdf_max_mm = 300.0f;
      p.depth_max_mm = 2000.0f;
      p.bilateral = 0;
      p.bilateral_radius = 3;
      p.icp_dist_mm = 20.0f;
      p.icp_angle_deg = 30.0f;
      break;
    }
    case KF_PRESET_ROOM_CPU:
    case KF_PRESET_ROOM:
      break;
    case KF_PRESET_ROOM_5MM_F16:
      v = {{720, 720, 600}, {-400, -400, -240}, 5.0f, KF_TSDF_F16};
      p.truncation_mm = 15.0f;
      p.sdf_min_mm = -15.0f;
      p.sdf_max_mm = 150.0f;
      p.icp_compensated = 1;
      break;
    case KF_PRESET_SPHERES_128: {
      v = {{128, 128, 128}, {-64, -64, 36}, 10.0f, KF_TSDF_F32};
      const float k3[9] = {262.5f, 0.0f, 159.5f, 0.0f, 262.5f, 119.5f, 0.0f, 0.0f, 1.0f};
      memcpy(k, k3, sizeof k);
      H = 240;
      W = 320;
      p.bilateral = 0;
      p.truncation_mm = 50.0f;  // covers the 2 cm pose jitter of the gradient checks
      p.sdf_min_mm = -50.0f;
      p.sdf_max_mm = 500.0f;
      p.raycast_mode = KF_RAYCAST_TRILINEAR;
      break;
    }
    case KF_PRESET_ROOM_HD_F16: {
      v = {{1024, 1024, 768}, {-512, -512, -300}, 4.0f, KF_TSDF_F16};
      const float k4[9] = {1000.0f, 0.0f, 639.5f, 0.0f, 1000.0f, 359.5f, 0.0f, 0.0f, 1.0f};
      memcpy(k, k4, sizeof k);
      H = 720;
      W = 1280;
      p.truncation_mm = 20.0f;
      p.sdf_min_mm = -20.0f;
      p.sdf_max_mm = 200.0f;
      p.icp_compensated = 1;
      p.raycast_skip = 1;  // hole rays walk up to ~2500 voxels
      break;
    }
    default:
      return kf_fail(KF_ERR_INVALID, "unknown preset");
  }
  if (vol) *vol = v;
  if (prm) *prm = p;
  if (K) memcpy(K, k, sizeof k);
  if (height) *height = H;
  if (width) *width = W;
  return KF_OK;
}

int kf_gate_thresholds(float icp_angle_deg, float* icp_lo, float* flip_hi) {
  if (icp_lo) *icp_lo = icp_cos_threshold(icp_angle_deg);
  if (flip_hi) *flip_hi = flip_cos_threshold();
  return KF_OK;
}

int kf_dist_gate_threshold(float icp_dist_mm, float* dist2_hi) {
  KF_REQUIRE(dist2_hi, KF_ERR_INVALID, "null pointer");
  *dist2_hi = icp_dist2_threshold(icp_dist_mm);
  return KF_OK;
}

float kf_expf_abi(float x) { return kf_expf(x); }

int kf_mat4_inverse(const float T[16], float Tinv[16]) {
  KF_REQUIRE(T && Tinv, KF_ERR_INVALID, "null pointer");
  return kf_m4_inverse((const float*)T, (float*)Tinv) ? kf_fail(KF_ERR_INVALID, "singular")
                                                         : KF_OK;
}

int kf_mat3_inverse(const float K[9], float Kinv[9]) {
  KF_REQUIRE(K && Kinv, KF_ERR_INVALID, "null pointer");
  return kf_m3_inverse((const float*)K, (float*)Kinv) ? kf_fail(KF_ERR_INVALID, "singular")
                                                         : KF_OK;
}

int kf_ldlt6(const double A[36], const double b[6], double x[6]) {
  KF_REQUIRE(A && b && x, KF_ERR_INVALID, "null pointer");
  return kf_ldlt6_solve(A, b, x);
}

int kf_spatial_kernel(int32_t radius, float sigma_s_px, float* weights) {
  KF_REQUIRE(weights && radius >= 1 && radius <= KF_MAX_BILATERAL_RADIUS, KF_ERR_INVALID,
             "bad spatial kernel arguments");
  host_spatial_kernel(radius, sigma_s_px, weights);
  return KF_OK;
}

int kf_destroy(kf_context* c) {
  if (!c) return KF_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  timing_drop(c);
  void* bufs[] = {c->ray_queue, c->ray_counts, c->integ_groups, c->integ_counts, c->depth,   c->depth_tmp, c->depth_u16, c->vmap,         c->nmap,
                  c->mask,    c->model_v,   c->model_n,   c->model_nt, c->hit, c->integ_tab,
                  c->integ_tiles,
                  c->spatial, c->range_lut, c->volume,    c->state,     c->icp_partials, c->icp_masks,
                  c->frames,  c->poses,     c->grad_tsdf, c->adj_depth,    c->adj_normal,
                  c->adj_partials, c->gn_partials, c->adj_touched, c->adj_map_v,
                  c->adj_map_n, c->adj_map_hit, c->model_aux, c->adj_map_aux,
                  c->color, c->jpeg_coef, c->jpeg_planes, c->jpeg_q, c->cloud,
                  c->skip_flags};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->jpeg_coef_host) (void)hipHostFree(c->jpeg_coef_host);
  if (c->jpeg_q_host) (void)hipHostFree(c->jpeg_q_host);
  if (c->jpeg_copied) (void)hipEventDestroy(c->jpeg_copied);
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  for (int i = 0; i < 2; ++i) {
    if (c->up_host[i]) (void)hipHostFree(c->up_host[i]);
    if (c->up_copied[i]) (void)hipEventDestroy(c->up_copied[i]);
    if (c->up_consumed[i]) (void)hipEventDestroy(c->up_consumed[i]);
  }
  if (c->up_dev[1]) (void)hipFree(c->up_dev[1]);
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  for (kf_pyr_level& P : c->pyr) {
    void* pb[] = {P.depth, P.vmap, P.nmap, P.mask, P.mv, P.mn, P.mt, P.ok};
    for (void* b : pb)
      if (b) (void)hipFree(b);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return KF_OK;
}

// Buffers of the coarse ICP levels 1 .. levels-1 (kept once allocated).
static int alloc_pyramid(kf_context* c, int levels) {
  int H = c->H, W = c->W;
  for (int l = 1; l < levels; ++l) {
    H /= 2;
    W /= 2;
    kf_pyr_level& P = c->pyr[l];
    if (P.depth) continue;
    KF_REQUIRE(H >= 2 && W >= 2, KF_ERR_INVALID, "image too small for this many ICP levels");
    const size_t n = (size_t)H * W;
    P.H = H;
    P.W = W;
    icp_grid(n, device_cus(c->device), &P.blocks, &P.bpx, &P.px, &P.threads);
    struct {
      void** p;
      size_t bytes;
    } a[] = {{(void**)&P.depth, n * 4}, {(void**)&P.vmap, n * 12}, {(void**)&P.nmap, n * 12},
             {(void**)&P.mask, n},      {(void**)&P.mv, n * 12},   {(void**)&P.mn, n * 12},
             {(void**)&P.mt, n * 4},    {(void**)&P.ok, n}};
    for (auto& b : a)
      if (hipMalloc(b.p, b.bytes) != hipSuccess) {
        *b.p = nullptr;
        for (auto& f : a) {
          if (*f.p) (void)hipFree(*f.p);
          *f.p = nullptr;
        }
        return kf_fail(KF_ERR_OOM, "cannot allocate the ICP pyramid");
      }
  }
  return KF_OK;
}

static int apply_params(kf_context* c, const kf_params* prm) {
  std::string why;
  KF_REQUIRE(params_valid(prm, &why), KF_ERR_INVALID, why);
  c->prm = *prm;
  float w[(2 * KF_MAX_BILATERAL_RADIUS + 1) * (2 * KF_MAX_BILATERAL_RADIUS + 1)];
  host_spatial_kernel(prm->bilateral_radius, prm->sigma_s_px, w);
  const int L = 2 * prm->bilateral_radius + 1;
  KF_HIP_CHECK(hipMemcpy(c->spatial, w, sizeof(float) * L * L, hipMemcpyHostToDevice));
  {
    // range weights of integer depth differences, with the filter's float
    // operations (the argument -(t*t)/den, then kf_expf); kf_expf returns
    // exactly 0 below its underflow bound, and the argument only decreases
    // with j, so the table ends where the bound is crossed
    std::vector<float> lut(KF_RANGE_LUT_MAX);
    const float sr = prm->sigma_r_mm;
    const float den = 2 * sr * sr;
    int n = 0;
    for (; n < KF_RANGE_LUT_MAX; ++n) {
      const float t = (float)n;
      const float x = -(t * t) / den;
      if (x < -103.972084045410156f) break;
      lut[n] = kf_expf(x);
    }
    c->range_lut_n = n < KF_RANGE_LUT_MAX ? n : 0;
    if (c->range_lut_n > 0)
      KF_HIP_CHECK(hipMemcpy(c->range_lut, lut.data(), sizeof(float) * n, hipMemcpyHostToDevice));
  }
  c->icp_cos_lo = icp_cos_threshold(prm->icp_angle_deg);
  c->icp_dist2_hi = icp_dist2_threshold(prm->icp_dist_mm);
  c->flip_cos_hi = flip_cos_threshold();
  if (prm->record_icp && !c->icp_masks) {
    const size_t bytes = (size_t)KF_MAX_ICP_ITERS * c->H * c->W;
    if (hipMalloc(&c->icp_masks, bytes) != hipSuccess) {
      c->icp_masks = nullptr;
      return kf_fail(KF_ERR_OOM, "cannot allocate ICP mask record");
    }
    KF_HIP_CHECK(hipMemset(c->icp_masks, 0, bytes));
  }
  return alloc_pyramid(c, prm->icp_levels);
}

int kf_create(const kf_volume_desc* vol, int32_t height, int32_t width, const kf_params* prm,
              kf_context** out) {
  KF_REQUIRE(vol && prm && out, KF_ERR_INVALID, "null argument");
  *out = nullptr;
  KF_REQUIRE(height > 1 && width > 1 && (int64_t)height * width < (1ll << 28), KF_ERR_INVALID,
             "bad image size");
  for (int i = 0; i < 3; ++i)
    KF_REQUIRE(vol->dims[i] >= 3 && vol->dims[i] <= 65536, KF_ERR_INVALID, "bad volume dims");
  const int64_t nv = (int64_t)vol->dims[0] * vol->dims[1] * vol->dims[2];
  {
    const int64_t bricks = (int64_t)((vol->dims[0] + 7) / 8) * ((vol->dims[1] + 7) / 8) *
                           ((vol->dims[2] + 7) / 8);
    KF_REQUIRE(bricks * 512 < (1ll << 31), KF_ERR_INVALID,
               "volume must hold fewer than 2^31 voxels (in 8^3 bricks)");
  }
  KF_REQUIRE(vol->voxel_mm > 0, KF_ERR_INVALID, "voxel size must be positive");
  KF_REQUIRE(vol->precision == KF_TSDF_F32 || vol->precision == KF_TSDF_F16, KF_ERR_INVALID,
             "unknown precision");
  std::string why;
  KF_REQUIRE(params_valid(prm, &why), KF_ERR_INVALID, why);

  kf_context* c = new kf_context();
  c->H = height;
  c->W = width;
  c->vol = *vol;
  c->n_voxels = nv;
  c->vg.nx = vol->dims[0];
  c->vg.ny = vol->dims[1];
  c->vg.nz = vol->dims[2];
  c->vg.sx = vol->start[0];
  c->vg.sy = vol->start[1];
  c->vg.sz = vol->start[2];
  c->vg.d = vol->voxel_mm;
  c->vg.bx = (vol->dims[0] + KF_BRICK - 1) / KF_BRICK;
  c->vg.by = (vol->dims[1] + KF_BRICK - 1) / KF_BRICK;
  c->vg.bz = (vol->dims[2] + KF_BRICK - 1) / KF_BRICK;
  c->n_bricks = c->vg.bx * c->vg.by * c->vg.bz;
  c->n_store = (int64_t)c->n_bricks * 512;

  const size_t n = (size_t)height * width;
  // ICP grid before the device is known (the block-count check)
  icp_grid(n, KF_ICP_MAX_BLOCKS, &c->icp_blocks, &c->icp_bpx, &c->icp_px, &c->icp_threads);
  c->volume_bytes = (size_t)c->n_store * (vol->precision == KF_TSDF_F32 ? 8 : 4);
  if (c->icp_blocks > KF_ICP_MAX_BLOCKS) {
    delete c;
    return kf_fail(KF_ERR_INVALID, "image too large for the ICP reduction (max 2^20 pixels)");
  }

  auto fail = [&](int code, const std::string& m) {
    kf_destroy(c);
    return kf_fail(code, m);
  };
  if (hipGetDevice(&c->device) != hipSuccess) return fail(KF_ERR_HIP, "no HIP device");
  icp_grid(n, device_cus(c->device), &c->icp_blocks, &c->icp_bpx, &c->icp_px,
           &c->icp_threads);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(KF_ERR_HIP, "hipStreamCreate failed");
  struct Alloc {
    void** p;
    size_t bytes;
  } allocs[] = {
      {(void**)&c->depth, n * 4},       {(void**)&c->depth_tmp, n * 4},
      {(void**)&c->depth_u16, n * 2},   {(void**)&c->vmap, n * 12},
      {(void**)&c->nmap, n * 12},       {(void**)&c->mask, n},
      {(void**)&c->model_v, n * 12},    {(void**)&c->model_n, n * 12},
      {(void**)&c->model_nt, n * 4},
      {(void**)&c->model_aux, n * 8},
      {(void**)&c->hit, n * 4},         {(void**)&c->integ_tab, n * 8},
      {(void**)&c->color, n * 3},
      {(void**)&c->integ_tiles, (size_t)((width + 15) / 16) * ((height + 15) / 16) * 8},
      {&c->ray_queue, kf_ray_queue_entries(height, width) * kf_ray_state_bytes()},
      {(void**)&c->ray_counts, (size_t)2 * KF_RAY_PARTS * KF_RAY_COUNT_STRIDE * 4},
      {(void**)&c->integ_groups, kf_integ_group_entries(c->vg) * 4},
      {(void**)&c->integ_counts, (size_t)4 * KF_RAY_PARTS * KF_RAY_COUNT_STRIDE * 4},
      {(void**)&c->spatial, 4 * (2 * KF_MAX_BILATERAL_RADIUS + 1) * (2 * KF_MAX_BILATERAL_RADIUS + 1)},
      {(void**)&c->range_lut, 4 * KF_RANGE_LUT_MAX},
      {(void**)&c->state, sizeof(kf_dev_state)},
      {(void**)&c->icp_partials, (size_t)2 * c->icp_blocks * KF_ICP_SUMS * 2 * 4},
      {(void**)&c->gn_partials, kf_refine_partials(c) * 4},
      {(void**)&c->skip_flags, (size_t)c->n_bricks},
      {&c->volume, c->volume_bytes},
  };
  for (auto& a : allocs) {
    if (hipMalloc(a.p, a.bytes) != hipSuccess) {
      *a.p = nullptr;
      return fail(KF_ERR_OOM, "device allocation failed");
    }
  }
  c->tsdf = c->volume;
  c->weight = (char*)c->volume + c->volume_bytes / 2;
  int rc = apply_params(c, prm);
  if (rc != KF_OK) {
    const std::string m = g_last_error;
    return fail(rc, m);
  }
  void* zero_bufs[] = {c->depth,   c->depth_tmp, c->vmap,      c->nmap, c->mask,
                       c->model_v, c->model_n,   c->model_aux, c->color};
  size_t zero_sizes[] = {n * 4, n * 4, n * 12, n * 12, n, n * 12, n * 12, n * 8, n * 3};
  for (int i = 0; i < 9; ++i)
    if (hipMemsetAsync(zero_bufs[i], 0, zero_sizes[i], c->stream) != hipSuccess)
      return fail(KF_ERR_HIP, "hipMemsetAsync failed");
  if (hipMemsetAsync(c->hit, 0xff, n * 4, c->stream) != hipSuccess ||
      hipMemsetAsync(c->state, 0, sizeof(kf_dev_state), c->stream) != hipSuccess ||
      hipMemsetAsync(c->ray_counts, 0, (size_t)2 * KF_RAY_PARTS * KF_RAY_COUNT_STRIDE * 4,
                     c->stream) != hipSuccess ||
      hipMemsetAsync(c->integ_counts, 0, (size_t)4 * KF_RAY_PARTS * KF_RAY_COUNT_STRIDE * 4,
                     c->stream) != hipSuccess)
    return fail(KF_ERR_HIP, "hipMemsetAsync failed");
  if (kf_launch_set_icp_stop(c->state, c->prm, c->stream) != hipSuccess ||
      kf_launch_fill_f32(c->model_nt, (int64_t)n, INFINITY, c->stream) != hipSuccess)
    return fail(KF_ERR_HIP, "kernel launch failed");
  *out = c;
  rc = kf_reset(c);
  if (rc != KF_OK) {
    *out = nullptr;
    const std::string m = g_last_error;
    return fail(rc, m);
  }
  return KF_OK;
}

int kf_set_params(kf_context* c, const kf_params* prm) {
  KF_REQUIRE(c && prm, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  const int rc = apply_params(c, prm);
  if (rc != KF_OK) return rc;
  KF_LAUNCH(kf_launch_set_icp_stop(c->state, c->prm, c->stream));
  return KF_OK;
}

int kf_get_params(const kf_context* c, kf_params* prm) {
  KF_REQUIRE(c && prm, KF_ERR_INVALID, "null argument");
  *prm = c->prm;
  return KF_OK;
}

int kf_set_pose(kf_context* c, const float T[16]) {
  KF_REQUIRE(c && T, KF_ERR_INVALID, "null argument");
  kf_pose_arg a;
  memcpy(a.T, T, sizeof a.T);
  KF_REQUIRE(kf_m4_inverse(a.T, a.Tinv) == 0, KF_ERR_INVALID, "pose is singular");
  k_set_pose<<<1, 64, 0, c->stream>>>(c->state, a);
  KF_HIP_CHECK(hipGetLastError());
  return KF_OK;
}

int kf_reset(kf_context* c) {
  KF_REQUIRE(c, KF_ERR_INVALID, "null context");
  c->saved_ok = false;
  KF_LAUNCH(kf_launch_volume_fill(c));
  KF_LAUNCH(kf_launch_skip_flags_clear(c));
  const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  return kf_set_pose(c, I);
}

int kf_synchronize(kf_context* c) {
  KF_REQUIRE(c, KF_ERR_INVALID, "null context");
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  return KF_OK;
}

int kf_device_bytes(const kf_context* c, size_t* bytes) {
  KF_REQUIRE(c && bytes, KF_ERR_INVALID, "null argument");
  const size_t n = (size_t)c->H * c->W;
  // per pixel: depth, filtered depth, u16 staging, vertex and normal maps,
  // mask, model vertex, normal and normal factor, trilinear plane, hit index,
  // integration table, colour
  size_t b = c->volume_bytes + n * (4 + 4 + 2 + 12 + 12 + 1 + 12 + 12 + 4 + 8 + 4 + 8 + 3) +
             sizeof(kf_dev_state) +
             (size_t)2 * c->icp_blocks * KF_ICP_SUMS * 8 +
             kf_ray_queue_entries(c->H, c->W) * kf_ray_state_bytes() +
             kf_integ_group_entries(c->vg) * 4 + (size_t)c->n_bricks +
             (size_t)6 * KF_RAY_PARTS * KF_RAY_COUNT_STRIDE * 4;
  if (c->icp_masks) b += (size_t)KF_MAX_ICP_ITERS * n;
  for (const kf_pyr_level& P : c->pyr)
    if (P.depth) b += (size_t)P.H * P.W * (4 + 12 + 12 + 1 + 12 + 12 + 4 + 1);
  if (c->frames) b += (size_t)c->n_frames * n * 2 + (size_t)c->n_frames * 64;
  if (c->up_dev[1]) b += n * 2;
  *bytes = b;
  return KF_OK;
}

// ---- A1 -------------------------------------------------------------------

// Before the first reader of a host upload is enqueued: its copy to the
// device is done.
static int wait_upload(kf_context* c) {
  KF_HIP_CHECK(hipEventSynchronize(c->up_copied[c->up_cur]));
  return KF_OK;
}

// Runs the range gate of a host-uploaded frame (up_dev[up_cur] -> depth) if
// it is still pending.
static int flush_depth_gate(kf_context* c) {
  if (!c->depth_gate_pending) return KF_OK;
  c->depth_gate_pending = false;
  const int rc = wait_upload(c);
  if (rc != KF_OK) return rc;
  kf_stage_begin(c, KF_STAGE_DEPTH);
  KF_LAUNCH(kf_launch_depth_u16(c, c->up_dev[c->up_cur]));
  kf_stage_end(c, KF_STAGE_DEPTH);
  KF_HIP_CHECK(hipEventRecord(c->up_consumed[c->up_cur], c->stream));
  return KF_OK;
}

// The two upload slots, their pinned buffers, events and copy stream, all
// or none (created on the first host upload).
static int upload_slots(kf_context* c) {
  if (c->copy_stream) return KF_OK;
  const size_t bytes = (size_t)c->H * c->W * 2;
  bool ok = hipMalloc((void**)&c->up_dev[1], bytes) == hipSuccess;
  if (!ok) c->up_dev[1] = nullptr;
  for (int i = 0; i < 2 && ok; ++i) {
    ok = hipHostMalloc((void**)&c->up_host[i], bytes, hipHostMallocDefault) == hipSuccess;
    if (!ok) c->up_host[i] = nullptr;
  }
  for (int i = 0; i < 2 && ok; ++i) {
    ok = hipEventCreateWithFlags(&c->up_copied[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) c->up_copied[i] = nullptr;
    if (ok) ok = hipEventCreateWithFlags(&c->up_consumed[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) c->up_consumed[i] = nullptr;
  }
  hipStream_t s = nullptr;
  if (ok) ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
  // every event starts recorded, so the first waits on it return at once
  for (int i = 0; i < 2 && ok; ++i)
    ok = hipEventRecord(c->up_copied[i], s) == hipSuccess &&
         hipEventRecord(c->up_consumed[i], c->stream) == hipSuccess;
  if (!ok) {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 2; ++i) {
      if (c->up_host[i]) (void)hipHostFree(c->up_host[i]);
      if (c->up_copied[i]) (void)hipEventDestroy(c->up_copied[i]);
      if (c->up_consumed[i]) (void)hipEventDestroy(c->up_consumed[i]);
      c->up_host[i] = nullptr;
      c->up_copied[i] = c->up_consumed[i] = nullptr;
    }
    if (c->up_dev[1]) (void)hipFree(c->up_dev[1]);
    c->up_dev[1] = nullptr;
    (void)hipGetLastError();
    return kf_fail(KF_ERR_OOM, "cannot allocate the host upload slots");
  }
  c->up_dev[0] = c->depth_u16;
  c->copy_stream = s;
  return KF_OK;
}

// The working depth must be loaded; a pending range gate runs first.
static int need_depth(kf_context* c) {
  KF_REQUIRE(c->have_depth, KF_ERR_STATE, "no depth frame loaded");
  return flush_depth_gate(c);
}
#define KF_NEED_DEPTH(c)            \
  do {                              \
    const int rc_ = need_depth(c);  \
    if (rc_ != KF_OK) return rc_;   \
  } while (0)

int kf_upload_depth_u16(kf_context* c, const uint16_t* depth, int32_t mem_kind) {
  KF_REQUIRE(c && depth, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  if (mem_kind == KF_MEM_HOST) {
    // The copy is ours: the range gate waits for its first reader, so that
    // kf_pipeline_incoming(KF_FRAME_CURRENT) can gate and filter in one pass.
    int rc = upload_slots(c);
    if (rc != KF_OK) return rc;
    const size_t bytes = (size_t)c->H * c->W * 2;
    const int i = c->up_next;
    // The two streams are ordered from the host (a stream waiting on another
    // stream's event costs the host more than the copy itself): the slot's
    // pinned buffer is free once its previous copy is done, and its device
    // frame once the previous frame in it has been read; the reader waits
    // for this copy (wait_upload) before it is enqueued.
    KF_HIP_CHECK(hipEventSynchronize(c->up_copied[i]));
    memcpy(c->up_host[i], depth, bytes);
    KF_HIP_CHECK(hipEventSynchronize(c->up_consumed[i]));
    KF_HIP_CHECK(hipMemcpyAsync(c->up_dev[i], c->up_host[i], bytes, hipMemcpyHostToDevice,
                                c->copy_stream));
    KF_HIP_CHECK(hipEventRecord(c->up_copied[i], c->copy_stream));
    c->up_cur = i;
    c->up_next = i ^ 1;
    c->depth_gate_pending = true;
  } else {
    // the caller's device buffer may change after this call: gate it now
    c->depth_gate_pending = false;
    kf_stage_begin(c, KF_STAGE_DEPTH);
    KF_LAUNCH(kf_launch_depth_u16(c, depth));
    kf_stage_end(c, KF_STAGE_DEPTH);
  }
  c->have_depth = true;
  return KF_OK;
}

int kf_upload_depth_f32(kf_context* c, const float* depth_mm, int32_t mem_kind) {
  KF_REQUIRE(c && depth_mm, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  c->depth_gate_pending = false;
  KF_HIP_CHECK(hipMemcpyAsync(c->depth, depth_mm, (size_t)c->H * c->W * 4,
                              mem_kind == KF_MEM_HOST ? hipMemcpyHostToDevice
                                                      : hipMemcpyDeviceToDevice,
                              c->stream));
  c->have_depth = true;
  return KF_OK;
}

int kf_read_depth_file(const char* path, int32_t height, int32_t width, uint16_t* out) {
  KF_REQUIRE(path && out && height > 0 && width > 0, KF_ERR_INVALID, "bad argument");
  FILE* fp = fopen(path, "rb");
  KF_REQUIRE(fp, KF_ERR_IO, std::string("cannot open ") + path);
  const size_t n = (size_t)height * width;
  std::vector<unsigned char> raw(n * 2);
  const size_t got = fread(raw.data(), 1, n * 2, fp);
  fclose(fp);
  KF_REQUIRE(got == n * 2, KF_ERR_IO, std::string("short read from ") + path);
  for (size_t i = 0; i < n; ++i) out[i] = (uint16_t)(raw[2 * i] | (raw[2 * i + 1] << 8));
  return KF_OK;
}

int kf_download_depth(kf_context* c, float* depth_mm) {
  KF_REQUIRE(c && depth_mm, KF_ERR_INVALID, "null argument");
  const int rc = flush_depth_gate(c);
  if (rc != KF_OK) return rc;
  KF_HIP_CHECK(hipMemcpyAsync(depth_mm, c->depth, (size_t)c->H * c->W * 4, hipMemcpyDeviceToHost,
                              c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  return KF_OK;
}

// ---- A4 -------------------------------------------------------------------

int kf_bilateral(kf_context* c) {
  KF_REQUIRE(c, KF_ERR_INVALID, "null context");
  c->saved_ok = false;
  KF_NEED_DEPTH(c);
  kf_stage_begin(c, KF_STAGE_BILATERAL);
  KF_LAUNCH(kf_launch_bilateral(c));
  kf_stage_end(c, KF_STAGE_BILATERAL);
  return KF_OK;
}

// ---- A3 -------------------------------------------------------------------

int kf_vertex_normals(kf_context* c, const float K[9]) {
  KF_REQUIRE(c && K, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  KF_NEED_DEPTH(c);
  kf_cam cam;
  kf_make_cam(K, &cam);
  kf_stage_begin(c, KF_STAGE_VERTEX);
  KF_LAUNCH(kf_launch_vertex(c, cam));
  kf_stage_end(c, KF_STAGE_VERTEX);
  return KF_OK;
}

int kf_download_vertex_maps(kf_context* c, float* vx, float* vy, float* vz, float* nx, float* ny,
                            float* nz, uint8_t* mask) {
  KF_REQUIRE(c, KF_ERR_INVALID, "null context");
  const size_t n = (size_t)c->H * c->W;
  // the vertex map's z is 0 where the mask is 0; Vz proper is the depth
  std::vector<kf_p3> V(n), N(n);
  KF_HIP_CHECK(hipMemcpyAsync(V.data(), c->vmap, n * 12, hipMemcpyDeviceToHost, c->stream));
  KF_HIP_CHECK(hipMemcpyAsync(N.data(), c->nmap, n * 12, hipMemcpyDeviceToHost, c->stream));
  if (mask)
    KF_HIP_CHECK(hipMemcpyAsync(mask, c->mask, n, hipMemcpyDeviceToHost, c->stream));
  if (vz) {
    const int rc = flush_depth_gate(c);
    if (rc != KF_OK) return rc;
    KF_HIP_CHECK(hipMemcpyAsync(vz, c->depth, n * 4, hipMemcpyDeviceToHost, c->stream));
  }
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < n; ++i) {
    if (vx) vx[i] = V[i].x;
    if (vy) vy[i] = V[i].y;
    if (nx) nx[i] = N[i].x;
    if (ny) ny[i] = N[i].y;
    if (nz) nz[i] = N[i].z;
  }
  return KF_OK;
}

int kf_upload_vertex_maps(kf_context* c, const float* vx, const float* vy, const float* vz,
                          const float* nx, const float* ny, const float* nz,
                          const uint8_t* mask) {
  KF_REQUIRE(c && vx && vy && vz && nx && ny && nz && mask, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  const size_t n = (size_t)c->H * c->W;
  std::vector<kf_p3> V(n), N(n);
  for (size_t i = 0; i < n; ++i) {
    V[i] = kf_p3{vx[i], vy[i], mask[i] ? vz[i] : 0.0f};
    N[i] = kf_p3{nx[i], ny[i], nz[i]};
  }
  KF_HIP_CHECK(hipMemcpyAsync(c->vmap, V.data(), n * 12, hipMemcpyHostToDevice, c->stream));
  KF_HIP_CHECK(hipMemcpyAsync(c->nmap, N.data(), n * 12, hipMemcpyHostToDevice, c->stream));
  KF_HIP_CHECK(hipMemcpyAsync(c->mask, mask, n, hipMemcpyHostToDevice, c->stream));
  c->depth_gate_pending = false;
  KF_HIP_CHECK(hipMemcpyAsync(c->depth, vz, n * 4, hipMemcpyHostToDevice, c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  c->have_depth = true;
  return KF_OK;
}

// ---- A5-A7 ----------------------------------------------------------------

int kf_estimate_pose(kf_context* c, const float K[9], float T_inout[16]) {
  KF_REQUIRE(c && K && T_inout, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  int rc = kf_set_pose(c, T_inout);
  if (rc != KF_OK) return rc;
  kf_cam cam;
  kf_make_cam(K, &cam);
  kf_stage_begin(c, KF_STAGE_ICP);
  KF_LAUNCH(kf_launch_icp(c, cam));
  kf_stage_end(c, KF_STAGE_ICP);
  KF_HIP_CHECK(hipMemcpyAsync(T_inout, c->state->T, 64, hipMemcpyDeviceToHost, c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  return KF_OK;
}

int kf_icp_history(kf_context* c, int32_t* n_iters, float* T_iters, double* sums,
                   int32_t* status, uint8_t* masks) {
  KF_REQUIRE(c, KF_ERR_INVALID, "null context");
  kf_dev_state* h = new kf_dev_state;
  hipError_t e = hipMemcpyAsync(h, c->state, sizeof(kf_dev_state), hipMemcpyDeviceToHost,
                                c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    delete h;
    return kf_fail(KF_ERR_HIP, hipGetErrorString(e));
  }
  const int n = h->icp_iters_done;
  if (n_iters) *n_iters = n;
  if (T_iters) memcpy(T_iters, h->iter_T, sizeof(float) * 16 * n);
  if (sums) memcpy(sums, h->iter_sums, sizeof(double) * KF_ICP_SUMS * n);
  if (status) memcpy(status, h->iter_status, sizeof(int32_t) * n);
  delete h;
  if (masks) {
    KF_REQUIRE(c->icp_masks && c->prm.record_icp, KF_ERR_STATE, "record_icp is not enabled");
    KF_HIP_CHECK(hipMemcpy(masks, c->icp_masks, (size_t)n * c->H * c->W, hipMemcpyDeviceToHost));
  }
  return KF_OK;
}

}  // extern "C"

// ---- ICP with no device allocation (kf_estimate_pose_static) ---------------
// Scratch of the allocation-free ICP entry point: the block partials of two
// iteration parities, then the pose record, one copy per GPU that the
// library runs on, held until the process ends.  In host code `scratch` is
// a stand-in for the symbol, not a device address: host code reaches the
// array only through the per-device address hipGetSymbolAddress reports.
static constexpr size_t kIcpStaticPartialBytes = (size_t)2 * KF_ICP_MAX_BLOCKS * KF_ICP_SUMS * 8;
static_assert(kIcpStaticPartialBytes % 256 == 0, "the pose record starts on a 256-byte line");
static_assert(kIcpStaticPartialBytes + sizeof(kf_dev_state) <= KF_ICP_STATIC_SCRATCH_BYTES,
              "partials and pose record fit the static scratch");
static __device__ __align__(256) unsigned char scratch[KF_ICP_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_dev[kMaxDevices] = {};  // the array's address on each device

__global__ void k_pose_out(const kf_dev_state* st, float* T_out) {
  if (threadIdx.x < 16) T_out[threadIdx.x] = st->T[threadIdx.x];
}

// The current device's copy of `scratch`, looked up once per device with
// that device current.
int scratch_of_current_device(int* device, char** ptr) {
  KF_HIP_CHECK(hipGetDevice(device));
  KF_REQUIRE(*device >= 0 && *device < kMaxDevices, KF_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[*device]) {
    void* p = nullptr;
    KF_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    KF_REQUIRE(p, KF_ERR_HIP, "hipGetSymbolAddress returned no address for the ICP scratch");
    g_scratch_dev[*device] = static_cast<char*>(p);
  }
  *ptr = g_scratch_dev[*device];
  return KF_OK;
}
}  // namespace

extern "C" {

int kf_estimate_pose_static(const float* vmap, const float* nmap, const float* model_v,
                            const float* model_n, int32_t height, int32_t width,
                            const float K[9], const kf_params* prm, const float T_prev[16],
                            float* T_out, void* stream) {
  KF_REQUIRE(vmap && nmap && model_v && model_n && K && prm && T_prev && T_out, KF_ERR_INVALID,
             "null argument");
  KF_REQUIRE(height > 1 && width > 1 && (int64_t)height * width < (1ll << 28), KF_ERR_INVALID,
             "bad image size");
  KF_REQUIRE(prm->icp_iterations >= 0 && prm->icp_iterations <= KF_MAX_ICP_ITERS,
             KF_ERR_INVALID, "icp_iterations must be in [0, 64]");
  KF_REQUIRE((prm->icp_robust == 0 || prm->icp_robust == 1) && prm->icp_min_inliers >= 0,
             KF_ERR_INVALID, "icp_robust must be 0 or 1 and icp_min_inliers non-negative");
  KF_REQUIRE(prm->icp_levels == 1, KF_ERR_INVALID,
             "the allocation-free ICP runs one level (icp_levels = 1)");
  int device = 0;
  char* base = nullptr;
  int rc = scratch_of_current_device(&device, &base);
  if (rc != KF_OK) return rc;
  // a host-side view of the ICP's buffers: the caller's maps, the static
  // scratch and the caller's stream (nothing here owns device memory)
  kf_context c;
  c.device = device;
  c.H = height;
  c.W = width;
  c.prm = *prm;
  c.prm.record_icp = 0;
  c.stream = static_cast<hipStream_t>(stream);
  icp_grid((size_t)height * width, device_cus(device), &c.icp_blocks, &c.icp_bpx, &c.icp_px,
           &c.icp_threads);
  KF_REQUIRE(c.icp_blocks <= KF_ICP_MAX_BLOCKS, KF_ERR_INVALID,
             "image too large for the static ICP scratch (at most 2^20 pixels)");
  c.vmap = reinterpret_cast<kf_p3*>(const_cast<float*>(vmap));
  c.nmap = reinterpret_cast<kf_p3*>(const_cast<float*>(nmap));
  c.model_v = reinterpret_cast<kf_p3*>(const_cast<float*>(model_v));
  c.model_n = reinterpret_cast<kf_p3*>(const_cast<float*>(model_n));
  c.icp_partials = base;
  c.state = reinterpret_cast<kf_dev_state*>(base + kIcpStaticPartialBytes);
  c.icp_cos_lo = icp_cos_threshold(prm->icp_angle_deg);
  c.icp_dist2_hi = icp_dist2_threshold(prm->icp_dist_mm);
  kf_pose_arg a;
  memcpy(a.T, T_prev, sizeof a.T);
  KF_REQUIRE(kf_m4_inverse(a.T, a.Tinv) == 0, KF_ERR_INVALID, "pose is singular");
  kf_cam cam;
  kf_make_cam(K, &cam);
  k_set_pose<<<1, 64, 0, c.stream>>>(c.state, a);
  KF_HIP_CHECK(hipGetLastError());
  KF_LAUNCH(kf_launch_set_icp_stop(c.state, c.prm, c.stream));
  KF_LAUNCH(kf_launch_icp(&c, cam));
  k_pose_out<<<1, 64, 0, c.stream>>>(c.state, T_out);
  KF_HIP_CHECK(hipGetLastError());
  return KF_OK;
}

int kf_pyramid_level(kf_context* c, int32_t level, int32_t* height, int32_t* width, float K[9],
                     float* depth, float* vx, float* vy, float* vz, float* nx, float* ny,
                     float* nz, uint8_t* mask, float* mx, float* my, float* mz, float* mnx,
                     float* mny, float* mnz, uint8_t* model_ok) {
  KF_REQUIRE(c && level >= 1 && level < KF_MAX_ICP_LEVELS, KF_ERR_INVALID, "bad level");
  const kf_pyr_level& P = c->pyr[level];
  KF_REQUIRE(P.depth, KF_ERR_STATE, "the pyramid level is not allocated (icp_levels)");
  if (height) *height = P.H;
  if (width) *width = P.W;
  const size_t n = (size_t)P.H * P.W;
  std::vector<kf_p3> V(n), N(n), MV(n), MN(n);
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  KF_HIP_CHECK(hipMemcpy(V.data(), P.vmap, n * 12, hipMemcpyDeviceToHost));
  KF_HIP_CHECK(hipMemcpy(N.data(), P.nmap, n * 12, hipMemcpyDeviceToHost));
  KF_HIP_CHECK(hipMemcpy(MV.data(), P.mv, n * 12, hipMemcpyDeviceToHost));
  KF_HIP_CHECK(hipMemcpy(MN.data(), P.mn, n * 12, hipMemcpyDeviceToHost));
  if (depth) KF_HIP_CHECK(hipMemcpy(depth, P.depth, n * 4, hipMemcpyDeviceToHost));
  if (vz) KF_HIP_CHECK(hipMemcpy(vz, P.depth, n * 4, hipMemcpyDeviceToHost));
  if (mask) KF_HIP_CHECK(hipMemcpy(mask, P.mask, n, hipMemcpyDeviceToHost));
  if (model_ok) KF_HIP_CHECK(hipMemcpy(model_ok, P.ok, n, hipMemcpyDeviceToHost));
  float* out[10] = {vx, vy, nx, ny, nz, mx, my, mz, mnx, mny};
  for (size_t i = 0; i < n; ++i) {
    const float val[10] = {V[i].x, V[i].y, N[i].x, N[i].y, N[i].z,
                           MV[i].x, MV[i].y, MV[i].z, MN[i].x, MN[i].y};
    for (int k = 0; k < 10; ++k)
      if (out[k]) out[k][i] = val[k];
    if (mnz) mnz[i] = MN[i].z;
  }
  if (K) {
    // the intrinsics of the last ICP's level 0 are not kept: K is the level's
    // scaling of the K passed in
    float Kl[9];
    for (int i = 0; i < 9; ++i) Kl[i] = K[i];
    for (int l = 0; l < level; ++l) {
      float Kc[9];
      kf_pyramid_intrinsics(Kl, Kc);
      for (int i = 0; i < 9; ++i) Kl[i] = Kc[i];
    }
    for (int i = 0; i < 9; ++i) K[i] = Kl[i];
  }
  return KF_OK;
}

// ---- A8 -------------------------------------------------------------------

static int integrate_at_current_pose(kf_context* c, const kf_cam& cam, int first) {
  kf_stage_begin(c, KF_STAGE_INTEG_PREP);
  KF_LAUNCH(kf_launch_integ_table(c, cam));
  kf_stage_end(c, KF_STAGE_INTEG_PREP);
  kf_stage_begin(c, KF_STAGE_INTEGRATE);
  KF_LAUNCH(kf_launch_integrate(c, cam, first));
  kf_stage_end(c, KF_STAGE_INTEGRATE);
  return KF_OK;
}

int kf_integrate(kf_context* c, const float K[9], const float T[16], int32_t first_frame) {
  KF_REQUIRE(c && K && T, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  KF_NEED_DEPTH(c);
  int rc = kf_set_pose(c, T);
  if (rc != KF_OK) return rc;
  kf_cam cam;
  kf_make_cam(K, &cam);
  return integrate_at_current_pose(c, cam, first_frame ? 1 : 0);
}

}  // extern "C"

// Conversions between the brick-tiled device arrays and the reference's
// linear layout (index ix*Ny*Nz + iy*Nz + iz, z fastest).
template <class Fn>
static void for_each_voxel(const kf_context* c, Fn fn) {
  const int Nx = c->vg.nx, Ny = c->vg.ny, Nz = c->vg.nz;
  int64_t lin = 0;
  for (int x = 0; x < Nx; ++x)
    for (int y = 0; y < Ny; ++y)
      for (int z = 0; z < Nz; ++z, ++lin) fn(lin, (int64_t)kf_store_index(c->vg, x, y, z));
}

extern "C" {

int kf_download_volume(kf_context* c, float* tsdf, float* weight) {
  KF_REQUIRE(c && (tsdf || weight), KF_ERR_INVALID, "null argument");
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  const size_t ns = (size_t)c->n_store;
  if (c->vol.precision == KF_TSDF_F32) {
    std::vector<float> T(tsdf ? ns : 0), Wt(weight ? ns : 0);
    if (tsdf) KF_HIP_CHECK(hipMemcpy(T.data(), c->tsdf, ns * 4, hipMemcpyDeviceToHost));
    if (weight) KF_HIP_CHECK(hipMemcpy(Wt.data(), c->weight, ns * 4, hipMemcpyDeviceToHost));
    for_each_voxel(c, [&](int64_t lin, int64_t s) {
      if (tsdf) tsdf[lin] = T[s];
      if (weight) weight[lin] = Wt[s];
    });
  } else {
    std::vector<uint16_t> T(tsdf ? ns : 0), Wt(weight ? ns : 0);
    if (tsdf) KF_HIP_CHECK(hipMemcpy(T.data(), c->tsdf, ns * 2, hipMemcpyDeviceToHost));
    if (weight) KF_HIP_CHECK(hipMemcpy(Wt.data(), c->weight, ns * 2, hipMemcpyDeviceToHost));
    for_each_voxel(c, [&](int64_t lin, int64_t s) {
      if (tsdf) tsdf[lin] = kf_f16_bits_to_f32(T[s]);
      if (weight) weight[lin] = Wt[s] ? (float)Wt[s] : NAN;
    });
  }
  return KF_OK;
}

int kf_download_volume_box(kf_context* c, const int32_t origin[3], const int32_t size[3],
                           float* tsdf, float* weight) {
  KF_REQUIRE(c && origin && size && (tsdf || weight), KF_ERR_INVALID, "null argument");
  const int N[3] = {c->vg.nx, c->vg.ny, c->vg.nz};
  for (int a = 0; a < 3; ++a)
    KF_REQUIRE(size[a] > 0 && origin[a] >= 0 && origin[a] + size[a] <= N[a], KF_ERR_INVALID,
               "box outside the volume");
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  // copy the whole bricks that cover the box, one brick column (all z) at a time
  const int isz = c->vol.precision == KF_TSDF_F32 ? 4 : 2;
  const int bz0 = origin[2] >> 3, bz1 = (origin[2] + size[2] - 1) >> 3;
  const size_t col = (size_t)(bz1 - bz0 + 1) * 512;
  std::vector<uint8_t> T(col * isz), Wt(col * isz);
  const char* tb = static_cast<const char*>(c->tsdf);
  const char* wb = static_cast<const char*>(c->weight);
  for (int bx = origin[0] >> 3; bx <= (origin[0] + size[0] - 1) >> 3; ++bx)
    for (int by = origin[1] >> 3; by <= (origin[1] + size[1] - 1) >> 3; ++by) {
      const size_t first = (((size_t)bx * c->vg.by + by) * c->vg.bz + bz0) * 512;
      if (tsdf) KF_HIP_CHECK(hipMemcpy(T.data(), tb + first * isz, col * isz, hipMemcpyDeviceToHost));
      if (weight)
        KF_HIP_CHECK(hipMemcpy(Wt.data(), wb + first * isz, col * isz, hipMemcpyDeviceToHost));
      for (int x = std::max(bx * 8, origin[0]); x < std::min(bx * 8 + 8, origin[0] + size[0]); ++x)
        for (int y = std::max(by * 8, origin[1]); y < std::min(by * 8 + 8, origin[1] + size[1]);
             ++y)
          for (int z = origin[2]; z < origin[2] + size[2]; ++z) {
            const size_t s = (size_t)kf_store_index(c->vg, x, y, z) - first;
            const size_t o = ((size_t)(x - origin[0]) * size[1] + (y - origin[1])) * size[2] +
                             (z - origin[2]);
            if (c->vol.precision == KF_TSDF_F32) {
              if (tsdf) memcpy(&tsdf[o], &T[s * 4], 4);
              if (weight) memcpy(&weight[o], &Wt[s * 4], 4);
            } else {
              uint16_t t16, w16;
              if (tsdf) {
                memcpy(&t16, &T[s * 2], 2);
                tsdf[o] = kf_f16_bits_to_f32(t16);
              }
              if (weight) {
                memcpy(&w16, &Wt[s * 2], 2);
                weight[o] = w16 ? (float)w16 : NAN;
              }
            }
          }
    }
  return KF_OK;
}

int kf_upload_volume(kf_context* c, const float* tsdf, const float* weight) {
  KF_REQUIRE(c && tsdf && weight, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  c->skip_flags_stale = true;
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  const size_t ns = (size_t)c->n_store;
  if (c->vol.precision == KF_TSDF_F32) {
    std::vector<float> T(ns, NAN), Wt(ns, NAN);
    for_each_voxel(c, [&](int64_t lin, int64_t s) {
      T[s] = tsdf[lin];
      Wt[s] = weight[lin];
    });
    KF_HIP_CHECK(hipMemcpy(c->tsdf, T.data(), ns * 4, hipMemcpyHostToDevice));
    KF_HIP_CHECK(hipMemcpy(c->weight, Wt.data(), ns * 4, hipMemcpyHostToDevice));
  } else {
    std::vector<uint16_t> T(ns, 0x7e00u), Wt(ns, 0);
    for_each_voxel(c, [&](int64_t lin, int64_t s) {
      T[s] = kf_f32_to_f16_bits(tsdf[lin]);
      const float w = weight[lin];
      Wt[s] = (w == w && w > 0) ? (w >= 65535.0f ? 65535u : (uint16_t)w) : 0;
    });
    KF_HIP_CHECK(hipMemcpy(c->tsdf, T.data(), ns * 2, hipMemcpyHostToDevice));
    KF_HIP_CHECK(hipMemcpy(c->weight, Wt.data(), ns * 2, hipMemcpyHostToDevice));
  }
  return KF_OK;
}

int kf_download_volume_raw(kf_context* c, void* out, size_t bytes) {
  KF_REQUIRE(c && out && bytes == c->volume_bytes, KF_ERR_INVALID, "size mismatch");
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  KF_HIP_CHECK(hipMemcpy(out, c->volume, bytes, hipMemcpyDeviceToHost));
  return KF_OK;
}

int kf_upload_volume_raw(kf_context* c, const void* in, size_t bytes) {
  KF_REQUIRE(c && in && bytes == c->volume_bytes, KF_ERR_INVALID, "size mismatch");
  c->saved_ok = false;
  c->skip_flags_stale = true;
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  KF_HIP_CHECK(hipMemcpy(c->volume, in, bytes, hipMemcpyHostToDevice));
  return KF_OK;
}

int kf_volume_device_ptr(kf_context* c, void** ptr, size_t* bytes) {
  KF_REQUIRE(c && ptr, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  c->skip_flags_stale = true;  // the next raycast rebuilds the skip flags
  *ptr = c->volume;
  if (bytes) *bytes = c->volume_bytes;
  return KF_OK;
}

// ---- A9-A10 ---------------------------------------------------------------

static int raycast_at_current_pose(kf_context* c, const kf_cam& cam) {
  kf_stage_begin(c, KF_STAGE_RAYCAST);
  KF_LAUNCH(kf_launch_raycast(c, cam));
  kf_stage_end(c, KF_STAGE_RAYCAST);
  return KF_OK;
}

int kf_raycast(kf_context* c, const float K[9], const float T[16]) {
  KF_REQUIRE(c && K && T, KF_ERR_INVALID, "null argument");
  KF_NEED_DEPTH(c);
  int rc = kf_set_pose(c, T);
  if (rc != KF_OK) return rc;
  kf_cam cam;
  kf_make_cam(K, &cam);
  c->saved_ok = false;
  rc = raycast_at_current_pose(c, cam);
  if (rc == KF_OK && c->prm.raycast_mode == KF_RAYCAST_TRILINEAR) {
    memcpy(c->saved_K, K, sizeof(c->saved_K));
    memcpy(c->saved_T, T, sizeof(c->saved_T));
    c->saved_ok = true;
  }
  return rc;
}

int kf_download_model_maps(kf_context* c, float* x, float* y, float* z, float* nx, float* ny,
                           float* nz, int32_t* hit_index) {
  KF_REQUIRE(c, KF_ERR_INVALID, "null context");
  const size_t n = (size_t)c->H * c->W;
  std::vector<kf_p3> V(n), N(n);
  KF_HIP_CHECK(hipMemcpyAsync(V.data(), c->model_v, n * 12, hipMemcpyDeviceToHost, c->stream));
  KF_HIP_CHECK(hipMemcpyAsync(N.data(), c->model_n, n * 12, hipMemcpyDeviceToHost, c->stream));
  if (hit_index)
    KF_HIP_CHECK(hipMemcpyAsync(hit_index, c->hit, n * 4, hipMemcpyDeviceToHost, c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < n; ++i) {
    if (x) x[i] = V[i].x;
    if (y) y[i] = V[i].y;
    if (z) z[i] = V[i].z;
    if (nx) nx[i] = N[i].x;
    if (ny) ny[i] = N[i].y;
    if (nz) nz[i] = N[i].z;
  }
  return KF_OK;
}

int kf_refine_pose(kf_context* c, const float K[9], int32_t steps, float T_inout[16]) {
  KF_REQUIRE(c && K && steps >= 0 && steps <= KF_MAX_REFINE_STEPS, KF_ERR_INVALID,
             "bad argument");
  c->saved_ok = false;
  KF_NEED_DEPTH(c);
  if (T_inout) {
    int rc = kf_set_pose(c, T_inout);
    if (rc != KF_OK) return rc;
  }
  kf_cam cam;
  kf_make_cam(K, &cam);
  kf_stage_begin(c, KF_STAGE_REFINE);
  KF_LAUNCH(kf_launch_refine(c, cam, steps));
  kf_stage_end(c, KF_STAGE_REFINE);
  c->refine_steps_done = steps;
  if (T_inout) return kf_get_pose(c, T_inout);
  return KF_OK;
}

int kf_refine_history(kf_context* c, int32_t* n_steps, double* msr, int32_t* count,
                      int32_t* status) {
  KF_REQUIRE(c, KF_ERR_INVALID, "null context");
  const int n = c->refine_steps_done;
  if (n_steps) *n_steps = n;
  if (n == 0) return KF_OK;
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (msr)
    KF_HIP_CHECK(hipMemcpy(msr, c->state->refine_msr, n * 8, hipMemcpyDeviceToHost));
  if (count)
    KF_HIP_CHECK(hipMemcpy(count, c->state->refine_count, n * 4, hipMemcpyDeviceToHost));
  if (status)
    KF_HIP_CHECK(hipMemcpy(status, c->state->refine_status, n * 4, hipMemcpyDeviceToHost));
  return KF_OK;
}

int kf_export_model_ply(kf_context* c, const char* path, const uint8_t* bgr, int32_t binary) {
  KF_REQUIRE(c && path, KF_ERR_INVALID, "null argument");
  const size_t n = (size_t)c->H * c->W;
  std::vector<float> P(6 * n);
  int rc = kf_download_model_maps(c, &P[0], &P[n], &P[2 * n], &P[3 * n], &P[4 * n], &P[5 * n],
                                  nullptr);
  if (rc != KF_OK) return rc;
  rc = kf_write_ply(path, (int64_t)n, &P[0], &P[n], &P[2 * n], &P[3 * n], &P[4 * n], &P[5 * n],
                    bgr, binary);
  return rc == KF_OK ? KF_OK : kf_fail(rc, "cannot write the PLY file");
}

int kf_download_rendered_depth(kf_context* c, float* depth_mm) {
  KF_REQUIRE(c && depth_mm, KF_ERR_INVALID, "null argument");
  const size_t n = (size_t)c->H * c->W;
  if (!c->maps_trilinear) {
    // the last raycast into the model maps was a REFERENCE one
    memset(depth_mm, 0, n * 4);
    return KF_OK;
  }
  std::vector<float2> A(n);
  KF_HIP_CHECK(hipMemcpyAsync(A.data(), c->model_aux, n * 8, hipMemcpyDeviceToHost, c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < n; ++i) depth_mm[i] = A[i].x;
  return KF_OK;
}

}  // extern "C"

// The adjoint's buffers, allocated together on the first kf_raycast_adjoint.
// Either all of them exist (grad_tsdf set last) or none: a failed allocation
// frees the others, so a later call allocates again instead of launching on
// null buffers.
static int adjoint_buffers(kf_context* c) {
  if (c->grad_tsdf) return KF_OK;
  const size_t n = (size_t)c->H * c->W;
  const size_t touched_bytes = (size_t)((c->n_bricks + 31) / 32) * 4;
  float* grad = nullptr;
  struct Buf {
    void** p;
    size_t bytes;
  } bufs[] = {
      {(void**)&c->adj_touched, touched_bytes},  {(void**)&c->adj_depth, n * 4},
      {(void**)&c->adj_normal, n * 16},           {(void**)&c->adj_partials, kf_adjoint_partials(c) * 8},
      {(void**)&c->adj_map_v, n * 12},            {(void**)&c->adj_map_n, n * 12},
      {(void**)&c->adj_map_aux, n * 8},
      {(void**)&c->adj_map_hit, n * 4},           {(void**)&grad, (size_t)c->n_store * 4},
  };
  bool ok = true;
  for (auto& b : bufs)
    if (ok && hipMalloc(b.p, b.bytes) != hipSuccess) {
      *b.p = nullptr;
      ok = false;
    }
  // all zero once; afterwards each adjoint zeroes the bricks the previous
  // one wrote
  ok = ok && hipMemsetAsync(grad, 0, (size_t)c->n_store * 4, c->stream) == hipSuccess &&
       hipMemsetAsync(c->adj_touched, 0, touched_bytes, c->stream) == hipSuccess;
  if (!ok) {
    (void)hipStreamSynchronize(c->stream);
    for (auto& b : bufs) {
      if (*b.p) (void)hipFree(*b.p);
      *b.p = nullptr;
    }
    return kf_fail(KF_ERR_OOM, "cannot allocate the raycast adjoint's buffers");
  }
  c->grad_tsdf = grad;
  return KF_OK;
}

extern "C" {

int kf_raycast_adjoint(kf_context* c, const float K[9], const float T[16], const float* dl_ddepth,
                       const float* dl_dnormal, float* dl_dtsdf, double dl_dxi[6]) {
  KF_REQUIRE(c && K && T && dl_ddepth, KF_ERR_INVALID, "null argument");
  KF_NEED_DEPTH(c);
  const size_t n = (size_t)c->H * c->W;
  int rc = adjoint_buffers(c);
  if (rc != KF_OK) return rc;
  rc = kf_set_pose(c, T);
  if (rc != KF_OK) return rc;
  KF_HIP_CHECK(hipMemcpyAsync(c->adj_depth, dl_ddepth, n * 4, hipMemcpyHostToDevice, c->stream));
  c->adj_normal_set = dl_dnormal != nullptr;
  c->adj_want_tsdf = dl_dtsdf != nullptr;
  std::vector<float4> gn;
  if (dl_dnormal) {
    gn.resize(n);
    for (size_t i = 0; i < n; ++i)
      gn[i] = make_float4(dl_dnormal[3 * i], dl_dnormal[3 * i + 1], dl_dnormal[3 * i + 2], 0.f);
    KF_HIP_CHECK(hipMemcpyAsync(c->adj_normal, gn.data(), n * 16, hipMemcpyHostToDevice,
                                c->stream));
  }
  kf_cam cam;
  kf_make_cam(K, &cam);
  const bool saved = c->saved_ok && memcmp(K, c->saved_K, sizeof(c->saved_K)) == 0 &&
                     memcmp(T, c->saved_T, sizeof(c->saved_T)) == 0;
  c->adj_used_saved = saved;
  kf_stage_begin(c, KF_STAGE_RAYCAST_ADJOINT);
  KF_LAUNCH(kf_launch_raycast_adjoint(c, cam, saved));
  kf_stage_end(c, KF_STAGE_RAYCAST_ADJOINT);
  double xi[6];
  KF_HIP_CHECK(hipMemcpyAsync(xi, c->state->xi_grad, sizeof(xi), hipMemcpyDeviceToHost,
                              c->stream));
  std::vector<float> G(dl_dtsdf ? (size_t)c->n_store : 0);
  if (dl_dtsdf)
    KF_HIP_CHECK(hipMemcpyAsync(G.data(), c->grad_tsdf, G.size() * 4, hipMemcpyDeviceToHost,
                                c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (dl_dxi)
    for (int i = 0; i < 6; ++i) dl_dxi[i] = xi[i];
  if (dl_dtsdf) for_each_voxel(c, [&](int64_t lin, int64_t s) { dl_dtsdf[lin] = G[s]; });
  return KF_OK;
}

int kf_adjoint_used_saved(kf_context* c, int32_t* used_saved) {
  KF_REQUIRE(c && used_saved, KF_ERR_INVALID, "null argument");
  *used_saved = c->adj_used_saved ? 1 : 0;
  return KF_OK;
}

int kf_upload_model_maps(kf_context* c, const float* x, const float* y, const float* z,
                         const float* nx, const float* ny, const float* nz) {
  KF_REQUIRE(c && x && y && z && nx && ny && nz, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  const size_t n = (size_t)c->H * c->W;
  std::vector<kf_p3> V(n), N(n);
  std::vector<float> Nt(n);
  for (size_t i = 0; i < n; ++i) {
    V[i] = kf_p3{x[i], y[i], z[i]};
    N[i] = kf_p3{nx[i], ny[i], nz[i]};
    Nt[i] = kf_inv_norm3(nx[i], ny[i], nz[i]);
  }
  c->maps_trilinear = false;
  KF_HIP_CHECK(hipMemcpyAsync(c->model_v, V.data(), n * 12, hipMemcpyHostToDevice, c->stream));
  KF_HIP_CHECK(hipMemcpyAsync(c->model_n, N.data(), n * 12, hipMemcpyHostToDevice, c->stream));
  KF_HIP_CHECK(hipMemcpyAsync(c->model_nt, Nt.data(), n * 4, hipMemcpyHostToDevice, c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  return KF_OK;
}

// ---- colour: LoadColorData and SetUpPointClouds ----------------------------

// Grows the JPEG decode's buffers to n_blocks coefficient blocks and
// plane_bytes of component samples (the staging buffers are idle: the
// caller has waited for their last copy).
static int jpeg_buffers(kf_context* c, int64_t n_blocks, int64_t plane_bytes) {
  if (!c->jpeg_copied)
    KF_HIP_CHECK(hipEventCreateWithFlags(&c->jpeg_copied, hipEventDisableTiming));
  if (!c->jpeg_q) {
    KF_REQUIRE(hipMalloc((void**)&c->jpeg_q, 4 * 64 * 2) == hipSuccess, KF_ERR_OOM,
               "cannot allocate the quantisation tables");
    if (hipHostMalloc((void**)&c->jpeg_q_host, 4 * 64 * 2, hipHostMallocDefault) != hipSuccess) {
      c->jpeg_q_host = nullptr;
      return kf_fail(KF_ERR_OOM, "cannot allocate pinned staging");
    }
  }
  if (n_blocks > c->jpeg_blocks_cap) {
    KF_HIP_CHECK(hipStreamSynchronize(c->stream));  // the old buffer's last reader
    if (c->jpeg_coef) (void)hipFree(c->jpeg_coef);
    if (c->jpeg_coef_host) (void)hipHostFree(c->jpeg_coef_host);
    c->jpeg_coef = nullptr;
    c->jpeg_coef_host = nullptr;
    c->jpeg_blocks_cap = 0;
    const size_t bytes = (size_t)n_blocks * 64 * 2;
    if (hipMalloc((void**)&c->jpeg_coef, bytes) != hipSuccess) {
      c->jpeg_coef = nullptr;
      return kf_fail(KF_ERR_OOM, "cannot allocate the coefficient blocks");
    }
    if (hipHostMalloc((void**)&c->jpeg_coef_host, bytes, hipHostMallocDefault) != hipSuccess) {
      c->jpeg_coef_host = nullptr;
      return kf_fail(KF_ERR_OOM, "cannot allocate pinned staging");
    }
    c->jpeg_blocks_cap = n_blocks;
  }
  if (plane_bytes > c->jpeg_planes_cap) {
    KF_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (c->jpeg_planes) (void)hipFree(c->jpeg_planes);
    c->jpeg_planes_cap = 0;
    if (hipMalloc((void**)&c->jpeg_planes, (size_t)plane_bytes) != hipSuccess) {
      c->jpeg_planes = nullptr;
      return kf_fail(KF_ERR_OOM, "cannot allocate the sample planes");
    }
    c->jpeg_planes_cap = plane_bytes;
  }
  return KF_OK;
}

int kf_upload_color_jpeg(kf_context* c, const uint8_t* jpeg, size_t n) {
  KF_REQUIRE(c && jpeg && n > 0, KF_ERR_INVALID, "bad argument");
  kf_jpeg_frame f;
  std::string err;
  if (kf_jpeg_decode_entropy(jpeg, n, &f, nullptr, 0, &err) != 0)
    return kf_fail(KF_ERR_INVALID, "JPEG: " + err);
  KF_REQUIRE(f.H == c->H && f.W == c->W, KF_ERR_INVALID,
             "colour image is " + std::to_string(f.W) + "x" + std::to_string(f.H) +
                 ", the context " + std::to_string(c->W) + "x" + std::to_string(c->H));
  kf_jpeg_dev d;
  d.H = f.H;
  d.W = f.W;
  d.nc = f.nc;
  d.rgb = f.rgb ? 1 : 0;
  d.n_blocks = f.n_blocks;
  int64_t plane_bytes = 0;
  for (int i = 0; i < f.nc; ++i) {
    const kf_jpeg_component& s = f.comp[i];
    kf_jpeg_dev_comp& o = d.c[i];
    o.block0 = s.block0;
    o.plane0 = plane_bytes;
    o.bw = s.bw;
    o.stride = s.bw * 8;
    o.dw = s.dw;
    o.dh = s.dh;
    o.fh = f.hmax / s.h;
    o.fv = f.vmax / s.v;
    o.tq = s.tq;
    plane_bytes += (int64_t)s.bw * 8 * s.bh * 8;
  }
  // the staging buffers may still be read by the previous upload's copies
  if (c->jpeg_copied) KF_HIP_CHECK(hipEventSynchronize(c->jpeg_copied));
  int rc = jpeg_buffers(c, f.n_blocks, plane_bytes);
  if (rc != KF_OK) return rc;
  if (kf_jpeg_decode_entropy(jpeg, n, &f, c->jpeg_coef_host, c->jpeg_blocks_cap, &err) != 0)
    return kf_fail(KF_ERR_INVALID, "JPEG: " + err);
  memcpy(c->jpeg_q_host, f.q, sizeof f.q);
  KF_HIP_CHECK(hipMemcpyAsync(c->jpeg_coef, c->jpeg_coef_host, (size_t)f.n_blocks * 64 * 2,
                              hipMemcpyHostToDevice, c->stream));
  KF_HIP_CHECK(hipMemcpyAsync(c->jpeg_q, c->jpeg_q_host, sizeof f.q, hipMemcpyHostToDevice,
                              c->stream));
  KF_HIP_CHECK(hipEventRecord(c->jpeg_copied, c->stream));
  kf_stage_begin(c, KF_STAGE_COLOR);
  KF_LAUNCH(kf_launch_jpeg_decode(c, d));
  kf_stage_end(c, KF_STAGE_COLOR);
  return KF_OK;
}

int kf_jpeg_info(const uint8_t* jpeg, size_t n, int32_t* height, int32_t* width,
                 int32_t* components) {
  KF_REQUIRE(jpeg && n > 0, KF_ERR_INVALID, "bad argument");
  kf_jpeg_frame f;
  std::string err;
  if (kf_jpeg_decode_entropy(jpeg, n, &f, nullptr, 0, &err) != 0)
    return kf_fail(KF_ERR_INVALID, "JPEG: " + err);
  if (height) *height = f.H;
  if (width) *width = f.W;
  if (components) *components = f.nc;
  return KF_OK;
}

int kf_jpeg_coefficients(const uint8_t* jpeg, size_t n, int16_t* coef, int64_t capacity,
                         int64_t* n_blocks) {
  KF_REQUIRE(jpeg && n > 0, KF_ERR_INVALID, "bad argument");
  kf_jpeg_frame f;
  std::string err;
  if (kf_jpeg_decode_entropy(jpeg, n, &f, nullptr, 0, &err) != 0)
    return kf_fail(KF_ERR_INVALID, "JPEG: " + err);
  if (n_blocks) *n_blocks = f.n_blocks;
  if (!coef) return KF_OK;
  KF_REQUIRE(capacity >= f.n_blocks, KF_ERR_INVALID, "coefficient buffer too small");
  if (kf_jpeg_decode_entropy(jpeg, n, &f, coef, capacity, &err) != 0)
    return kf_fail(KF_ERR_INVALID, "JPEG: " + err);
  return KF_OK;
}

int kf_load_color_file(kf_context* c, const char* path) {
  KF_REQUIRE(c && path, KF_ERR_INVALID, "null argument");
  FILE* fp = fopen(path, "rb");
  KF_REQUIRE(fp, KF_ERR_IO, std::string("cannot open ") + path);
  std::vector<uint8_t> data;
  uint8_t chunk[1 << 16];
  size_t got;
  while ((got = fread(chunk, 1, sizeof chunk, fp)) > 0) data.insert(data.end(), chunk, chunk + got);
  const bool bad = ferror(fp) != 0;
  fclose(fp);
  KF_REQUIRE(!bad && !data.empty(), KF_ERR_IO, std::string("cannot read ") + path);
  return kf_upload_color_jpeg(c, data.data(), data.size());
}

int kf_upload_color_bgr(kf_context* c, const uint8_t* bgr, int32_t mem_kind) {
  KF_REQUIRE(c && bgr, KF_ERR_INVALID, "null argument");
  KF_HIP_CHECK(hipMemcpyAsync(c->color, bgr, (size_t)c->H * c->W * 3,
                              mem_kind == KF_MEM_HOST ? hipMemcpyHostToDevice
                                                      : hipMemcpyDeviceToDevice,
                              c->stream));
  if (mem_kind == KF_MEM_HOST) KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  return KF_OK;
}

int kf_download_color(kf_context* c, uint8_t* bgr) {
  KF_REQUIRE(c && bgr, KF_ERR_INVALID, "null argument");
  KF_HIP_CHECK(hipMemcpyAsync(bgr, c->color, (size_t)c->H * c->W * 3, hipMemcpyDeviceToHost,
                              c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  return KF_OK;
}

int kf_point_cloud(kf_context* c, float* xyzrgb, int32_t mem_kind) {
  KF_REQUIRE(c && xyzrgb, KF_ERR_INVALID, "null argument");
  const size_t bytes = (size_t)c->H * c->W * 16;
  float4* out = reinterpret_cast<float4*>(xyzrgb);
  if (mem_kind == KF_MEM_DEVICE) {
    KF_REQUIRE(((uintptr_t)xyzrgb & 15) == 0, KF_ERR_INVALID,
               "device output must be 16-byte aligned");
  } else {
    if (!c->cloud && hipMalloc((void**)&c->cloud, bytes) != hipSuccess) {
      c->cloud = nullptr;
      return kf_fail(KF_ERR_OOM, "cannot allocate the point cloud");
    }
    out = c->cloud;
  }
  KF_LAUNCH(kf_launch_point_cloud(c, out));
  if (mem_kind != KF_MEM_DEVICE) {
    KF_HIP_CHECK(hipMemcpyAsync(xyzrgb, c->cloud, bytes, hipMemcpyDeviceToHost, c->stream));
    KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  return KF_OK;
}

// ---- frame sequence ---------------------------------------------------------

int kf_stream_upload_u16(kf_context* c, const uint16_t* frames, int32_t n_frames,
                         int32_t mem_kind) {
  KF_REQUIRE(c && frames && n_frames > 0, KF_ERR_INVALID, "bad argument");
  c->saved_ok = false;
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (c->frames) (void)hipFree(c->frames);
  if (c->poses) (void)hipFree(c->poses);
  c->frames = nullptr;
  c->poses = nullptr;
  c->n_frames = 0;
  const size_t bytes = (size_t)n_frames * c->H * c->W * 2;
  if (hipMalloc(&c->frames, bytes) != hipSuccess) {
    c->frames = nullptr;
    return kf_fail(KF_ERR_OOM, "cannot allocate the frame stream");
  }
  if (hipMalloc(&c->poses, (size_t)n_frames * 64) != hipSuccess) {
    c->poses = nullptr;
    return kf_fail(KF_ERR_OOM, "cannot allocate the pose record");
  }
  KF_HIP_CHECK(hipMemcpy(c->frames, frames, bytes,
                         mem_kind == KF_MEM_HOST ? hipMemcpyHostToDevice
                                                 : hipMemcpyDeviceToDevice));
  KF_HIP_CHECK(hipMemset(c->poses, 0, (size_t)n_frames * 64));
  c->n_frames = n_frames;
  return KF_OK;
}

static int load_stream_frame(kf_context* c, int32_t frame) {
  if (frame == KF_FRAME_CURRENT) {
    // the depth the caller uploaded
    KF_REQUIRE(c->have_depth, KF_ERR_STATE, "no depth frame loaded");
    if (c->depth_gate_pending && c->prm.bilateral) {
      // a uint16 frame from the host: range gate and filter in one kernel,
      // as for a frame of the resident stream
      c->depth_gate_pending = false;
      const int rc = wait_upload(c);
      if (rc != KF_OK) return rc;
      kf_stage_begin(c, KF_STAGE_BILATERAL);
      KF_LAUNCH(kf_launch_depth_bilateral(c, c->up_dev[c->up_cur]));
      kf_stage_end(c, KF_STAGE_BILATERAL);
      KF_HIP_CHECK(hipEventRecord(c->up_consumed[c->up_cur], c->stream));
      return KF_OK;
    }
    const int rc = flush_depth_gate(c);
    if (rc != KF_OK) return rc;
    if (c->prm.bilateral) {
      kf_stage_begin(c, KF_STAGE_BILATERAL);
      KF_LAUNCH(kf_launch_bilateral(c));
      kf_stage_end(c, KF_STAGE_BILATERAL);
    }
    return KF_OK;
  }
  c->depth_gate_pending = false;
  KF_REQUIRE(c->frames && frame >= 0 && frame < c->n_frames, KF_ERR_INVALID,
             "frame outside the uploaded stream");
  const uint16_t* src = c->frames + (size_t)frame * c->H * c->W;
  if (c->prm.bilateral) {
    // range gate and filter in one kernel
    kf_stage_begin(c, KF_STAGE_BILATERAL);
    KF_LAUNCH(kf_launch_depth_bilateral(c, src));
    kf_stage_end(c, KF_STAGE_BILATERAL);
  } else {
    kf_stage_begin(c, KF_STAGE_DEPTH);
    KF_LAUNCH(kf_launch_depth_u16(c, src));
    kf_stage_end(c, KF_STAGE_DEPTH);
  }
  c->have_depth = true;
  return KF_OK;
}

static int record_pose(kf_context* c, int32_t frame) {
  if (frame == KF_FRAME_CURRENT) return KF_OK;
  KF_HIP_CHECK(hipMemcpyAsync(c->poses + (size_t)frame * 16, c->state->T, 64,
                              hipMemcpyDeviceToDevice, c->stream));
  return KF_OK;
}

int kf_pipeline_initial(kf_context* c, const float K[9], int32_t frame) {
  KF_REQUIRE(c && K, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  int rc = load_stream_frame(c, frame);
  if (rc != KF_OK) return rc;
  kf_cam cam;
  kf_make_cam(K, &cam);
  if ((rc = integrate_at_current_pose(c, cam, 1)) != KF_OK) return rc;
  if ((rc = raycast_at_current_pose(c, cam)) != KF_OK) return rc;
  return record_pose(c, frame);
}

int kf_pipeline_incoming(kf_context* c, const float K[9], int32_t frame) {
  KF_REQUIRE(c && K, KF_ERR_INVALID, "null argument");
  c->saved_ok = false;
  int rc = load_stream_frame(c, frame);
  if (rc != KF_OK) return rc;
  kf_cam cam;
  kf_make_cam(K, &cam);
  // the vertex and normal maps are computed by the first full-resolution ICP
  // iteration (raw_vertices.cu:3-51 in that kernel); with no iteration, by
  // the vertex kernel
  const bool fused = c->prm.icp_iterations > 0;
  if (!fused) {
    kf_stage_begin(c, KF_STAGE_VERTEX);
    KF_LAUNCH(kf_launch_vertex(c, cam));
    kf_stage_end(c, KF_STAGE_VERTEX);
  }
  kf_stage_begin(c, KF_STAGE_ICP);
  KF_LAUNCH(kf_launch_icp(c, cam, fused));
  kf_stage_end(c, KF_STAGE_ICP);
  if (c->prm.refine_steps > 0) {
    kf_stage_begin(c, KF_STAGE_REFINE);
    KF_LAUNCH(kf_launch_refine(c, cam, c->prm.refine_steps));
    kf_stage_end(c, KF_STAGE_REFINE);
    c->refine_steps_done = c->prm.refine_steps;
  }
  if ((rc = integrate_at_current_pose(c, cam, 0)) != KF_OK) return rc;
  if ((rc = raycast_at_current_pose(c, cam)) != KF_OK) return rc;
  return record_pose(c, frame);
}

int kf_pipeline_run(kf_context* c, const float K[9], int32_t first, int32_t count) {
  KF_REQUIRE(c && K && count > 0 && first >= 0, KF_ERR_INVALID, "bad argument");
  c->saved_ok = false;
  int rc = kf_pipeline_initial(c, K, first);
  for (int32_t f = first + 1; rc == KF_OK && f < first + count; ++f)
    rc = kf_pipeline_incoming(c, K, f);
  return rc;
}

int kf_get_pose(kf_context* c, float T[16]) {
  KF_REQUIRE(c && T, KF_ERR_INVALID, "null argument");
  KF_HIP_CHECK(hipMemcpyAsync(T, c->state->T, 64, hipMemcpyDeviceToHost, c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  return KF_OK;
}

int kf_download_poses(kf_context* c, float* poses, int32_t n_frames) {
  KF_REQUIRE(c && poses && c->poses && n_frames > 0 && n_frames <= c->n_frames, KF_ERR_INVALID,
             "bad argument");
  KF_HIP_CHECK(hipMemcpyAsync(poses, c->poses, (size_t)n_frames * 64, hipMemcpyDeviceToHost,
                              c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  return KF_OK;
}

// ---- measurement ----------------------------------------------------------

int kf_stage_times(kf_context* c, double* ms, int64_t* counts) {
  KF_REQUIRE(c, KF_ERR_INVALID, "null context");
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  kf_timing* t = timing_of(c);
  for (auto& p : t->pending) {
    float e = 0.f;
    if (hipEventElapsedTime(&e, p.a, p.b) == hipSuccess) {
      c->stage_ms[p.stage] += e;
      c->stage_count[p.stage] += 1;
    }
    t->pool.push_back(p.a);
    t->pool.push_back(p.b);
  }
  t->pending.clear();
  for (int i = 0; i < KF_STAGE_COUNT; ++i) {
    if (ms) ms[i] = c->stage_ms[i];
    if (counts) counts[i] = c->stage_count[i];
  }
  return KF_OK;
}

int kf_reset_stage_times(kf_context* c) {
  KF_REQUIRE(c, KF_ERR_INVALID, "null context");
  int rc = kf_stage_times(c, nullptr, nullptr);
  if (rc != KF_OK) return rc;
  for (int i = 0; i < KF_STAGE_COUNT; ++i) {
    c->stage_ms[i] = 0.0;
    c->stage_count[i] = 0;
  }
  return KF_OK;
}

int kf_count_raycast_work(kf_context* c, int32_t enable) {
  KF_REQUIRE(c, KF_ERR_INVALID, "null context");
  KF_HIP_CHECK(hipMemsetAsync(&c->state->ray_steps, 0, 2 * sizeof(unsigned long long), c->stream));
  KF_HIP_CHECK(hipMemsetAsync(&c->state->ray_skipped, 0, sizeof(unsigned long long), c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  c->count_rays = enable != 0;
  return KF_OK;
}

int kf_integration_work(kf_context* c, int64_t* groups_kept, int64_t* bricks_kept) {
  KF_REQUIRE(c && groups_kept && bricks_kept, KF_ERR_INVALID, "null argument");
  // the last integration appended to slot (integ_slot ^ 1) of the group and
  // brick counters; the next integration clears that slot
  const int s = c->integ_slot ^ 1;
  const size_t lines = (size_t)KF_RAY_PARTS * KF_RAY_COUNT_STRIDE;
  std::vector<uint32_t> g(lines), b(lines);
  KF_HIP_CHECK(hipMemcpyAsync(g.data(), c->integ_counts + (size_t)s * lines, lines * 4,
                              hipMemcpyDeviceToHost, c->stream));
  KF_HIP_CHECK(hipMemcpyAsync(b.data(), c->integ_counts + (size_t)(2 + s) * lines, lines * 4,
                              hipMemcpyDeviceToHost, c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  int64_t ng = 0, nb = 0;
  for (int p = 0; p < KF_RAY_PARTS; ++p) {
    ng += g[(size_t)p * KF_RAY_COUNT_STRIDE];
    nb += b[(size_t)p * KF_RAY_COUNT_STRIDE];
  }
  *groups_kept = ng;
  *bricks_kept = nb;
  return KF_OK;
}

int kf_raycast_work(kf_context* c, int64_t* voxels_visited, int64_t* hits) {
  KF_REQUIRE(c && voxels_visited && hits, KF_ERR_INVALID, "null argument");
  unsigned long long v[2];
  KF_HIP_CHECK(hipMemcpyAsync(v, &c->state->ray_steps, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  *voxels_visited = (int64_t)v[0];
  *hits = (int64_t)v[1];
  return KF_OK;
}

int kf_raycast_skipped(kf_context* c, int64_t* voxels_skipped) {
  KF_REQUIRE(c && voxels_skipped, KF_ERR_INVALID, "null argument");
  unsigned long long v = 0;
  KF_HIP_CHECK(hipMemcpyAsync(&v, &c->state->ray_skipped, sizeof(v), hipMemcpyDeviceToHost,
                              c->stream));
  KF_HIP_CHECK(hipStreamSynchronize(c->stream));
  *voxels_skipped = (int64_t)v;
  return KF_OK;
}

}  // extern "C"
