// kf_internal.h - context, device state and kernel launchers of the gfx950
// KinectFusion core.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/kinfu_amd.h"
#include "kf_hd_math.h"

// Per-frame state that lives in device memory so that the whole frame
// sequence (ICP -> integrate -> raycast) runs without host round trips.
struct kf_dev_state {
  float T[16];          // current camera-to-world pose (the reference's extrinsic_matrix)
  float Tinv[16];       // inverse of T
  float Tprev[16];      // pose the model maps were raycast at
  float Tprev_inv[16];  // its inverse (sensor_pose_estimation.cu:256)
  float F[16];          // Tprev_inv * T (frame_to_frame, sensor_pose_estimation.cu:343)
  float center_dir[4];  // integration front-test direction (projective_tsdf.cu:57-71)
  int32_t icp_iters_done;
  unsigned long long ray_steps;  // raycast work counters (voxels visited, hits)
  unsigned long long ray_hits;
  double xi_grad[6];             // raycast adjoint: dL/dxi (omega, v)
  double refine_msr[KF_MAX_REFINE_STEPS];     // refinement: mean squared residual
  int32_t refine_count[KF_MAX_REFINE_STEPS];  //   pixels used
  int32_t refine_status[KF_MAX_REFINE_STEPS]; //   0 solved, else no update
  alignas(64) float iter_T[KF_MAX_ICP_ITERS][16];  // one 64-byte line per iteration
  double iter_sums[KF_MAX_ICP_ITERS][KF_ICP_SUMS];
  int32_t iter_status[KF_MAX_ICP_ITERS];
  int32_t icp_stop_at;    // opt-in early stop: the first iteration that did not run (0: none)
  int32_t icp_stop_from;  //   the first iteration whose kernel may stop
  float icp_stop_rot;     //   thresholds (kf_icp_converged)
  float icp_stop_trans;
  unsigned long long ray_skipped;  // raycast work counter: steps taken by skip rounds
};

// A packed 3-vector of an image map (12 bytes, 4-byte aligned): a wave's 64
// consecutive records are 768 contiguous bytes, read as one dwordx3 load per
// lane, with no unused fourth lane as in a float4.
struct kf_p3 {
  float x, y, z;
};

// Camera constants passed by value to every kernel.
struct kf_cam {
  float K[9];
  float Kinv[9];
};

// Volume geometry passed by value.
//
// Storage is brick-tiled: the volume is cut into 8x8x8 bricks, bricks are
// laid out x-major (z fastest), and inside a brick the voxel (lx, ly, lz)
// sits at lx*64 + ly*8 + lz.  TSDF and weight are separate arrays of
// bx*by*bz*512 elements (float, or half/uint16 for the FP16 volume); voxels
// of partial bricks beyond the volume are padding that nothing reads.  A
// wave's rays then touch a few cache lines per step instead of one line per
// ray (x neighbours of the reference's linear layout are Ny*Nz elements
// apart), and the raycast reads a dense TSDF-only array.
struct kf_vol {
  int32_t nx, ny, nz;
  int32_t sx, sy, sz;  // start indices
  float d;             // voxel edge (mm)
  int32_t bx, by, bz;  // bricks per axis
};

#define KF_BRICK 8

// Ray queue between the two raycast passes: KF_RAY_PARTS partitions, each
// with its own append counter on its own 256-byte line (a single counter
// serialises the appends of every wave), two counter slots used by alternate
// raycasts (each clears the other's).  First-pass block b appends to
// partition b % KF_RAY_PARTS, whose region holds the rays of all its blocks.
// Bilateral range weights of integer depth differences j (u16 frames):
// entry j = expf(-(j * j) / (2 sigma_r^2)) with the filter's own float
// operations; every j >= the table length gives exactly 0.  Used when the
// zero tail starts within this many entries.
#define KF_RANGE_LUT_MAX 2048

#define KF_RAY_PARTS 64
#define KF_RAY_COUNT_STRIDE 64  // uint32 per counter line
__host__ __device__ inline int kf_ray_part_capacity(int n_blocks) {
  return (n_blocks + KF_RAY_PARTS - 1) / KF_RAY_PARTS * 256;
}

__host__ __device__ inline int kf_store_index(const kf_vol& v, int x, int y, int z) {
  return ((((x >> 3) * v.by + (y >> 3)) * v.bz + (z >> 3)) << 9) | ((x & 7) << 6) |
         ((y & 7) << 3) | (z & 7);
}

// ICP reduction geometry.
#ifndef KF_ICP_BLOCK
#define KF_ICP_BLOCK 512
#endif
// Block partials one lane reads per row when an iteration is finished: the
// pixels per thread (1 to 4) are chosen so that at most 64 * 8 blocks cover
// the image (up to 2^20 pixels).
#define KF_ICP_MAX_PART_PER_LANE 8
#define KF_ICP_MAX_BLOCKS (64 * KF_ICP_MAX_PART_PER_LANE)

// A coarse level of the opt-in ICP pyramid (levels 1 .. icp_levels-1): its
// image and ICP grid, the frame's depth, vertex, normal and mask maps, and
// the model maps with their surface flags.
struct kf_pyr_level {
  int H = 0, W = 0, blocks = 0, bpx = 0, px = 0, threads = KF_ICP_BLOCK;
  float* depth = nullptr;
  kf_p3* vmap = nullptr;
  kf_p3* nmap = nullptr;
  uint8_t* mask = nullptr;
  kf_p3* mv = nullptr;
  kf_p3* mn = nullptr;
  float* mt = nullptr;  // kf_inv_norm3 of mn
  uint8_t* ok = nullptr;
};

// The colour decode's frame as the device kernels take it: per component
// its first coefficient block, its sample plane (byte offset, row stride =
// 8 x blocks per row), its real samples and its upsampling factors.
struct kf_jpeg_dev_comp {
  int64_t block0 = 0, plane0 = 0;
  int bw = 0, stride = 0;
  int dw = 0, dh = 0;
  int fh = 1, fv = 1;
  int tq = 0;
};
struct kf_jpeg_dev {
  int H = 0, W = 0, nc = 0, rgb = 0;
  int64_t n_blocks = 0;
  kf_jpeg_dev_comp c[4];
};

struct kf_context {
  int device = 0;
  int H = 0, W = 0;
  kf_volume_desc vol{};
  kf_params prm{};
  kf_vol vg{};
  hipStream_t stream = nullptr;

  // images (H*W)
  float* depth = nullptr;       // working depth, float mm
  float* depth_tmp = nullptr;   // bilateral output
  uint16_t* depth_u16 = nullptr;// device frame of uint16 uploads (slot 0)
  kf_p3* vmap = nullptr;        // {Vx, Vy, Vz}, Vz = depth where the mask is 1, else 0
  kf_p3* nmap = nullptr;        // {Nx, Ny, Nz}
  uint8_t* mask = nullptr;      // vertex mask (reference layout)
  kf_p3* model_v = nullptr;     // raycast vertices {x, y, z}
  kf_p3* model_n = nullptr;     // raycast normals {nx, ny, nz}
  float* model_nt = nullptr;    // kf_inv_norm3 of each model normal (the factor the
                                // ICP's normal gate normalises it by; +inf for 0)
  float2* model_aux = nullptr;  // trilinear raycasts only: {camera depth, crossing te}
  bool maps_trilinear = false;  // model_aux belongs to the maps (last raycast trilinear)
  int32_t* hit = nullptr;       // raycast hit voxel index
  void* ray_queue = nullptr;    // KF_RAY_PARTS x kf_ray_part_capacity ray states
  uint32_t* ray_counts = nullptr; // [2][KF_RAY_PARTS * KF_RAY_COUNT_STRIDE] append counters
  float2* integ_tab = nullptr;  // per-pixel {predicted depth, 1/|R lambda|}
  float2* integ_tiles = nullptr;// per 16x16 tile: {min, max} camera distance a voxel
                                // projecting there can be updated at (NaN-free)
  int* integ_groups = nullptr;  // groups, then bricks, that pass the culls (partitioned lists)
  uint32_t* integ_counts = nullptr; // [group, brick][2 slots][KF_RAY_PARTS * KF_RAY_COUNT_STRIDE]
  int integ_slot = 0;           // integ_counts slot of the next integration
  float* spatial = nullptr;     // bilateral spatial weights
  float* range_lut = nullptr;   // bilateral range weights by integer |depth difference|
  int range_lut_n = 0;          // entries (0: not usable, the filter evaluates expf)

  // volume (one allocation: TSDF array, then weight array, brick-tiled)
  void* volume = nullptr;
  void* tsdf = nullptr;
  void* weight = nullptr;
  size_t volume_bytes = 0;
  int64_t n_voxels = 0;   // Nx*Ny*Nz
  int64_t n_store = 0;    // bricks * 512
  int n_bricks = 0;
  // Second-pass skip flags, one byte per brick (set where the brick or a
  // neighbour holds a negative TSDF value; a superset between resets).
  // skip_flags_stale: the volume was handed out by address since the flags
  // were last rebuilt, so the next raycast rebuilds them first.
  uint8_t* skip_flags = nullptr;
  bool skip_flags_stale = false;

  // ICP
  kf_dev_state* state = nullptr;
  void* icp_partials = nullptr;    // 2 parities x [KF_ICP_SUMS][blocks] doubles (or
                                   // [2 * KF_ICP_SUMS][blocks] floats, compensated)
  int icp_blocks = 0;
  int icp_bpx = 0;                 // pixels per block
  int icp_px = 4;                  // pixels per thread, at most
  int icp_threads = KF_ICP_BLOCK;  // threads per block (icp_grid)
  uint8_t* icp_masks = nullptr;    // [KF_MAX_ICP_ITERS][H*W] when record_icp
  float icp_cos_lo = 0.f;          // ICP normal gate as a cosine threshold
  float icp_dist2_hi = 0.f;        // ICP distance gate as a squared-distance threshold
  kf_pyr_level pyr[KF_MAX_ICP_LEVELS]; // [1 ..] coarse levels (allocated when enabled)
  float flip_cos_hi = 0.f;         // raycast normal flip threshold

  // stream of frames
  uint16_t* frames = nullptr;
  int32_t n_frames = 0;
  float* poses = nullptr;          // [n_frames][16]

  // timing
  double stage_ms[KF_STAGE_COUNT] = {};
  int64_t stage_count[KF_STAGE_COUNT] = {};
  bool have_depth = false;
  // depth_u16 holds a frame uploaded from the host whose range gate into
  // depth has not run yet: the host-fed pipeline gates and filters it in
  // one kernel; any other reader of depth runs the gate first.
  bool depth_gate_pending = false;
  // Uploads of uint16 frames from host memory use two slots in turn: the
  // caller's frame is copied into a pinned buffer, then onto the device by
  // copy_stream, so that the copy runs beside the previous frame's kernels.
  // up_copied[i]: slot i's copy done (recorded on copy_stream);
  // up_consumed[i]: the last reader of slot i's device frame done (recorded
  // on stream).  The host waits for these events; neither stream waits on
  // the other's.  Created on the first host upload.
  hipStream_t copy_stream = nullptr;
  uint16_t* up_host[2] = {nullptr, nullptr};
  uint16_t* up_dev[2] = {nullptr, nullptr};  // up_dev[0] = depth_u16
  hipEvent_t up_copied[2] = {nullptr, nullptr};
  hipEvent_t up_consumed[2] = {nullptr, nullptr};
  int up_next = 0;  // slot of the next host upload
  int up_cur = 0;   // slot of the frame whose gate is pending
  bool count_rays = false;  // raycast work counters on
  int ray_slot = 0;         // ray_counts slot of the next raycast
  // raycast adjoint buffers (allocated on first use)
  float* grad_tsdf = nullptr;   // dL/dF, n_store floats, native order
  uint32_t* adj_touched = nullptr; // one bit per brick that holds non-zero dL/dF
                                   // (zeroed before the next adjoint)
  float* adj_depth = nullptr;   // dL/d(rendered depth), H*W
  float4* adj_normal = nullptr; // dL/d(normal), H*W
  kf_p3* adj_map_v = nullptr;   // the adjoint's own trilinear raycast maps (when the
  kf_p3* adj_map_n = nullptr;   //   model maps' saved crossings do not apply), so
  float2* adj_map_aux = nullptr;//   that the context's model maps stay as they are
  int32_t* adj_map_hit = nullptr;
  bool adj_normal_set = false;
  bool adj_want_tsdf = false;   // the current adjoint call returns dL/dTSDF
  // The model maps hold a trilinear raycast's crossings (model_aux.y = te where
  // hit >= 0) made at saved_K / saved_T with the current depth, volume and
  // parameters: every call that changes one of those clears saved_ok.
  bool saved_ok = false;
  float saved_K[9] = {};
  float saved_T[16] = {};
  bool adj_used_saved = false;  // the last adjoint started from the saved crossings
  // colour (LoadColorData / SetUpPointClouds): the H x W x 3 B, G, R image
  // (zeros until one is loaded), and the JPEG decode's buffers, grown on
  // demand: coefficient blocks (device, and pinned host staging), the
  // component sample planes and the 4 quantisation tables; jpeg_copied marks
  // the staging buffers' last copy to the device
  uint8_t* color = nullptr;
  int16_t* jpeg_coef = nullptr;
  int16_t* jpeg_coef_host = nullptr;
  int64_t jpeg_blocks_cap = 0;
  uint8_t* jpeg_planes = nullptr;
  int64_t jpeg_planes_cap = 0;
  uint16_t* jpeg_q = nullptr;
  uint16_t* jpeg_q_host = nullptr;
  hipEvent_t jpeg_copied = nullptr;
  float4* cloud = nullptr;        // point cloud staging for host output
  double* adj_partials = nullptr; // dL/dxi block partials
  float* gn_partials = nullptr;   // refinement block partials
  int refine_steps_done = 0;      // steps of the last refinement
};

// error reporting ---------------------------------------------------------
void kf_set_error(const std::string& msg);
int kf_fail(int code, const std::string& msg);
#define KF_HIP_CHECK(expr)                                                    \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      return kf_fail(KF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// launchers (each enqueues on ctx->stream and returns hipGetLastError) ------
hipError_t kf_launch_depth_u16(kf_context* c, const uint16_t* src);
hipError_t kf_launch_bilateral(kf_context* c);
hipError_t kf_launch_depth_bilateral(kf_context* c, const uint16_t* src);
hipError_t kf_launch_vertex(kf_context* c, const kf_cam& cam);
// maps_from_depth: the first full-resolution iteration computes the frame's
// vertex, normal and mask maps from the working depth (instead of
// kf_launch_vertex) and writes them to the context's maps.
hipError_t kf_launch_icp(kf_context* c, const kf_cam& cam, bool maps_from_depth = false);
// Opt-in ICP early stop (kf_params icp_stop_rot / icp_stop_trans_mm, both > 0):
// whether it is on, and the first iteration whose kernel may stop it (the
// coarse pyramid levels always run in full).  kf_launch_set_icp_stop stores
// the thresholds and that iteration in the device state.
inline bool kf_icp_may_stop(const kf_params& p) {
  return p.icp_stop_rot > 0 && p.icp_stop_trans_mm > 0;
}
int kf_icp_stop_from(const kf_params& p);
hipError_t kf_launch_set_icp_stop(kf_dev_state* st, const kf_params& p, hipStream_t s);
// Intrinsics of the next coarser pyramid level (K and Kc: 3x3 row-major).
void kf_pyramid_intrinsics(const float* K, float* Kc);
hipError_t kf_launch_integ_table(kf_context* c, const kf_cam& cam);
hipError_t kf_launch_integrate(kf_context* c, const kf_cam& cam, int first);
hipError_t kf_launch_raycast(kf_context* c, const kf_cam& cam);
// saved: start from the crossings the last (trilinear) raycast saved, which
// was made at the same pose, camera, depth and volume (kf_context::saved_ok).
hipError_t kf_launch_raycast_adjoint(kf_context* c, const kf_cam& cam, bool saved);
size_t kf_adjoint_partials(const kf_context* c);  // doubles
size_t kf_refine_partials(const kf_context* c);   // floats
hipError_t kf_launch_refine(kf_context* c, const kf_cam& cam, int steps);
hipError_t kf_launch_volume_fill(kf_context* c);
// n floats at p set to v.
hipError_t kf_launch_fill_f32(float* p, int64_t n, float v, hipStream_t s);
// Skip flags: cleared (a reset volume), or rebuilt from the whole volume.
hipError_t kf_launch_skip_flags_clear(kf_context* c);
hipError_t kf_launch_skip_flags_rebuild(kf_context* c);
// dequantisation + inverse DCT of the coefficient blocks in jpeg_coef, then
// upsampling and colour conversion into color
hipError_t kf_launch_jpeg_decode(kf_context* c, const kf_jpeg_dev& f);
// the point cloud of the model maps and the colour image into out (H*W)
hipError_t kf_launch_point_cloud(kf_context* c, float4* out);
hipError_t kf_launch_set_pose(kf_context* c, const float* T_host);
hipError_t kf_launch_record_pose(kf_context* c, int frame);

size_t kf_ray_state_bytes();
size_t kf_ray_queue_entries(int H, int W);
size_t kf_integ_group_entries(const kf_vol& vg);

// shared host helpers
void kf_make_cam(const float* K, kf_cam* cam);
void kf_stage_begin(kf_context* c, int stage);
void kf_stage_end(kf_context* c, int stage);
