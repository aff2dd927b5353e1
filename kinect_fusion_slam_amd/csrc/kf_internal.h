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
  uint32_t ticket;      // last-block-done counter of the ICP reduction
  int32_t pad[2];
  float iter_T[KF_MAX_ICP_ITERS][16];
  double iter_sums[KF_MAX_ICP_ITERS][KF_ICP_SUMS];
  int32_t iter_status[KF_MAX_ICP_ITERS];
};

// Camera constants passed by value to every kernel.
struct kf_cam {
  float K[9];
  float Kinv[9];
};

// Volume geometry passed by value.
struct kf_vol {
  int32_t nx, ny, nz;
  int32_t sx, sy, sz;  // start indices
  float d;             // voxel edge (mm)
  // floor division by nz and ny for linear voxel ids < 2^31:
  // q = (umulhi(n, m) + n) >> s
  uint32_t mz, my;
  int32_t shz, shy;
};

// ICP reduction geometry.
#define KF_ICP_BLOCK 256
#define KF_ICP_PIX_PER_THREAD 4

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
  uint16_t* depth_u16 = nullptr;// staging for host uint16 uploads
  float4* vmap = nullptr;       // {Vx, Vy, Vz(=depth), mask}
  float4* nmap = nullptr;       // {Nx, Ny, Nz, 0}
  uint8_t* mask = nullptr;      // vertex mask (reference layout)
  float4* model_v = nullptr;    // raycast vertices {x, y, z, 0}
  float4* model_n = nullptr;    // raycast normals
  int32_t* hit = nullptr;       // raycast hit voxel index
  float2* integ_tab = nullptr;  // per-pixel {predicted depth, 1/|R lambda|}
  float* spatial = nullptr;     // bilateral spatial weights

  // volume
  void* volume = nullptr;
  size_t volume_bytes = 0;
  int64_t n_voxels = 0;

  // ICP
  kf_dev_state* state = nullptr;
  float* icp_partials = nullptr;   // [blocks][KF_ICP_SUMS][2]
  int icp_blocks = 0;
  uint8_t* icp_masks = nullptr;    // [KF_MAX_ICP_ITERS][H*W] when record_icp
  float icp_cos_lo = 0.f;          // ICP normal gate as a cosine threshold
  float flip_cos_hi = 0.f;         // raycast normal flip threshold

  // stream of frames
  uint16_t* frames = nullptr;
  int32_t n_frames = 0;
  float* poses = nullptr;          // [n_frames][16]

  // timing
  hipEvent_t ev[2 * KF_STAGE_COUNT] = {};
  double stage_ms[KF_STAGE_COUNT] = {};
  int64_t stage_count[KF_STAGE_COUNT] = {};
  bool have_depth = false;
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
hipError_t kf_launch_vertex(kf_context* c, const kf_cam& cam);
hipError_t kf_launch_icp(kf_context* c, const kf_cam& cam);
hipError_t kf_launch_integ_table(kf_context* c, const kf_cam& cam);
hipError_t kf_launch_integrate(kf_context* c, const kf_cam& cam, int first);
hipError_t kf_launch_raycast(kf_context* c, const kf_cam& cam);
hipError_t kf_launch_volume_fill(kf_context* c);
hipError_t kf_launch_set_pose(kf_context* c, const float* T_host);
hipError_t kf_launch_record_pose(kf_context* c, int frame);

// shared host helpers
void kf_make_cam(const float* K, kf_cam* cam);
void kf_stage_begin(kf_context* c, int stage);
void kf_stage_end(kf_context* c, int stage);
