/*
 * kinfu_amd.h - C ABI of the MI355X (gfx950) KinectFusion core.
 *
 * The reference exposes its GPU stages as five host wrappers in
 * src/parallel_processing/parallel_processing.h:19-53 (Projective_TSDF,
 * Ray_Casting, Calculate_Vertices_And_Normals, Estimate_Sensor_Pose,
 * Bilateral_Filtering).  Those take cv::Mat& and C++ references, allocate and
 * free every device buffer on each call and copy the whole volume across PCIe
 * twice per frame.  This ABI replaces them with plain C types and an opaque
 * context that owns every device buffer for its lifetime: the volume, the
 * depth/vertex/normal maps, the raycast model maps and the ICP reduction
 * scratch all stay resident in HBM.
 *
 * Conventions (identical to the reference):
 *   - Units are millimetres.
 *   - K is 3x3 row-major (9 floats), T is 4x4 row-major camera-to-world
 *     (16 floats): p_world = T * [p_cam; 1]                 (projective_tsdf.cu:43-45)
 *   - Pixels are 1-indexed when projected: K^-1 is applied to (x+1, y+1, 1)
 *     and a projection (u, v) reads depth[(v-1)*W + (u-1)]  (raw_vertices.cu:25-30)
 *   - The volume is dense, x-major / z-fastest:
 *     index = ix*(Ny*Nz) + iy*Nz + iz; voxel centre = (i+start)*d + d/2
 *     (kinfu_pipeline.cpp:47-58).  Unobserved voxels hold NaN TSDF.
 *   - Images are H x W row-major.
 *
 * Every entry point returns KF_OK (0) or a negative KF_ERR_* code and never
 * aborts; kf_last_error() gives the message of the last failure on the
 * calling thread.  Work is enqueued on the context's own HIP stream; entry
 * points that hand results to host memory synchronise that stream.  A
 * context is used by one host thread at a time.
 */
#ifndef KINFU_AMD_H_
#define KINFU_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KF_OK 0
#define KF_ERR_INVALID (-1) /* bad argument / size / pointer            */
#define KF_ERR_HIP (-2)     /* HIP runtime error                        */
#define KF_ERR_OOM (-3)     /* device allocation failed                 */
#define KF_ERR_STATE (-4)   /* call out of order (e.g. no frame loaded) */
#define KF_ERR_IO (-5)      /* file could not be read or written        */

#define KF_INTRINSIC_ELEMENTS 9  /* parallel_processing.h:16 */
#define KF_EXTRINSIC_ELEMENTS 16 /* parallel_processing.h:17 */

#define KF_TSDF_F32 0 /* float TSDF + float weight (the reference's storage)  */
#define KF_TSDF_F16 1 /* half TSDF + uint16 weight (precision variant)        */

#define KF_MEM_HOST 0   /* pointer argument is host memory                   */
#define KF_MEM_DEVICE 1 /* pointer argument is device memory (this GPU)      */

#define KF_MAX_ICP_ITERS 64
#define KF_MAX_BILATERAL_RADIUS 8
#define KF_ICP_SUMS 28 /* 21 unique J^TJ, 6 J^Tr, inlier count          */
#define KF_MAX_REFINE_STEPS 64
#define KF_MAX_ICP_LEVELS 3

/* Stage indices for kf_stage_times(). */
#define KF_STAGE_DEPTH 0     /* uint16 -> float mm + range gate                */
#define KF_STAGE_BILATERAL 1 /* bilateral filter                              */
#define KF_STAGE_VERTEX 2    /* vertex + normal maps                          */
#define KF_STAGE_ICP 3       /* all ICP iterations of one frame               */
#define KF_STAGE_INTEGRATE 4 /* projective TSDF integration, per-voxel kernel  */
#define KF_STAGE_RAYCAST 5   /* voxel-traversal raycast                       */
#define KF_STAGE_INTEG_PREP 6 /* integration per-pixel table                  */
#define KF_STAGE_RAYCAST_ADJOINT 7 /* kf_raycast_adjoint kernels              */
#define KF_STAGE_REFINE 8    /* pose refinement through the trilinear raycast  */
#define KF_STAGE_ICP_ITER 9  /* the iteration kernels of one ICP, back to back: KF_STAGE_ICP
                                without its last (finishing) kernel; one count per ICP */
#define KF_STAGE_COLOR 10     /* colour JPEG: inverse DCT + upsampling/colour kernels */
#define KF_STAGE_COUNT 11

/* Parameter presets for kf_preset(). */
#define KF_PRESET_REFERENCE 100 /* the reference's compiled-in defaults          */
#define KF_PRESET_ROOM_CPU 0     /* 640x480, 360x360x300 @10mm, FP32 (CPU case)    */
#define KF_PRESET_ROOM 1     /* same geometry, GPU stream                      */
#define KF_PRESET_ROOM_5MM_F16 2     /* 720x720x600 @5mm, FP16+u16                     */
#define KF_PRESET_SPHERES_128 3     /* 128^3 @10mm, 320x240 differentiable render     */
#define KF_PRESET_ROOM_HD_F16 4     /* 1024x1024x768 @4mm, FP16+u16, 1280x720         */

typedef struct kf_context kf_context;

/* Dense volume.  Replaces the _voxel_* constants of kinfu_pipeline.h:42-49. */
typedef struct kf_volume_desc {
  int32_t dims[3];   /* voxel_length (x), voxel_width (y), voxel_height (z) */
  int32_t start[3];  /* voxel_grid_{x,y,z}_start, in voxels                 */
  float voxel_mm;    /* voxel_distance                                      */
  int32_t precision; /* KF_TSDF_F32 or KF_TSDF_F16                          */
} kf_volume_desc;

/* Per-context parameters.  In the reference every one of these is a
 * compile-time constant (kinfu_pipeline.h:52-62, system_interface.cpp:10-11,
 * sensor_pose_estimation.cu:48,69,339). */
typedef struct kf_params {
  float truncation_mm;      /* _truncated_distance                       */
  float sdf_min_mm;         /* _sdf_minimum                              */
  float sdf_max_mm;         /* _sdf_maximum                              */
  float depth_min_mm;       /* min_depth (inclusive)                     */
  float depth_max_mm;       /* max_depth (inclusive)                     */
  int32_t bilateral;        /* 1: filter each frame before vertex maps   */
  int32_t bilateral_radius; /* _bw_radius, <= KF_MAX_BILATERAL_RADIUS    */
  float sigma_s_px;         /* _sigma_d                                  */
  float sigma_r_mm;         /* _sigma_r                                  */
  int32_t icp_iterations;   /* <= KF_MAX_ICP_ITERS                       */
  float icp_dist_mm;        /* point distance gate                       */
  float icp_angle_deg;      /* normal angle gate                         */
  int32_t icp_compensated;  /* 1: Kahan/TwoSum compensated FP32 sums     */
  int32_t record_icp;       /* 1: keep per-iteration inlier masks        */
  int32_t stage_timing;     /* 1: hipEvent timing of every stage         */
  int32_t raycast_fix_bounds; /* 0: the reference's zero-crossing bound
                                 checks (voxel_traversal.cu:233-281, which
                                 test ZX/ZY against voxel_height); 1: test
                                 every index against its own axis        */
  int32_t raycast_mode;     /* KF_RAYCAST_REFERENCE or KF_RAYCAST_TRILINEAR */
  int32_t refine_steps;     /* pipeline: Gauss-Newton pose refinement steps
                               after ICP (0 = off, <= KF_MAX_REFINE_STEPS) */
  float refine_max_residual_mm; /* refinement ignores |residual| above this  */
  int32_t icp_robust;       /* 0: the reference's ICP update (a failed solve
                               applies the right-hand side); 1: an iteration
                               whose solve fails, or that has fewer than
                               icp_min_inliers inliers, leaves the pose
                               unchanged (no reference counterpart)       */
  int32_t icp_min_inliers;  /* icp_robust: inlier count below which an
                               iteration is not applied                    */
  int32_t icp_levels;       /* 1: the reference's ICP, every iteration at full
                               resolution (sensor_pose_estimation.cu:339);
                               2 or 3: coarse to fine, icp_coarse_iters at
                               half (and quarter) resolution first, then
                               icp_iterations at full resolution (no
                               reference counterpart)                      */
  int32_t icp_coarse_iters[2]; /* iterations at half and quarter resolution */
  float pyramid_max_step_mm; /* a coarse depth pixel averages its 2x2 block
                                only where the four depths differ by at most
                                this, else it is 0                          */
  float icp_stop_rot;       /* opt-in early stop (0 = off, the reference's
                               fixed iteration count): the full-resolution
                               iterations end after an increment whose
                               rotation parameters are all below this (rad)
                               and whose translation is below
                               icp_stop_trans_mm (no reference counterpart) */
  float icp_stop_trans_mm;
  int32_t raycast_skip;     /* 1: the raycast's second pass takes runs of
                               steps through voxels that the volume's skip
                               flags show hold no negative TSDF value in one
                               round, without loading them (visits, hits and
                               maps are the same as without it; it pays off
                               on long walks: on in KF_PRESET_ROOM_HD_F16);
                               0: every voxel is walked (no reference
                               counterpart either way)                    */
} kf_params;

/* Raycast hit handling.  REFERENCE: the reference's per-axis interpolation
 * and central-difference normal (voxel_traversal.cu:222-377).  TRILINEAR:
 * the same walk, then the zero crossing of the trilinearly interpolated TSDF
 * between samples spaced one voxel along the ray from the skip point, and
 * the normal as the normalised trilinear gradient there.  TRILINEAR is
 * differentiable (kf_raycast_adjoint) and has no reference counterpart. */
#define KF_RAYCAST_REFERENCE 0
#define KF_RAYCAST_TRILINEAR 1

/* ---- context ------------------------------------------------------------ */

/* Fills vol/prm/K/height/width with a preset (KF_PRESET_*).  Any output
 * pointer may be NULL. */
int kf_preset(int32_t which, kf_volume_desc* vol, kf_params* prm, float K[9],
              int32_t* height, int32_t* width);

/* Creates a context on the current HIP device, allocates every buffer for an
 * H x W image and the volume, and fills the volume with "never observed"
 * (NaN TSDF; weight NaN for F32 / 0 for F16) as kinfu_pipeline.cpp:61-64. */
int kf_create(const kf_volume_desc* vol, int32_t height, int32_t width,
              const kf_params* prm, kf_context** out);
int kf_destroy(kf_context* ctx);
const char* kf_last_error(void);
/* sizes[0..2] = sizeof(kf_params), sizeof(kf_volume_desc) and
 * sizeof(kf_synth_scene) as the library was built.  A binding or program
 * built against another version of this header compares them with its own
 * and refuses to run on a mismatch: the library would otherwise write a
 * grown struct past the end of the caller's copy.  No reference counterpart
 * (the reference builds its pipeline into one program). */
int kf_struct_sizes(size_t sizes[3]);
int kf_set_params(kf_context* ctx, const kf_params* prm);
int kf_get_params(const kf_context* ctx, kf_params* prm);
/* Re-fills the volume with "never observed" and the pose with identity. */
int kf_reset(kf_context* ctx);
int kf_synchronize(kf_context* ctx);
/* Bytes of device memory the context holds. */
int kf_device_bytes(const kf_context* ctx, size_t* bytes);

/* ---- A1: depth input (system_utility.cpp:58-68, 77-86) -------------------
 * Converts uint16 mm to float mm and zeroes values outside
 * [depth_min_mm, depth_max_mm] (inclusive both ends, cv::inRange).
 * mem_kind KF_MEM_HOST: the frame is copied into the context and the range
 * gate is enqueued by the next call that reads the depth; with the bilateral
 * filter on, kf_pipeline_initial / _incoming(KF_FRAME_CURRENT) gate and
 * filter it in one kernel.  KF_MEM_DEVICE: gated at once (the caller's
 * buffer is not kept). */
int kf_upload_depth_u16(kf_context* ctx, const uint16_t* depth, int32_t mem_kind);
/* Loads an already-gated float depth map unchanged (stage tests, callers
 * holding float depth as the reference's cv::Mat CV_32F does). */
int kf_upload_depth_f32(kf_context* ctx, const float* depth_mm, int32_t mem_kind);
/* Reads a raw 640x480-style uint16 little-endian .dat frame
 * (system_utility.cpp:58-68 reads "../data/dep%04d.dat") into host memory. */
int kf_read_depth_file(const char* path, int32_t height, int32_t width, uint16_t* out);
int kf_download_depth(kf_context* ctx, float* depth_mm);

/* ---- colour (system_utility.cpp:70-75, threeD_viewer.cpp:20-55) ---------
 * The reference loads each frame's colour image with cv::imread(path)
 * (LoadColorData, 8-bit B, G, R) and only uses it for its viewer, which
 * copies the raycast surface with each pixel's colour into a PCL cloud
 * (ThreeDViewer::SetUpPointClouds).  Here the context keeps an H x W x 3
 * B, G, R image on the device (zeros until one is loaded, as the
 * reference's cv::Mat::zeros at system_utility.cpp:13).
 *
 * kf_upload_color_jpeg decodes a baseline JPEG (SOF0/SOF1, 8-bit, Huffman;
 * 1 or 3 components; sampling 4:4:4, 4:2:2, 4:2:0 or other integral
 * factors except 1:2 vertical-only; restart intervals) to the values
 * cv::imread gives with libjpeg-turbo's default decompression: ISLOW
 * inverse DCT, "fancy" chroma upsampling, 16-bit fixed-point YCbCr -> RGB.
 * The markers and the Huffman decoding run on the host (the serial part);
 * dequantisation, inverse DCT, upsampling and colour conversion run on the
 * device.  The image must be H x W.  Progressive, arithmetic-coded,
 * lossless and 12-bit JPEGs, and 4-component (CMYK) images, give
 * KF_ERR_INVALID.  Returns once the host buffers may be reused; the device
 * work is enqueued. */
int kf_upload_color_jpeg(kf_context* ctx, const uint8_t* jpeg, size_t n_bytes);
/* LoadColorData: the JPEG file at path (the reference's
 * ../data/img%04d.jpg); KF_ERR_IO when it cannot be read. */
int kf_load_color_file(kf_context* ctx, const char* path);
/* Host only (no context): a baseline JPEG's size and component count. */
int kf_jpeg_info(const uint8_t* jpeg, size_t n_bytes, int32_t* height, int32_t* width,
                 int32_t* components);
/* Host only: the Huffman-decoded coefficient blocks kf_upload_color_jpeg
 * sends to the device, 64 int16 each in natural (row-major) order, not
 * dequantised: component after component, each over whole MCUs in raster
 * order (MCUs per row x horizontal sampling factor blocks per row).  coef
 * NULL: *n_blocks only; else capacity_blocks must be at least that. */
int kf_jpeg_coefficients(const uint8_t* jpeg, size_t n_bytes, int16_t* coef,
                         int64_t capacity_blocks, int64_t* n_blocks);
/* An already decoded H x W x 3 B, G, R image (mem_kind KF_MEM_*). */
int kf_upload_color_bgr(kf_context* ctx, const uint8_t* bgr, int32_t mem_kind);
int kf_download_color(kf_context* ctx, uint8_t* bgr);
/* SetUpPointClouds: H*W records of 4 floats {x, y, z, rgb}, one per pixel
 * (index = row * W + column): where the last raycast's vertex has z != 0,
 * the vertex and the colour of the same pixel packed as r << 16 | g << 8 |
 * b and stored as the bits of a float (PCL's PointXYZRGB rgb field);
 * elsewhere four zeros.  mem_kind KF_MEM_DEVICE: xyzrgb is device memory,
 * 16-byte aligned. */
int kf_point_cloud(kf_context* ctx, float* xyzrgb, int32_t mem_kind);

/* ---- A4: bilateral filter (bilateral_filtering.cu:3-108) -----------------
 * Replaces the working depth by its bilateral filtering with the context's
 * radius/sigmas.  Pixels outside [r, H-r] x [r, W-r] become 0, as the
 * reference's zero-initialised output copied over the depth does. */
int kf_bilateral(kf_context* ctx);
/* The (2r+1)^2 spatial weights exp(-(dx^2+dy^2)/(2 sigma^2)) the filter uses
 * (system_utility.cpp:88-113), row-major by (dy+r, dx+r). */
int kf_spatial_kernel(int32_t radius, float sigma_s_px, float* weights);

/* ---- A3: vertex and normal maps (raw_vertices.cu:3-126) ------------------ */
int kf_vertex_normals(kf_context* ctx, const float K[9]);
/* vz is the working depth.  Any pointer may be NULL. */
int kf_download_vertex_maps(kf_context* ctx, float* vx, float* vy, float* vz,
                            float* nx, float* ny, float* nz, uint8_t* mask);
/* Loads vertex/normal/mask planes (ICP stage tests); vz becomes the depth. */
int kf_upload_vertex_maps(kf_context* ctx, const float* vx, const float* vy,
                          const float* vz, const float* nx, const float* ny,
                          const float* nz, const uint8_t* mask);

/* ---- A5-A7: point-to-plane ICP (sensor_pose_estimation.cu:7-470) ---------
 * T_inout holds the previous pose on entry (the pose the model maps were
 * raycast at) and the refined pose on return. */
int kf_estimate_pose(kf_context* ctx, const float K[9], float T_inout[16]);
/* Per-iteration record of the last kf_estimate_pose / pipeline frame:
 * pose after each iteration [n][16], FP64 sums [n][KF_ICP_SUMS] (order:
 * upper triangle of J^TJ row-major, J^Tr, count), solver status [n]
 * (0 = LDLT solved, 1 = not positive definite: increment = right-hand side,
 * or none with icp_robust; KF_ICP_STATUS_GATED = solved but fewer than
 * icp_min_inliers inliers, no increment), inlier masks [n][H*W] when
 * record_icp is set.  Any pointer may be NULL. */
#define KF_ICP_STATUS_GATED 2
int kf_icp_history(kf_context* ctx, int32_t* n_iters, float* T_iters,
                   double* sums, int32_t* status, uint8_t* masks);
/* Coarse-to-fine ICP (icp_levels > 1; no reference counterpart).  Level l
 * has floor(H/2^l) x floor(W/2^l) pixels, each covering a 2x2 block of level
 * l-1, and intrinsics fx/2, skew/2, (cx + 0.5)/2, fy/2, (cy + 0.5)/2 of level
 * l-1 (1-indexed pixels).  Its depth is the block mean ((d00 + d01) +
 * (d10 + d11)) * 0.25 where all four are > 0 and within pyramid_max_step_mm,
 * else 0; its vertex, normal and mask maps follow raw_vertices.cu:3-51 on
 * that depth; its model maps are the block mean of the vertices and the
 * normalised sum of the normals where all four pixels hold a surface
 * (model_ok), else 0.  The iterations of the levels run coarsest first and
 * are numbered across levels in kf_icp_history (masks of level l use its
 * first H_l*W_l bytes).  kf_pyramid_level returns level l (1 or 2) as the
 * last ICP built it; K (nullable) holds the full-resolution intrinsics on
 * entry and level l's on return; any pointer may be NULL. */
int kf_pyramid_level(kf_context* ctx, int32_t level, int32_t* height, int32_t* width,
                     float K[9], float* depth, float* vx, float* vy, float* vz, float* nx,
                     float* ny, float* nz, uint8_t* mask, float* mx, float* my, float* mz,
                     float* mnx, float* mny, float* mnz, uint8_t* model_ok);

/* ICP without a context and without device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory, such
 * as code under HIP graph capture or a tight serving loop.  It runs the same
 * kernels as kf_estimate_pose, on the caller's device arrays of height*width
 * packed {x, y, z} float triplets (12 bytes per pixel):
 *   vmap     frame vertices, with z = 0 where the vertex mask is 0 (so a pixel
 *            takes part exactly where the reference's Vz != 0 && mask holds),
 *   nmap     frame normals,
 *   model_v, model_n   the model vertices and normals raycast at T_prev.
 * T_prev (host, read during the call) is the pose the model maps were
 * raycast at and the starting estimate; the refined pose (16 floats,
 * row-major) is written to the device array T_out.  Only the ICP fields of
 * prm are used (icp_iterations, icp_dist_mm, icp_angle_deg,
 * icp_compensated, icp_robust, icp_min_inliers).  The work is enqueued on
 * `stream` (a hipStream_t; NULL for the null stream) and the call returns
 * without waiting.  The block partials and the pose record live in a
 * static device array of KF_ICP_STATIC_SCRATCH_BYTES on each GPU, shared by
 * every call on that GPU and kept until the process ends: calls that
 * overlap on one GPU race (enqueue them on one stream).  Nothing in it has
 * to start at zero.  Images of more than 2^20 pixels are rejected
 * (KF_ERR_INVALID). */
#define KF_ICP_STATIC_SCRATCH_BYTES (256 * 1024)
int kf_estimate_pose_static(const float* vmap, const float* nmap, const float* model_v,
                            const float* model_n, int32_t height, int32_t width,
                            const float K[9], const kf_params* prm, const float T_prev[16],
                            float* T_out, void* stream);

/* ---- A8: projective TSDF integration (projective_tsdf.cu:5-214) ---------- */
int kf_integrate(kf_context* ctx, const float K[9], const float T[16],
                 int32_t first_frame);
/* Volume in the reference's layout: two float arrays (F32 contexts copy, F16
 * contexts widen; F16 weight 0 is returned as NaN, "never observed"). */
int kf_download_volume(kf_context* ctx, float* tsdf, float* weight);
int kf_upload_volume(kf_context* ctx, const float* tsdf, const float* weight);
/* The voxels [origin, origin + size) of the volume, in the reference layout of
 * that box (x-major, z fastest), converted as kf_download_volume does. */
int kf_download_volume_box(kf_context* ctx, const int32_t origin[3], const int32_t size[3],
                           float* tsdf, float* weight);
/* Native storage, two arrays back to back: all TSDF values, then all
 * weights (F32: float + float; F16: IEEE half bits + uint16).  Each array is
 * brick-tiled: voxel (x, y, z) sits at
 *   ((((x>>3)*By + (y>>3))*Bz + (z>>3)) << 9) | ((x&7)<<6) | ((y&7)<<3) | (z&7)
 * with By = ceil(Ny/8), Bz = ceil(Nz/8), and each array holds
 * ceil(Nx/8)*By*Bz*512 entries (padding voxels stay "never observed").
 * bytes must equal that size; kf_volume_device_ptr reports it. */
int kf_download_volume_raw(kf_context* ctx, void* out, size_t bytes);
int kf_upload_volume_raw(kf_context* ctx, const void* in, size_t bytes);
/* Device address of the native volume (for zero-copy consumers).  The call
 * drops a trilinear raycast's saved crossings (see kf_raycast_adjoint) and
 * has the next raycast rebuild its skip flags from the volume: after writing
 * the volume through the address, call it again before the next raycast or
 * adjoint. */
int kf_volume_device_ptr(kf_context* ctx, void** ptr, size_t* bytes);

/* ---- A9-A10: voxel-traversal raycast (voxel_traversal.cu:4-544) ---------
 * Skips along each pixel ray to (depth - truncation), then walks the
 * 1-indexed Amanatides-Woo traversal to the first +/- zero crossing. */
int kf_raycast(kf_context* ctx, const float K[9], const float T[16]);
/* Model maps in world mm; zero where no surface was found.  hit_index is the
 * 0-based linear index of the voxel where the crossing was detected, -1 for
 * no hit.  Any pointer may be NULL. */
int kf_download_model_maps(kf_context* ctx, float* x, float* y, float* z,
                           float* nx, float* ny, float* nz, int32_t* hit_index);
int kf_upload_model_maps(kf_context* ctx, const float* x, const float* y,
                         const float* z, const float* nx, const float* ny,
                         const float* nz);
/* Point cloud of the last raycast as a PLY file (the reference builds the
 * same cloud for its PCL viewer, threeD_viewer.cpp:20-55): one vertex per
 * pixel whose model z is non-zero, with its normal, and with colour from a
 * H x W x 3 B,G,R image when bgr is non-NULL.  binary: 1 little-endian
 * binary, 0 ASCII. */
int kf_export_model_ply(kf_context* ctx, const char* path, const uint8_t* bgr, int32_t binary);
/* The writer on host arrays of n points (x, y, z required; normals and
 * colour nullable); points with z == 0 are skipped. */
int kf_write_ply(const char* path, int64_t n, const float* x, const float* y, const float* z,
                 const float* nx, const float* ny, const float* nz, const uint8_t* bgr,
                 int32_t binary);
/* Camera-frame depth (mm) of the raycast vertices, 0 for no hit; written by
 * KF_RAYCAST_TRILINEAR raycasts (0 everywhere after a REFERENCE raycast). */
int kf_download_rendered_depth(kf_context* ctx, float* depth_mm);
/* Adjoint of the TRILINEAR raycast at pose T with the context's depth and
 * volume (no reference counterpart; the differentiable renderer).  Given
 * dL/d(rendered depth) [H*W] and optionally dL/d(normal) [H*W*3, xyz
 * interleaved, world frame], returns dL/dTSDF in the reference layout
 * [Nx*Ny*Nz] (nullable) and dL/dxi (nullable) for the left perturbation
 * T(xi) = exp([omega, v]^) T, xi = (omega_x, omega_y, omega_z, v_x, v_y,
 * v_z) in rad and mm.  When the last call that changed the context was a
 * kf_raycast in KF_RAYCAST_TRILINEAR mode with the same K and T (bitwise),
 * the adjoint starts from the crossing that raycast saved per pixel (the
 * voxel-entry parameter, in the unused w lane of its normal map) and skips
 * the walk; otherwise it first runs that raycast into buffers of its own
 * (the context's model maps are left as they are).  Both give the same sums
 * up to the order of float additions.  dL/dTSDF is accumulated with float
 * atomics, so its low bits depend on the order of arrival; with dl_dtsdf
 * NULL it is not computed at all (pose-only adjoint). */
int kf_raycast_adjoint(kf_context* ctx, const float K[9], const float T[16],
                       const float* dl_ddepth, const float* dl_dnormal, float* dl_dtsdf,
                       double dl_dxi[6]);
/* 1 if the last kf_raycast_adjoint started from a raycast's saved crossings,
 * 0 if it ran its own trilinear raycast first (or no adjoint ran yet). */
int kf_adjoint_used_saved(kf_context* ctx, int32_t* used_saved);
/* Pose refinement through the differentiable raycast (no reference
 * counterpart).  Minimises sum_p (depth_p(T) - d_p)^2 over the pixels whose
 * trilinear raycast hits and whose |residual| <= refine_max_residual_mm,
 * where d is the context's working depth (the observed frame) and depth(T)
 * the TSDF rendered at T: `steps` Gauss-Newton steps with J = d depth / d xi
 * from the raycast's derivative, FP32 sums per block, FP64 sums, solve and
 * T <- exp(-x^) T.  T_inout (nullable) sets the starting pose and receives
 * the result; otherwise the context pose is refined in place. */
int kf_refine_pose(kf_context* ctx, const float K[9], int32_t steps, float T_inout[16]);
/* Per step of the last refinement: mean squared residual (mm^2) before the
 * step, pixels used, status (0 solved; 2 too few pixels; 1 or 3 no update).
 * Any pointer may be NULL. */
int kf_refine_history(kf_context* ctx, int32_t* n_steps, double* msr, int32_t* count,
                      int32_t* status);

/* ---- frame sequence (KinfuPipeline, kinfu_pipeline.cpp:91-193) ----------- */
/* Keeps n frames of uint16 depth resident in HBM for the pipeline calls. */
int kf_stream_upload_u16(kf_context* ctx, const uint16_t* frames, int32_t n_frames,
                         int32_t mem_kind);
/* InitialProcessing: depth(frame) -> [bilateral] -> integrate(first) ->
 * raycast, at the context pose. */
int kf_pipeline_initial(kf_context* ctx, const float K[9], int32_t frame);
/* IncomingFrameProcessing: depth(frame) -> [bilateral] -> vertex/normal ->
 * ICP (updates the pose on the device; its first iteration computes the
 * vertex and normal maps) -> integrate -> raycast.  Enqueued only; nothing
 * is copied to the host.  frame = KF_FRAME_CURRENT (here and in
 * kf_pipeline_initial) takes the depth last loaded by kf_upload_depth_u16 /
 * _f32 instead of a stream frame, as the reference loads each frame from the
 * host (kinfu_pipeline.cpp:128-131); the pose is then not recorded. */
#define KF_FRAME_CURRENT (-1)
int kf_pipeline_incoming(kf_context* ctx, const float K[9], int32_t frame);
/* initial(first) then incoming(first+1 .. first+count-1); enqueued only. */
int kf_pipeline_run(kf_context* ctx, const float K[9], int32_t first, int32_t count);
int kf_get_pose(kf_context* ctx, float T[16]);
int kf_set_pose(kf_context* ctx, const float T[16]);
/* Pose after each processed stream frame, [n][16] (frame order). */
int kf_download_poses(kf_context* ctx, float* poses, int32_t n_frames);

/* ---- measurement -------------------------------------------------------- */
/* Accumulated milliseconds and launch counts per KF_STAGE_*, from hipEvents
 * recorded on the context stream while stage_timing is set. */
int kf_stage_times(kf_context* ctx, double* ms, int64_t* counts);
int kf_reset_stage_times(kf_context* ctx);
/* Raycast work counters (no reference counterpart; a measurement aid).  While
 * enabled, every raycast adds the voxels its walks visit (iterations of the
 * loop at voxel_traversal.cu:203-421) and its hit count into two device
 * counters; enabling or disabling zeroes them.  Counting uses separate kernel
 * instantiations, so it is left off when timing. */
int kf_count_raycast_work(kf_context* ctx, int32_t enable);
int kf_raycast_work(kf_context* ctx, int64_t* voxels_visited, int64_t* hits);
/* Of the counted visits, those taken by the second pass's skip rounds
 * (kf_params raycast_skip; no reference counterpart, a measurement aid). */
int kf_raycast_skipped(kf_context* ctx, int64_t* voxels_skipped);
/* Work of the last integration (no reference counterpart; a measurement
 * aid): groups of 2x2x2 bricks and 8^3-voxel bricks that passed its culls
 * (projective_tsdf.cu:5-140 visits every voxel; the culled bricks hold no
 * voxel that passes the reference's per-voxel tests).  Valid until the next
 * integration. */
int kf_integration_work(kf_context* ctx, int64_t* groups_kept, int64_t* bricks_kept);

/* ---- conventions exposed for tests ---------------------------------------
 * The angle gates are evaluated as cosine thresholds that are exactly
 * equivalent to the reference's acos tests for a correctly rounded acos:
 *   ICP  (sensor_pose_estimation.cu:63-69): acosf(c)*(180/pi_f) <= gate
 *        <=>  icp_lo <= c <= 1
 *   flip (voxel_traversal.cu:361-364): acos((double)c)*(180/pi_f) > 90
 *        <=>  -1 <= c < flip_hi                                          */
int kf_gate_thresholds(float icp_angle_deg, float* icp_lo, float* flip_hi);
/* The ICP distance gate (sensor_pose_estimation.cu:44-48) is evaluated on the
 * squared distance d2 (the reference's float sum of squares) as
 *   sqrtf(d2) <= gate && sqrtf(d2) >= 0   <=>   d2 <= dist2_hi
 * for a correctly rounded sqrtf; dist2_hi is -1 when no distance passes. */
int kf_dist_gate_threshold(float icp_dist_mm, float* dist2_hi);
/* The exp used by the bilateral filter and its spatial kernel. */
float kf_expf_abi(float x);
/* FP64 helpers rounded to FP32, as the kernels use them. */
int kf_mat4_inverse(const float T[16], float Tinv[16]);
int kf_mat3_inverse(const float K[9], float Kinv[9]);
int kf_ldlt6(const double A[36], const double b[6], double x[6]);

/* ---- synthetic RGB-D stream ----------------------------------------------
 * Box room with spheres and axis-aligned boxes, analytic raycast, Kinect
 * axial noise sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m, random holes,
 * depth zeroed beyond max_depth. Replaces the reference's data/ frames. */
#define KF_SYNTH_MAX_SPHERES 16
#define KF_SYNTH_MAX_BOXES 16
typedef struct kf_synth_scene {
  float room_min[3], room_max[3];               /* mm, seen from inside */
  int32_t n_spheres;
  float spheres[KF_SYNTH_MAX_SPHERES][4];       /* cx, cy, cz, r (mm)   */
  int32_t n_boxes;
  float boxes[KF_SYNTH_MAX_BOXES][6];           /* min xyz, max xyz     */
} kf_synth_scene;

int kf_synth_build_scene(int32_t n_spheres, int32_t n_boxes, uint64_t seed,
                         const float room_min[3], const float room_max[3],
                         kf_synth_scene* out);
/* Ground-truth camera-to-world pose of the orbit trajectory (frame 0 =
 * identity, <= 1 cm and 0.5 deg per frame). */
int kf_synth_orbit_pose(int32_t frame, float T[16]);
/* Renders one uint16 mm depth frame.  noise_scale multiplies sigma(z)
 * (0 = noiseless); clean_mm (nullable) receives the exact float z-depth. */
int kf_synth_render(const kf_synth_scene* scene, const float K[9], const float T[16],
                    int32_t height, int32_t width, float noise_scale, float hole_frac,
                    float max_depth_mm, uint64_t noise_seed, uint16_t* depth,
                    float* clean_mm, int32_t n_threads);
/* Renders frames [first, first+n) of the orbit: noise seed 1000+frame. */
int kf_synth_stream(const kf_synth_scene* scene, const float K[9], int32_t height,
                    int32_t width, int32_t first, int32_t n, float noise_scale,
                    float hole_frac, float max_depth_mm, uint16_t* frames,
                    float* gt_poses, int32_t n_threads);
/* The scene each preset uses (6 spheres + 4 boxes for the 640x480 room
 * presets, 12 + 8 for KF_PRESET_ROOM_HD_F16, 3 spheres for
 * KF_PRESET_SPHERES_128; seed 0; room inset from the volume by truncation +
 * 2 voxels). */
int kf_synth_preset_scene(int32_t which, kf_synth_scene* out);

#ifdef __cplusplus
}
#endif

#endif /* KINFU_AMD_H_ */
