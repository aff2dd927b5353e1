// kf_pipeline.cpp - Kinfu::KinfuPipeline over the C ABI (kinfu_pipeline.cpp:6-193).
#include "kf_pipeline.h"

#include <stdio.h>
#include <string.h>

#include <chrono>

namespace Kinfu {

KinfuPipeline::KinfuPipeline(int height, int width, float max_depth, float min_depth) {
  kf_volume_desc vol;
  kf_params prm;
  float K[9];
  kf_preset(KF_PRESET_REFERENCE, &vol, &prm, K, nullptr, nullptr);
  prm.depth_max_mm = max_depth;
  prm.depth_min_mm = min_depth;
  status_ = init(vol, prm, K, height, width);
}

KinfuPipeline::KinfuPipeline(int preset) {
  kf_volume_desc vol;
  kf_params prm;
  float K[9];
  int h = 0, w = 0;
  status_ = kf_preset(preset, &vol, &prm, K, &h, &w);
  if (status_ == KF_OK) status_ = init(vol, prm, K, h, w);
}

KinfuPipeline::KinfuPipeline(const kf_volume_desc& vol, const kf_params& prm, const float K[9],
                             int height, int width) {
  status_ = init(vol, prm, K, height, width);
}

KinfuPipeline::~KinfuPipeline() { kf_destroy(ctx_); }

int KinfuPipeline::init(const kf_volume_desc& vol, const kf_params& prm, const float K[9], int h,
                        int w) {
  h_ = h;
  w_ = w;
  prm_ = prm;
  vol_ = vol;
  memcpy(intrinsic_matrix, K, sizeof intrinsic_matrix);
  const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  memcpy(extrinsic_matrix, I, sizeof extrinsic_matrix);
  frame_.resize((size_t)h * w);
  return kf_create(&vol, h, w, &prm, &ctx_);
}

int KinfuPipeline::load(int frame) {
  int rc;
  if (frame_source) {
    rc = frame_source(frame, frame_.data());
  } else {
    char path[512];
    snprintf(path, sizeof path, "%s/dep%04d.dat", data_dir.c_str(), frame);
    rc = kf_read_depth_file(path, h_, w_, frame_.data());
  }
  if (rc != KF_OK) return rc;
  rc = kf_upload_depth_u16(ctx_, frame_.data(), KF_MEM_HOST);
  if (rc == KF_OK && prm_.bilateral) rc = kf_bilateral(ctx_);
  return rc;
}

int KinfuPipeline::InitialProcessing(int start_frame) {
  if (status_ != KF_OK) return status_;
  int rc = load(start_frame);
  if (rc == KF_OK) rc = kf_integrate(ctx_, intrinsic_matrix, extrinsic_matrix, 1);
  if (rc == KF_OK) rc = kf_raycast(ctx_, intrinsic_matrix, extrinsic_matrix);
  if (rc == KF_OK) rc = kf_synchronize(ctx_);
  return rc;
}

int KinfuPipeline::IncomingFrameProcessing(int current_frame, int /*per_nth_frame*/) {
  if (status_ != KF_OK) return status_;
  const auto t0 = std::chrono::steady_clock::now();
  int rc = load(current_frame);
  if (rc == KF_OK) rc = kf_vertex_normals(ctx_, intrinsic_matrix);
  if (rc == KF_OK) rc = kf_estimate_pose(ctx_, intrinsic_matrix, extrinsic_matrix);
  if (rc == KF_OK && verbose)
    printf("frame %d\ncam_pos_x %g\ncam_pos_y %g\ncam_pos_z %g\n", current_frame,
           extrinsic_matrix[3], extrinsic_matrix[7], extrinsic_matrix[11]);
  if (rc == KF_OK) rc = kf_integrate(ctx_, intrinsic_matrix, extrinsic_matrix, 0);
  if (rc == KF_OK) rc = kf_raycast(ctx_, intrinsic_matrix, extrinsic_matrix);
  if (rc == KF_OK) rc = kf_synchronize(ctx_);
  last_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rc == KF_OK && verbose) printf("elapsed time: %g s\n", last_ms_ / 1000.0);
  return rc;
}

int KinfuPipeline::download_volume(std::vector<float>& tsdf, std::vector<float>& weight) {
  if (status_ != KF_OK) return status_;
  const size_t n = (size_t)vol_.dims[0] * vol_.dims[1] * vol_.dims[2];
  tsdf.resize(n);
  weight.resize(n);
  return kf_download_volume(ctx_, tsdf.data(), weight.data());
}

}  // namespace Kinfu
