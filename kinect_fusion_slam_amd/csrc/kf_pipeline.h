// kf_pipeline.h - C++ host mirror of the reference's Kinfu::KinfuPipeline
// (src/kinfu_pipeline.h:14-35), driving the gfx950 stages through the C ABI.
//
// Same public surface: intrinsic_matrix (3x3 row-major), extrinsic_matrix
// (4x4 row-major camera-to-world, updated in place by ICP),
// InitialProcessing(start_frame) and IncomingFrameProcessing(frame, n).
// cv::Mat members become plain float arrays; the TSDF lives on the GPU and
// is read back with download_volume() instead of being a host array.
#pragma once

#include <stdint.h>

#include <functional>
#include <string>
#include <vector>

#include "../../include/kinfu_amd.h"

namespace Kinfu {

class KinfuPipeline {
 public:
  // The reference's constructor (kinfu_pipeline.cpp:6-65): its compiled-in
  // volume, intrinsics and truncation, with the given depth range.
  KinfuPipeline(int height, int width, float max_depth, float min_depth);
  // Any preset (KF_PRESET_*), or explicit descriptors.
  explicit KinfuPipeline(int preset);
  KinfuPipeline(const kf_volume_desc& vol, const kf_params& prm, const float K[9], int height,
                int width);
  ~KinfuPipeline();
  KinfuPipeline(const KinfuPipeline&) = delete;
  KinfuPipeline& operator=(const KinfuPipeline&) = delete;

  float intrinsic_matrix[9];
  float extrinsic_matrix[16];

  // Supplies uint16 mm depth for a frame index; returns KF_OK or an error.
  // Default: reads data_dir + "/dep%04d.dat" (system_utility.cpp:58-68).
  std::function<int(int, uint16_t*)> frame_source;
  std::string data_dir = "../data";
  bool verbose = false;  // print frame, cam_pos_* and elapsed time per frame

  // kinfu_pipeline.cpp:91-119 / :121-193.  Return KF_OK or a KF_ERR_* code.
  int InitialProcessing(int start_frame);
  int IncomingFrameProcessing(int current_frame, int per_nth_frame);

  int height() const { return h_; }
  int width() const { return w_; }
  kf_context* context() { return ctx_; }
  int status() const { return status_; }
  // The volume in the reference's layout: Nx*Ny*Nz floats each (index
  // ix*Ny*Nz + iy*Nz + iz), NaN TSDF where never observed, as the
  // reference's global_tsdf / global_weight_tsdf (kinfu_pipeline.cpp:61-64).
  int download_volume(std::vector<float>& tsdf, std::vector<float>& weight);
  const kf_volume_desc& volume_desc() const { return vol_; }
  double last_frame_ms() const { return last_ms_; }

 private:
  int init(const kf_volume_desc& vol, const kf_params& prm, const float K[9], int h, int w);
  int load(int frame);
  kf_context* ctx_ = nullptr;
  int h_ = 0, w_ = 0;
  int status_ = KF_OK;
  double last_ms_ = 0.0;
  kf_params prm_{};
  kf_volume_desc vol_{};
  std::vector<uint16_t> frame_;
};

}  // namespace Kinfu
