// kinfu_shim.cpp - test-only C entry points over Kinfu::KinfuPipeline, the C++
// host mirror of the reference's KinfuPipeline (kinfu_pipeline.cpp:6-193).
//
// TEST INFRASTRUCTURE ONLY.  A pytest process loads this library with ctypes
// (so no other process is started), and it drives the mirror the way the
// reference's main loop drives KinfuPipeline (system_interface.cpp:21-60):
// the reference's constructor (height, width, max_depth, min_depth), then
// InitialProcessing(start_frame) and IncomingFrameProcessing for each later
// frame, each frame read from data_dir/dep%04d.dat by kf_read_depth_file.
#include <string.h>

#include <vector>

#include "kf_pipeline.h"

extern "C" {

// Runs frames first .. first+count-1; poses receives extrinsic_matrix after
// each frame ([count][16]), tsdf and weight the final volume in the
// reference's layout (n_voxels floats each).  Returns KF_OK or a KF_ERR_*.
int shim_run_reference_pipeline(const char* data_dir, int height, int width, float max_depth,
                                float min_depth, int first, int count, float* poses,
                                float* tsdf, float* weight, int64_t n_voxels) {
  if (!data_dir || !poses || !tsdf || !weight || count < 1) return KF_ERR_INVALID;
  // KinfuPipeline holds a kf_params and a kf_volume_desc: a library built
  // from another header would construct it past the end of this frame.
  size_t sz[3];
  if (kf_struct_sizes(sz) != KF_OK || sz[0] != sizeof(kf_params) ||
      sz[1] != sizeof(kf_volume_desc))
    return KF_ERR_INVALID;
  Kinfu::KinfuPipeline p(height, width, max_depth, min_depth);
  if (p.status() != KF_OK) return p.status();
  p.data_dir = data_dir;
  int rc = p.InitialProcessing(first);
  if (rc != KF_OK) return rc;
  memcpy(poses, p.extrinsic_matrix, sizeof p.extrinsic_matrix);
  for (int i = 1; i < count; ++i) {
    rc = p.IncomingFrameProcessing(first + i, 1);
    if (rc != KF_OK) return rc;
    memcpy(poses + 16 * i, p.extrinsic_matrix, sizeof p.extrinsic_matrix);
  }
  std::vector<float> t, w;
  rc = p.download_volume(t, w);
  if (rc != KF_OK) return rc;
  if ((int64_t)t.size() != n_voxels || (int64_t)w.size() != n_voxels) return KF_ERR_INVALID;
  memcpy(tsdf, t.data(), t.size() * sizeof(float));
  memcpy(weight, w.data(), w.size() * sizeof(float));
  return KF_OK;
}

}  // extern "C"
