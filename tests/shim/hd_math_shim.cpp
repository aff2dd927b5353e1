// hd_math_shim.cpp - test-only C entry points over the host build of
// kinect_fusion_slam_amd/csrc/kf_hd_math.h, the arithmetic that the gfx950
// kernels and the CPU oracle share.
//
// TEST INFRASTRUCTURE ONLY.  Because the kernels and the oracle both use
// these functions, a mistake in one of them would not show in a GPU-vs-oracle
// comparison; tests/test_hd_math.py loads this library with ctypes and pins
// each function against numpy, exact rational arithmetic or closed forms.
// Built with g++ -ffp-contract=off, as the oracle is.
#include "kf_hd_math.h"

extern "C" {

int hd_round_to_int(float v) { return kf_round_to_int(v); }
int hd_floor_to_int(float v) { return kf_floor_to_int(v); }
void hd_mat4_mul(const float* A, const float* B, float* C) { kf_mat4_mul(A, B, C); }
int hd_sym_index(int i, int j) { return kf_sym_index(i, j); }
int hd_icp_increment(const double* sums, float* inc) { return kf_icp_increment(sums, inc); }
int hd_icp_increment_robust(const double* sums, float* inc, double min_inliers) {
  return kf_icp_increment_robust(sums, inc, min_inliers);
}
int hd_icp_update_gated(const double* sums, float* T, double min_inliers) {
  return kf_icp_update_gated(sums, T, min_inliers);
}
int hd_icp_converged(const float* inc, float stop_rot, float stop_trans) {
  return kf_icp_converged(inc, stop_rot, stop_trans) ? 1 : 0;
}
float hd_inv_norm3(float x, float y, float z) { return kf_inv_norm3(x, y, z); }

}  // extern "C"
