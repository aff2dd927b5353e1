"""Writes tests/golden/room_small.npz: a small seeded case run through every
oracle stage, so that the CPU tests notice any change to the oracle and the
GPU tests can compare the HIP stages with fixed outputs.

The reference publishes no test vectors and cannot be built here, so these
vectors come from the oracle (whose stages are pinned separately by closed
forms and brute force).  Regenerate only on purpose:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import kinect_fusion_slam_amd as kf  # noqa: E402
from oracle import kf_oracle as oracle  # noqa: E402

# geometry: 64x48 camera, 40x40x36 voxels at 20 mm in front of it
H, W = 48, 64
DIMS, START, VOXEL = (40, 40, 36), (-20, -20, 4), 20.0
K = np.array([60.0, 0, 31.5, 0, 60.0, 23.5, 0, 0, 1], np.float32)
TRUNC, SMIN, SMAX = 60.0, -60.0, 600.0
DMIN, DMAX = 300.0, 4000.0
BIL_R, SIG_S, SIG_R = 2, 4.5, 30.0
ICP_ITERS, ICP_DIST, ICP_ANGLE = 3, 100.0, 20.0
N_FRAMES = 3


def descs():
    vol, prm, _, _, _ = kf.preset(kf.PRESET_ROOM)
    vol.dims[:] = list(DIMS)
    vol.start[:] = list(START)
    vol.voxel_mm = VOXEL
    prm.truncation_mm, prm.sdf_min_mm, prm.sdf_max_mm = TRUNC, SMIN, SMAX
    prm.depth_min_mm, prm.depth_max_mm = DMIN, DMAX
    prm.bilateral, prm.bilateral_radius = 1, BIL_R
    prm.sigma_s_px, prm.sigma_r_mm = SIG_S, SIG_R
    prm.icp_iterations, prm.icp_dist_mm, prm.icp_angle_deg = ICP_ITERS, ICP_DIST, ICP_ANGLE
    return vol, prm


def build():
    vol, prm = descs()
    scene = kf.preset_scene(kf.PRESET_ROOM)
    frames, gt = kf.stream(scene, K, H, W, 0, N_FRAMES)
    out = {"frames": frames, "gt": gt, "K": K}

    def gated(f):
        return oracle.bilateral(oracle.range_gate(f, DMIN, DMAX), BIL_R, SIG_S, SIG_R)

    out["gate1"] = oracle.range_gate(frames[1], DMIN, DMAX)
    d1 = gated(frames[1])
    out["bilateral1"] = d1
    vm = oracle.vertex_normals(d1, kf.mat3_inverse(K).ravel())
    for k, v in vm.items():
        out["vertex1_" + k] = v
    t, w = oracle.new_volume(DIMS, "f32")
    for i in (0, 1):
        oracle.integrate(gated(frames[i]), K, gt[i], vol, prm, i == 0, t, w)
    out["tsdf01"], out["weight01"] = t, w
    th, wh = oracle.new_volume(DIMS, "f16")
    for i in (0, 1):
        oracle.integrate(gated(frames[i]), K, gt[i], vol, prm, i == 0, th, wh)
    out["tsdf01_f16"], out["weight01_f16"] = th, wh
    rc = oracle.raycast(d1, K, gt[1], t, vol, TRUNC, 0)
    for k in ("x", "y", "z", "nx", "ny", "nz", "hit"):
        out["raycast1_" + k] = rc[k]
    d2 = gated(frames[2])
    fr = oracle.vertex_normals(d2, kf.mat3_inverse(K).ravel())
    fr["vz"] = d2
    icp = oracle.icp(fr, rc, K, gt[1], ICP_ITERS, ICP_DIST, ICP_ANGLE, reduce_mode=0)
    out["icp2_T"], out["icp2_T_iters"] = icp["T"], icp["T_iters"]
    out["icp2_masks"], out["icp2_sums"] = icp["masks"], icp["sums"]
    poses, pt, pw = oracle.pipeline(frames, K, vol, prm, reduce_mode=0)
    out["pipeline_poses"], out["pipeline_tsdf"], out["pipeline_weight"] = poses, pt, pw
    return out


if __name__ == "__main__":
    data = build()
    path = os.path.join(HERE, "room_small.npz")
    np.savez_compressed(path, **data)
    print(path, os.path.getsize(path), "bytes")
