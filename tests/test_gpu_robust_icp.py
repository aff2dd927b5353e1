"""GPU ICP with the tracking-robust update (icp_robust = 1; no reference
counterpart), through the C ABI, against the oracle's mirror of it:

* gate 0 on a well-posed frame pair: every solve succeeds, so the robust
  update is the reference's, iteration for iteration, on the GPU too;
* gate above the inlier count: no iteration is applied (status
  KF_ICP_STATUS_GATED) and the pose returned is exactly the starting pose;
* a frame with one inlier (rank-1 J^T J): the reference's update applies the
  right-hand side (checked against the oracle), the robust one keeps the pose
  (status 1).
"""
import math

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

ROT_TOL = 1e-5      # rad
TRANS_TOL_MM = 1e-2


def rot_err(A, B):
    R = A[:3, :3].astype(np.float64) @ B[:3, :3].astype(np.float64).T
    return float(np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1)))


def pose(t, rx, ry, rz):
    cx, sx, cy, sy, cz, sz = (math.cos(rx), math.sin(rx), math.cos(ry), math.sin(ry),
                              math.cos(rz), math.sin(rz))
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    T = np.eye(4)
    T[:3, :3] = Rz @ Ry @ Rx
    T[:3, 3] = t
    return T.astype(np.float32)


def context(kf, H, W, K):
    vol, prm, _, _, _ = kf.preset(kf.PRESET_ROOM)
    vol.dims[:] = [16, 16, 16]
    vol.start[:] = [-8, -8, 0]
    vol.voxel_mm = 100.0
    return kf.KinfuContext(vol, prm, H, W), prm


def frame_pair(kf, oracle, K, H, W, T1):
    scene = kf.preset_scene(kf.PRESET_ROOM)
    I4 = np.eye(4, dtype=np.float32)
    d0 = kf.render(scene, K, I4, H, W, noise_scale=0.0, hole_frac=0.0).astype(np.float32)
    d1 = kf.render(scene, K, T1, H, W, noise_scale=0.0, hole_frac=0.0).astype(np.float32)
    Ki = kf.mat3_inverse(K).ravel()
    m = oracle.vertex_normals(d0, Ki)
    model = {"x": m["vx"], "y": m["vy"], "z": d0, "nx": m["nx"], "ny": m["ny"], "nz": m["nz"]}
    fr = oracle.vertex_normals(d1, Ki)
    fr["vz"] = d1
    return fr, model


def upload(c, fr, model):
    c.upload_vertex_maps(fr["vx"], fr["vy"], fr["vz"], fr["nx"], fr["ny"], fr["nz"], fr["mask"])
    c.upload_model_maps(*(model[k] for k in ("x", "y", "z", "nx", "ny", "nz")))


def test_robust_icp_gate_zero_and_gate_above_count(kf, oracle):
    H, W = 240, 320
    K = np.array([262.5, 0, 159.5, 0, 262.5, 119.5, 0, 0, 1], np.float32)
    I4 = np.eye(4, dtype=np.float32)
    fr, model = frame_pair(kf, oracle, K, H, W, pose((8.0, -5.0, 6.0), 0.004, -0.006, 0.003))
    say("[gpu] robust ICP: gate 0 and gate above the inlier count, 320x240")
    c, prm = context(kf, H, W, K)
    with c:
        upload(c, fr, model)
        T_ref = c.estimate_pose(K, I4)
        h_ref = c.icp_history()
        c.set_params(icp_robust=1, icp_min_inliers=0)
        upload(c, fr, model)
        T_rob = c.estimate_pose(K, I4)
        h_rob = c.icp_history()
        n0 = int(h_ref["sums"][0][27])
        c.set_params(icp_robust=1, icp_min_inliers=n0 + 1)
        upload(c, fr, model)
        T_gated = c.estimate_pose(K, I4)
        h_gated = c.icp_history()
    n = prm.icp_iterations
    assert np.all(h_ref["status"] == 0)
    assert np.array_equal(T_rob, T_ref) and np.array_equal(h_rob["T"], h_ref["T"])
    assert np.array_equal(h_rob["status"], h_ref["status"])
    want = oracle.icp(fr, model, K, I4, n, prm.icp_dist_mm, prm.icp_angle_deg, reduce_mode=1,
                      min_inliers=0)
    assert rot_err(T_rob, want["T"]) < ROT_TOL
    assert np.abs(T_rob[:3, 3] - want["T"][:3, 3]).max() < TRANS_TOL_MM
    assert np.all(h_gated["status"] == kf.KF_ICP_STATUS_GATED)
    assert np.all(h_gated["sums"][:, 27] == n0)
    assert np.array_equal(T_gated, I4)
    gw = oracle.icp(fr, model, K, I4, n, prm.icp_dist_mm, prm.icp_angle_deg, reduce_mode=1,
                    min_inliers=n0 + 1)
    assert np.array_equal(gw["status"], h_gated["status"]) and np.array_equal(gw["T"], I4)


def test_robust_icp_failed_solve_keeps_the_pose(kf, oracle):
    """4x4 frame on z = 1000 with normals (0, 0, -1); the model is 1000 mm
    away except one pixel 3 mm away: one inlier, rank-1 J^T J."""
    H, W = 4, 4
    K = np.array([4.0, 0, 2.5, 0, 4.0, 2.5, 0, 0, 1], np.float32)
    y, x = np.mgrid[0:H, 0:W]
    Ki = np.linalg.inv(K.reshape(3, 3).astype(np.float64))
    vx = (1000.0 * (Ki[0, 0] * (x + 1) + Ki[0, 2])).astype(np.float32)
    vy = (1000.0 * (Ki[1, 1] * (y + 1) + Ki[1, 2])).astype(np.float32)
    vz = np.full((H, W), 1000.0, np.float32)
    zero = np.zeros((H, W), np.float32)
    fr = {"vx": vx, "vy": vy, "vz": vz, "nx": zero, "ny": zero,
          "nz": np.full((H, W), -1.0, np.float32), "mask": np.ones((H, W), np.uint8)}
    mz = (vz + 1000.0).astype(np.float32)
    mz[1, 2] = vz[1, 2] + 3.0
    model = {"x": vx, "y": vy, "z": mz, "nx": zero, "ny": zero,
             "nz": np.full((H, W), -1.0, np.float32)}
    I4 = np.eye(4, dtype=np.float32)
    say("[gpu] robust ICP: one inlier, rank-1 system, 4x4")
    c, prm = context(kf, H, W, K)
    with c:
        c.set_params(icp_iterations=2)
        upload(c, fr, model)
        T_ref = c.estimate_pose(K, I4)
        h_ref = c.icp_history()
        c.set_params(icp_robust=1, icp_min_inliers=0)
        upload(c, fr, model)
        T_rob = c.estimate_pose(K, I4)
        h_rob = c.icp_history()
    want = oracle.icp(fr, model, K, I4, 2, prm.icp_dist_mm, prm.icp_angle_deg, reduce_mode=1)
    assert np.all(h_ref["sums"][:, 27] == want["sums"][:, 27])
    assert np.array_equal(h_ref["status"], want["status"]) and np.all(h_ref["status"] == 1)
    np.testing.assert_allclose(T_ref, want["T"], rtol=1e-6, atol=1e-6)
    assert not np.array_equal(T_ref, I4)
    assert np.all(h_rob["status"] == 1) and np.all(h_rob["sums"][:, 27] == 1)
    assert np.array_equal(T_rob, I4)
