"""Pose refinement through the differentiable raycast (kf_refine_pose).

A three-sphere TSDF is rendered (FP32 oracle, trilinear mode) at a known
pose; that rendering is the observed depth.  Starting 0.8 deg and 5 mm away,
the Gauss-Newton refinement must return to within 0.2 mm and 0.05 deg of the
known pose, with the mean squared depth residual falling by more than 1000x.
(The floor is not zero: the observation was rendered with another skip
depth, so its samples along each ray sit elsewhere on the trilinear field.)
"""
import numpy as np
import pytest

from conftest import say
from se3 import se3_exp, sphere_tsdf

pytestmark = pytest.mark.gpu


def rot_deg(A, B):
    R = A[:3, :3].astype(np.float64) @ B[:3, :3].astype(np.float64).T
    return float(np.degrees(np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1))))


def test_refinement_recovers_a_known_pose(kf, oracle):
    vol, prm, _, _, _ = kf.preset(kf.PRESET_SPHERES_128)
    vol.dims[:] = [64, 64, 64]
    vol.start[:] = [-32, -32, 0]
    vol.voxel_mm = 10.0
    H, W = 120, 160
    K = np.array([120.0, 0, 79.5, 0, 120.0, 59.5, 0, 0, 1], np.float32)
    trunc = prm.truncation_mm
    F = sphere_tsdf(vol, [(0, 0, 420.0, 120.0), (110, -60, 330.0, 70.0),
                          (-100, 80, 360.0, 80.0)], trunc).astype(np.float32)
    T_true = np.eye(4)
    T_true[:3, :3] = se3_exp([0.01, -0.02, 0.005, 0, 0, 0])[:3, :3]
    T_true[:3, 3] = (5.0, -4.0, 10.0)
    T_true = T_true.astype(np.float32)
    r = oracle.raycast_tri(np.full((H, W), 250.0, np.float32), K, T_true, F, vol, trunc)
    obs = np.where(r["hit"] >= 0, r["depth"], 0).astype(np.float32)
    assert (obs > 0).sum() > 5000
    T0 = (se3_exp([np.radians(0.6), np.radians(-0.5), np.radians(0.3), 4.0, -3.0, 2.0])
          @ T_true.astype(np.float64)).astype(np.float32)
    say("[gpu] refinement: 64^3 three-sphere volume, 160x120")
    with kf.KinfuContext(vol, prm, H, W) as c:
        c.upload_volume(F, np.where(np.isnan(F), np.nan, 1.0).astype(np.float32))
        c.upload_depth_f32(obs)
        T = c.refine_pose(K, 12, T0)
        hist = c.refine_history()
    say(f"[gpu] refinement: msr {hist['msr'][0]:.3g} -> {hist['msr'][-1]:.3g} mm^2, "
        f"pixels {hist['count'][0]} -> {hist['count'][-1]}, "
        f"error {np.abs(T[:3, 3] - T_true[:3, 3]).max():.4f} mm {rot_deg(T, T_true):.5f} deg")
    assert np.all(hist["status"] == 0)
    assert hist["msr"][-1] < 1e-3 * hist["msr"][0]
    assert np.abs(T[:3, 3] - T_true[:3, 3]).max() < 0.2
    assert rot_deg(T, T_true) < 0.05


def test_refinement_pixels_match_oracle_with_far_starts(kf, oracle):
    """One refinement step's used pixels (trilinear hit with |depth - observed|
    <= refine_max_residual_mm) and mean squared residual against the FP32
    oracle's trilinear render at the same pose.  A third of the observed
    depths are halved, so those rays start far in front of the surface and
    walk past the parameter beyond which no crossing can pass the residual
    gate (where the refinement's walk stops); the 15 mm gate splits the
    other pixels' residuals on both sides."""
    vol, prm, _, _, _ = kf.preset(kf.PRESET_SPHERES_128)
    vol.dims[:] = [64, 64, 64]
    vol.start[:] = [-32, -32, 0]
    vol.voxel_mm = 10.0
    prm.refine_max_residual_mm = 15.0
    H, W = 120, 160
    K = np.array([120.0, 0, 79.5, 0, 120.0, 59.5, 0, 0, 1], np.float32)
    trunc = prm.truncation_mm
    F = sphere_tsdf(vol, [(0, 0, 420.0, 120.0), (110, -60, 330.0, 70.0),
                          (-100, 80, 360.0, 80.0)], trunc).astype(np.float32)
    T_true = np.eye(4, dtype=np.float32)
    r = oracle.raycast_tri(np.full((H, W), 250.0, np.float32), K, T_true, F, vol, trunc)
    obs = np.where(r["hit"] >= 0, r["depth"], 0).astype(np.float32)
    obs[:, : W // 3] *= np.float32(0.5)
    T0 = se3_exp([np.radians(1.0), np.radians(-0.8), 0.0, 6.0, -5.0, 8.0]).astype(np.float32)
    want = oracle.raycast_tri(obs, K, T0, F, vol, trunc)
    res = want["depth"] - obs
    used = (want["hit"] >= 0) & (obs > 0) & (np.abs(res) <= np.float32(15.0))
    far = (obs > 0) & (np.arange(W)[None, :] < W // 3)
    assert used.sum() > 2000 and (~used & (want["hit"] >= 0) & (obs > 0)).sum() > 500
    assert not (used & far).any()
    say("[gpu] refinement pixels vs oracle: 64^3 three-sphere volume, 160x120, far starts")
    with kf.KinfuContext(vol, prm, H, W) as c:
        c.upload_volume(F, np.where(np.isnan(F), np.nan, 1.0).astype(np.float32))
        c.upload_depth_f32(obs)
        c.refine_pose(K, 1, T0)
        hist = c.refine_history()
    msr = float((res[used].astype(np.float64) ** 2).mean())
    say(f"[gpu] refinement step 0: pixels {hist['count'][0]} (oracle {int(used.sum())}), "
        f"msr {hist['msr'][0]:.6g} (oracle {msr:.6g})")
    assert hist["count"][0] == int(used.sum())
    assert abs(hist["msr"][0] - msr) <= 1e-5 * msr
