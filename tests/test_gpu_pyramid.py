"""The opt-in coarse-to-fine ICP pyramid on the GPU (icp_levels = 3).

No reference counterpart: the reference runs its 10 ICP iterations at full
resolution (sensor_pose_estimation.cu:339), which stays the default.

* parity: the half- and quarter-resolution levels the GPU builds (depth,
  vertex and normal maps, mask, model maps and their surface flags) are
  bit-exact against the oracle's restatement; the inlier mask and inlier
  count of every iteration (4 + 5 + 10 over the three levels) match the
  oracle replayed from the GPU's own starting pose of that iteration; the
  final pose is within the ICP tests' tolerance of the oracle's pyramid ICP;
* the room stream taken every 4th frame (up to 4 cm and 2 deg between
  frames): trajectory error with and without the pyramid;
* 1280x720 into 1024x1024x768 voxels at 4 mm (FP16): trajectory error and
  frames/s with and without the pyramid.
"""
import time

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    same = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b
    assert same.all(), f"{what}: {int((~same).sum())} of {a.size} differ"


def gated(oracle, frame, prm):
    d = oracle.range_gate(frame, prm.depth_min_mm, prm.depth_max_mm)
    if prm.bilateral:
        d = oracle.bilateral(d, prm.bilateral_radius, prm.sigma_s_px, prm.sigma_r_mm)
    return d


def ate_mm(poses, gt):
    e = np.linalg.norm(poses[:, :3, 3].astype(np.float64) - gt[:, :3, 3], axis=1)
    return float(np.sqrt((e ** 2).mean()))


def kinv(kf):
    return lambda K: kf.mat3_inverse(np.asarray(K, np.float32).reshape(3, 3)).ravel()


@pytest.mark.parametrize("compensated", [0, 1])
def test_pyramid_levels_masks_and_pose_match_oracle(kf, room_stream, oracle, compensated):
    vol, prm, K, H, W, frames, gt = room_stream
    t, w = oracle.new_volume(vol.dims)
    for i in (0, 1):
        oracle.integrate(gated(oracle, frames[i], prm), K, gt[i], vol, prm, i == 0, t, w)
    d1 = gated(oracle, frames[1], prm)
    model = oracle.raycast(d1, K, gt[1], t, vol, prm.truncation_mm, 0)
    d2 = gated(oracle, frames[2], prm)
    iters = (prm.icp_iterations, 5, 4)
    say(f"[gpu] ICP pyramid 3 levels, compensated={compensated}")
    with kf.KinfuContext.from_preset(kf.PRESET_ROOM, record_icp=1, icp_levels=3,
                                     icp_coarse_iters=[5, 4],
                                     icp_compensated=compensated) as c:
        c.upload_volume(t, w)
        c.upload_depth_f32(d1)
        c.raycast(K, gt[1])
        c.upload_depth_f32(d2)
        c.vertex_normals(K)
        T = c.estimate_pose(K, gt[1])
        hist = c.icp_history(masks=True)
        got = {l: c.pyramid_level(l, K) for l in (1, 2)}
        step = c.params.pyramid_max_step_mm
    levels = oracle.pyramid_levels(d2, model, model["hit"], K, kinv(kf), 3, step)
    for l in (1, 2):
        Kl, fr, mo = levels[l]
        g = got[l]
        assert (g["H"], g["W"]) == fr["vz"].shape == (H >> l, W >> l)
        assert_same(g["K"], Kl, f"level {l} K")
        assert_same(g["depth"], fr["vz"], f"level {l} depth")
        for k in ("vx", "vy", "vz", "nx", "ny", "nz", "mask"):
            assert_same(g["frame"][k], fr[k], f"level {l} frame {k}")
        for k in ("x", "y", "z", "nx", "ny", "nz", "ok"):
            assert_same(g["model"][k], mo[k], f"level {l} model {k}")
        assert g["model"]["ok"].mean() > 0.5 and g["frame"]["mask"].mean() > 0.5
    # every iteration, coarsest level first, replayed from the GPU's start pose
    order = [l for l in (2, 1, 0) for _ in range(iters[l])]
    assert hist["T"].shape[0] == len(order) == sum(iters)
    start = gt[1]
    for i, l in enumerate(order):
        Kl, fr, mo = levels[l]
        r = oracle.icp(fr, mo, Kl, gt[1], 1, prm.icp_dist_mm, prm.icp_angle_deg,
                       reduce_mode=1, T_start=start)
        h, w_ = fr["vz"].shape
        mask = hist["masks"][i].ravel()[:h * w_].reshape(h, w_)
        assert_same(mask, r["masks"][0], f"inlier mask iteration {i} (level {l})")
        assert hist["sums"][i][27] == r["sums"][0][27]
        start = hist["T"][i]
    want, _ = oracle.icp_pyramid(levels, gt[1], iters, prm.icp_dist_mm, prm.icp_angle_deg)
    R = T[:3, :3].astype(np.float64) @ want[:3, :3].astype(np.float64).T
    assert np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1)) < 1e-5
    assert np.abs(T[:3, 3] - want[:3, 3]).max() < 1e-2


def test_pyramid_on_a_fast_room_stream(kf):
    """Every 4th frame of the room orbit: the pyramid's coarse levels see a
    wider basin.  Report both trajectories' error; the pyramid must not be
    worse."""
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    frames, gt = kf.stream(kf.preset_scene(kf.PRESET_ROOM), K, H, W, 0, 300, threads=16)
    frames, gt = frames[::4], gt[::4]
    n = frames.shape[0]
    res = {}
    for levels in (1, 3):
        say(f"[gpu] room stream every 4th frame ({n} frames), icp_levels={levels}")
        with kf.KinfuContext.from_preset(kf.PRESET_ROOM, icp_levels=levels) as c:
            c.stream_upload(frames)
            c.pipeline_run(K, 0, n)
            res[levels] = ate_mm(c.poses(n), gt)
    say(f"[gpu] fast room stream: ATE single level {res[1]:.2f} mm, pyramid {res[3]:.2f} mm")
    assert res[3] <= res[1] + 0.5


def test_room_hd_with_and_without_pyramid(kf):
    """1280x720 into 1024x1024x768 @4mm FP16, 200 frames: ICP with the
    reference's 10 full-resolution iterations, then with 5 + 4 coarse
    iterations first."""
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM_HD_F16)
    n = 200
    frames, gt = kf.stream(kf.preset_scene(kf.PRESET_ROOM_HD_F16), K, H, W, 0, n, threads=16)
    res = {}
    for levels in (1, 3):
        say(f"[gpu] 1280x720 HD F16: {n} frames, icp_levels={levels}")
        with kf.KinfuContext.from_preset(kf.PRESET_ROOM_HD_F16, icp_levels=levels) as c:
            c.stream_upload(frames)
            c.synchronize()
            t0 = time.perf_counter()
            c.pipeline_run(K, 0, n)
            c.synchronize()
            el = time.perf_counter() - t0
            res[levels] = (ate_mm(c.poses(n), gt), n / el)
    say(f"[gpu] 1280x720 HD: ICP single level ATE {res[1][0]:.3f} mm at {res[1][1]:.1f} "
        f"frames/s; pyramid ATE {res[3][0]:.3f} mm at {res[3][1]:.1f} frames/s")
    assert res[3][0] < 10.0 and res[1][0] < 10.0
