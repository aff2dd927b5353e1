"""Opt-in ICP early stop (icp_stop_rot / icp_stop_trans_mm; no reference
counterpart, off in every preset): through the C ABI, against the oracle's
mirror (oracle.icp(stop=...)).

* a well-posed 320x240 frame pair: the GPU's own run without the stop gives
  the increments of its 10 iterations; thresholds far from the first small
  one make the stopped run end right after it, with the same poses, sums
  and inlier masks as the full run up to there (the masks bit-exact against
  the oracle replayed from each GPU starting pose), the oracle stopping
  after the same iteration;
* the frame sequence with the stop: every pose within the parity tolerance
  of the full run's, fewer iterations in total;
* with the coarse-to-fine pyramid the coarse levels always run in full.
"""
import numpy as np
import pytest

from conftest import say
from test_gpu_robust_icp import TRANS_TOL_MM, ROT_TOL, context, frame_pair, pose, rot_err, upload

pytestmark = pytest.mark.gpu

STOP = (5e-6, 5e-3)  # rad, mm


def increments(T_iters, T_start):
    """x (rotation, translation) of each increment, from consecutive poses."""
    prev = [T_start] + list(T_iters)
    out = []
    for i in range(len(T_iters)):
        D = prev[i + 1].astype(np.float64) @ np.linalg.inv(prev[i].astype(np.float64))
        out.append((np.array([D[1, 2], D[2, 0], D[0, 1]]), D[:3, 3]))
    return out


def first_small(incs):
    small = [bool(np.abs(r).max() < STOP[0] and np.abs(t).max() < STOP[1]) for r, t in incs]
    assert not small[0] and any(small)
    k = small.index(True)
    r, t = incs[k]
    assert np.abs(r).max() < 0.5 * STOP[0] and np.abs(t).max() < 0.5 * STOP[1]
    r, t = incs[k - 1]
    assert np.abs(r).max() > 2 * STOP[0] or np.abs(t).max() > 2 * STOP[1]
    return k


def test_icp_early_stop_frame_pair(kf, oracle):
    H, W = 240, 320
    K = np.array([262.5, 0, 159.5, 0, 262.5, 119.5, 0, 0, 1], np.float32)
    I4 = np.eye(4, dtype=np.float32)
    fr, model = frame_pair(kf, oracle, K, H, W, pose((8.0, -5.0, 6.0), 0.004, -0.006, 0.003))
    say("[gpu] ICP early stop: 320x240 frame pair, full run then stopped run")
    c, prm = context(kf, H, W, K)
    with c:
        c.set_params(record_icp=1)
        upload(c, fr, model)
        T_full = c.estimate_pose(K, I4)
        h_full = c.icp_history(masks=True)
        c.set_params(icp_stop_rot=STOP[0], icp_stop_trans_mm=STOP[1])
        upload(c, fr, model)
        T_stop = c.estimate_pose(K, I4)
        h_stop = c.icp_history(masks=True)
    n = prm.icp_iterations
    assert h_full["T"].shape[0] == n
    k = first_small(increments(h_full["T"], I4))
    assert h_stop["T"].shape[0] == k + 1
    assert np.array_equal(h_stop["T"], h_full["T"][:k + 1])
    assert np.array_equal(h_stop["sums"], h_full["sums"][:k + 1])
    assert np.array_equal(h_stop["masks"], h_full["masks"][:k + 1])
    assert np.array_equal(T_stop, h_full["T"][k])
    start = I4
    for i in range(k + 1):
        r = oracle.icp(fr, model, K, I4, 1, prm.icp_dist_mm, prm.icp_angle_deg, reduce_mode=1,
                       T_start=start)
        assert np.array_equal(h_stop["masks"][i], r["masks"][0]), f"inlier mask iteration {i}"
        start = h_stop["T"][i]
    want = oracle.icp(fr, model, K, I4, n, prm.icp_dist_mm, prm.icp_angle_deg, reduce_mode=1,
                      stop=STOP)
    assert want["iters_done"] == k + 1
    assert rot_err(T_stop, want["T"]) < ROT_TOL
    assert np.abs(T_stop[:3, 3] - want["T"][:3, 3]).max() < TRANS_TOL_MM
    assert rot_err(T_stop, T_full) < ROT_TOL
    assert np.abs(T_stop[:3, 3] - T_full[:3, 3]).max() < TRANS_TOL_MM


@pytest.mark.parametrize("levels", [1, 2])
def test_icp_early_stop_frame_sequence(kf, room_stream, levels):
    vol, prm, K, H, W, frames, gt = room_stream
    n = frames.shape[0]
    say(f"[gpu] ICP early stop: {n}-frame sequence, {levels} level(s)")
    poses, iters = {}, {}
    for stop in (False, True):
        v, p, _, _, _ = kf.preset(kf.PRESET_ROOM)
        p.icp_levels = levels
        if stop:
            p.icp_stop_rot, p.icp_stop_trans_mm = STOP
        with kf.KinfuContext(v, p, H, W) as c:
            c.stream_upload(frames)
            c.pipeline_initial(K, 0)
            done = []
            for f in range(1, n):
                c.pipeline_incoming(K, f)
                done.append(c.icp_history()["T"].shape[0])
            poses[stop] = c.poses(n)
            iters[stop] = np.array(done)
    coarse = p.icp_coarse_iters[0] if levels == 2 else 0
    assert np.all(iters[False] == coarse + p.icp_iterations)
    assert np.all(iters[True] > coarse) and np.all(iters[True] <= coarse + p.icp_iterations)
    assert iters[True].sum() < iters[False].sum()
    for i in range(n):
        assert rot_err(poses[True][i], poses[False][i]) < ROT_TOL, i
        assert np.abs(poses[True][i][:3, 3] - poses[False][i][:3, 3]).max() < TRANS_TOL_MM, i
