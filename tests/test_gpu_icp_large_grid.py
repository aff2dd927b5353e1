"""The ICP's large-image grid: above 256 blocks of 2048 pixels, one block of
1024 threads per compute unit (3 or 4 pixels per thread) instead of more
512-thread blocks (icp_grid; k_icp_iter<..., BLK = 1024>).

At 1280x720 (921 600 pixels: 256 blocks of 3600), in plain and compensated
sums: the context ICP (its model normal factor plane, NT = true) and the
allocation-free entry point on the same maps (normalize3 in the kernel,
NT = false) run the same grid, so the poses are bit-exact; both agree with
the oracle's ICP to the pose tolerance of the other ICP tests (1e-2 mm,
1e-5 rad).
"""
import numpy as np
import pytest

from conftest import say
from test_zz_no_alloc import context_pose, icp_case, static_pose

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("compensated", [0, 1])
def test_icp_1280x720_context_and_static_grids_agree(kf, oracle, compensated):
    import torch
    H, W = 720, 1280
    K = np.array([1000.0, 0, 639.5, 0, 1000.0, 359.5, 0, 0, 1], np.float32)
    vol, prm, gt, fr, model = icp_case(kf, oracle, K, H, W)
    prm.icp_compensated = compensated
    say(f"[gpu] ICP 1280x720 (compensated {compensated}): allocation-free entry point")
    T_static = static_pose(kf, torch, prm, K, H, W, fr, model, gt[1])
    say(f"[gpu] ICP 1280x720 (compensated {compensated}): context")
    T_ctx = context_pose(kf, vol, prm, K, H, W, fr, model, gt[1])
    assert np.isfinite(T_ctx).all()
    assert np.array_equal(T_static, T_ctx), (T_static, T_ctx)
    assert not np.array_equal(T_ctx, gt[1])
    want = oracle.icp(fr, model, K, gt[1], prm.icp_iterations, prm.icp_dist_mm,
                      prm.icp_angle_deg, reduce_mode=1)
    assert np.abs(T_ctx[:3, 3] - want["T"][:3, 3]).max() < 1e-2
    assert np.abs(T_ctx[:3, :3] - want["T"][:3, :3]).max() < 1e-5
