"""GPU stages against the committed golden case (tests/golden/room_small.npz),
without the oracle: bit-exact for maps, masks, indices and volumes; poses to
1e-5 rad / 0.01 mm (the ICP sums are reduced in another order)."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, say

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_golden as mg  # noqa: E402

pytestmark = pytest.mark.gpu
PATH = os.path.join(ROOT, "tests", "golden", "room_small.npz")


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    if a.dtype.kind == "f":
        ok = (a == b) | (np.isnan(a) & np.isnan(b))
    else:
        ok = a == b
    assert ok.all(), f"{what}: {int((~ok).sum())} of {a.size} differ"


def rot_err(A, B):
    R = A[:3, :3].astype(np.float64) @ B[:3, :3].astype(np.float64).T
    return float(np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1)))


@pytest.fixture(scope="module")
def golden():
    return np.load(PATH)


def make_ctx(kf, precision=None, record_icp=0):
    vol, prm = mg.descs()
    if precision is not None:
        vol.precision = precision
    prm.record_icp = record_icp
    return kf.KinfuContext(vol, prm, mg.H, mg.W)


def test_golden_stages_f32(kf, golden, tmp_path):
    g = golden
    K, gt = g["K"], g["gt"]
    say("[gpu] golden case: depth gate, bilateral, vertex maps")
    with make_ctx(kf, record_icp=1) as c:
        c.upload_depth_u16(g["frames"][1])
        assert_same(c.depth(), g["gate1"], "range gate")
        c.bilateral()
        assert_same(c.depth(), g["bilateral1"], "bilateral")
        c.vertex_normals(K)
        vm = c.vertex_maps()
        for k in ("vx", "vy", "nx", "ny", "nz", "mask"):
            assert_same(vm[k], g["vertex1_" + k], k)
        say("[gpu] golden case: integration of frames 0, 1")
        c.reset()
        for i in (0, 1):
            c.upload_depth_u16(g["frames"][i])
            c.bilateral()
            c.integrate(K, gt[i], i == 0)
        t, w = c.volume()
        assert_same(t, g["tsdf01"], "tsdf")
        assert_same(w, g["weight01"], "weight")
        say("[gpu] golden case: raycast at frame 1")
        c.upload_depth_f32(g["bilateral1"])
        c.raycast(K, gt[1])
        m = c.model_maps()
        for k in ("x", "y", "z", "nx", "ny", "nz", "hit"):
            assert_same(m[k], g["raycast1_" + k], "raycast " + k)
        # the point cloud of that raycast
        bgr = np.full((mg.H, mg.W, 3), (10, 20, 30), np.uint8)
        c.export_ply(tmp_path / "model.ply", bgr)
        ply = kf.read_ply(tmp_path / "model.ply")
        keep = g["raycast1_z"] != 0
        for k in ("x", "y", "z", "nx", "ny", "nz"):
            assert_same(ply[k], g["raycast1_" + k][keep], "ply " + k)
        assert np.all(ply["red"] == 30) and np.all(ply["blue"] == 10)
        say("[gpu] golden case: ICP of frame 2")
        c.upload_depth_u16(g["frames"][2])
        c.bilateral()
        c.vertex_normals(K)
        T = c.estimate_pose(K, gt[1])
        hist = c.icp_history(masks=True)
        assert_same(hist["masks"][0], g["icp2_masks"][0], "ICP inlier mask iteration 0")
        assert rot_err(T, g["icp2_T"]) < 1e-5
        assert np.abs(T[:3, 3] - g["icp2_T"][:3, 3]).max() < 1e-2


def test_golden_integration_f16(kf, golden):
    g = golden
    say("[gpu] golden case: FP16 integration")
    with make_ctx(kf, precision=kf.KF_TSDF_F16) as c:
        for i in (0, 1):
            c.upload_depth_u16(g["frames"][i])
            c.bilateral()
            c.integrate(g["K"], g["gt"][i], i == 0)
        t, w = c.volume()
    th, wh = g["tsdf01_f16"], g["weight01_f16"]
    assert_same(t, th.view(np.float16).astype(np.float32), "half tsdf")
    assert_same(w, np.where(wh == 0, np.nan, wh.astype(np.float32)), "u16 weight")


def test_golden_pipeline_poses(kf, golden):
    g = golden
    say("[gpu] golden case: 3-frame pipeline")
    with make_ctx(kf) as c:
        c.stream_upload(g["frames"])
        c.pipeline_run(g["K"], 0, g["frames"].shape[0])
        poses = c.poses(g["frames"].shape[0])
    for i in range(poses.shape[0]):
        assert rot_err(poses[i], g["pipeline_poses"][i]) < 1e-5, i
        assert np.abs(poses[i][:3, 3] - g["pipeline_poses"][i][:3, 3]).max() < 1e-2, i
