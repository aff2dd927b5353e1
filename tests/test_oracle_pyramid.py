"""Pins for the oracle's coarse-to-fine ICP pyramid (no GPU needed).

The pyramid has no reference counterpart (the reference runs its ICP at full
resolution, sensor_pose_estimation.cu:339), so it is pinned by what the
geometry fixes:

* the intrinsics convention: on a fronto-parallel plane every coarse vertex
  is the mean of the four fine vertices it covers (a wrong principal point
  or focal length shifts it by a fraction of a pixel), at both levels;
* the depth rule: exact plane depths, and zero for a hole, a NaN or a step
  above the limit (inclusive at the limit);
* the model rule: block means of the vertices, renormalised normal sums,
  and no surface unless all four children hold one;
* the ICP: a known rigid motion is recovered coarse to fine, with each
  level's iterations starting from the level before.
"""
import math

import numpy as np

from conftest import ROOT  # noqa: F401
from test_oracle_pins import intr, pose


def kinv(kf):
    return lambda K: kf.mat3_inverse(np.asarray(K, np.float32).reshape(3, 3)).ravel()


def test_pyramid_intrinsics_map_fine_vertices_to_their_block_mean(kf, oracle):
    K = intr(525.0, 319.5, 239.5)
    H, W, D = 480, 640, 1500.0
    d = np.full((H, W), D, np.float32)
    fine = oracle.vertex_normals(d, kinv(kf)(K))
    Kl, dl = K, d
    for level in (1, 2):
        Kl = oracle.pyramid_intrinsics(Kl)
        dl = oracle.pyramid_depth(dl, 30.0)
        assert dl.shape == (H >> level, W >> level)
        assert np.all(dl == np.float32(D))  # the mean of four equal depths is exact
        coarse = oracle.vertex_normals(dl, kinv(kf)(Kl))
        s = 1 << level
        h, w = dl.shape
        for key in ("vx", "vy"):
            # mean of the s x s fine block (the coarse pixel's footprint)
            blocks = fine[key][:h * s, :w * s].reshape(h, s, w, s).astype(np.float64)
            want = blocks.mean(axis=(1, 3))
            ok = coarse["mask"].astype(bool)
            ok[-1, :] = ok[:, -1] = False
            ok &= np.all(fine["mask"][:h * s, :w * s].reshape(h, s, w, s), axis=(1, 3))
            # the last fine row/column carries no vertex (mask rule)
            ok[-1, :] = False
            ok[:, -1] = False
            err = np.abs(coarse[key][ok] - want[ok]).max()
            assert err < 2e-3, (level, key, err)  # mm; a half-pixel shift is ~1.4 mm
        nrm = coarse["nz"][:-1, :-1]
        assert np.all(nrm == -1.0)  # plane facing the camera: normal (0, 0, -1) exactly


def test_pyramid_depth_rule(oracle):
    d = np.array([[1000, 1000, 1000, 0, 1000, 1030, 1000, np.nan],
                  [1000, 1000, 1000, 1000, 1000, 1000, 1000, 1000],
                  [1000, 1031, 1000, 1004, 1000, 1000, 10, 10],
                  [1000, 1000, 1000, 1002, 1000, 1000, 10, 10]], np.float32)
    out = oracle.pyramid_depth(d, 30.0)
    assert out.shape == (2, 4)
    np.testing.assert_array_equal(out[0], [1000.0, 0.0, 1007.5, 0.0])  # hole, step 30, NaN
    np.testing.assert_array_equal(out[1], [0.0, ((1000 + 1004) + (1000 + 1002)) * 0.25,
                                           1000.0, 10.0])  # step 31 > 30
    # odd sizes drop the last row and column
    assert oracle.pyramid_depth(np.ones((5, 7), np.float32), 1.0).shape == (2, 3)


def test_pyramid_model_rule(oracle):
    rng = np.random.default_rng(3)
    H, W = 6, 8
    m = {k: rng.normal(size=(H, W)).astype(np.float32) for k in ("x", "y", "z")}
    n = rng.normal(size=(3, H, W)).astype(np.float32)
    n /= np.linalg.norm(n, axis=0)
    m.update(nx=n[0], ny=n[1], nz=n[2])
    ok = np.ones((H, W), np.uint8)
    ok[1, 3] = 0
    out = oracle.pyramid_model(m, ok)
    want_ok = np.ones((3, 4), np.uint8)
    want_ok[0, 1] = 0
    np.testing.assert_array_equal(out["ok"], want_ok)
    for k in ("x", "y", "z", "nx", "ny", "nz"):
        assert np.all(out[k][want_ok == 0] == 0)
    good = want_ok.astype(bool)
    for k in ("x", "y", "z"):
        b = m[k].reshape(3, 2, 4, 2)
        want = ((b[:, 0, :, 0] + b[:, 0, :, 1]) + (b[:, 1, :, 0] + b[:, 1, :, 1])) * np.float32(0.25)
        np.testing.assert_array_equal(out[k][good], want[good])
    s = np.stack([n[i].reshape(3, 2, 4, 2).astype(np.float64).sum(axis=(1, 3)) for i in range(3)])
    s /= np.linalg.norm(s, axis=0)
    for i, k in enumerate(("nx", "ny", "nz")):
        assert np.abs(out[k][good] - s[i][good]).max() < 1e-6


def test_pyramid_icp_recovers_a_known_rigid_motion(kf, oracle):
    """The room seen from two poses 25 mm and 1.5 deg apart, noiseless: the
    model maps from frame 0's depth at identity, the frame at T1.  Coarse to
    fine (4, 5, 10 iterations at quarter, half and full resolution) the pose
    is recovered to the depth quantisation."""
    scene = kf.preset_scene(kf.PRESET_ROOM)
    K = intr(525.0, 319.5, 239.5)
    H, W = 480, 640
    I4 = np.eye(4, dtype=np.float32)
    T1 = pose(t=(18.0, -12.0, 12.0), rx=0.012, ry=-0.018, rz=0.01)
    d0 = kf.render(scene, K, I4, H, W, noise_scale=0.0, hole_frac=0.0).astype(np.float32)
    d1 = kf.render(scene, K, T1, H, W, noise_scale=0.0, hole_frac=0.0).astype(np.float32)
    m = oracle.vertex_normals(d0, kinv(kf)(K))
    model = {"x": m["vx"], "y": m["vy"], "z": d0, "nx": m["nx"], "ny": m["ny"], "nz": m["nz"]}
    hit = np.where((d0 > 0) & (m["nz"] != 0), 0, -1)
    levels = oracle.pyramid_levels(d1, model, hit, K, kinv(kf), 3, 30.0)
    assert [lv[1]["vz"].shape for lv in levels] == [(480, 640), (240, 320), (120, 160)]
    T, record = oracle.icp_pyramid(levels, I4, (10, 5, 4), 100.0, 20.0)
    assert [l for l, _ in record] == [2, 1, 0]
    # each level starts where the coarser one ended
    np.testing.assert_array_equal(record[1][1]["T_iters"][0].shape, (4, 4))
    T = T.astype(np.float64)
    assert np.abs(T[:3, 3] - T1[:3, 3]).max() < 1.0
    R = T[:3, :3] @ T1[:3, :3].T.astype(np.float64)
    assert math.degrees(math.acos(min(1.0, (np.trace(R) - 1) / 2))) < 0.05
