"""Long raycast walks.  The rays still walking after the first raycast pass
(a few thousand per 640x480 frame, up to ~400 steps: bilateral-filtered hole
pixels whose depth of a few 1e-5 mm starts them behind the camera) are
finished by the second pass, one wave per ray, 64 steps per round.  These
tests check that it walks exactly the reference's voxels:

  * after a GPU pipeline run, the GPU's own model maps equal the oracle's
    raycast of the GPU's own volume, bit for bit, with the same number of
    visited voxels, in F32 and F16;
  * adversarial volumes: free space of NaN, positive values, +0 and -0,
    isolated negative voxels in a fraction of the bricks, and rays that start
    behind the camera so that they walk the whole volume.
Bar: bit-exact hit indices and maps and exact work counters, as in the main
parity tests."""
import numpy as np
import pytest

from conftest import say, small_setup

pytestmark = pytest.mark.gpu


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    ok = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b
    assert ok.all(), f"{what}: {int((~ok).sum())} of {a.size} differ"


@pytest.mark.parametrize("precision", ["f32", "f16"])
def test_pipeline_maps_equal_oracle_raycast_of_own_volume(kf, oracle, room_stream, precision):
    vol, prm, K, H, W, frames, gt = room_stream
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    vol.precision = kf.KF_TSDF_F32 if precision == "f32" else kf.KF_TSDF_F16
    say(f"[gpu] long walks: 10-frame pipeline, {precision}")
    with kf.KinfuContext(vol, prm, H, W) as c:
        c.stream_upload(frames)
        c.pipeline_run(K, 0, len(frames))
        got = c.model_maps()
        d = c.depth()
        T = c.pose()
        t, _ = c.volume()
        c.count_raycast_work(True)
        c.raycast(K, T)
        visited, _ = c.raycast_work()
        c.count_raycast_work(False)
    want = oracle.raycast(d, K, T, np.ascontiguousarray(t, np.float32), vol,
                          prm.truncation_mm, 0)
    assert_same(got["hit"], want["hit"], "hit index")
    for k in ("x", "y", "z", "nx", "ny", "nz"):
        assert_same(got[k], want[k], k)
    assert visited == want["steps"]
    # the stream has hole pixels whose filtered depth is tiny: long walks
    assert (got["hit"] >= 0).mean() > 0.5


def adversarial_volume(rng, dims, neg_brick_frac):
    nx, ny, nz = dims
    u = rng.random((nx, ny, nz))
    t = np.where(u < 0.45, np.nan, rng.uniform(1e-3, 1.0, (nx, ny, nz))).astype(np.float32)
    t[(u >= 0.90) & (u < 0.95)] = 0.0
    t[u >= 0.95] = -0.0
    # isolated negative voxels in some bricks
    bx, by, bz = (nx + 7) // 8, (ny + 7) // 8, (nz + 7) // 8
    pick = rng.random((bx, by, bz)) < neg_brick_frac
    for (i, j, k) in zip(*np.nonzero(pick)):
        for _ in range(int(rng.integers(1, 4))):
            x = min(8 * i + int(rng.integers(0, 8)), nx - 1)
            y = min(8 * j + int(rng.integers(0, 8)), ny - 1)
            z = min(8 * k + int(rng.integers(0, 8)), nz - 1)
            t[x, y, z] = -rng.uniform(1e-3, 1.0)
    w = np.where(np.isnan(t), np.nan, 1.0).astype(np.float32)
    return t, w


@pytest.mark.parametrize("neg_brick_frac,camera", [(0.0, "generic"), (0.05, "generic"),
                                                    (0.3, "generic"), (0.05, "axis"),
                                                    (0.3, "axis")])
def test_adversarial_volume_long_walks(kf, oracle, neg_brick_frac, camera):
    """camera "axis": fx = fy = 128 and the principal point (64, 48) make
    K^-1 exact, so the rays of pixel column 63 and row 47 (1-indexed 64, 48)
    have a direction component of exactly +0 at the identity pose (an axis
    that never steps), and pixel (63, 47) walks along z alone."""
    rng = np.random.default_rng(int(neg_brick_frac * 1000) + 7)
    vol, prm, K, H, W = small_setup(kf, H=96, W=128, dims=(70, 66, 75), start=(-35, -33, -10),
                                    voxel=20.0)
    if camera == "generic":
        K = np.array([110.0, 0, 63.5, 0, 110.0, 47.5, 0, 0, 1], np.float32)
    else:
        K = np.array([128.0, 0, 64.0, 0, 128.0, 48.0, 0, 0, 1], np.float32)
        Ki = kf.mat3_inverse(K).ravel()
        assert Ki[0] * 64 + Ki[2] == 0 and Ki[4] * 48 + Ki[5] == 0
    t, w = adversarial_volume(rng, tuple(vol.dims), neg_brick_frac)
    d = rng.uniform(300.0, 1500.0, (H, W)).astype(np.float32)
    d[rng.random((H, W)) < 0.5] = np.float32(4.8e-5)  # rays from behind the camera
    d[rng.random((H, W)) < 0.05] = 0.0
    T = np.eye(4, dtype=np.float32)
    if camera == "axis":
        # long walks on the zero-component column and row
        assert (d[:, 63] == np.float32(4.8e-5)).sum() > 10
        assert (d[47, :] == np.float32(4.8e-5)).sum() > 10
        d[47, 63] = np.float32(4.8e-5)
    say(f"[gpu] long walks: adversarial volume, negative brick fraction {neg_brick_frac}, "
        f"{camera} camera")
    with kf.KinfuContext(vol, prm, H, W) as c:
        c.upload_volume(t, w)
        c.upload_depth_f32(d)
        c.count_raycast_work(True)
        c.raycast(K, T)
        visited, hits = c.raycast_work()
        got = c.model_maps()
    want = oracle.raycast(d, K, T, t, vol, prm.truncation_mm, 0)
    assert_same(got["hit"], want["hit"], "hit index")
    for k in ("x", "y", "z", "nx", "ny", "nz"):
        assert_same(got[k], want[k], k)
    assert (visited, hits) == (want["steps"], want["hits"])
    assert want["steps"] > 20 * H * W // 2  # long walks were exercised
