"""Empty-space skipping of the raycast's second pass (kf_params raycast_skip,
no reference counterpart).  A skip round advances the walk over steps whose
voxels the per-brick skip flags show cannot hold a negative TSDF value, in
one pass over the round's boundary sequences and without loading them; the
reference's walk (voxel_traversal.cu:203-421) visits each of those voxels
and takes no decision there.  The walk must end in the same state either way.

Bar: with raycast_skip = 1 and 0, bit-identical poses, volumes, model maps
and hit indices, and identical visit and hit counts (kf_raycast_work counts
the skipped visits as visits); the skip rounds must actually be taken
(kf_raycast_skipped > 0).  The flags come from integration (room stream), or
are rebuilt from an uploaded volume (adversarial volumes with isolated
negative voxels, against the oracle's raycast) or after the volume's device
address was handed out."""
import numpy as np
import pytest

from conftest import say, small_setup
from test_gpu_long_walks import adversarial_volume

pytestmark = pytest.mark.gpu

MAP_KEYS = ("x", "y", "z", "nx", "ny", "nz", "hit")


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    ok = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b
    assert ok.all(), f"{what}: {int((~ok).sum())} of {a.size} differ"


def run_stream(kf, vol, prm, K, H, W, frames, skip, mode):
    prm.raycast_skip = skip
    prm.raycast_mode = mode
    with kf.KinfuContext(vol, prm, H, W) as c:
        c.stream_upload(frames)
        c.pipeline_run(K, 0, len(frames))
        poses = c.poses(len(frames))
        t, w = c.volume()
        maps = c.model_maps()
        # one more raycast of the final volume at the last pose, counted
        T = c.pose()
        c.count_raycast_work(True)
        c.raycast(K, T)
        visited, hits = c.raycast_work()
        skipped = c.raycast_skipped()
        c.count_raycast_work(False)
    return poses, t, w, maps, (visited, hits), skipped


@pytest.mark.parametrize("precision,mode", [("f32", 0), ("f16", 0), ("f32", 1)])
def test_room_stream_skip_on_equals_off(kf, precision, mode):
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    vol.precision = kf.KF_TSDF_F32 if precision == "f32" else kf.KF_TSDF_F16
    n = 30
    frames, _ = kf.stream(kf.preset_scene(kf.PRESET_ROOM), K, H, W, 0, n, threads=16)
    say(f"[gpu] raycast skip: room stream, {n} frames, {precision}, mode {mode}, skip on")
    on = run_stream(kf, vol, prm, K, H, W, frames, 1, mode)
    say(f"[gpu] raycast skip: room stream, {n} frames, {precision}, mode {mode}, skip off")
    off = run_stream(kf, vol, prm, K, H, W, frames, 0, mode)
    assert_same(on[0], off[0], "poses")
    assert_same(on[2], off[2], "weight")
    assert_same(on[1], off[1], "tsdf")
    for k in MAP_KEYS:
        assert_same(on[3][k], off[3][k], f"model {k}")
    assert on[4] == off[4], f"visits/hits {on[4]} vs {off[4]}"
    assert off[5] == 0
    say(f"[gpu] raycast skip: {on[5]} of {on[4][0]} visits taken by skip rounds")
    assert on[5] > 0.02 * on[4][0], "skip rounds were not taken"


def adversarial_case(kf, precision, neg_brick_frac):
    rng = np.random.default_rng(int(neg_brick_frac * 1000) + 101)
    vol, prm, K, H, W = small_setup(kf, H=96, W=128, dims=(90, 84, 120), start=(-45, -42, -10),
                                    voxel=20.0)
    vol.precision = kf.KF_TSDF_F32 if precision == "f32" else kf.KF_TSDF_F16
    K = np.array([110.0, 0, 63.5, 0, 110.0, 47.5, 0, 0, 1], np.float32)
    t, w = adversarial_volume(rng, tuple(vol.dims), neg_brick_frac)
    if precision == "f16":
        t = t.astype(np.float16).astype(np.float32)
    d = rng.uniform(300.0, 1500.0, (H, W)).astype(np.float32)
    d[rng.random((H, W)) < 0.6] = np.float32(4.8e-5)  # rays from behind the camera
    d[rng.random((H, W)) < 0.05] = 0.0
    a = 0.3
    T = np.array([[np.cos(a), 0, np.sin(a), 40.0], [0, 1, 0, -30.0],
                  [-np.sin(a), 0, np.cos(a), 15.0], [0, 0, 0, 1]], np.float32)
    return vol, prm, K, H, W, t, w, d, T


def raycast_counted(kf, vol, prm, H, W, t, w, d, K, T, skip, mode):
    prm.raycast_skip = skip
    prm.raycast_mode = mode
    with kf.KinfuContext(vol, prm, H, W) as c:
        c.upload_volume(t, w)
        c.upload_depth_f32(d)
        c.count_raycast_work(True)
        c.raycast(K, T)
        visited, hits = c.raycast_work()
        skipped = c.raycast_skipped()
        got = c.model_maps()
        if mode == 1:
            got["depth"] = c.rendered_depth()
    return got, (visited, hits), skipped


@pytest.mark.parametrize("precision,neg_brick_frac,mode",
                         [("f32", 0.0, 0), ("f32", 0.01, 0), ("f32", 0.05, 0),
                          ("f16", 0.0, 0), ("f16", 0.01, 0), ("f16", 0.05, 0),
                          ("f32", 0.0, 1), ("f32", 0.05, 1), ("f16", 0.01, 1)])
def test_adversarial_volume_skip_equals_oracle(kf, oracle, precision, neg_brick_frac, mode):
    """Isolated negative voxels in a fraction of the bricks, free space of NaN,
    positive values, +0 and -0, and rays that start behind the camera: the
    skip rounds must stop before every negative voxel a walk meets.  The
    uploaded volume's flags are rebuilt before the raycast.  F16: the
    volume is rounded to half first, and the oracle walks the rounded one.
    Mode 1 (trilinear hit handling, the same walk): against the oracle's
    trilinear raycast, and the visit counts against the walk without skip
    rounds."""
    vol, prm, K, H, W, t, w, d, T = adversarial_case(kf, precision, neg_brick_frac)
    say(f"[gpu] raycast skip: adversarial {precision} volume, negative bricks {neg_brick_frac}, "
        f"mode {mode}")
    got, work, skipped = raycast_counted(kf, vol, prm, H, W, t, w, d, K, T, 1, mode)
    if mode == 0:
        want = oracle.raycast(d, K, T, t, vol, prm.truncation_mm, 0)
        assert work == (want["steps"], want["hits"])
        keys = MAP_KEYS
    else:
        want = oracle.raycast_tri(d, K, T, t, vol, prm.truncation_mm)
        _, work_off, _ = raycast_counted(kf, vol, prm, H, W, t, w, d, K, T, 0, mode)
        assert work == work_off
        keys = MAP_KEYS + ("depth",)
    for k in keys:
        assert_same(got[k], want[k], k)
    say(f"[gpu] raycast skip: {skipped} of {work[0]} visits skipped")
    if neg_brick_frac <= 0.01:
        assert skipped > work[0] // 4, "skip rounds were not taken"


def test_flags_rebuilt_after_device_address(kf):
    """The volume's device address handed out (kf_volume_device_ptr) marks the
    skip flags stale; the next raycast rebuilds them from the volume, and the
    walk ends as without skipping."""
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    n = 12
    frames, _ = kf.stream(kf.preset_scene(kf.PRESET_ROOM), K, H, W, 0, n, threads=16)
    say("[gpu] raycast skip: flags rebuilt after the volume's address was handed out")
    out = {}
    for skip in (1, 0):
        prm.raycast_skip = skip
        with kf.KinfuContext(vol, prm, H, W) as c:
            c.stream_upload(frames)
            c.pipeline_run(K, 0, n)
            c.volume_raw()  # takes the device address
            T = c.pose()
            c.count_raycast_work(True)
            c.raycast(K, T)
            out[skip] = (c.model_maps(), c.raycast_work(), c.raycast_skipped())
            c.count_raycast_work(False)
    for k in MAP_KEYS:
        assert_same(out[1][0][k], out[0][0][k], k)
    assert out[1][1] == out[0][1]
    assert out[1][2] > 0


def test_skip_switched_on_after_integration_without_flags(kf):
    """Integration keeps the skip flags only while raycast_skip is on; a
    raycast that skips after frames integrated with it off rebuilds them."""
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    n = 12
    frames, _ = kf.stream(kf.preset_scene(kf.PRESET_ROOM), K, H, W, 0, n, threads=16)
    say("[gpu] raycast skip: switched on after integration without flags")
    prm.raycast_skip = 0
    with kf.KinfuContext(vol, prm, H, W) as c:
        c.stream_upload(frames)
        c.pipeline_run(K, 0, n)
        T = c.pose()
        c.count_raycast_work(True)
        c.raycast(K, T)
        want = (c.model_maps(), c.raycast_work())
        c.set_params(raycast_skip=1)
        c.count_raycast_work(True)
        c.raycast(K, T)
        got = (c.model_maps(), c.raycast_work(), c.raycast_skipped())
    for k in MAP_KEYS:
        assert_same(got[0][k], want[0][k], k)
    assert got[1] == want[1]
    assert got[2] > 0


def test_refinement_with_skip_rounds_equals_without(kf):
    """The pose refinement's raycasts (mode 3, walks bounded by the residual
    gate) and the pipeline's trilinear raycast with skip rounds on and off:
    the same poses after every frame."""
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    n = 8
    frames, _ = kf.stream(kf.preset_scene(kf.PRESET_ROOM), K, H, W, 0, n, threads=16)
    poses = {}
    for skip in (1, 0):
        say(f"[gpu] raycast skip: {n} frames with 2 refinement steps, skip {skip}")
        prm.raycast_skip = skip
        prm.raycast_mode = kf.KF_RAYCAST_TRILINEAR
        prm.refine_steps = 2
        with kf.KinfuContext(vol, prm, H, W) as c:
            c.stream_upload(frames)
            c.pipeline_run(K, 0, n)
            poses[skip] = c.poses(n)
    assert_same(poses[1], poses[0], "poses")
