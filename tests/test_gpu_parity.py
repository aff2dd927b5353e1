"""GPU parity: every HIP stage against the CPU oracle on identical inputs.

Bar (from the kernels' design, every stage evaluates the oracle's float
expressions in the same order with -ffp-contract=off):
  * depth gate, bilateral, vertex/normal maps, integration (F32 and F16),
    raycast maps and hit-voxel indices, ICP inlier masks: bit-exact;
  * ICP poses, whose sums are reduced in a different order: 1e-5 rad in
    rotation and 1e-5 m (0.01 mm) in translation.
"""
import numpy as np
import pytest

from conftest import say, small_setup

pytestmark = pytest.mark.gpu

ROT_TOL = 1e-5      # rad
TRANS_TOL_MM = 1e-2  # 1e-5 m


def rot_err(A, B):
    R = A[:3, :3].astype(np.float64) @ B[:3, :3].astype(np.float64).T
    return float(np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1)))


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    if a.dtype.kind == "f":
        same = (a == b) | (np.isnan(a) & np.isnan(b))
    else:
        same = a == b
    bad = int((~same).sum())
    assert bad == 0, f"{what}: {bad} of {a.size} elements differ"


def kf_sym(i, j):
    """Index of J^TJ[i][j] (i <= j) in the 28 ICP sums."""
    return i * 6 - (i * (i - 1)) // 2 + (j - i)


def gated(oracle, frame, prm):
    d = oracle.range_gate(frame, prm.depth_min_mm, prm.depth_max_mm)
    if prm.bilateral:
        d = oracle.bilateral(d, prm.bilateral_radius, prm.sigma_s_px, prm.sigma_r_mm)
    return d


@pytest.fixture(scope="module")
def ctx_room(kf):
    say("[gpu] create 640x480 room context (38.88M voxels, F32)")
    c = kf.KinfuContext.from_preset(kf.PRESET_ROOM, record_icp=1)
    yield c
    c.close()


def test_depth_gate_bitexact(ctx_room, room_stream, oracle):
    vol, prm, K, H, W, frames, gt = room_stream
    for f in (0, 5):
        say(f"[gpu] depth gate frame {f}")
        ctx_room.upload_depth_u16(frames[f])
        assert_same(ctx_room.depth(), oracle.range_gate(frames[f], prm.depth_min_mm,
                                                    prm.depth_max_mm), "depth")


@pytest.mark.parametrize("radius", [1, 2, 3, 5, 7])
def test_bilateral_bitexact(ctx_room, room_stream, oracle, radius):
    vol, prm, K, H, W, frames, gt = room_stream
    say(f"[gpu] bilateral r={radius}")
    ctx_room.set_params(bilateral_radius=radius)
    try:
        ctx_room.upload_depth_u16(frames[2])
        ctx_room.bilateral()
        got = ctx_room.depth()
    finally:
        ctx_room.set_params(bilateral_radius=prm.bilateral_radius)
    d = oracle.range_gate(frames[2], prm.depth_min_mm, prm.depth_max_mm)
    want = oracle.bilateral(d, radius, prm.sigma_s_px, prm.sigma_r_mm)
    assert_same(got, want, "bilateral")
    assert np.all(got[:radius] == 0) and np.all(got[:, :radius] == 0)


@pytest.mark.parametrize("radius", [2, 5])
@pytest.mark.parametrize("sigma_r", [5.0, 30.0, 150.0])
def test_pipeline_gate_bilateral_bitexact(kf, room_stream, oracle, radius, sigma_r):
    """The pipeline's per-frame path: fused range gate + bilateral on u16
    frames, then the vertex/normal maps.  Its range weights of integer depth differences come from a table while the exact-zero
    tail starts within 2048 entries (sigma_r 5 and 30 mm: 73 and 433 entries)
    and from expf otherwise (150 mm)."""
    vol, prm, K, H, W, frames, gt = room_stream
    vs, ps, _, _, _ = kf.preset(kf.PRESET_ROOM)
    vs.dims[:] = [40, 40, 36]
    ps.bilateral, ps.bilateral_radius, ps.sigma_r_mm = 1, radius, sigma_r
    say(f"[gpu] pipeline gate + bilateral + maps r={radius} sigma_r={sigma_r}")
    with kf.KinfuContext(vs, ps, H, W) as c:
        c.stream_upload(frames[:3])
        c.pipeline_run(K, 0, 3)
        got = c.depth()
        maps = c.vertex_maps()
    d = oracle.range_gate(frames[2], prm.depth_min_mm, prm.depth_max_mm)
    want = oracle.bilateral(d, radius, ps.sigma_s_px, sigma_r)
    assert_same(got, want, "bilateral")
    wm = oracle.vertex_normals(want, kf.mat3_inverse(K).ravel())
    for k in ("vx", "vy", "nx", "ny", "nz", "mask"):
        assert_same(maps[k], wm[k], k)
    assert_same(maps["vz"], want, "vz")


def test_vertex_normals_bitexact(kf, ctx_room, room_stream, oracle):
    vol, prm, K, H, W, frames, gt = room_stream
    say("[gpu] vertex/normal maps")
    d = gated(oracle, frames[4], prm)
    ctx_room.upload_depth_f32(d)
    ctx_room.vertex_normals(K)
    got = ctx_room.vertex_maps()
    want = oracle.vertex_normals(d, kf.mat3_inverse(K).ravel())
    for k in ("vx", "vy", "nx", "ny", "nz", "mask"):
        assert_same(got[k], want[k], k)
    assert_same(got["vz"], d, "vz")


def _integrate_both(kf, ctx, oracle, vol, prm, K, depths, poses, kind):
    t, w = oracle.new_volume(vol.dims, kind)
    ctx.reset()
    for i, (d, T) in enumerate(zip(depths, poses)):
        say(f"[gpu] integrate {kind} frame {i}")
        ctx.upload_depth_f32(d)
        ctx.integrate(K, T, i == 0)
        oracle.integrate(d, K, T, vol, prm, i == 0, t, w)
    return t, w


def test_integrate_f32_bitexact(kf, ctx_room, room_stream, oracle):
    vol, prm, K, H, W, frames, gt = room_stream
    depths = [gated(oracle, frames[i], prm) for i in (0, 3, 6)]
    t, w = _integrate_both(kf, ctx_room, oracle, vol, prm, K, depths, gt[[0, 3, 6]], "f32")
    gt_t, gt_w = ctx_room.volume()
    assert_same(gt_t, t, "tsdf")
    assert_same(gt_w, w, "weight")
    obs = np.isfinite(w)
    assert obs.sum() > 100000
    assert np.all(np.abs(t[obs]) <= 1.0)
    assert np.all(np.isnan(t[~obs]))


def test_integrate_f16_bitexact(kf, room_stream, oracle):
    vol, prm, K, H, W, frames, gt = room_stream
    say("[gpu] create 640x480 room context in F16")
    ctx = kf.KinfuContext.from_preset(kf.PRESET_ROOM, precision=kf.KF_TSDF_F16)
    try:
        vol16 = ctx.vol
        depths = [gated(oracle, frames[i], prm) for i in (0, 2, 4)]
        t, w = _integrate_both(kf, ctx, oracle, vol16, prm, K, depths, gt[[0, 2, 4]], "f16")
        # the reference layout widens half -> float exactly, weight 0 -> NaN
        got_t, got_w = ctx.volume()
        assert_same(got_t, oracle.f16_bits_to_f32(t), "half tsdf")
        assert_same(got_w, np.where(w == 0, np.nan, w.astype(np.float32)), "u16 weight")
        assert int((w > 0).sum()) > 100000
    finally:
        ctx.close()


@pytest.fixture(scope="module")
def model_volume(kf, room_stream, oracle):
    """Oracle F32 volume after frames 0 and 1 at ground-truth poses."""
    vol, prm, K, H, W, frames, gt = room_stream
    t, w = oracle.new_volume(vol.dims, "f32")
    for i in (0, 1):
        oracle.integrate(gated(oracle, frames[i], prm), K, gt[i], vol, prm, i == 0, t, w)
    return t, w


@pytest.mark.parametrize("fix", [0, 1])
def test_raycast_f32_bitexact(kf, ctx_room, room_stream, oracle, model_volume, fix):
    vol, prm, K, H, W, frames, gt = room_stream
    t, w = model_volume
    say(f"[gpu] raycast F32 fix_bounds={fix}")
    ctx_room.set_params(raycast_fix_bounds=fix)
    try:
        ctx_room.upload_volume(t, w)
        d = gated(oracle, frames[1], prm)
        ctx_room.upload_depth_f32(d)
        ctx_room.raycast(K, gt[1])
        got = ctx_room.model_maps()
    finally:
        ctx_room.set_params(raycast_fix_bounds=0)
    want = oracle.raycast(d, K, gt[1], t, vol, prm.truncation_mm, fix)
    assert_same(got["hit"], want["hit"], "hit index")
    for k in ("x", "y", "z", "nx", "ny", "nz"):
        assert_same(got[k], want[k], k)
    assert (got["hit"] >= 0).mean() > 0.5


def test_raycast_work_counters_match_oracle(kf, ctx_room, room_stream, oracle, model_volume):
    """The counting kernels visit exactly the oracle's voxels and hits, and
    produce the same maps as the plain ones."""
    vol, prm, K, H, W, frames, gt = room_stream
    t, w = model_volume
    say("[gpu] raycast work counters")
    ctx_room.upload_volume(t, w)
    d = gated(oracle, frames[1], prm)
    ctx_room.upload_depth_f32(d)
    ctx_room.count_raycast_work(True)
    try:
        ctx_room.raycast(K, gt[1])
        ctx_room.raycast(K, gt[1])
        visited, hits = ctx_room.raycast_work()
        got = ctx_room.model_maps()
    finally:
        ctx_room.count_raycast_work(False)
    want = oracle.raycast(d, K, gt[1], t, vol, prm.truncation_mm, 0)
    assert (visited, hits) == (2 * want["steps"], 2 * want["hits"])
    assert_same(got["hit"], want["hit"], "hit index")
    assert_same(got["z"], want["z"], "z")


def test_raycast_f16_bitexact(kf, room_stream, oracle):
    vol, prm, K, H, W, frames, gt = room_stream
    say("[gpu] raycast F16")
    ctx = kf.KinfuContext.from_preset(kf.PRESET_ROOM, precision=kf.KF_TSDF_F16)
    try:
        t, w = oracle.new_volume(vol.dims, "f16")
        for i in (0, 1):
            oracle.integrate(gated(oracle, frames[i], prm), K, gt[i], ctx.vol, prm, i == 0, t, w)
        ctx.upload_volume(oracle.f16_bits_to_f32(t), np.where(w == 0, np.nan, w.astype(np.float32)))
        d = gated(oracle, frames[1], prm)
        ctx.upload_depth_f32(d)
        ctx.raycast(K, gt[1])
        got = ctx.model_maps()
        want = oracle.raycast(d, K, gt[1], t, ctx.vol, prm.truncation_mm, 0)
        assert_same(got["hit"], want["hit"], "hit index")
        for k in ("x", "y", "z", "nx", "ny", "nz"):
            assert_same(got[k], want[k], k)
    finally:
        ctx.close()


@pytest.mark.parametrize("compensated", [0, 1])
def test_icp_masks_bitexact_and_pose(kf, ctx_room, room_stream, oracle, model_volume,
                                     compensated):
    vol, prm, K, H, W, frames, gt = room_stream
    t, w = model_volume
    d1 = gated(oracle, frames[1], prm)
    model = oracle.raycast(d1, K, gt[1], t, vol, prm.truncation_mm, 0)
    d2 = gated(oracle, frames[2], prm)
    fr = oracle.vertex_normals(d2, kf.mat3_inverse(K).ravel())
    fr["vz"] = d2
    say(f"[gpu] ICP compensated={compensated}")
    ctx_room.set_params(icp_compensated=compensated)
    try:
        ctx_room.upload_vertex_maps(fr["vx"], fr["vy"], d2, fr["nx"], fr["ny"], fr["nz"], fr["mask"])
        ctx_room.upload_model_maps(*(model[k] for k in ("x", "y", "z", "nx", "ny", "nz")))
        T = ctx_room.estimate_pose(K, gt[1])
        hist = ctx_room.icp_history(masks=True)
    finally:
        ctx_room.set_params(icp_compensated=0)
    n = prm.icp_iterations
    assert hist["T"].shape[0] == n
    # replay each GPU iteration's starting estimate through the oracle: the
    # association and gates must select exactly the same pixels
    start = gt[1]
    for i in range(n):
        r = oracle.icp(fr, model, K, gt[1], 1, prm.icp_dist_mm, prm.icp_angle_deg,
                       reduce_mode=1, T_start=start)
        assert_same(hist["masks"][i], r["masks"][0], f"inlier mask iteration {i}")
        assert hist["sums"][i][27] == r["sums"][0][27]
        # Float sums in another order differ by a few ulps of the sum of
        # |terms|, bounded by Cauchy-Schwarz: sum|a_i a_j| <= sqrt(A_ii A_jj)
        # and sum|a_i b| <= sqrt(A_ii * N) * dist_gate (|b| <= point distance).
        ref = r["sums"][0]
        diag = np.array([ref[kf_sym(i_, i_)] for i_ in range(6)])
        scale = np.array([np.sqrt(diag[i_] * diag[j_]) for i_ in range(6) for j_ in range(i_, 6)]
                         + [np.sqrt(diag[i_] * ref[27]) * prm.icp_dist_mm for i_ in range(6)])
        err = np.abs(hist["sums"][i][:27] - ref[:27]) / scale
        assert err.max() < (1e-6 if compensated else 1e-5), err.max()
        start = hist["T"][i]
    want = oracle.icp(fr, model, K, gt[1], n, prm.icp_dist_mm, prm.icp_angle_deg, reduce_mode=1)
    assert rot_err(T, want["T"]) < ROT_TOL
    assert np.abs(T[:3, 3] - want["T"][:3, 3]).max() < TRANS_TOL_MM


def test_pipeline_first_frames_match_oracle(kf, ctx_room, room_stream, oracle):
    vol, prm, K, H, W, frames, gt = room_stream
    n = frames.shape[0]
    say(f"[gpu] pipeline {n} frames")
    ctx_room.reset()
    ctx_room.stream_upload(frames)
    ctx_room.pipeline_run(K, 0, n)
    poses = ctx_room.poses(n)
    want, t, w = oracle.pipeline(frames, K, vol, prm, reduce_mode=0)
    for i in range(n):
        assert rot_err(poses[i], want[i]) < ROT_TOL, i
        assert np.abs(poses[i][:3, 3] - want[i][:3, 3]).max() < TRANS_TOL_MM, i
