"""Edge cases through the C ABI against the oracle: ragged image and volume
sizes (nothing a multiple of the 16x16 tiles or 8^3 bricks), empty and
out-of-range frames, a tiny image and a 3x3x3 volume.  Same bar as the main
parity tests: bit-exact maps, masks, indices and volumes; poses to 1e-5 rad
and 0.01 mm."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    ok = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b
    assert ok.all(), f"{what}: {int((~ok).sum())} of {a.size} differ"


def rot_err(A, B):
    R = A[:3, :3].astype(np.float64) @ B[:3, :3].astype(np.float64).T
    return float(np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1)))


def geometry(kf, H, W, dims, start, voxel, f):
    vol, prm, _, _, _ = kf.preset(kf.PRESET_ROOM)
    vol.dims[:] = list(dims)
    vol.start[:] = list(start)
    vol.voxel_mm = voxel
    prm.truncation_mm, prm.sdf_min_mm, prm.sdf_max_mm = 3 * voxel, -3 * voxel, 30 * voxel
    K = np.array([f, 0, (W - 1) / 2, 0, f, (H - 1) / 2, 0, 0, 1], np.float32)
    return vol, prm, K


def run_stages(kf, oracle, vol, prm, K, H, W, frames, gt, precision):
    """Every stage on GPU and oracle for frames 0..2, then a 4-frame pipeline."""
    vol.precision = precision
    kind = "f32" if precision == kf.KF_TSDF_F32 else "f16"
    prm.record_icp = 1

    def gated(f):
        d = oracle.range_gate(f, prm.depth_min_mm, prm.depth_max_mm)
        return oracle.bilateral(d, prm.bilateral_radius, prm.sigma_s_px, prm.sigma_r_mm)

    with kf.KinfuContext(vol, prm, H, W) as c:
        c.upload_depth_u16(frames[1])
        assert_same(c.depth(), oracle.range_gate(frames[1], prm.depth_min_mm, prm.depth_max_mm),
                    "range gate")
        c.bilateral()
        d1 = gated(frames[1])
        assert_same(c.depth(), d1, "bilateral")
        c.vertex_normals(K)
        got = c.vertex_maps()
        want = oracle.vertex_normals(d1, kf.mat3_inverse(K).ravel())
        for k in ("vx", "vy", "nx", "ny", "nz", "mask"):
            assert_same(got[k], want[k], k)
        t, w = oracle.new_volume(vol.dims, kind)
        c.reset()
        for i in (0, 1):
            c.upload_depth_f32(gated(frames[i]))
            c.integrate(K, gt[i], i == 0)
            oracle.integrate(gated(frames[i]), K, gt[i], vol, prm, i == 0, t, w)
        gt_t, gt_w = c.volume()
        if kind == "f32":
            assert_same(gt_t, t, "tsdf")
            assert_same(gt_w, w, "weight")
        else:
            assert_same(gt_t, oracle.f16_bits_to_f32(t), "half tsdf")
            assert_same(gt_w, np.where(w == 0, np.nan, w.astype(np.float32)), "u16 weight")
        for mode in (0, 1):
            c.set_params(raycast_mode=mode)
            c.upload_depth_f32(d1)
            c.raycast(K, gt[1])
            m = c.model_maps()
            if mode == 0:
                r = oracle.raycast(d1, K, gt[1], t, vol, prm.truncation_mm, 0)
            else:
                r = oracle.raycast_tri(d1, K, gt[1], t, vol, prm.truncation_mm)
            for k in ("x", "y", "z", "nx", "ny", "nz", "hit"):
                assert_same(m[k], r[k], f"raycast mode {mode} {k}")
        c.set_params(raycast_mode=0)
        # ICP of frame 2 against the mode-0 model maps of frame 1
        model = oracle.raycast(d1, K, gt[1], t, vol, prm.truncation_mm, 0)
        c.upload_depth_f32(d1)
        c.raycast(K, gt[1])
        d2 = gated(frames[2])
        c.upload_depth_f32(d2)
        c.vertex_normals(K)
        c.estimate_pose(K, gt[1])
        hist = c.icp_history(masks=True)
        fr = oracle.vertex_normals(d2, kf.mat3_inverse(K).ravel())
        fr["vz"] = d2
        want = oracle.icp(fr, model, K, gt[1], 1, prm.icp_dist_mm, prm.icp_angle_deg,
                          reduce_mode=1)
        assert_same(hist["masks"][0], want["masks"][0], "ICP mask iteration 0")
        # the frame sequence, with compensated ICP sums against the oracle's
        # FP64 sums: these small scenes give ill-conditioned normal equations,
        # where plain FP32 sums in two different orders drift apart
        n = frames.shape[0]
        c.set_params(icp_compensated=1)
        c.reset()
        c.stream_upload(frames)
        c.pipeline_run(K, 0, n)
        poses = c.poses(n)
    if kind == "f32":
        want_poses, _, _ = oracle.pipeline(frames, K, vol, prm, reduce_mode=1)
        for i in range(n):
            assert rot_err(poses[i], want_poses[i]) < 1e-5, i
            assert np.abs(poses[i][:3, 3] - want_poses[i][:3, 3]).max() < 1e-2, i
    return model


def small_scene(kf):
    """Two spheres and a box inside a 740x820x580 mm volume at z 100..680."""
    import ctypes as C
    s = kf.SynthScene()
    rmin = (C.c_float * 3)(-3000, -3000, -3000)
    rmax = (C.c_float * 3)(3000, 3000, 3000)
    kf.lib.kf_synth_build_scene(0, 0, 0, rmin, rmax, C.byref(s))
    s.n_spheres = 2
    s.spheres[0][:] = [0.0, 20.0, 470.0, 130.0]
    s.spheres[1][:] = [-170.0, -120.0, 420.0, 70.0]
    s.n_boxes = 1
    s.boxes[0][:] = [90.0, -260.0, 380.0, 300.0, -60.0, 560.0]
    return s


@pytest.mark.parametrize("precision", ["f32", "f16"])
def test_ragged_image_and_volume(kf, oracle, precision):
    H, W = 75, 101
    vol, prm, K = geometry(kf, H, W, (37, 41, 29), (-18, -20, 5), 20.0, 70.0)
    frames, gt = kf.stream(small_scene(kf), K, H, W, 0, 4)
    say(f"[gpu] ragged 101x75 image, 37x41x29 volume, {precision}")
    prec = kf.KF_TSDF_F32 if precision == "f32" else kf.KF_TSDF_F16
    model = run_stages(kf, oracle, vol, prm, K, H, W, frames, gt, prec)
    assert (model["hit"] >= 0).sum() > 500


def test_ragged_wide_truncation_multi_batch_first_pass(kf, oracle):
    """Truncation 8 voxels (160 mm at 20 mm): the first raycast pass runs
    ceil((1.3 * 8 + 2) / 8) = 2 batches per ray (the multi-batch kernels), and
    every stage stays bit-exact against the oracle."""
    H, W = 75, 101
    vol, prm, K = geometry(kf, H, W, (37, 41, 29), (-18, -20, 5), 20.0, 70.0)
    prm.truncation_mm, prm.sdf_min_mm, prm.sdf_max_mm = 160.0, -160.0, 1600.0
    frames, gt = kf.stream(small_scene(kf), K, H, W, 0, 4)
    say("[gpu] ragged 101x75 image, 37x41x29 volume, truncation 8 voxels")
    model = run_stages(kf, oracle, vol, prm, K, H, W, frames, gt, kf.KF_TSDF_F32)
    assert (model["hit"] >= 0).sum() > 500


def test_empty_and_out_of_range_frames(kf, oracle):
    H, W = 48, 64
    vol, prm, K = geometry(kf, H, W, (40, 40, 36), (-20, -20, 4), 20.0, 60.0)
    say("[gpu] empty and out-of-range frames")
    with kf.KinfuContext(vol, prm, H, W) as c:
        for fill in (0, 65535, 100):
            frame = np.full((H, W), fill, np.uint16)
            c.upload_depth_u16(frame)
            assert np.all(c.depth() == 0)
            c.bilateral()
            assert np.all(c.depth() == 0)
            c.vertex_normals(K)
            vm = c.vertex_maps()
            assert np.all(vm["mask"] == 0) and np.all(vm["vx"] == 0)
            c.integrate(K, np.eye(4, dtype=np.float32), True)
            t, w = c.volume()
            assert np.all(np.isnan(t)) and np.all(np.isnan(w))
            for mode in (0, 1):
                c.set_params(raycast_mode=mode)
                c.raycast(K, np.eye(4, dtype=np.float32))
                m = c.model_maps()
                assert np.all(m["hit"] == -1) and np.all(m["z"] == 0)
            c.set_params(raycast_mode=0)
            T = c.estimate_pose(K, np.eye(4, dtype=np.float32))
            hist = c.icp_history()
            # no inliers: the right-hand side (zero) is the increment
            assert np.all(hist["sums"][:, 27] == 0) and np.all(hist["status"] != 0)
            assert np.array_equal(T, np.eye(4, dtype=np.float32))
        frames = np.zeros((3, H, W), np.uint16)
        c.reset()
        c.stream_upload(frames)
        c.pipeline_run(K, 0, 3)
        assert np.array_equal(c.poses(3), np.broadcast_to(np.eye(4, dtype=np.float32), (3, 4, 4)))


def test_tiny_image_and_3x3x3_volume(kf, oracle):
    H, W = 8, 8
    vol, prm, K = geometry(kf, H, W, (3, 3, 3), (-1, -1, 2), 100.0, 8.0)
    prm.bilateral_radius = 1
    frames = np.full((3, H, W), 450, np.uint16)
    frames[:, 2:5, 3:6] = 380
    gt = np.broadcast_to(np.eye(4, dtype=np.float32), (3, 4, 4)).copy()
    say("[gpu] 8x8 image, 3x3x3 volume")
    run_stages(kf, oracle, vol, prm, K, H, W, frames, gt, kf.KF_TSDF_F32)
