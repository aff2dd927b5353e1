"""The larger workloads on the GPU, as parity and accuracy reports.

* 720x720x600 voxels at 5 mm, FP16 TSDF + u16 weight, 300 frames with the
  sensor noise halved, compensated ICP sums: trajectory error against ground
  truth (and against the same run with an FP32 volume), and the voxel-wise
  TSDF error of the FP16 integration against the FP64 oracle on a 40-voxel
  slab through the volume.
"""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def ate_mm(poses, gt):
    e = np.linalg.norm(poses[:, :3, 3].astype(np.float64) - gt[:, :3, 3], axis=1)
    return float(np.sqrt((e ** 2).mean()))


def test_precision_variant_5mm_f16(kf, oracle):
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM_5MM_F16)
    assert vol.precision == kf.KF_TSDF_F16 and prm.icp_compensated == 1
    n = 300
    frames, gt = kf.stream(kf.preset_scene(kf.PRESET_ROOM_5MM_F16), K, H, W, 0, n,
                           noise_scale=0.5, threads=16)
    ate = {}
    for name, prec, mode in (("f16", kf.KF_TSDF_F16, 0), ("f32", kf.KF_TSDF_F32, 0),
                             ("f16_trilinear", kf.KF_TSDF_F16, 1)):
        say(f"[gpu] 720x720x600 @5mm: {n} frames, {name}")
        v, p, _, _, _ = kf.preset(kf.PRESET_ROOM_5MM_F16)
        v.precision = prec
        p.raycast_mode = mode
        with kf.KinfuContext(v, p, H, W) as c:
            c.stream_upload(frames)
            c.pipeline_run(K, 0, n)
            ate[name] = ate_mm(c.poses(n), gt)
    # TSDF of the first three frames at ground-truth poses: FP16 GPU vs FP64 oracle
    x0, nx = 340, 40
    sub = type("V", (), {})()
    sub.dims = [nx, vol.dims[1], vol.dims[2]]
    sub.start = [vol.start[0] + x0, vol.start[1], vol.start[2]]
    sub.voxel_mm = vol.voxel_mm
    t64, w64 = oracle.new_volume(sub.dims, "f64")
    t32, w32 = oracle.new_volume(sub.dims, "f32")
    say("[gpu] 720x720x600 @5mm: FP16 integration of 3 frames vs the FP64 oracle")
    with kf.KinfuContext(vol, prm, H, W) as c:
        for i in range(3):
            d = oracle.bilateral(oracle.range_gate(frames[i], prm.depth_min_mm, prm.depth_max_mm),
                                 prm.bilateral_radius, prm.sigma_s_px, prm.sigma_r_mm)
            c.upload_depth_f32(d)
            c.integrate(K, gt[i], i == 0)
            oracle.integrate(d, K, gt[i], sub, prm, i == 0, t64, w64)
            oracle.integrate(d, K, gt[i], sub, prm, i == 0, t32, w32)
        t16, w16 = c.volume_box([x0, 0, 0], sub.dims)
    obs16, obs64 = np.isfinite(w16), np.isfinite(w64)
    both = obs16 & obs64
    err = np.abs(t16[both].astype(np.float64) - t64[both])
    err32 = np.abs(t16[both].astype(np.float64) - t32[both])
    q = np.quantile(err, [0.5, 0.99, 0.999])
    say(f"[gpu] 5mm F16: ATE f16 {ate['f16']:.3f} mm, f32 {ate['f32']:.3f} mm, f16 trilinear "
        f"{ate['f16_trilinear']:.3f} mm; TSDF |F16 - FP64| (x trunc {prm.truncation_mm} mm) "
        f"mean {err.mean():.2e} median {q[0]:.2e} p99 {q[1]:.2e} p99.9 {q[2]:.2e} max "
        f"{err.max():.2e}; |F16 - FP32 update| max {err32.max():.2e}; {both.sum()} voxels, "
        f"observed-set mismatch {(obs16 != obs64).sum()}")
    assert both.sum() > 100000
    assert (obs16 != obs64).sum() <= 1e-3 * obs64.sum()
    # storage: three running-average updates, each rounded to half (2^-12
    # below 1), against the same updates in FP32
    assert err32.max() <= 2.0 ** -11
    # against FP64: the same bound except where FP32 and FP64 project a voxel
    # to different pixels in some frame (a different depth sample)
    assert q[2] <= 2.0 ** -11 and (err > 2.0 ** -11).mean() < 1e-3
    assert abs(ate["f16"] - ate["f32"]) <= 1.0


def test_room_hd_icp_then_refinement(kf):
    """1024x1024x768 voxels at 4 mm (FP16 + u16), 200 frames at 1280x720
    (f = 1000), 12 spheres and 8 boxes: ICP alone, then ICP followed by 20
    Gauss-Newton steps through the differentiable raycast."""
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM_HD_F16)
    n = 200
    frames, gt = kf.stream(kf.preset_scene(kf.PRESET_ROOM_HD_F16), K, H, W, 0, n, threads=16)
    res = {}
    for steps in (0, 20):
        say(f"[gpu] 1280x720 into 1024x1024x768 @4mm F16: {n} frames, refine_steps={steps}")
        v, p, _, _, _ = kf.preset(kf.PRESET_ROOM_HD_F16)
        p.refine_steps = steps
        p.stage_timing = 1
        import time
        with kf.KinfuContext(v, p, H, W) as c:
            c.stream_upload(frames)
            c.synchronize()
            t0 = time.perf_counter()
            c.pipeline_run(K, 0, n)
            c.synchronize()
            el = time.perf_counter() - t0
            st = c.stage_times()
            res[steps] = (ate_mm(c.poses(n), gt), n / el, st)
    gain = res[0][0] - res[20][0]
    say(f"[gpu] 1280x720 HD: ICP only ATE {res[0][0]:.3f} mm at {res[0][1]:.1f} frames/s; "
        f"ICP + 20 refinement steps ATE {res[20][0]:.3f} mm at {res[20][1]:.1f} frames/s; "
        f"refine {res[20][2]['refine'][0] / max(res[20][2]['refine'][1], 1):.3f} ms/frame")
    assert res[20][0] < res[0][0]
    assert res[20][0] < 2.0
