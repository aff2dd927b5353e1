"""Whole-stream parity: the GPU pipeline's volume and model maps against the
oracle replayed at the GPU's own poses.

The integration kernels cull groups and bricks by FP64 corner tests, margins
and per-tile depth bands before any voxel is touched.  Each cull is argued
conservative; these tests check that argument at every pose a long stream
visits.  For each frame k the GPU pipeline records the pose it integrated
and raycast at (kf_download_poses).  The oracle then takes the same frames
through its own range gate and bilateral filter and integrates each one at
exactly that pose (projective_tsdf.cu:5-140).  Finally it raycasts the
replayed volume at the checkpoint frames' poses (voxel_traversal.cu:75-424).
ICP differences cannot enter: both sides use the GPU's poses.

Bar: TSDF and weight volumes bit-exact at every checkpoint (NaN = never
observed on both sides).  Model maps and hit-voxel indices of the checkpoint
frame are bit-exact as well.

* room stream (640x480, 360x360x300 @10 mm, FP32, trunc 30 mm, bilateral
  5x5): all 300 frames, checkpoints after frames 150 and 300;
* the reference's compiled-in configuration (KF_PRESET_REFERENCE: K
  524.639 / 523.503 / 316.625 / 256.2318, trunc 300 mm, band [-30, 300],
  depth [300, 2000] mm, ICP 20 mm / 30 deg, bilateral off;
  kinfu_pipeline.h:42-62, kinfu_pipeline.cpp:14-24, 141-145,
  sensor_pose_estimation.cu:48, 69): 30 frames of the room stream rendered
  with that K, checkpoints after frames 10, 20 and 30.  At 300 mm / 10 mm the
  raycast's first pass runs ceil((1.3 * 30 + 2) / 8) = 6 batches of 8 steps,
  its multi-batch instantiation;
* the 5 mm precision workload (720x720x600 @5 mm = 311M voxels, FP16 TSDF +
  u16 weight, trunc 15 mm, band [-15, 150], compensated ICP sums; the stream
  with half the sensor noise): all 300 frames, checkpoints after frames 150
  and 300.  Bit-exact here means the same half and u16 bits;
* the 1280x720 workload (1024x1024x768 @4 mm = 805M voxels, FP16 + u16,
  fx = fy = 1000, trunc 20 mm, 12 spheres and 8 boxes, raycast skip rounds
  on): all 200 frames, checkpoints after frames 100 and 200.  The oracle has
  no skip rounds, so this also checks that they leave the maps unchanged
  over a whole stream.
"""
import time

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

MAP_KEYS = ("x", "y", "z", "nx", "ny", "nz")


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    same = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b
    bad = int((~same).sum())
    assert bad == 0, f"{what}: {bad} of {a.size} elements differ"


def gated(oracle, frame, prm):
    d = oracle.range_gate(frame, prm.depth_min_mm, prm.depth_max_mm)
    if prm.bilateral:
        d = oracle.bilateral(d, prm.bilateral_radius, prm.sigma_s_px, prm.sigma_r_mm)
    return d


def run_gpu(kf, vol, prm, K, H, W, frames, checkpoints):
    """Runs the pipeline over all frames; after frame c - 1 for each c in
    checkpoints, downloads the volume and the model maps."""
    snaps = {}
    with kf.KinfuContext(vol, prm, H, W) as ctx:
        ctx.stream_upload(frames)
        done = 0
        for c in checkpoints:
            say(f"[gpu] pipeline frames {done}..{c - 1}")
            if done == 0:
                ctx.pipeline_run(K, 0, c)
            else:
                for f in range(done, c):
                    ctx.pipeline_incoming(K, f)
            done = c
            t, w = ctx.volume()
            snaps[c] = (t, w, ctx.model_maps())
        poses = ctx.poses(len(frames))
    return poses, snaps


def replay(oracle, vol, prm, K, frames, poses, snaps, what, kind="f32"):
    """Integrates every frame at the GPU's pose and compares each checkpoint.

    kind "f16": the oracle keeps half bits and u16 counts; the GPU's
    reference-layout download widens half to float exactly and gives NaN for
    a weight of 0, so the oracle's volume is widened the same way."""
    t, w = oracle.new_volume(vol.dims, kind)
    t0 = time.perf_counter()
    last = max(snaps)
    for k in range(last):
        d = gated(oracle, frames[k], prm)
        oracle.integrate(d, K, poses[k], vol, prm, k == 0, t, w)
        if (k + 1) % 50 == 0 and k + 1 not in snaps:
            say(f"[gpu] {what}: oracle replay at frame {k + 1} "
                f"({time.perf_counter() - t0:.1f} s)")
        if k + 1 in snaps:
            gt_t, gt_w, maps = snaps[k + 1]
            say(f"[gpu] {what}: checkpoint after frame {k + 1} "
                f"({time.perf_counter() - t0:.1f} s of oracle replay)")
            if kind == "f16":
                assert_same(gt_w, np.where(w == 0, np.float32(np.nan), w.astype(np.float32)),
                            f"{what} u16 weight after frame {k + 1}")
                assert_same(gt_t, oracle.f16_bits_to_f32(t), f"{what} half tsdf after frame {k + 1}")
            else:
                assert_same(gt_w, w, f"{what} weight after frame {k + 1}")
                assert_same(gt_t, t, f"{what} tsdf after frame {k + 1}")
            want = oracle.raycast(d, K, poses[k], t, vol, prm.truncation_mm,
                                  prm.raycast_fix_bounds)
            assert_same(maps["hit"], want["hit"], f"{what} hit index at frame {k}")
            for key in MAP_KEYS:
                assert_same(maps[key], want[key], f"{what} model {key} at frame {k}")
            assert (want["hit"] >= 0).mean() > 0.3, "too few raycast hits to be a test"
    return t, w


def test_room_stream_300_frames_volume_and_maps_bitexact(kf, oracle):
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    n = 300
    frames, gt = kf.stream(kf.preset_scene(kf.PRESET_ROOM), K, H, W, 0, n, threads=16)
    say(f"[gpu] room stream: {n} frames 640x480 into 360x360x300 @10mm (F32)")
    poses, snaps = run_gpu(kf, vol, prm, K, H, W, frames, (150, 300))
    err = np.linalg.norm(poses[:, :3, 3] - gt[:, :3, 3], axis=1)
    assert err.max() < 60.0, f"tracking lost (max {err.max():.1f} mm)"
    t, w = replay(oracle, vol, prm, K, frames, poses, snaps, "room")
    # the stream keeps adding observations: the mid-stream volume is a
    # strict subset of the final one
    assert np.isfinite(snaps[150][1]).sum() < np.isfinite(w).sum()


def test_reference_preset_30_frames_volume_and_maps_bitexact(kf, oracle):
    vol, prm, K, H, W = kf.preset(kf.PRESET_REFERENCE)
    assert (prm.truncation_mm, prm.sdf_min_mm, prm.sdf_max_mm) == (300.0, -30.0, 300.0)
    assert prm.bilateral == 0 and (prm.icp_dist_mm, prm.icp_angle_deg) == (20.0, 30.0)
    n = 30
    frames, gt = kf.stream(kf.preset_scene(kf.PRESET_ROOM), K, H, W, 0, n, threads=16)
    say(f"[gpu] reference preset: {n} frames 640x480, trunc 300 mm, 6-batch first pass")
    poses, snaps = run_gpu(kf, vol, prm, K, H, W, frames, (10, 20, 30))
    err = np.linalg.norm(poses[:, :3, 3] - gt[:, :3, 3], axis=1)
    assert err.max() < 60.0, f"tracking lost (max {err.max():.1f} mm)"
    t, w = replay(oracle, vol, prm, K, frames, poses, snaps, "reference preset")
    # the wide band updates voxels up to 300 mm in front of the surface
    assert np.isfinite(w).sum() > 200_000


def observed(w):
    return int((w > 0).sum()) if w.dtype == np.uint16 else int(np.isfinite(w).sum())


@pytest.mark.parametrize("preset,n,checkpoints,noise", [
    ("PRESET_ROOM_5MM_F16", 300, (150, 300), 0.5),
    ("PRESET_ROOM_HD_F16", 200, (100, 200), 1.0),
], ids=["5mm_f16_300", "hd_f16_200"])
def test_f16_workload_stream_volume_and_maps_bitexact(kf, oracle, preset, n, checkpoints, noise):
    code = getattr(kf, preset)
    vol, prm, K, H, W = kf.preset(code)
    assert vol.precision == kf.KF_TSDF_F16
    frames, gt = kf.stream(kf.preset_scene(code), K, H, W, 0, n, noise_scale=noise, threads=16)
    say(f"[gpu] {preset}: {n} frames {W}x{H} into {list(vol.dims)} @{vol.voxel_mm}mm (F16)")
    poses, snaps = run_gpu(kf, vol, prm, K, H, W, frames, checkpoints)
    err = np.linalg.norm(poses[:, :3, 3] - gt[:, :3, 3], axis=1)
    assert err.max() < 60.0, f"tracking lost (max {err.max():.1f} mm)"
    t, w = replay(oracle, vol, prm, K, frames, poses, snaps, preset, kind="f16")
    assert np.isfinite(snaps[checkpoints[0]][1]).sum() < observed(w)
