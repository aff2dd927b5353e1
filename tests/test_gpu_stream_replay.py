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
  its multi-batch instantiation.
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


def replay(oracle, vol, prm, K, frames, poses, snaps, what):
    """Integrates every frame at the GPU's pose and compares each checkpoint."""
    t, w = oracle.new_volume(vol.dims, "f32")
    t0 = time.perf_counter()
    last = max(snaps)
    for k in range(last):
        d = gated(oracle, frames[k], prm)
        oracle.integrate(d, K, poses[k], vol, prm, k == 0, t, w)
        if k + 1 in snaps:
            gt_t, gt_w, maps = snaps[k + 1]
            say(f"[gpu] {what}: checkpoint after frame {k + 1} "
                f"({time.perf_counter() - t0:.1f} s of oracle replay)")
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
