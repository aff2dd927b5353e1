"""Full-pipeline throughput on one MI355X.

Workload: a synthetic 640x480 uint16 depth stream (box room, 6 spheres,
4 boxes, Kinect axial noise, 2% holes) into a 360x360x300 volume at 10 mm
(38.88 M voxels), truncation 30 mm.  One step is one IncomingFrameProcessing
frame: depth gate -> 5x5 bilateral -> vertex and normal maps -> 10
point-to-plane ICP iterations -> TSDF integration -> raycast, all enqueued on
the device with no host round trip.  Frames are resident in HBM before the
timed region.

Runs, each on a fresh context over the same frames:
  1. FP32 volume, timed                         -> value, ms_per_step
  2. FP16 TSDF + u16 weight, timed              -> fp16
  3. FP32 with per-stage hipEvents               -> stage_ms, roofline
  4. FP32 with the raycast work counters on      -> algorithmic bytes
  5. FP32 with the trilinear raycast, timed      -> trilinear.value, ATE
  6. the same with stage events, then on the final state, after one untimed
     pair each, 20 raycast + pose-only adjoint pairs, 10 raycast +
     pose-and-TSDF adjoint pairs (the adjoint from the raycast's saved
     crossings) and 10 such pairs with the depth uploaded again in between
     (the adjoint walks every ray)               -> adjoint cost
and the CPU oracle on a bounded sample of the same workload.

Prints one JSON line.  The roofline object is for the raycast stage (the
largest share of a frame): k_raycast + k_raycast_rest, timed with hipEvents
on the context stream, against its algorithmic bytes (voxels visited x
sizeof(TSDF) + hits x 13 taps x sizeof(TSDF) + depth in + 6 maps and the hit
index out).
"""
import argparse
import json
import statistics
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s full pipeline 640x480 38.9M voxels, fp16 vs fp32 TSDF at ATE<=1mm"
HBM_PEAK_GBS = 8000.0

# HBM bytes per frame of the raycast stage (k_raycast + k_raycast_rest), from
# two rocprofv3 --pmc passes over `python3 bench.py --steps 30 --warmup 5
# --no-cpu` (FETCH_SIZE x 2 + WRITE_SIZE x 1; the factors are what the same
# counters read for a device-to-device copy of 256 MiB).  The passes and the
# per-kernel figures are in profiles/raycast_room_640x480_pmc.txt.
RAYCAST_TRAFFIC_BYTES = 26294477


def ensure_built():
    lib = os.path.join(ROOT, "kinect_fusion_slam_amd", "libkinfu_amd.so")
    orc = os.path.join(ROOT, "oracle", "libkf_oracle.so")
    if not (os.path.exists(lib) and os.path.exists(orc)):
        import __graft_entry__
        __graft_entry__.build()


def ate_mm(poses, gt):
    err = np.linalg.norm(poses[:, :3, 3].astype(np.float64) - gt[:, :3, 3], axis=1)
    return float(np.sqrt((err ** 2).mean()))


def run_gpu(kf, precision, frames, K, warmup, steps, timing=False, count=False, mode=0,
            adjoint_reps=0):
    """One pass over frames[0 .. warmup+steps]; times frames warmup+1 ...
    With adjoint_reps, then times that many (raycast, pose-only raycast
    adjoint) pairs and half as many (raycast, adjoint with dL/dTSDF) pairs on
    the final volume, pose and depth, each after one untimed pair."""
    vol, prm, _, H, W = kf.preset(kf.PRESET_ROOM)
    vol.precision = precision
    prm.stage_timing = 1 if timing else 0
    prm.raycast_mode = mode
    ctx = kf.KinfuContext(vol, prm, H, W)
    try:
        ctx.stream_upload(frames)
        ctx.pipeline_initial(K, 0)
        for f in range(1, warmup + 1):
            ctx.pipeline_incoming(K, f)
        ctx.synchronize()
        ctx.reset_stage_times()
        if count:
            ctx.count_raycast_work(True)
        t0 = time.perf_counter()
        for f in range(warmup + 1, warmup + 1 + steps):
            ctx.pipeline_incoming(K, f)
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        stages = ctx.stage_times()
        work = ctx.raycast_work() if count else None
        poses = ctx.poses(warmup + 1 + steps)
        if adjoint_reps:
            T = ctx.pose()
            d = ctx.depth()
            ones = np.ones((H, W), np.float32)
            work = {}
            for key in ("pose_only", "full", "full_walk"):
                # one untimed pair first, then the timed pairs (a full adjoint
                # also copies dL/dTSDF to the host, outside the timed stage);
                # full_walk uploads the same depth again between the two,
                # which drops the raycast's saved crossings
                full = key != "pose_only"
                for rep in range(1 + (adjoint_reps if not full else max(1, adjoint_reps // 2))):
                    if rep == 1:
                        ctx.reset_stage_times()
                    ctx.raycast(K, T)
                    if key == "full_walk":
                        ctx.upload_depth_f32(d)
                    ctx.raycast_adjoint(K, T, ones, want_tsdf=full)
                adj = ctx.stage_times()
                if key != "full_walk":
                    work["raycast"] = adj["raycast"][0] / max(adj["raycast"][1], 1)
                work[key] = adj["raycast_adjoint"][0] / max(adj["raycast_adjoint"][1], 1)
    finally:
        ctx.close()
    return elapsed, stages, poses, work


def cpu_baseline(kf, repeats=3):
    from oracle import kf_oracle as oracle
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM_CPU)
    scene = kf.preset_scene(kf.PRESET_ROOM_CPU)
    n = 40
    frames, _ = kf.stream(scene, K, H, W, 0, n)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        oracle.pipeline(frames, K, vol, prm, reduce_mode=0)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": round(n / med, 3), "unit": "frames/s", "cores": oracle.max_threads(),
            "kind": "port",
            "sample": f"CPU oracle (C++/OpenMP restatement of the reference kernels, g++ -O2, "
                      f"FP32): the same stream's first {n} frames, 640x480 into 360x360x300 "
                      f"@10mm, full frame sequence incl. initial frame; median of {repeats} "
                      f"runs of {med:.2f} s"}


def copy_calibration():
    """Three 256 MiB device-to-device copies: the known traffic against which
    rocprofv3's FETCH_SIZE and WRITE_SIZE counters are scaled."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=280)
    ap.add_argument("--warmup", type=int, default=19)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the known-size copies used to scale the HBM counters")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus above 1 is not "
                                   "supported (the volume is not sharded across GPUs)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    ensure_built()
    import kinect_fusion_slam_amd as kf

    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    n_frames = 1 + args.warmup + args.steps
    t0 = time.perf_counter()
    frames, gt = kf.stream(kf.preset_scene(kf.PRESET_ROOM), K, H, W, 0, n_frames, threads=16)
    print(f"[bench] generated {n_frames} frames in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)

    el32, _, p32, _ = run_gpu(kf, kf.KF_TSDF_F32, frames, K, args.warmup, args.steps)
    print(f"[bench] fp32 {args.steps / el32:.1f} frames/s", file=sys.stderr, flush=True)
    el16, _, p16, _ = run_gpu(kf, kf.KF_TSDF_F16, frames, K, args.warmup, args.steps)
    print(f"[bench] fp16 {args.steps / el16:.1f} frames/s", file=sys.stderr, flush=True)
    _, st32, _, _ = run_gpu(kf, kf.KF_TSDF_F32, frames, K, args.warmup, args.steps, timing=True)
    _, st16, _, _ = run_gpu(kf, kf.KF_TSDF_F16, frames, K, args.warmup, args.steps, timing=True)
    _, _, _, work = run_gpu(kf, kf.KF_TSDF_F32, frames, K, args.warmup, args.steps, count=True)
    print("[bench] stage timing and work counters done", file=sys.stderr, flush=True)
    elt, _, ptri, _ = run_gpu(kf, kf.KF_TSDF_F32, frames, K, args.warmup, args.steps,
                              mode=kf.KF_RAYCAST_TRILINEAR)
    _, sttri, _, adj = run_gpu(kf, kf.KF_TSDF_F32, frames, K, args.warmup, args.steps,
                               timing=True, mode=kf.KF_RAYCAST_TRILINEAR, adjoint_reps=20)
    print(f"[bench] trilinear {args.steps / elt:.1f} frames/s", file=sys.stderr, flush=True)

    def per_stage(st):
        return {k: round(ms / max(cnt, 1), 4) for k, (ms, cnt) in st.items() if cnt}

    def avg_s(st, stage):
        ms, cnt = st[stage]
        return ms / max(cnt, 1) / 1e3

    npx = H * W
    nvox = int(vol.dims[0]) * int(vol.dims[1]) * int(vol.dims[2])
    visited, hits = work[0] / args.steps, work[1] / args.steps
    ray_bytes = visited * 4 + hits * 13 * 4 + 8 * npx * 4
    ray_s = avg_s(st32, "raycast")
    achieved = ray_bytes / ray_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": RAYCAST_TRAFFIC_BYTES,
                "kernel": "raycast stage: k_raycast<F32> + k_raycast_rest<F32>",
                "algorithmic_bytes": int(ray_bytes), "avg_launch_us": round(ray_s * 1e6, 2),
                "voxels_visited_per_frame": int(visited), "hits_per_frame": int(hits)}

    # the other stages against the same peak (algorithmic bytes per launch)
    icp_iters = prm.icp_iterations
    other = {
        "integrate_dense_model": {
            "bytes": nvox * 16, "us": round(avg_s(st32, "integrate") * 1e6, 2),
            "note": "dense read-modify-write of every voxel (16 B); the kernel culls bricks "
                    "outside the view frustum, so it can run above this figure"},
        "icp": {"bytes": icp_iters * 13 * npx * 4, "us": round(avg_s(st32, "icp") * 1e6, 2)},
        "vertex": {"bytes": 6 * npx * 4 + npx, "us": round(avg_s(st32, "vertex") * 1e6, 2)},
        "bilateral": {"bytes": 2 * npx * 4, "us": round(avg_s(st32, "bilateral") * 1e6, 2)},
    }
    for v in other.values():
        v["GBs"] = round(v["bytes"] / (v["us"] * 1e-6) / 1e9, 1) if v["us"] else None

    a32, a16 = ate_mm(p32, gt), ate_mm(p16, gt)
    out = {
        "metric": METRIC,
        "value": round(args.steps / el32, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el32 / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "640x480 synthetic room stream into 360x360x300 voxels @10mm "
                               "(38.88M), FP32 TSDF+weight, trunc 30mm, bilateral 5x5, "
                               "10 ICP iterations, raycast per frame",
                   "frames_timed": args.steps, "image": [H, W], "volume": list(vol.dims),
                   "parallelism": "single GPU"},
        "fp16": {"value": round(args.steps / el16, 2), "unit": "frames/s",
                 "ratio_vs_fp32": round(el32 / el16, 3), "stage_ms": per_stage(st16)},
        "ate_mm": {"fp32": round(a32, 3), "fp16": round(a16, 3),
                   "fp16_minus_fp32": round(a16 - a32, 3)},
        "stage_ms": per_stage(st32),
        "trilinear": {
            "note": "same stream and precision with the trilinear zero-crossing raycast "
                    "(differentiable mode) instead of the reference interpolation",
            "value": round(args.steps / elt, 2), "unit": "frames/s",
            "ate_mm": round(ate_mm(ptri, gt), 3), "stage_ms": per_stage(sttri),
            "raycast_forward_ms": round(adj["raycast"], 4),
            "adjoint_pose_only_ms": round(adj["pose_only"], 4),
            "adjoint_pose_and_tsdf_ms": round(adj["full"], 4),
            "adjoint_vs_forward": round(adj["full"] / adj["raycast"], 3),
            "adjoint_pose_only_vs_forward": round(adj["pose_only"] / adj["raycast"], 3),
            "adjoint_note": "the adjoint right after the raycast at the same inputs starts from "
                            "the raycast's saved per-pixel crossings; with any input changed "
                            "in between it walks every ray again (walk figures)",
            "adjoint_walk_pose_and_tsdf_ms": round(adj["full_walk"], 4),
            "adjoint_walk_vs_forward": round(adj["full_walk"] / adj["raycast"], 3)},
        "roofline": roofline,
        "stage_roofline": other,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(kf)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
