"""Full-pipeline throughput on one MI355X.

Workload: a synthetic 640x480 uint16 depth stream (box room, 6 spheres,
4 boxes, Kinect axial noise, 2% holes) into a 360x360x300 volume at 10 mm
(38.88 M voxels), truncation 30 mm.  One step is one IncomingFrameProcessing
frame: depth gate -> 5x5 bilateral -> 10 point-to-plane ICP iterations (the
first computes the vertex and normal maps) -> TSDF integration -> raycast,
all enqueued on the device with no host round trip.  Frames are resident in HBM before the
timed region.

Runs, each on a fresh context over the same frames:
  1. FP32 volume, timed                         -> value, ms_per_step
  2. FP16 TSDF + u16 weight, timed              -> fp16
  3. FP32 with per-stage hipEvents, including the ICP iteration kernels
     without the finishing kernel              -> stage_ms, roofline
  4. FP16 with per-stage hipEvents              -> fp16.stage_ms
  5. FP32 with the raycast work counters on, and the bricks each integration
     kept read after every frame               -> algorithmic bytes
  6. FP32 fed from host memory frame by frame (kf_upload_depth_u16 from host
     then kf_pipeline_incoming on that depth, as the reference loads each
     frame), timed                             -> host_fed
  6b. FP32 with the opt-in ICP early stop, timed, then once more reading the
     iterations each frame ran              -> icp_early_stop
  7. FP32 with the trilinear raycast, timed    -> trilinear.value, ATE
  8. the same with stage events, then on the final state, after one untimed
     pair each, 20 raycast + pose-only adjoint pairs, 10 raycast +
     pose-and-TSDF adjoint pairs (the adjoint from the raycast's saved
     crossings) and 10 such pairs with the depth uploaded again in between
     (no saved crossings: the adjoint first runs its own trilinear raycast)
                                                -> adjoint cost
and the CPU oracle on a bounded sample of the same workload.

Prints one JSON line.  The roofline object is for the dominant kernel,
k_icp_iter (10 launches per frame): its algorithmic bytes per launch (every
pixel's vertex, normal and mask read, and the model vertex and normal
gathered: 6 x 4 + 1 + 6 x 4 = 49 B per pixel; the first launch of a frame
reads the depth and writes those maps instead: 4 + 25 + 24 = 53 B; 49.4 B
per pixel on average over the 10) over its mean launch time,
measured with one hipEvent pair around a frame's 10 launches on the context
stream.  Its
traffic is the HBM bytes per dispatch that rocprofv3's FETCH_SIZE and
WRITE_SIZE report (profiles/icp_room_640x480_pmc.txt), when that summary is
present.
"""
import argparse
import json
import os
import re
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s full pipeline 640x480 38.9M voxels, fp16 vs fp32 TSDF at ATE<=1mm"
HBM_PEAK_GBS = 8000.0
ICP_PMC_SUMMARY = os.path.join(ROOT, "profiles", "icp_room_640x480_pmc.txt")


def ensure_built():
    lib = os.path.join(ROOT, "kinect_fusion_slam_amd", "libkinfu_amd.so")
    orc = os.path.join(ROOT, "oracle", "libkf_oracle.so")
    if not (os.path.exists(lib) and os.path.exists(orc)):
        import __graft_entry__
        __graft_entry__.build()


def ate_mm(poses, gt):
    err = np.linalg.norm(poses[:, :3, 3].astype(np.float64) - gt[:, :3, 3], axis=1)
    return float(np.sqrt((err ** 2).mean()))


def icp_traffic():
    """HBM bytes per k_icp_iter dispatch from the committed rocprofv3 summary
    (its 'bench-traffic' line: FETCH_SIZE and WRITE_SIZE already scaled by the
    copy calibration), or None."""
    try:
        text = open(ICP_PMC_SUMMARY).read()
    except OSError:
        return None
    m = re.search(r"^bench-traffic k_icp_iter read_bytes=(\d+) write_bytes=(\d+)", text, re.M)
    return (int(m.group(1)), int(m.group(2))) if m else None


# opt-in ICP early stop (no reference counterpart): rotation (rad) and
# translation (mm) thresholds of the increment after which the ICP ends
ICP_STOP = (5e-6, 5e-3)


def new_context(kf, precision, timing=False, mode=0, icp_stop=None):
    vol, prm, _, H, W = kf.preset(kf.PRESET_ROOM)
    vol.precision = precision
    prm.stage_timing = 1 if timing else 0
    prm.raycast_mode = mode
    if icp_stop:
        prm.icp_stop_rot, prm.icp_stop_trans_mm = icp_stop
    return kf.KinfuContext(vol, prm, H, W)


def run_gpu(kf, precision, frames, K, warmup, steps, timing=False, count=False, mode=0,
            adjoint_reps=0, icp_stop=None):
    """One pass over frames[0 .. warmup+steps]; times frames warmup+1 ...
    count: raycast work counters on, and the integration's kept bricks read
    after every timed frame (a host round trip per frame: not timed).  With
    adjoint_reps, then times that many (raycast, pose-only raycast adjoint)
    pairs and half as many (raycast, adjoint with dL/dTSDF) pairs on the
    final volume, pose and depth, each after one untimed pair."""
    ctx = new_context(kf, precision, timing, mode, icp_stop)
    _, _, _, H, W = kf.preset(kf.PRESET_ROOM)
    try:
        ctx.stream_upload(frames)
        ctx.pipeline_initial(K, 0)
        for f in range(1, warmup + 1):
            ctx.pipeline_incoming(K, f)
        ctx.synchronize()
        ctx.reset_stage_times()
        work = None
        if count:
            ctx.count_raycast_work(True)
            bricks = groups = 0
            for f in range(warmup + 1, warmup + 1 + steps):
                ctx.pipeline_incoming(K, f)
                g, b = ctx.integration_work()
                groups += g
                bricks += b
            visited, hits = ctx.raycast_work()
            work = {"visited": visited, "hits": hits, "groups": groups, "bricks": bricks}
            elapsed = None
        else:
            t0 = time.perf_counter()
            for f in range(warmup + 1, warmup + 1 + steps):
                ctx.pipeline_incoming(K, f)
            ctx.synchronize()
            elapsed = time.perf_counter() - t0
        stages = ctx.stage_times()
        poses = ctx.poses(warmup + 1 + steps)
        if adjoint_reps:
            T = ctx.pose()
            d = ctx.depth()
            ones = np.ones((H, W), np.float32)
            work = {}
            for key in ("pose_only", "full", "full_no_saved"):
                # one untimed pair first, then the timed pairs (a full adjoint
                # also copies dL/dTSDF to the host, outside the timed stage);
                # full_no_saved uploads the same depth again between the two,
                # which drops the raycast's saved crossings
                full = key != "pose_only"
                for rep in range(1 + (adjoint_reps if not full else max(1, adjoint_reps // 2))):
                    if rep == 1:
                        ctx.reset_stage_times()
                    ctx.raycast(K, T)
                    if key == "full_no_saved":
                        ctx.upload_depth_f32(d)
                    ctx.raycast_adjoint(K, T, ones, want_tsdf=full)
                adj = ctx.stage_times()
                if key != "full_no_saved":
                    work["raycast"] = adj["raycast"][0] / max(adj["raycast"][1], 1)
                work[key] = adj["raycast_adjoint"][0] / max(adj["raycast_adjoint"][1], 1)
    finally:
        ctx.close()
    return elapsed, stages, poses, work


def run_host_fed(kf, frames, K, warmup, steps):
    """FP32, each frame handed over from host memory: kf_upload_depth_u16
    (pageable host array, KF_MEM_HOST) then kf_pipeline_incoming on that
    depth (KF_FRAME_CURRENT), as IncomingFrameProcessing loads and uploads
    every frame (kinfu_pipeline.cpp:128-131).  Returns the elapsed time of
    the timed frames."""
    ctx = new_context(kf, kf.KF_TSDF_F32)
    try:
        ctx.upload_depth_u16(frames[0])
        ctx.pipeline_initial(K, kf.KF_FRAME_CURRENT)
        for f in range(1, warmup + 1):
            ctx.upload_depth_u16(frames[f])
            ctx.pipeline_incoming(K, kf.KF_FRAME_CURRENT)
        ctx.synchronize()
        t0 = time.perf_counter()
        for f in range(warmup + 1, warmup + 1 + steps):
            ctx.upload_depth_u16(frames[f])
            ctx.pipeline_incoming(K, kf.KF_FRAME_CURRENT)
        ctx.synchronize()
        return time.perf_counter() - t0
    finally:
        ctx.close()


def icp_iterations_per_frame(kf, frames, K, warmup, steps, icp_stop):
    """ICP iterations run by each of the timed frames with the early stop
    (the per-iteration record read after every frame; not timed)."""
    ctx = new_context(kf, kf.KF_TSDF_F32, icp_stop=icp_stop)
    try:
        ctx.stream_upload(frames)
        ctx.pipeline_initial(K, 0)
        for f in range(1, warmup + 1):
            ctx.pipeline_incoming(K, f)
        its = []
        for f in range(warmup + 1, warmup + 1 + steps):
            ctx.pipeline_incoming(K, f)
            its.append(ctx.icp_history()["T"].shape[0])
        return np.array(its)
    finally:
        ctx.close()


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "model": model}


def cpu_baseline(kf, repeats=3):
    """The oracle's frame sequence (C++/OpenMP restatement of the reference
    kernels) on config 0: the first 10 frames of the same stream into the
    same volume, in the -O3 x86-64-v3 timing build, at all OpenMP threads and
    at one thread; median of `repeats` runs each."""
    from oracle import kf_oracle as oracle
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM_CPU)
    scene = kf.preset_scene(kf.PRESET_ROOM_CPU)
    n = 10
    frames, _ = kf.stream(scene, K, H, W, 0, n)
    n_threads = oracle.max_threads()

    def timed(threads):
        oracle.set_threads(threads)
        times, stage = [], np.zeros(6)
        for _ in range(repeats):
            ms = np.zeros(6)
            t0 = time.perf_counter()
            oracle.pipeline(frames, K, vol, prm, reduce_mode=0, stage_ms=ms, timing_build=True)
            times.append(time.perf_counter() - t0)
            stage = ms
        return statistics.median(times), stage

    med_n, stage_n = timed(n_threads)
    med_1, _ = timed(1)
    oracle.set_threads(n_threads)
    names = ("depth_gate", "bilateral", "vertex", "icp", "integrate", "raycast")
    return {"value": round(n / med_n, 3), "unit": "frames/s", "cores": n_threads,
            "kind": "port",
            "sample": f"CPU oracle (C++/OpenMP restatement of the reference kernels, FP32): "
                      f"the stream's first {n} frames (config 0: 640x480 into 360x360x300 "
                      f"@10mm, full frame sequence incl. the initial frame), median of "
                      f"{repeats} runs: {med_n:.2f} s at {n_threads} threads, {med_1:.2f} s "
                      f"at 1 thread",
            "single_thread_value": round(n / med_1, 3),
            "flags": "g++ -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off",
            "stage_ms_per_frame": {k: round(v / n, 2) for k, v in zip(names, stage_n)},
            "host": host_info()}


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
    elh = run_host_fed(kf, frames, K, args.warmup, args.steps)
    print(f"[bench] host-fed fp32 {args.steps / elh:.1f} frames/s", file=sys.stderr, flush=True)
    els, _, pstop, _ = run_gpu(kf, kf.KF_TSDF_F32, frames, K, args.warmup, args.steps,
                               icp_stop=ICP_STOP)
    its = icp_iterations_per_frame(kf, frames, K, args.warmup, args.steps, ICP_STOP)
    print(f"[bench] fp32 with the ICP early stop {args.steps / els:.1f} frames/s, "
          f"{its.mean():.2f} iterations per frame", file=sys.stderr, flush=True)
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
    steps = args.steps
    # dominant kernel: one ICP iteration (10 per frame).  Iterations 1..9
    # read the frame's vertex, normal and mask and gather the model's vertex
    # and normal (49 B per pixel); iteration 0 reads the depth instead and
    # writes the frame's maps (4 + 25 + 24 = 53 B per pixel)
    icp_frame_bytes = npx * ((prm.icp_iterations - 1) * (6 * 4 + 1 + 6 * 4)
                             + (4 + 6 * 4 + 1 + 6 * 4))
    icp_iter_bytes = icp_frame_bytes // prm.icp_iterations
    # the icp_iter stage brackets the iteration kernels of one frame's ICP
    # with one event pair (events between the kernels would add their own
    # gaps): its mean over frames divided by the iteration count
    icp_iter_s = avg_s(st32, "icp_iter") / prm.icp_iterations
    achieved = icp_iter_bytes / icp_iter_s / 1e9
    traffic = icp_traffic()
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": (traffic[0] + traffic[1]) if traffic else None,
                "kernel": "k_icp_iter (FP32 sums, 640x480, 10 launches per frame)",
                "algorithmic_bytes": icp_iter_bytes, "avg_launch_us": round(icp_iter_s * 1e6, 3),
                "launches_per_frame": prm.icp_iterations}
    if traffic:
        roofline["traffic_read_write"] = list(traffic)

    # the other stages against the same peak (algorithmic bytes per launch)
    visited, hits = work["visited"] / steps, work["hits"] / steps
    bricks = work["bricks"] / steps
    ray_bytes = visited * 4 + hits * 13 * 4 + 8 * npx * 4
    other = {
        "raycast": {"bytes": int(ray_bytes), "us": round(avg_s(st32, "raycast") * 1e6, 2),
                    "voxels_visited_per_frame": int(visited), "hits_per_frame": int(hits),
                    "note": "k_raycast + k_raycast_rest: voxels visited x 4 B + hits x 13 "
                            "taps x 4 B + depth in and 6 maps + hit index out"},
        "integrate": {"bytes": int(bricks * 512 * 16),
                      "us": round(avg_s(st32, "integrate") * 1e6, 2),
                      "bricks_kept_per_frame": round(bricks, 1),
                      "groups_kept_per_frame": round(work["groups"] / steps, 1),
                      "note": "k_integ_cull + k_integrate: TSDF and weight (4 + 4 B) read and "
                              "written for every voxel of the bricks that pass the culls"},
        "icp_stage": {"bytes": icp_frame_bytes, "us": round(avg_s(st32, "icp") * 1e6, 2),
                      "note": "vertex and normal maps computed by the first iteration"},
        "bilateral": {"bytes": 2 * npx * 4, "us": round(avg_s(st32, "bilateral") * 1e6, 2)},
    }
    for v in other.values():
        v["GBs"] = round(v["bytes"] / (v["us"] * 1e-6) / 1e9, 1) if v["us"] else None
        v["frac"] = round(v["GBs"] / HBM_PEAK_GBS, 5) if v["GBs"] else None

    a32, a16 = ate_mm(p32, gt), ate_mm(p16, gt)
    out = {
        "metric": METRIC,
        "value": round(steps / el32, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(el32 / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "640x480 synthetic room stream into 360x360x300 voxels @10mm "
                               "(38.88M), FP32 TSDF+weight, trunc 30mm, bilateral 5x5, "
                               "10 ICP iterations, raycast per frame",
                   "frames_timed": steps, "image": [H, W], "volume": list(vol.dims),
                   "parallelism": "single GPU"},
        "fp16": {"value": round(steps / el16, 2), "unit": "frames/s",
                 "ratio_vs_fp32": round(el32 / el16, 3), "stage_ms": per_stage(st16)},
        "ate_mm": {"fp32": round(a32, 3), "fp16": round(a16, 3),
                   "fp16_minus_fp32": round(a16 - a32, 3)},
        "stage_ms": per_stage(st32),
        "icp_early_stop": {
            "value": round(steps / els, 2), "unit": "frames/s",
            "ate_mm": round(ate_mm(pstop, gt), 3),
            "iterations_per_frame": {"mean": round(float(its.mean()), 3),
                                     "min": int(its.min()), "max": int(its.max())},
            "stop_rot_rad": ICP_STOP[0], "stop_trans_mm": ICP_STOP[1],
            "note": "opt-in (no reference counterpart): FP32, the ICP ends after an increment "
                    "below both thresholds (at most the reference's 10 iterations); not the "
                    "metric, which runs the fixed 10"},
        "host_fed": {"value": round(steps / elh, 2), "unit": "frames/s",
                     "note": "FP32, each frame uploaded from pageable host memory "
                             "(kf_upload_depth_u16) before kf_pipeline_incoming on it, as the "
                             "reference loads every frame; PCIe-inclusive, not the metric"},
        "trilinear": {
            "note": "same stream and precision with the trilinear zero-crossing raycast "
                    "(differentiable mode) instead of the reference interpolation",
            "value": round(steps / elt, 2), "unit": "frames/s",
            "ate_mm": round(ate_mm(ptri, gt), 3), "stage_ms": per_stage(sttri),
            "raycast_forward_ms": round(adj["raycast"], 4),
            "adjoint_pose_only_ms": round(adj["pose_only"], 4),
            "adjoint_pose_and_tsdf_ms": round(adj["full"], 4),
            "adjoint_vs_forward": round(adj["full"] / adj["raycast"], 3),
            "adjoint_pose_only_vs_forward": round(adj["pose_only"] / adj["raycast"], 3),
            "adjoint_note": "the adjoint right after the raycast at the same inputs starts from "
                            "the raycast's saved per-pixel crossings; with an input changed in "
                            "between it first runs the trilinear raycast into buffers of its "
                            "own and starts from those crossings (no_saved figures)",
            "adjoint_no_saved_pose_and_tsdf_ms": round(adj["full_no_saved"], 4),
            "adjoint_no_saved_vs_forward": round(adj["full_no_saved"] / adj["raycast"], 3)},
        "roofline": roofline,
        "stage_roofline": other,
    }
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(kf)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
