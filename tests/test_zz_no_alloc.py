"""kf_estimate_pose_static, the ICP entry point that allocates no device
memory, against kf_estimate_pose on a context with the same inputs.

Two calls in a row, the second with a smaller image (640x480, then 320x240),
so that the second runs on the static scratch the first one left behind.
Both run the same kernels on the same grid, so the poses must be bit-exact;
the first is also checked against the oracle's ICP to the pose tolerance of
the other ICP tests.  Device arrays are torch tensors (plumbing only); the
calls go through the C ABI.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def gated(oracle, frame, prm):
    d = oracle.range_gate(frame, prm.depth_min_mm, prm.depth_max_mm)
    if prm.bilateral:
        d = oracle.bilateral(d, prm.bilateral_radius, prm.sigma_s_px, prm.sigma_r_mm)
    return d


def icp_case(kf, oracle, K, H, W):
    """Frame 2's vertex maps against the raycast of frames 0-1 at gt[1]."""
    vol, prm, _, _, _ = kf.preset(kf.PRESET_ROOM)
    frames, gt = kf.stream(kf.preset_scene(kf.PRESET_ROOM), K, H, W, 0, 3, threads=16)
    t, w = oracle.new_volume(vol.dims)
    for i in (0, 1):
        oracle.integrate(gated(oracle, frames[i], prm), K, gt[i], vol, prm, i == 0, t, w)
    d1 = gated(oracle, frames[1], prm)
    model = oracle.raycast(d1, K, gt[1], t, vol, prm.truncation_mm, 0)
    d2 = gated(oracle, frames[2], prm)
    fr = oracle.vertex_normals(d2, kf.mat3_inverse(K).ravel())
    fr["vz"] = d2
    return vol, prm, gt, fr, model


def packed(*planes):
    return np.ascontiguousarray(np.stack(planes, axis=-1).astype(np.float32))


def static_pose(kf, torch, prm, K, H, W, fr, model, T_prev):
    vz_used = np.where(fr["mask"] != 0, fr["vz"], 0).astype(np.float32)
    host = [packed(fr["vx"], fr["vy"], vz_used), packed(fr["nx"], fr["ny"], fr["nz"]),
            packed(model["x"], model["y"], model["z"]),
            packed(model["nx"], model["ny"], model["nz"])]
    dev = [torch.from_numpy(a).cuda() for a in host]
    T_out = torch.full((16,), float("nan"), dtype=torch.float32, device="cuda")
    Kc = np.ascontiguousarray(K, np.float32)
    Tp = np.ascontiguousarray(T_prev, np.float32).reshape(16)
    stream = torch.cuda.current_stream().cuda_stream
    rc = kf.lib.kf_estimate_pose_static(
        *(C.c_void_p(a.data_ptr()) for a in dev), H, W,
        Kc.ctypes.data_as(C.POINTER(C.c_float)), C.byref(prm),
        Tp.ctypes.data_as(C.POINTER(C.c_float)), C.c_void_p(T_out.data_ptr()),
        C.c_void_p(stream))
    assert rc == 0, kf.lib.kf_last_error().decode()
    torch.cuda.synchronize()
    return T_out.cpu().numpy().reshape(4, 4)


def context_pose(kf, vol, prm, K, H, W, fr, model, T_prev):
    with kf.KinfuContext(vol, prm, H, W) as c:
        c.upload_vertex_maps(fr["vx"], fr["vy"], fr["vz"], fr["nx"], fr["ny"], fr["nz"],
                             fr["mask"])
        c.upload_model_maps(*(model[k] for k in ("x", "y", "z", "nx", "ny", "nz")))
        return c.estimate_pose(K, T_prev)


def test_static_icp_equals_context_icp_two_calls_second_smaller(kf, oracle):
    import torch
    K1 = np.array([525.0, 0, 319.5, 0, 525.0, 239.5, 0, 0, 1], np.float32)
    K2 = np.array([262.5, 0, 159.5, 0, 262.5, 119.5, 0, 0, 1], np.float32)
    results = []
    for (H, W, K) in ((480, 640, K1), (240, 320, K2)):
        vol, prm, gt, fr, model = icp_case(kf, oracle, K, H, W)
        say(f"[gpu] kf_estimate_pose_static {W}x{H}")
        T_static = static_pose(kf, torch, prm, K, H, W, fr, model, gt[1])
        say(f"[gpu] kf_estimate_pose (context) {W}x{H}")
        T_ctx = context_pose(kf, vol, prm, K, H, W, fr, model, gt[1])
        assert np.isfinite(T_static).all()
        assert np.array_equal(T_static, T_ctx), (T_static, T_ctx)
        assert not np.array_equal(T_static, gt[1])  # the iterations moved the pose
        results.append((prm, K, gt, fr, model, T_static))
    # the first case against the oracle (the other ICP tests' tolerance)
    prm, K, gt, fr, model, T = results[0]
    want = oracle.icp(fr, model, K, gt[1], prm.icp_iterations, prm.icp_dist_mm,
                      prm.icp_angle_deg, reduce_mode=1)
    assert np.abs(T[:3, 3] - want["T"][:3, 3]).max() < 1e-2
    assert np.abs(T[:3, :3] - want["T"][:3, :3]).max() < 1e-5


def test_static_icp_rejects_bad_arguments(kf):
    """A null input, and an image beyond the scratch (2^21 pixels), are
    rejected before any GPU work.  The large case still passes real device
    arrays of that size, so that no kernel could read past them."""
    import torch
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    Kc = np.ascontiguousarray(K, np.float32)
    T = np.eye(4, dtype=np.float32)
    f = C.POINTER(C.c_float)
    rc = kf.lib.kf_estimate_pose_static(None, None, None, None, H, W, Kc.ctypes.data_as(f),
                                        C.byref(prm), T.ctypes.data_as(f), None, None)
    assert rc == -1
    Hb, Wb = 1024, 2048
    maps = [torch.zeros(Hb * Wb * 3, dtype=torch.float32, device="cuda") for _ in range(4)]
    T_out = torch.zeros(16, dtype=torch.float32, device="cuda")
    rc = kf.lib.kf_estimate_pose_static(*(C.c_void_p(m.data_ptr()) for m in maps), Hb, Wb,
                                        Kc.ctypes.data_as(f), C.byref(prm),
                                        T.ctypes.data_as(f), C.c_void_p(T_out.data_ptr()), None)
    torch.cuda.synchronize()
    assert rc == -1 and b"2^20" in kf.lib.kf_last_error()
