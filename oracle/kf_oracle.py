"""ctypes view of libkf_oracle.so, the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product (kinect_fusion_slam_amd) never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkf_oracle.so")
# the same source built -O3 for x86-64-v3, for the benchmark's CPU timing only
FAST_LIB_PATH = os.path.join(_HERE, "libkf_oracle_fast.so")

_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)
_U16 = C.POINTER(C.c_uint16)


def _load(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built (make -C oracle)")
    lib = C.CDLL(path)
    sig = {
        "oracle_range_gate": (None, [_U16, _F, C.c_int64, C.c_float, C.c_float]),
        "oracle_spatial_kernel": (None, [C.c_int, C.c_float, _F]),
        "oracle_bilateral": (None, [_F, _F, C.c_int, C.c_int, _F, C.c_int, C.c_float]),
        "oracle_vertex_normals": (None, [_F, _F, C.c_int, C.c_int, _F, _F, _F, _F, _F, _U8]),
        "oracle_integrate_f32": (None, [_F, C.c_int, C.c_int, _F, _F, _I, _I, C.c_float,
                                        C.c_float, C.c_float, C.c_float, C.c_int, _F, _F]),
        "oracle_integrate_f64": (None, [_F, C.c_int, C.c_int, _F, _F, _I, _I, C.c_float,
                                        C.c_float, C.c_float, C.c_float, C.c_int, _D, _D]),
        "oracle_integrate_f16": (None, [_F, C.c_int, C.c_int, _F, _F, _I, _I, C.c_float,
                                        C.c_float, C.c_float, C.c_float, C.c_int, _U16, _U16]),
        "oracle_raycast_tri_f32": (None, [_F, C.c_int, C.c_int, _F, _F, _F, _I, _I, C.c_float,
                                          C.c_float, C.POINTER(_F), _I32, _I32]),
        "oracle_raycast_tri_f16": (None, [_F, C.c_int, C.c_int, _F, _F, _U16, _I, _I, C.c_float,
                                          C.c_float, C.POINTER(_F), _I32, _I32]),
        "oracle_raycast_tri_f64": (None, [_F, C.c_int, C.c_int, _F, _D, _D, _I, _I, C.c_float,
                                          C.c_float, C.POINTER(_D), _I32, _I32, _D, _D, _D,
                                          _D]),
        "oracle_raycast_f32": (None, [_F, C.c_int, C.c_int, _F, _F, _F, _I, _I, C.c_float,
                                      C.c_float, C.c_int, _F, _F, _F, _F, _F, _F, _I32, _I64]),
        "oracle_raycast_f16": (None, [_F, C.c_int, C.c_int, _F, _F, _U16, _I, _I, C.c_float,
                                      C.c_float, C.c_int, _F, _F, _F, _F, _F, _F, _I32, _I64]),
        "oracle_raycast_f64": (None, [_F, C.c_int, C.c_int, _F, _F, _D, _I, _I, C.c_float,
                                      C.c_float, C.c_int, _D, _D, _D, _D, _D, _D, _I32, _I64]),
        "oracle_icp_angle_gate": (C.c_int, [C.c_float, C.c_float]),
        "oracle_normal_flip": (C.c_int, [C.c_float]),
        "oracle_icp": (None, [_F, _F, _F, _F, _F, _F, _U8, _F, _F, _F, _F, _F, _F, C.c_int,
                              C.c_int, _F, _F, _F, C.c_int, C.c_float, C.c_float, C.c_int, _F,
                              _D, _I32, _U8, C.c_double, C.c_float, C.c_float, _I32]),
        "oracle_pipeline": (None, [_U16, C.c_int, C.c_int, C.c_int, _F, _I, _I, C.c_float,
                                   C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_int, C.c_float, C.c_float, C.c_int, C.c_float,
                                   C.c_float, C.c_int, C.c_int, _F, _F, _F, _D, C.c_float,
                                   C.c_float]),
        "oracle_f32_to_f16": (None, [_F, _U16, C.c_int64]),
        "oracle_f16_to_f32": (None, [_U16, _F, C.c_int64]),
        "oracle_max_threads": (C.c_int, []),
        "oracle_pyramid_depth": (None, [_F, C.c_int, C.c_int, C.c_float, _F]),
        "oracle_pyramid_model": (None, [C.POINTER(_F), _U8, C.c_int, C.c_int, C.POINTER(_F),
                                        _U8]),
        "oracle_pyramid_intrinsics": (None, [_F, _F]),
        "oracle_set_threads": (None, [C.c_int]),
        "oracle_jpeg_decode": (C.c_int, [_U8, C.c_int64, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), _U8]),
        "oracle_point_cloud": (None, [_F, _F, _F, _U8, C.c_int, C.c_int, _F]),
        "oracle_jpeg_coefficients": (C.c_int, [_U8, C.c_int64, C.POINTER(C.c_int16),
                                               C.POINTER(C.c_int64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()
_fast = None


def timing_lib():
    """The -O3 x86-64-v3 build of the same oracle (the CPU baseline's timing)."""
    global _fast
    if _fast is None:
        _fast = _load(FAST_LIB_PATH)
    return _fast


def _f(a):
    return a.ctypes.data_as(_F)


def _c32(a, n=None):
    a = np.ascontiguousarray(a, np.float32)
    if n is not None:
        assert a.size == n, (a.size, n)
    return a


def _ivec(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def set_threads(n):
    lib.oracle_set_threads(int(n))
    if _fast is not None:
        _fast.oracle_set_threads(int(n))


def max_threads():
    return lib.oracle_max_threads()


def range_gate(depth_u16, dmin, dmax):
    d = np.ascontiguousarray(depth_u16, np.uint16)
    out = np.zeros(d.shape, np.float32)
    lib.oracle_range_gate(d.ctypes.data_as(_U16), _f(out), d.size, dmin, dmax)
    return out


def spatial_kernel(r, sigma):
    w = np.zeros((2 * r + 1, 2 * r + 1), np.float32)
    lib.oracle_spatial_kernel(r, sigma, _f(w))
    return w


def bilateral(depth, r, sigma_s, sigma_r):
    d = _c32(depth)
    H, W = d.shape
    ws = spatial_kernel(r, sigma_s)
    out = np.zeros_like(d)
    lib.oracle_bilateral(_f(d), _f(out), H, W, _f(ws), r, sigma_r)
    return out


def vertex_normals(depth, Kinv):
    d = _c32(depth)
    H, W = d.shape
    Ki = _c32(Kinv, 9)
    outs = [np.zeros_like(d) for _ in range(5)]
    m = np.zeros(d.shape, np.uint8)
    lib.oracle_vertex_normals(_f(d), _f(Ki), H, W, *(_f(o) for o in outs), m.ctypes.data_as(_U8))
    return dict(zip(("vx", "vy", "nx", "ny", "nz"), outs), mask=m)


def new_volume(dims, kind="f32"):
    n = int(np.prod(dims))
    if kind == "f32":
        return np.full(n, np.nan, np.float32), np.full(n, np.nan, np.float32)
    if kind == "f64":
        return np.full(n, np.nan, np.float64), np.full(n, np.nan, np.float64)
    if kind == "f16":
        return np.full(n, 0x7e00, np.uint16), np.zeros(n, np.uint16)
    raise ValueError(kind)


def integrate(depth, K, T, vol, prm, first, tsdf, weight):
    """In-place update of (tsdf, weight); dtype selects the f32/f64/f16 variant."""
    d = _c32(depth)
    H, W = d.shape
    args = (_f(d), H, W, _f(_c32(K, 9)), _f(_c32(T, 16)), _ivec(vol.dims), _ivec(vol.start),
            vol.voxel_mm, prm.truncation_mm, prm.sdf_min_mm, prm.sdf_max_mm, int(first))
    if tsdf.dtype == np.float32:
        lib.oracle_integrate_f32(*args, _f(tsdf), _f(weight))
    elif tsdf.dtype == np.float64:
        lib.oracle_integrate_f64(*args, tsdf.ctypes.data_as(_D), weight.ctypes.data_as(_D))
    elif tsdf.dtype == np.uint16:
        lib.oracle_integrate_f16(*args, tsdf.ctypes.data_as(_U16), weight.ctypes.data_as(_U16))
    else:
        raise TypeError(tsdf.dtype)


def raycast(depth, K, T, tsdf, vol, trunc, fix_bounds=0):
    d = _c32(depth)
    H, W = d.shape
    stats = np.zeros(2, np.int64)
    hit = np.zeros((H, W), np.int32)
    dt = np.float64 if tsdf.dtype == np.float64 else np.float32
    outs = [np.zeros((H, W), dt) for _ in range(6)]
    args = (_f(d), H, W, _f(_c32(K, 9)), _f(_c32(T, 16)))
    tail = (_ivec(vol.dims), _ivec(vol.start), vol.voxel_mm, trunc, int(fix_bounds))
    if tsdf.dtype == np.float32:
        lib.oracle_raycast_f32(*args, _f(tsdf), *tail, *(_f(o) for o in outs),
                               hit.ctypes.data_as(_I32), stats.ctypes.data_as(_I64))
    elif tsdf.dtype == np.uint16:
        lib.oracle_raycast_f16(*args, tsdf.ctypes.data_as(_U16), *tail, *(_f(o) for o in outs),
                               hit.ctypes.data_as(_I32), stats.ctypes.data_as(_I64))
    else:
        lib.oracle_raycast_f64(*args, tsdf.ctypes.data_as(_D), *tail,
                               *(o.ctypes.data_as(_D) for o in outs), hit.ctypes.data_as(_I32),
                               stats.ctypes.data_as(_I64))
    res = dict(zip(("x", "y", "z", "nx", "ny", "nz"), outs))
    res["hit"] = hit
    res["steps"], res["hits"] = int(stats[0]), int(stats[1])
    return res


TRI_KEYS = ("x", "y", "z", "nx", "ny", "nz", "depth")


def raycast_tri(depth, K, T, tsdf, vol, trunc, grad_depth=None, grad_normal=None):
    """Trilinear raycast.  tsdf float32 / uint16 (half bits): FP32 forward with
    a float32 pose.  tsdf float64: FP64 forward with a float64 pose, plus the
    adjoint when grad_depth is given (returns "grad_tsdf", "grad_xi")."""
    d = _c32(depth)
    H, W = d.shape
    hit = np.zeros((H, W), np.int32)
    disc = np.zeros((H, W, 4), np.int32)
    tail = (_ivec(vol.dims), _ivec(vol.start), vol.voxel_mm, trunc)
    if tsdf.dtype == np.float64:
        outs = [np.zeros((H, W), np.float64) for _ in range(7)]
        arr = (_D * 7)(*(o.ctypes.data_as(_D) for o in outs))
        Td = np.ascontiguousarray(T, np.float64).reshape(16)
        gF = gxi = gD = gN = None
        if grad_depth is not None:
            gD = np.ascontiguousarray(grad_depth, np.float64).reshape(H, W)
            gN = (np.ascontiguousarray(grad_normal, np.float64).reshape(H, W, 3)
                  if grad_normal is not None else None)
            gF = np.zeros(tsdf.size, np.float64)
            gxi = np.zeros(6, np.float64)
        ptr = lambda a: a.ctypes.data_as(_D) if a is not None else None  # noqa: E731
        lib.oracle_raycast_tri_f64(_f(d), H, W, _f(_c32(K, 9)), Td.ctypes.data_as(_D),
                                   tsdf.ctypes.data_as(_D), *tail, arr,
                                   hit.ctypes.data_as(_I32), disc.ctypes.data_as(_I32),
                                   ptr(gD), ptr(gN), ptr(gF), ptr(gxi))
    else:
        outs = [np.zeros((H, W), np.float32) for _ in range(7)]
        arr = (_F * 7)(*(_f(o) for o in outs))
        args = (_f(d), H, W, _f(_c32(K, 9)), _f(_c32(T, 16)))
        if tsdf.dtype == np.float32:
            lib.oracle_raycast_tri_f32(*args, _f(tsdf), *tail, arr, hit.ctypes.data_as(_I32),
                                       disc.ctypes.data_as(_I32))
        else:
            lib.oracle_raycast_tri_f16(*args, tsdf.ctypes.data_as(_U16), *tail, arr,
                                       hit.ctypes.data_as(_I32), disc.ctypes.data_as(_I32))
        gF = gxi = None
    res = dict(zip(TRI_KEYS, outs))
    res["hit"], res["disc"] = hit, disc
    if gF is not None:
        res["grad_tsdf"], res["grad_xi"] = gF, gxi
    return res


def icp(frame, model, K, T_prev, iters, dist_gate, angle_gate, reduce_mode=0, masks=True,
        T_start=None, min_inliers=None, stop=None):
    """frame: dict vx, vy, vz(depth), nx, ny, nz, mask; model: dict x, y, z, nx, ny, nz.
    min_inliers None: the reference's update (sensor_pose_estimation.cu:389-429);
    a number: the robust update (a failed solve or fewer inliers: no increment).
    stop (rot rad, trans mm): the opt-in early stop after an increment below
    both (kf_icp_converged; no reference counterpart); "iters_done" is the
    number of iterations run."""
    H, W = frame["vz"].shape
    fr = [_c32(frame[k]) for k in ("vx", "vy", "vz", "nx", "ny", "nz")]
    m = np.ascontiguousarray(frame["mask"], np.uint8)
    mo = [_c32(model[k]) for k in ("x", "y", "z", "nx", "ny", "nz")]
    T = _c32(T_prev, 16).copy()
    Ts = np.zeros((max(iters, 1), 16), np.float32)
    sums = np.zeros((max(iters, 1), 28), np.float64)
    st = np.zeros(max(iters, 1), np.int32)
    mk = np.zeros((max(iters, 1), H, W), np.uint8)
    done = np.zeros(1, np.int32)
    sr, stt = stop if stop is not None else (0.0, 0.0)
    lib.oracle_icp(*(_f(a) for a in fr), m.ctypes.data_as(_U8), *(_f(a) for a in mo), H, W,
                   _f(_c32(K, 9)), _f(T), _f(_c32(T_start, 16)) if T_start is not None else None,
                   iters, dist_gate, angle_gate, reduce_mode, _f(Ts),
                   sums.ctypes.data_as(_D), st.ctypes.data_as(_I32),
                   mk.ctypes.data_as(_U8) if masks else None,
                   -1.0 if min_inliers is None else float(min_inliers), float(sr), float(stt),
                   done.ctypes.data_as(_I32))
    n = int(done[0])
    return {"T": T.reshape(4, 4), "T_iters": Ts[:n].reshape(-1, 4, 4), "sums": sums[:n],
            "status": st[:n], "masks": mk[:n] if masks else None, "iters_done": n}


def f32_to_f16_bits(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape, np.uint16)
    lib.oracle_f32_to_f16(_f(x), out.ctypes.data_as(_U16), x.size)
    return out


def f16_bits_to_f32(h):
    h = np.ascontiguousarray(h, np.uint16)
    out = np.zeros(h.shape, np.float32)
    lib.oracle_f16_to_f32(h.ctypes.data_as(_U16), _f(out), h.size)
    return out


def icp_angle_gate(c, gate):
    return lib.oracle_icp_angle_gate(float(c), float(gate))


def normal_flip(c):
    return lib.oracle_normal_flip(float(c))


def pipeline(frames, K, vol, prm, reduce_mode=0, stage_ms=None, timing_build=False):
    """The whole frame sequence (kinfu_pipeline.cpp:91-193) on the CPU;
    timing_build runs it in the -O3 x86-64-v3 build of the same source."""
    lib_ = timing_lib() if timing_build else lib
    f = np.ascontiguousarray(frames, np.uint16)
    n, H, W = f.shape
    tsdf, weight = new_volume(vol.dims, "f32")
    poses = np.zeros((n, 16), np.float32)
    ms = np.zeros(6, np.float64)
    r = prm.bilateral_radius if prm.bilateral else 0
    lib_.oracle_pipeline(f.ctypes.data_as(_U16), n, H, W, _f(_c32(K, 9)), _ivec(vol.dims),
                        _ivec(vol.start), vol.voxel_mm, prm.truncation_mm, prm.sdf_min_mm,
                        prm.sdf_max_mm, prm.depth_min_mm, prm.depth_max_mm, r, prm.sigma_s_px,
                        prm.sigma_r_mm, prm.icp_iterations, prm.icp_dist_mm, prm.icp_angle_deg,
                        reduce_mode, prm.raycast_fix_bounds, _f(tsdf), _f(weight), _f(poses),
                        ms.ctypes.data_as(_D), prm.icp_stop_rot, prm.icp_stop_trans_mm)
    if stage_ms is not None:
        stage_ms[:] = ms
    return poses.reshape(n, 4, 4), tsdf, weight


# ---------------------------------------------------------------- ICP pyramid
# No reference counterpart: the reference's ICP runs at full resolution only
# (sensor_pose_estimation.cu:339).

def pyramid_depth(depth, max_step):
    d = _c32(depth)
    H, W = d.shape
    out = np.zeros((H // 2, W // 2), np.float32)
    lib.oracle_pyramid_depth(_f(d), H, W, max_step, _f(out))
    return out


def pyramid_model(model, ok):
    """model: dict x, y, z, nx, ny, nz (H x W); ok: H x W (1 = surface)."""
    keys = ("x", "y", "z", "nx", "ny", "nz")
    ins = [_c32(model[k]) for k in keys]
    H, W = ins[0].shape
    o = np.ascontiguousarray(ok, np.uint8)
    outs = [np.zeros((H // 2, W // 2), np.float32) for _ in keys]
    ok_out = np.zeros((H // 2, W // 2), np.uint8)
    lib.oracle_pyramid_model((_F * 6)(*(_f(a) for a in ins)), o.ctypes.data_as(_U8), H, W,
                             (_F * 6)(*(_f(a) for a in outs)), ok_out.ctypes.data_as(_U8))
    res = dict(zip(keys, outs))
    res["ok"] = ok_out
    return res


def pyramid_intrinsics(K):
    out = np.zeros(9, np.float32)
    lib.oracle_pyramid_intrinsics(_f(_c32(K, 9)), _f(out))
    return out


def pyramid_levels(depth, model, hit, K, Kinv_of, levels, max_step):
    """Levels 0 .. levels-1: [(K_l, frame maps dict incl. vz and mask, model
    maps dict incl. ok)], level 0 from the full-resolution inputs."""
    fr = vertex_normals(depth, Kinv_of(K))
    fr["vz"] = _c32(depth)
    mo = dict(model)
    mo["ok"] = (np.asarray(hit) >= 0).astype(np.uint8)
    out = [(np.asarray(K, np.float32).reshape(9), fr, mo)]
    d = _c32(depth)
    for _ in range(1, levels):
        Kl = pyramid_intrinsics(out[-1][0])
        d = pyramid_depth(d, max_step)
        fr = vertex_normals(d, Kinv_of(Kl))
        fr["vz"] = d
        mo = pyramid_model(out[-1][2], out[-1][2]["ok"])
        out.append((Kl, fr, mo))
    return out


def icp_pyramid(levels, T_prev, iters, dist_gate, angle_gate, reduce_mode=1):
    """Coarse to fine: iters[l] iterations at level l (coarsest first), each
    level starting from the pose the level before ended with; the model maps
    of every level were raycast at T_prev."""
    T = np.asarray(T_prev, np.float32).reshape(4, 4)
    record = []
    for l in range(len(levels) - 1, -1, -1):
        K, fr, mo = levels[l]
        if iters[l] == 0:
            continue
        r = icp(fr, mo, K, T_prev, iters[l], dist_gate, angle_gate, reduce_mode=reduce_mode,
                T_start=T)
        record.append((l, r))
        T = r["T"]
    return T, record


def jpeg_decode(data):
    """cv::imread(path, IMREAD_COLOR) of a baseline JPEG held in `data`
    (bytes): H x W x 3 uint8 in B, G, R order, or None where the oracle
    rejects the stream (system_utility.cpp:70-75)."""
    buf = np.frombuffer(bytes(data), np.uint8).copy()
    h, w = C.c_int(0), C.c_int(0)
    if lib.oracle_jpeg_decode(buf.ctypes.data_as(_U8), buf.size, C.byref(h), C.byref(w),
                              None) != 0:
        return None
    out = np.zeros((h.value, w.value, 3), np.uint8)
    if lib.oracle_jpeg_decode(buf.ctypes.data_as(_U8), buf.size, C.byref(h), C.byref(w),
                              out.ctypes.data_as(_U8)) != 0:
        return None
    return out


def point_cloud(x, y, z, bgr):
    """ThreeDViewer::SetUpPointClouds (threeD_viewer.cpp:20-55): H x W x 4
    float32 records {x, y, z, rgb}, rgb the bits of r << 16 | g << 8 | b."""
    z = _c32(z)
    H, W = z.shape
    out = np.zeros((H, W, 4), np.float32)
    b = np.ascontiguousarray(bgr, np.uint8)
    lib.oracle_point_cloud(_f(_c32(x)), _f(_c32(y)), _f(z), b.ctypes.data_as(_U8), H, W, _f(out))
    return out


def jpeg_coefficients(data):
    """The entropy-decoded coefficient blocks (n x 64 int16, natural order,
    not dequantised), component after component over whole MCUs."""
    buf = np.frombuffer(bytes(data), np.uint8).copy()
    nb = C.c_int64(0)
    if lib.oracle_jpeg_coefficients(buf.ctypes.data_as(_U8), buf.size, None, C.byref(nb)) != 0:
        return None
    out = np.zeros((nb.value, 64), np.int16)
    if lib.oracle_jpeg_coefficients(buf.ctypes.data_as(_U8), buf.size,
                                    out.ctypes.data_as(C.POINTER(C.c_int16)), C.byref(nb)) != 0:
        return None
    return out
