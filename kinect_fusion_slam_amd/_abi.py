"""ctypes binding of libkinfu_amd.so (the C ABI declared in include/kinfu_amd.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
kinect_fusion_slam_amd/csrc``).  Importing this module without the built
library raises immediately: there is no CPU fallback for any stage.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkinfu_amd.so")

KF_OK = 0
KF_TSDF_F32, KF_TSDF_F16 = 0, 1
KF_MEM_HOST, KF_MEM_DEVICE = 0, 1
KF_MAX_ICP_ITERS = 64
KF_ICP_SUMS = 28
KF_ICP_STATUS_GATED = 2
STAGES = ("depth", "bilateral", "vertex", "icp", "integrate", "raycast", "integrate_prep",
          "raycast_adjoint", "refine", "icp_iter", "color")
KF_FRAME_CURRENT = -1
KF_RAYCAST_REFERENCE, KF_RAYCAST_TRILINEAR = 0, 1
PRESET_REFERENCE = 100
PRESET_ROOM_CPU = 0       # 640x480, 360x360x300 @10mm FP32, 10-frame CPU case
PRESET_ROOM = 1           # same geometry, the GPU stream
PRESET_ROOM_5MM_F16 = 2   # 720x720x600 @5mm, FP16 TSDF + u16 weight
PRESET_SPHERES_128 = 3    # 128^3 @10mm, 320x240 differentiable render
PRESET_ROOM_HD_F16 = 4    # 1024x1024x768 @4mm FP16, 1280x720
PRESETS = (PRESET_ROOM_CPU, PRESET_ROOM, PRESET_ROOM_5MM_F16, PRESET_SPHERES_128,
           PRESET_ROOM_HD_F16)


class KinfuError(RuntimeError):
    pass


class VolumeDesc(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("start", C.c_int32 * 3), ("voxel_mm", C.c_float),
                ("precision", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("truncation_mm", C.c_float), ("sdf_min_mm", C.c_float),
                ("sdf_max_mm", C.c_float), ("depth_min_mm", C.c_float),
                ("depth_max_mm", C.c_float), ("bilateral", C.c_int32),
                ("bilateral_radius", C.c_int32), ("sigma_s_px", C.c_float),
                ("sigma_r_mm", C.c_float), ("icp_iterations", C.c_int32),
                ("icp_dist_mm", C.c_float), ("icp_angle_deg", C.c_float),
                ("icp_compensated", C.c_int32), ("record_icp", C.c_int32),
                ("stage_timing", C.c_int32), ("raycast_fix_bounds", C.c_int32),
                ("raycast_mode", C.c_int32), ("refine_steps", C.c_int32),
                ("refine_max_residual_mm", C.c_float), ("icp_robust", C.c_int32),
                ("icp_min_inliers", C.c_int32), ("icp_levels", C.c_int32),
                ("icp_coarse_iters", C.c_int32 * 2), ("pyramid_max_step_mm", C.c_float),
                ("icp_stop_rot", C.c_float), ("icp_stop_trans_mm", C.c_float),
                ("raycast_skip", C.c_int32)]


class SynthScene(C.Structure):
    _fields_ = [("room_min", C.c_float * 3), ("room_max", C.c_float * 3),
                ("n_spheres", C.c_int32), ("spheres", (C.c_float * 4) * 16),
                ("n_boxes", C.c_int32), ("boxes", (C.c_float * 6) * 16)]


_P = C.c_void_p
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)
_U16 = C.POINTER(C.c_uint16)

_SIGNATURES = {
    "kf_preset": (C.c_int, [C.c_int32, C.POINTER(VolumeDesc), C.POINTER(Params), _F, _I32, _I32]),
    "kf_create": (C.c_int, [C.POINTER(VolumeDesc), C.c_int32, C.c_int32, C.POINTER(Params),
                            C.POINTER(_P)]),
    "kf_destroy": (C.c_int, [_P]),
    "kf_last_error": (C.c_char_p, []),
    "kf_struct_sizes": (C.c_int, [C.POINTER(C.c_size_t)]),
    "kf_set_params": (C.c_int, [_P, C.POINTER(Params)]),
    "kf_get_params": (C.c_int, [_P, C.POINTER(Params)]),
    "kf_reset": (C.c_int, [_P]),
    "kf_synchronize": (C.c_int, [_P]),
    "kf_device_bytes": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "kf_upload_depth_u16": (C.c_int, [_P, _P, C.c_int32]),
    "kf_upload_depth_f32": (C.c_int, [_P, _P, C.c_int32]),
    "kf_read_depth_file": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, _U16]),
    "kf_download_depth": (C.c_int, [_P, _F]),
    "kf_bilateral": (C.c_int, [_P]),
    "kf_spatial_kernel": (C.c_int, [C.c_int32, C.c_float, _F]),
    "kf_vertex_normals": (C.c_int, [_P, _F]),
    "kf_download_vertex_maps": (C.c_int, [_P, _F, _F, _F, _F, _F, _F, _U8]),
    "kf_upload_vertex_maps": (C.c_int, [_P, _F, _F, _F, _F, _F, _F, _U8]),
    "kf_estimate_pose": (C.c_int, [_P, _F, _F]),
    "kf_icp_history": (C.c_int, [_P, _I32, _F, _D, _I32, _U8]),
    "kf_integrate": (C.c_int, [_P, _F, _F, C.c_int32]),
    "kf_download_volume": (C.c_int, [_P, _F, _F]),
    "kf_upload_volume": (C.c_int, [_P, _F, _F]),
    "kf_download_volume_raw": (C.c_int, [_P, _P, C.c_size_t]),
    "kf_upload_volume_raw": (C.c_int, [_P, _P, C.c_size_t]),
    "kf_volume_device_ptr": (C.c_int, [_P, C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "kf_raycast": (C.c_int, [_P, _F, _F]),
    "kf_download_model_maps": (C.c_int, [_P, _F, _F, _F, _F, _F, _F, _I32]),
    "kf_upload_color_jpeg": (C.c_int, [_P, _U8, C.c_size_t]),
    "kf_jpeg_info": (C.c_int, [_U8, C.c_size_t, _I32, _I32, _I32]),
    "kf_jpeg_coefficients": (C.c_int, [_U8, C.c_size_t, C.POINTER(C.c_int16), C.c_int64,
                                       C.POINTER(C.c_int64)]),
    "kf_load_color_file": (C.c_int, [_P, C.c_char_p]),
    "kf_upload_color_bgr": (C.c_int, [_P, _U8, C.c_int32]),
    "kf_download_color": (C.c_int, [_P, _U8]),
    "kf_point_cloud": (C.c_int, [_P, _F, C.c_int32]),
    "kf_upload_model_maps": (C.c_int, [_P, _F, _F, _F, _F, _F, _F]),
    "kf_stream_upload_u16": (C.c_int, [_P, _P, C.c_int32, C.c_int32]),
    "kf_pipeline_initial": (C.c_int, [_P, _F, C.c_int32]),
    "kf_pipeline_incoming": (C.c_int, [_P, _F, C.c_int32]),
    "kf_pipeline_run": (C.c_int, [_P, _F, C.c_int32, C.c_int32]),
    "kf_get_pose": (C.c_int, [_P, _F]),
    "kf_set_pose": (C.c_int, [_P, _F]),
    "kf_download_poses": (C.c_int, [_P, _F, C.c_int32]),
    "kf_stage_times": (C.c_int, [_P, _D, _I64]),
    "kf_reset_stage_times": (C.c_int, [_P]),
    "kf_count_raycast_work": (C.c_int, [_P, C.c_int32]),
    "kf_download_rendered_depth": (C.c_int, [_P, _F]),
    "kf_download_volume_box": (C.c_int, [_P, _I32, _I32, _F, _F]),
    "kf_refine_pose": (C.c_int, [_P, _F, C.c_int32, _F]),
    "kf_export_model_ply": (C.c_int, [_P, C.c_char_p, _U8, C.c_int32]),
    "kf_write_ply": (C.c_int, [C.c_char_p, C.c_int64, _F, _F, _F, _F, _F, _F, _U8, C.c_int32]),
    "kf_refine_history": (C.c_int, [_P, _I32, _D, _I32, _I32]),
    "kf_raycast_adjoint": (C.c_int, [_P, _F, _F, _F, _F, _F, _D]),
    "kf_raycast_work": (C.c_int, [_P, _I64, _I64]),
    "kf_raycast_skipped": (C.c_int, [_P, _I64]),
    "kf_adjoint_used_saved": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "kf_gate_thresholds": (C.c_int, [C.c_float, _F, _F]),
    "kf_dist_gate_threshold": (C.c_int, [C.c_float, _F]),
    "kf_integration_work": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "kf_pyramid_level": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _F,
                                   _F, _F, _F, _F, _F, _F, _F, _U8, _F, _F, _F, _F, _F, _F,
                                   _U8]),
    "kf_estimate_pose_static": (C.c_int, [_P, _P, _P, _P, C.c_int32, C.c_int32, _F,
                                          C.POINTER(Params), _F, _P, _P]),
    "kf_expf_abi": (C.c_float, [C.c_float]),
    "kf_mat4_inverse": (C.c_int, [_F, _F]),
    "kf_mat3_inverse": (C.c_int, [_F, _F]),
    "kf_ldlt6": (C.c_int, [_D, _D, _D]),
    "kf_synth_build_scene": (C.c_int, [C.c_int32, C.c_int32, C.c_uint64, _F, _F,
                                       C.POINTER(SynthScene)]),
    "kf_synth_orbit_pose": (C.c_int, [C.c_int32, _F]),
    "kf_synth_render": (C.c_int, [C.POINTER(SynthScene), _F, _F, C.c_int32, C.c_int32, C.c_float,
                                  C.c_float, C.c_float, C.c_uint64, _U16, _F, C.c_int32]),
    "kf_synth_stream": (C.c_int, [C.POINTER(SynthScene), _F, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_int32, C.c_float, C.c_float, C.c_float, _U16, _F,
                                  C.c_int32]),
    "kf_synth_preset_scene": (C.c_int, [C.c_int32, C.POINTER(SynthScene)]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP kernels are the only implementation of every stage)")
    lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    built = (C.c_size_t * 3)()
    lib.kf_struct_sizes(built)
    mine = (C.sizeof(Params), C.sizeof(VolumeDesc), C.sizeof(SynthScene))
    if tuple(built) != mine:
        raise ImportError(
            f"{LIB_PATH} was built with struct sizes (kf_params, kf_volume_desc, "
            f"kf_synth_scene) = {tuple(built)} but this binding declares {mine}; "
            "rebuild the library or update the binding")
    return lib


lib = _load()


def exported_symbols():
    return list(_SIGNATURES)


def check(rc):
    if rc != KF_OK:
        raise KinfuError(f"kinfu_amd error {rc}: {lib.kf_last_error().decode(errors='replace')}")
    return rc


def fptr(a):
    return a.ctypes.data_as(_F)


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} floats, got {a.size}")
    return a


def jpeg_info(data):
    """(height, width, components) of a baseline JPEG held in bytes."""
    buf = np.frombuffer(bytes(data), np.uint8)
    h, w, nc = C.c_int32(), C.c_int32(), C.c_int32()
    check(lib.kf_jpeg_info(buf.ctypes.data_as(_U8), buf.size, C.byref(h), C.byref(w),
                           C.byref(nc)))
    return h.value, w.value, nc.value


def jpeg_coefficients(data):
    """The Huffman-decoded coefficient blocks (n_blocks x 64 int16, natural
    order, not dequantised) that the colour upload sends to the device."""
    buf = np.frombuffer(bytes(data), np.uint8)
    nb = C.c_int64()
    check(lib.kf_jpeg_coefficients(buf.ctypes.data_as(_U8), buf.size, None, 0, C.byref(nb)))
    out = np.zeros((nb.value, 64), np.int16)
    check(lib.kf_jpeg_coefficients(buf.ctypes.data_as(_U8), buf.size,
                                   out.ctypes.data_as(C.POINTER(C.c_int16)), nb.value,
                                   C.byref(nb)))
    return out


def preset(which):
    vol, prm = VolumeDesc(), Params()
    K = np.zeros(9, np.float32)
    h, w = C.c_int32(), C.c_int32()
    check(lib.kf_preset(which, C.byref(vol), C.byref(prm), fptr(K), C.byref(h), C.byref(w)))
    return vol, prm, K, h.value, w.value


def gate_thresholds(angle_deg):
    lo, hi = C.c_float(), C.c_float()
    check(lib.kf_gate_thresholds(angle_deg, C.byref(lo), C.byref(hi)))
    return lo.value, hi.value


def read_depth_file(path, h, w):
    """A raw uint16 little-endian depth frame (system_utility.cpp:58-68)."""
    out = np.zeros((h, w), np.uint16)
    check(lib.kf_read_depth_file(str(path).encode(), h, w, out.ctypes.data_as(_U16)))
    return out


def dist_gate_threshold(dist_mm):
    """Largest squared distance the ICP distance gate passes (-1: none)."""
    v = C.c_float()
    check(lib.kf_dist_gate_threshold(dist_mm, C.byref(v)))
    return v.value


def mat4_inverse(T):
    T = _f32(T, 16)
    out = np.zeros(16, np.float32)
    check(lib.kf_mat4_inverse(fptr(T), fptr(out)))
    return out.reshape(4, 4)


def mat3_inverse(K):
    K = _f32(K, 9)
    out = np.zeros(9, np.float32)
    check(lib.kf_mat3_inverse(fptr(K), fptr(out)))
    return out.reshape(3, 3)


def ldlt6(A, b):
    A = np.ascontiguousarray(A, np.float64).reshape(36)
    b = np.ascontiguousarray(b, np.float64).reshape(6)
    x = np.zeros(6, np.float64)
    st = lib.kf_ldlt6(A.ctypes.data_as(_D), b.ctypes.data_as(_D), x.ctypes.data_as(_D))
    return st, x


def spatial_kernel(radius, sigma):
    w = np.zeros((2 * radius + 1) ** 2, np.float32)
    check(lib.kf_spatial_kernel(radius, sigma, fptr(w)))
    return w.reshape(2 * radius + 1, 2 * radius + 1)


def write_ply(path, x, y, z, nx=None, ny=None, nz=None, bgr=None, binary=True):
    x, y, z = (_f32(a) for a in (x, y, z))
    n = x.size
    nrm = [_f32(a) for a in (nx, ny, nz)] if nx is not None else None
    c = np.ascontiguousarray(bgr, np.uint8) if bgr is not None else None
    check(lib.kf_write_ply(str(path).encode(), n, fptr(x), fptr(y), fptr(z),
                           *((fptr(a) for a in nrm) if nrm else (None, None, None)),
                           c.ctypes.data_as(_U8) if c is not None else None,
                           1 if binary else 0))


def read_ply(path):
    """Minimal reader for the files write_ply produces: dict of arrays."""
    with open(path, "rb") as f:
        head = b""
        while not head.endswith(b"end_header\n"):
            head += f.readline()
        body = f.read()
    lines = head.decode().splitlines()
    n = int([l for l in lines if l.startswith("element vertex")][0].split()[-1])
    props = [(l.split()[1], l.split()[2]) for l in lines if l.startswith("property")]
    dt = np.dtype([(name, "<f4" if t == "float" else "u1") for t, name in props])
    if "binary_little_endian" in lines[1]:
        arr = np.frombuffer(body, dt, count=n)
    else:
        arr = np.loadtxt(body.decode().splitlines(), ndmin=2) if n else np.zeros((0, len(props)))
        arr = np.array([tuple(r) for r in arr], dtype=dt)
    return {name: arr[name] for _, name in props}


# ---------------------------------------------------------------- synthetic stream

def preset_scene(which):
    s = SynthScene()
    check(lib.kf_synth_preset_scene(which, C.byref(s)))
    return s


def orbit_pose(frame):
    T = np.zeros(16, np.float32)
    check(lib.kf_synth_orbit_pose(frame, fptr(T)))
    return T.reshape(4, 4)


def render(scene, K, T, h, w, noise_scale=1.0, hole_frac=0.02, max_depth=4000.0, seed=1000,
           threads=0, clean=False):
    K, T = _f32(K, 9), _f32(T, 16)
    d = np.zeros((h, w), np.uint16)
    c = np.zeros((h, w), np.float32) if clean else None
    check(lib.kf_synth_render(C.byref(scene), fptr(K), fptr(T), h, w, noise_scale, hole_frac,
                              max_depth, seed, d.ctypes.data_as(_U16),
                              fptr(c) if clean else None, threads))
    return (d, c) if clean else d


def stream(scene, K, h, w, first, n, noise_scale=1.0, hole_frac=0.02, max_depth=4000.0,
           threads=0):
    K = _f32(K, 9)
    frames = np.zeros((n, h, w), np.uint16)
    poses = np.zeros((n, 4, 4), np.float32)
    check(lib.kf_synth_stream(C.byref(scene), fptr(K), h, w, first, n, noise_scale, hole_frac,
                              max_depth, frames.ctypes.data_as(_U16), fptr(poses), threads))
    return frames, poses


def set_field(struct, name, value):
    """Sets a structure field; array fields take a sequence."""
    cur = getattr(struct, name)
    if isinstance(cur, C.Array):
        for i, v in enumerate(value):
            cur[i] = v
    else:
        setattr(struct, name, value)


# ---------------------------------------------------------------- context

class KinfuContext:
    """One device-resident KinectFusion context (the C ABI's kf_context)."""

    def __init__(self, vol, prm, height, width):
        self.vol, self.height, self.width = vol, height, width
        self.n = height * width
        self.nvox = vol.dims[0] * vol.dims[1] * vol.dims[2]
        h = _P()
        check(lib.kf_create(C.byref(vol), height, width, C.byref(prm), C.byref(h)))
        self.h = h

    @classmethod
    def from_preset(cls, which, **overrides):
        vol, prm, K, h, w = preset(which)
        for k, v in overrides.items():
            if hasattr(vol, k) and k in ("precision", "voxel_mm"):
                setattr(vol, k, v)
            else:
                set_field(prm, k, v)
        ctx = cls(vol, prm, h, w)
        ctx.K = K
        return ctx

    def close(self):
        if getattr(self, "h", None):
            lib.kf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # params
    @property
    def params(self):
        p = Params()
        check(lib.kf_get_params(self.h, C.byref(p)))
        return p

    def set_params(self, **kw):
        p = self.params
        for k, v in kw.items():
            set_field(p, k, v)
        check(lib.kf_set_params(self.h, C.byref(p)))

    def synchronize(self):
        check(lib.kf_synchronize(self.h))

    def reset(self):
        check(lib.kf_reset(self.h))

    # A1
    def upload_depth_u16(self, depth):
        d = np.ascontiguousarray(depth, np.uint16)
        assert d.size == self.n
        check(lib.kf_upload_depth_u16(self.h, d.ctypes.data_as(_P), KF_MEM_HOST))

    def upload_depth_f32(self, depth):
        d = _f32(depth, self.n)
        check(lib.kf_upload_depth_f32(self.h, d.ctypes.data_as(_P), KF_MEM_HOST))

    def depth(self):
        out = np.zeros((self.height, self.width), np.float32)
        check(lib.kf_download_depth(self.h, fptr(out)))
        return out

    # A4
    def bilateral(self):
        check(lib.kf_bilateral(self.h))

    # A3
    def vertex_normals(self, K):
        check(lib.kf_vertex_normals(self.h, fptr(_f32(K, 9))))

    def vertex_maps(self):
        shape = (self.height, self.width)
        out = {k: np.zeros(shape, np.float32) for k in ("vx", "vy", "vz", "nx", "ny", "nz")}
        m = np.zeros(shape, np.uint8)
        check(lib.kf_download_vertex_maps(self.h, *(fptr(out[k]) for k in
                                                    ("vx", "vy", "vz", "nx", "ny", "nz")),
                                          m.ctypes.data_as(_U8)))
        out["mask"] = m
        return out

    def upload_vertex_maps(self, vx, vy, vz, nx, ny, nz, mask):
        arrs = [_f32(a, self.n) for a in (vx, vy, vz, nx, ny, nz)]
        m = np.ascontiguousarray(mask, np.uint8)
        check(lib.kf_upload_vertex_maps(self.h, *(fptr(a) for a in arrs), m.ctypes.data_as(_U8)))

    # A5-A7
    def estimate_pose(self, K, T):
        T = _f32(T, 16).copy()
        check(lib.kf_estimate_pose(self.h, fptr(_f32(K, 9)), fptr(T)))
        return T.reshape(4, 4)

    def icp_history(self, masks=False):
        n = C.c_int32()
        Ts = np.zeros((KF_MAX_ICP_ITERS, 16), np.float32)
        sums = np.zeros((KF_MAX_ICP_ITERS, KF_ICP_SUMS), np.float64)
        st = np.zeros(KF_MAX_ICP_ITERS, np.int32)
        check(lib.kf_icp_history(self.h, C.byref(n), fptr(Ts), sums.ctypes.data_as(_D),
                                 st.ctypes.data_as(_I32), None))
        k = n.value
        out = {"T": Ts[:k].reshape(k, 4, 4), "sums": sums[:k], "status": st[:k]}
        if masks:
            m = np.zeros((k, self.height, self.width), np.uint8)
            check(lib.kf_icp_history(self.h, C.byref(n), None, None, None,
                                     m.ctypes.data_as(_U8)))
            out["masks"] = m
        return out

    # A8
    def integrate(self, K, T, first):
        check(lib.kf_integrate(self.h, fptr(_f32(K, 9)), fptr(_f32(T, 16)), int(first)))

    def volume(self):
        t = np.zeros(self.nvox, np.float32)
        w = np.zeros(self.nvox, np.float32)
        check(lib.kf_download_volume(self.h, fptr(t), fptr(w)))
        return t, w

    def volume_raw(self):
        """Native brick-tiled storage (TSDF array then weight array)."""
        ptr, nbytes = C.c_void_p(), C.c_size_t()
        check(lib.kf_volume_device_ptr(self.h, C.byref(ptr), C.byref(nbytes)))
        out = np.zeros(nbytes.value, np.uint8)
        check(lib.kf_download_volume_raw(self.h, out.ctypes.data_as(_P), out.size))
        return out

    def upload_volume(self, tsdf, weight):
        t, w = _f32(tsdf, self.nvox), _f32(weight, self.nvox)
        check(lib.kf_upload_volume(self.h, fptr(t), fptr(w)))

    def upload_volume_raw(self, raw):
        raw = np.ascontiguousarray(raw)
        check(lib.kf_upload_volume_raw(self.h, raw.ctypes.data_as(_P), raw.nbytes))

    # A9
    def raycast(self, K, T):
        check(lib.kf_raycast(self.h, fptr(_f32(K, 9)), fptr(_f32(T, 16))))

    def model_maps(self):
        shape = (self.height, self.width)
        out = {k: np.zeros(shape, np.float32) for k in ("x", "y", "z", "nx", "ny", "nz")}
        hit = np.zeros(shape, np.int32)
        check(lib.kf_download_model_maps(self.h, *(fptr(out[k]) for k in
                                                   ("x", "y", "z", "nx", "ny", "nz")),
                                         hit.ctypes.data_as(_I32)))
        out["hit"] = hit
        return out

    # colour (system_utility.cpp:70-75, threeD_viewer.cpp:20-55)
    def upload_color_jpeg(self, data):
        """Decodes a baseline JPEG (bytes) into the context's B, G, R image."""
        buf = np.frombuffer(bytes(data), np.uint8)
        check(lib.kf_upload_color_jpeg(self.h, buf.ctypes.data_as(_U8), buf.size))

    def load_color_file(self, path):
        check(lib.kf_load_color_file(self.h, os.fsencode(path)))

    def upload_color_bgr(self, bgr):
        a = np.ascontiguousarray(bgr, np.uint8)
        if a.size != self.n * 3:
            raise ValueError(f"expected {self.height}x{self.width}x3 bytes, got {a.shape}")
        check(lib.kf_upload_color_bgr(self.h, a.ctypes.data_as(_U8), KF_MEM_HOST))

    def color(self):
        out = np.zeros((self.height, self.width, 3), np.uint8)
        check(lib.kf_download_color(self.h, out.ctypes.data_as(_U8)))
        return out

    def point_cloud(self):
        """H x W x 4 float32 {x, y, z, rgb}; rgb holds the bits of
        r << 16 | g << 8 | b (PCL's PointXYZRGB layout)."""
        out = np.zeros((self.height, self.width, 4), np.float32)
        check(lib.kf_point_cloud(self.h, fptr(out), KF_MEM_HOST))
        return out

    def upload_model_maps(self, x, y, z, nx, ny, nz):
        arrs = [_f32(a, self.n) for a in (x, y, z, nx, ny, nz)]
        check(lib.kf_upload_model_maps(self.h, *(fptr(a) for a in arrs)))

    # frame sequence
    def stream_upload(self, frames):
        f = np.ascontiguousarray(frames, np.uint16)
        assert f.size % self.n == 0
        check(lib.kf_stream_upload_u16(self.h, f.ctypes.data_as(_P), f.size // self.n,
                                       KF_MEM_HOST))
        self.n_frames = f.size // self.n

    def pipeline_initial(self, K, frame):
        check(lib.kf_pipeline_initial(self.h, fptr(_f32(K, 9)), frame))

    def pipeline_incoming(self, K, frame):
        check(lib.kf_pipeline_incoming(self.h, fptr(_f32(K, 9)), frame))

    def pipeline_run(self, K, first, count):
        check(lib.kf_pipeline_run(self.h, fptr(_f32(K, 9)), first, count))

    def pose(self):
        T = np.zeros(16, np.float32)
        check(lib.kf_get_pose(self.h, fptr(T)))
        return T.reshape(4, 4)

    def set_pose(self, T):
        check(lib.kf_set_pose(self.h, fptr(_f32(T, 16))))

    def poses(self, n):
        out = np.zeros((n, 16), np.float32)
        check(lib.kf_download_poses(self.h, fptr(out), n))
        return out.reshape(n, 4, 4)

    def stage_times(self):
        ms = np.zeros(len(STAGES), np.float64)
        cnt = np.zeros(len(STAGES), np.int64)
        check(lib.kf_stage_times(self.h, ms.ctypes.data_as(_D), cnt.ctypes.data_as(_I64)))
        return {s: (float(ms[i]), int(cnt[i])) for i, s in enumerate(STAGES)}

    def reset_stage_times(self):
        check(lib.kf_reset_stage_times(self.h))

    def volume_box(self, origin, size):
        """(tsdf, weight) of the voxels [origin, origin+size), reference layout."""
        o = np.ascontiguousarray(origin, np.int32)
        z = np.ascontiguousarray(size, np.int32)
        n = int(np.prod(z))
        t, w = np.zeros(n, np.float32), np.zeros(n, np.float32)
        check(lib.kf_download_volume_box(self.h, o.ctypes.data_as(_I32), z.ctypes.data_as(_I32),
                                         fptr(t), fptr(w)))
        return t, w

    def refine_pose(self, K, steps, T=None):
        """Gauss-Newton refinement of the pose on the working depth; returns
        the refined pose (4x4)."""
        Tio = _f32(T, 16).copy() if T is not None else None
        check(lib.kf_refine_pose(self.h, fptr(_f32(K, 9)), int(steps),
                                 fptr(Tio) if Tio is not None else None))
        return Tio.reshape(4, 4) if Tio is not None else self.pose()

    def refine_history(self):
        n = C.c_int32()
        msr = np.zeros(64, np.float64)
        cnt = np.zeros(64, np.int32)
        st = np.zeros(64, np.int32)
        check(lib.kf_refine_history(self.h, C.byref(n), msr.ctypes.data_as(_D),
                                    cnt.ctypes.data_as(_I32), st.ctypes.data_as(_I32)))
        k = n.value
        return {"msr": msr[:k], "count": cnt[:k], "status": st[:k]}

    def export_ply(self, path, bgr=None, binary=True):
        c = np.ascontiguousarray(bgr, np.uint8) if bgr is not None else None
        check(lib.kf_export_model_ply(self.h, str(path).encode(),
                                      c.ctypes.data_as(_U8) if c is not None else None,
                                      1 if binary else 0))

    def rendered_depth(self):
        out = np.zeros((self.height, self.width), np.float32)
        check(lib.kf_download_rendered_depth(self.h, fptr(out)))
        return out

    def raycast_adjoint(self, K, T, grad_depth, grad_normal=None, want_tsdf=True):
        """Adjoint of the trilinear raycast: (dL/dTSDF in the reference
        layout or None, dL/dxi as 6 float64, omega then v)."""
        gD = _f32(grad_depth, self.n)
        gN = _f32(grad_normal, 3 * self.n) if grad_normal is not None else None
        gF = np.zeros(self.nvox, np.float32) if want_tsdf else None
        gxi = np.zeros(6, np.float64)
        check(lib.kf_raycast_adjoint(self.h, fptr(_f32(K, 9)), fptr(_f32(T, 16)), fptr(gD),
                                     fptr(gN) if gN is not None else None,
                                     fptr(gF) if gF is not None else None,
                                     gxi.ctypes.data_as(_D)))
        return gF, gxi

    def adjoint_used_saved(self):
        """True if the last raycast_adjoint started from the saved crossings
        of the raycast before it (same inputs) instead of walking the rays."""
        v = C.c_int32()
        check(lib.kf_adjoint_used_saved(self.h, C.byref(v)))
        return bool(v.value)

    def count_raycast_work(self, enable=True):
        check(lib.kf_count_raycast_work(self.h, 1 if enable else 0))

    def pyramid_level(self, level, K):
        """Level `level` (1 or 2) of the last coarse-to-fine ICP: a dict with
        H, W, K (that level's intrinsics), the frame's depth, vx .. nz and mask,
        and the model maps x .. nz with their model_ok flags."""
        h, w = C.c_int32(), C.c_int32()
        Kl = _f32(K, 9).copy()
        check(lib.kf_pyramid_level(self.h, level, C.byref(h), C.byref(w), None, *([None] * 7),
                                   None, *([None] * 6), None))
        shape = (h.value, w.value)
        f = {k: np.zeros(shape, np.float32) for k in ("depth", "vx", "vy", "vz", "nx", "ny",
                                                      "nz", "x", "y", "z", "mnx", "mny",
                                                      "mnz")}
        mask, ok = np.zeros(shape, np.uint8), np.zeros(shape, np.uint8)
        check(lib.kf_pyramid_level(self.h, level, None, None, fptr(Kl), fptr(f["depth"]),
                                   *(fptr(f[k]) for k in ("vx", "vy", "vz", "nx", "ny", "nz")),
                                   mask.ctypes.data_as(_U8),
                                   *(fptr(f[k]) for k in ("x", "y", "z", "mnx", "mny", "mnz")),
                                   ok.ctypes.data_as(_U8)))
        model = {"x": f["x"], "y": f["y"], "z": f["z"], "nx": f["mnx"], "ny": f["mny"],
                 "nz": f["mnz"], "ok": ok}
        frame = {k: f[k] for k in ("vx", "vy", "vz", "nx", "ny", "nz")}
        frame["mask"] = mask
        return {"H": h.value, "W": w.value, "K": Kl, "depth": f["depth"], "frame": frame,
                "model": model}

    def integration_work(self):
        """(groups, bricks) that passed the last integration's culls."""
        g, b = C.c_int64(), C.c_int64()
        check(lib.kf_integration_work(self.h, C.byref(g), C.byref(b)))
        return g.value, b.value

    def raycast_work(self):
        v, h = C.c_int64(), C.c_int64()
        check(lib.kf_raycast_work(self.h, C.byref(v), C.byref(h)))
        return v.value, h.value

    def raycast_skipped(self):
        """Counted visits that the second pass's skip rounds took."""
        v = C.c_int64()
        check(lib.kf_raycast_skipped(self.h, C.byref(v)))
        return v.value
