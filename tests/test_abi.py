"""The C ABI library: loads on a CPU host, exports every declared entry point,
validates arguments before touching the GPU, and its presets reproduce the
reference's compiled-in constants."""
import ctypes
import os
import re

import numpy as np

from conftest import ROOT


def declared_functions():
    hdr = open(os.path.join(ROOT, "include", "kinfu_amd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|float|const char\s*\*)\s*(kf_\w+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol(kf):
    names = declared_functions()
    assert len(names) >= 45
    lib = ctypes.CDLL(kf._abi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    unbound = sorted(set(names) - set(kf.exported_symbols()))
    assert not unbound, f"declared but not bound in the Python view: {unbound}"


def test_struct_sizes_match_the_binding(kf):
    """kf_struct_sizes reports the library's sizeof of every public struct;
    the ctypes view declares the same layouts (a stale build of either side
    would otherwise write past the other's copy)."""
    sizes = (ctypes.c_size_t * 3)()
    assert kf._abi.lib.kf_struct_sizes(sizes) == kf._abi.KF_OK
    assert tuple(sizes) == (ctypes.sizeof(kf._abi.Params), ctypes.sizeof(kf._abi.VolumeDesc),
                            ctypes.sizeof(kf._abi.SynthScene))
    assert kf._abi.lib.kf_struct_sizes(None) == -1  # KF_ERR_INVALID


def test_reference_preset_matches_reference_constants(kf):
    vol, prm, K, H, W = kf.preset(kf.PRESET_REFERENCE)
    # kinfu_pipeline.cpp:14-24
    np.testing.assert_array_equal(K, np.array([524.639, 0, 316.625, 0, 523.503, 256.2318, 0, 0, 1],
                                              np.float32))
    # kinfu_pipeline.h:42-62, system_interface.cpp:6-11, sensor_pose_estimation.cu:48,69,339
    assert (H, W) == (480, 640)
    assert list(vol.dims) == [360, 360, 300] and list(vol.start) == [-200, -200, -120]
    assert vol.voxel_mm == 10.0 and vol.precision == kf.KF_TSDF_F32
    assert (prm.truncation_mm, prm.sdf_min_mm, prm.sdf_max_mm) == (300.0, -30.0, 300.0)
    assert (prm.depth_min_mm, prm.depth_max_mm) == (300.0, 2000.0)
    assert (prm.icp_iterations, prm.icp_dist_mm, prm.icp_angle_deg) == (10, 20.0, 30.0)
    assert prm.bilateral == 0 and prm.bilateral_radius == 3
    assert (prm.sigma_s_px, prm.sigma_r_mm) == (4.5, 30.0)


def test_workload_presets(kf):
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    assert np.prod(list(vol.dims)) == 38_880_000 and (H, W) == (480, 640)
    assert K[0] == K[4] == 525.0 and K[2] == 319.5 and K[5] == 239.5
    assert prm.truncation_mm == 30.0 and prm.bilateral_radius == 2
    assert (prm.icp_dist_mm, prm.icp_angle_deg) == (100.0, 20.0)
    vol2, prm2, *_ = kf.preset(kf.PRESET_ROOM_5MM_F16)
    assert np.prod(list(vol2.dims)) == 311_040_000 and vol2.precision == kf.KF_TSDF_F16
    vol4, prm4, K4, H4, W4 = kf.preset(kf.PRESET_ROOM_HD_F16)
    assert np.prod(list(vol4.dims)) == 1024 * 1024 * 768 and (H4, W4) == (720, 1280)


def test_create_rejects_bad_arguments_without_a_gpu(kf):
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    h = ctypes.c_void_p()
    vol.dims[0] = 1
    rc = kf.lib.kf_create(ctypes.byref(vol), H, W, ctypes.byref(prm), ctypes.byref(h))
    assert rc == -1 and b"dims" in kf.lib.kf_last_error()
    vol.dims[:] = [2048, 2048, 1024]  # 2^32 voxels: above the int32 voxel index range
    rc = kf.lib.kf_create(ctypes.byref(vol), H, W, ctypes.byref(prm), ctypes.byref(h))
    assert rc == -1 and b"2^31" in kf.lib.kf_last_error()
    vol.dims[:] = [360, 360, 300]
    prm.bilateral_radius = 9
    rc = kf.lib.kf_create(ctypes.byref(vol), H, W, ctypes.byref(prm), ctypes.byref(h))
    assert rc == -1 and b"bilateral_radius" in kf.lib.kf_last_error()
    assert not h.value


def test_read_depth_file_roundtrip_and_missing(kf, tmp_path):
    # the reference reads raw little-endian uint16 with no header
    # (system_utility.cpp:58-68)
    H, W = 6, 8
    d = (np.arange(H * W, dtype=np.uint16) * 1237) % 65535
    p = tmp_path / "dep0001.dat"
    p.write_bytes(d.astype("<u2").tobytes())
    out = np.zeros(H * W, np.uint16)
    rc = kf.lib.kf_read_depth_file(str(p).encode(), H, W,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    assert rc == 0
    np.testing.assert_array_equal(out, d)
    rc = kf.lib.kf_read_depth_file(str(tmp_path / "none.dat").encode(), H, W,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    assert rc == -5
    p.write_bytes(d[:10].astype("<u2").tobytes())
    rc = kf.lib.kf_read_depth_file(str(p).encode(), H, W,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    assert rc == -5


def test_spatial_kernel_same_as_oracle_and_numpy(kf, oracle):
    for r, s in ((2, 4.5), (3, 4.5), (5, 2.0)):
        w = kf.spatial_kernel(r, s)
        np.testing.assert_array_equal(w, oracle.spatial_kernel(r, s))
        yy, xx = np.mgrid[-r:r + 1, -r:r + 1]
        np.testing.assert_allclose(w, np.exp(-(xx ** 2 + yy ** 2) / (2.0 * s * s)), rtol=3e-7)


def test_ply_writer_roundtrip(kf, tmp_path):
    # threeD_viewer.cpp:20-55 keeps pixels whose z != 0 and reads colour as B, G, R
    rng = np.random.default_rng(9)
    n = 500
    x, y, z = (rng.normal(size=n).astype(np.float32) * 100 for _ in range(3))
    z[::7] = 0.0
    nx, ny, nz = (rng.normal(size=n).astype(np.float32) for _ in range(3))
    bgr = rng.integers(0, 256, (n, 3)).astype(np.uint8)
    keep = z != 0
    for binary in (True, False):
        p = tmp_path / f"cloud_{int(binary)}.ply"
        kf.write_ply(p, x, y, z, nx, ny, nz, bgr, binary=binary)
        r = kf.read_ply(p)
        assert r["x"].size == keep.sum()
        for k, v in (("x", x), ("y", y), ("z", z), ("nx", nx), ("ny", ny), ("nz", nz)):
            np.testing.assert_array_equal(r[k], v[keep])
        np.testing.assert_array_equal(r["red"], bgr[keep, 2])
        np.testing.assert_array_equal(r["blue"], bgr[keep, 0])
    p = tmp_path / "xyz.ply"
    kf.write_ply(p, x, y, z)
    r = kf.read_ply(p)
    assert sorted(r) == ["x", "y", "z"] and r["z"].size == keep.sum()
    assert kf.lib.kf_write_ply(str(tmp_path / "no" / "dir.ply").encode(), n,
                               kf._abi.fptr(x), kf._abi.fptr(y), kf._abi.fptr(z),
                               None, None, None, None, 1) == -5


def test_host_mirror_shim_loads_and_rejects_bad_arguments(kf):
    """tests/shim/libkinfu_shim.so (the C entry of the C++ KinfuPipeline
    mirror, driven on the GPU by test_gpu_host_mirror.py) links the library
    and rejects a null data directory before creating a context."""
    import ctypes as C
    import os

    from conftest import ROOT
    shim = C.CDLL(os.path.join(ROOT, "tests", "shim", "libkinfu_shim.so"))
    fn = shim.shim_run_reference_pipeline
    fn.restype = C.c_int
    fn.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int,
                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    assert fn(None, 480, 640, 2000.0, 300.0, 25, 3, None, None, None, 0) == -1
