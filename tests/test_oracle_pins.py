"""Pins for the CPU oracle (no GPU needed).

The reference ships no golden vectors and cannot be built here, so every
oracle stage is pinned by what the mathematics fixes: closed forms on
analytic scenes (planes, spheres, integer-valued ICP rows), invariants
(symmetry, PSD normal equations, weight = observation count) and
brute-force enumerations on small inputs.  Each check is chosen so that a
plausible transcription slip (a dropped term, a swapped index or sign, a
transposed operand, an off-by-one bound) fails it.
"""
import math
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import ROOT  # noqa: F401  (puts the repo on sys.path)


def vol_desc(dims, start, voxel):
    return SimpleNamespace(dims=list(dims), start=list(start), voxel_mm=float(voxel))


def band(trunc, lo, hi):
    return SimpleNamespace(truncation_mm=float(trunc), sdf_min_mm=float(lo), sdf_max_mm=float(hi))


def centres(vol):
    """Voxel centres (i + start) * d + d / 2 in the reference layout (z fastest)."""
    d = vol.voxel_mm
    ax = [(np.arange(n) + s) * d + d / 2 for n, s in zip(vol.dims, vol.start)]
    return np.meshgrid(*ax, indexing="ij")


def pose(t=(0, 0, 0), rx=0.0, ry=0.0, rz=0.0):
    cx, sx, cy, sy, cz, sz = (math.cos(rx), math.sin(rx), math.cos(ry), math.sin(ry),
                              math.cos(rz), math.sin(rz))
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    T = np.eye(4)
    T[:3, :3] = Rz @ Ry @ Rx
    T[:3, 3] = t
    return T.astype(np.float32)


def intr(f, cx, cy):
    return np.array([f, 0, cx, 0, f, cy, 0, 0, 1], np.float32)


# ---------------------------------------------------------------- A1 range gate

def test_range_gate_is_inclusive_integer_compare(oracle):
    d = np.arange(0, 6000, 7, dtype=np.uint16)
    d = np.concatenate([d, np.uint16([299, 300, 301, 3999, 4000, 4001, 65535])])
    got = oracle.range_gate(d, 300.0, 4000.0)
    want = np.where((d >= 300) & (d <= 4000), d, 0).astype(np.float32)
    assert np.array_equal(got, want)


# ---------------------------------------------------------------- A4 bilateral

@pytest.mark.parametrize("r", [1, 2, 3])
def test_bilateral_constant_image_and_border_ring(oracle, r):
    H, W = 24, 32
    d = np.full((H, W), 1234.5, np.float32)
    out = oracle.bilateral(d, r, 4.5, 30.0)
    inner = out[r:H - r, r:W - r]  # rows r..H-r-1: every tap is in range
    np.testing.assert_allclose(inner, 1234.5, rtol=2e-7)
    ring = np.ones((H, W), bool)
    ring[r:H - r + 1, r:W - r + 1] = False  # processed: r <= y <= H-r, r <= x <= W-r
    assert np.all(out[ring] == 0)


def test_bilateral_preserves_a_linear_ramp(oracle):
    # symmetric taps with equal range weights cancel around the centre
    H, W, r = 20, 28, 2
    yy, xx = np.mgrid[0:H, 0:W]
    d = (1000.0 + 3.0 * xx + 2.0 * yy).astype(np.float32)
    out = oracle.bilateral(d, r, 4.5, 30.0)
    np.testing.assert_allclose(out[r:H - r, r:W - r], d[r:H - r, r:W - r], rtol=1e-6)


def test_bilateral_keeps_a_depth_edge(oracle):
    # the range term gives the far side ~exp(-1000^2/1800) weight: no blur
    H, W, r = 16, 16, 3
    d = np.full((H, W), 1000.0, np.float32)
    d[:, 8:] = 2000.0
    out = oracle.bilateral(d, r, 4.5, 30.0)
    np.testing.assert_allclose(out[r:H - r, r:8], 1000.0, atol=1e-3)
    np.testing.assert_allclose(out[r:H - r, 8:W - r], 2000.0, atol=1e-3)


def test_bilateral_fp64_brute_force_8x8(oracle):
    """Brute force of bilateral_filtering.cu:3-47 in FP64, including the rows
    and columns whose window runs off the image (taps read as 0) and a
    centre pixel of 0."""
    rng = np.random.default_rng(5)
    H, W, r, ss, sr = 8, 8, 2, 4.5, 30.0
    d = rng.uniform(800, 900, (H, W)).astype(np.float32)
    d[4, 4] = 0.0
    d[2, 5] = 0.0
    out = oracle.bilateral(d, r, ss, sr)
    flat = d.ravel().astype(np.float64)
    want = np.zeros((H, W))
    for y in range(r, H - r + 1):
        for x in range(r, W - r + 1):
            c = float(d[y, x])
            num = den = 0.0
            for dy in range(-r, r + 1):
                for dx in range(-r, r + 1):
                    i = (y + dy) * W + (x + dx)  # linear read, 0 past the end
                    v = flat[i] if 0 <= i < H * W else 0.0
                    w = math.exp(-(dx * dx + dy * dy) / (2 * ss * ss)) * \
                        math.exp(-(v - c) ** 2 / (2 * sr * sr)) + 1e-9
                    num += w * v
                    den += w
            want[y, x] = num / den
    np.testing.assert_allclose(out, want, rtol=2e-6, atol=1e-6)


# ---------------------------------------------------------------- A3 vertex/normal

def test_vertex_normals_fronto_parallel_plane(kf, oracle):
    H, W, Z = 12, 16, 1500.0
    K = intr(100.0, 7.5, 5.5)
    d = np.full((H, W), Z, np.float32)
    d[6, 9] = 0.0  # a hole
    r = oracle.vertex_normals(d, kf.mat3_inverse(K).ravel())
    y, x = np.mgrid[0:H, 0:W]
    vx = Z * (x + 1 - 7.5) / 100.0
    vy = Z * (y + 1 - 5.5) / 100.0
    mask = np.ones((H, W), np.uint8)
    mask[-1, :] = 0
    mask[:, -1] = 0
    mask[6, 9] = 0
    assert np.array_equal(r["mask"], mask)
    # computed where the pixel and its right and lower neighbours are all != 0
    full = mask.astype(bool).copy()
    full[6, 8] = full[5, 9] = False
    np.testing.assert_allclose(r["vx"][full], vx[full], rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(r["vy"][full], vy[full], rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(r["nz"][full], -1.0, atol=1e-6)
    np.testing.assert_allclose(r["nx"][full], 0.0, atol=1e-6)
    np.testing.assert_allclose(r["ny"][full], 0.0, atol=1e-6)
    # mask set before the neighbour test, vertex and normal left 0
    for (yy, xx) in ((6, 8), (5, 9)):
        assert r["mask"][yy, xx] == 1 and r["vx"][yy, xx] == 0 and r["nz"][yy, xx] == 0


def test_vertex_normals_tilted_plane_orientation(kf, oracle):
    # plane z = c + s*x seen from the origin; the normal is (s, 0, -1)/|.|
    H, W, c, s = 10, 14, 1200.0, 0.5
    fx, cx, cy = 80.0, 6.5, 4.5
    K = intr(fx, cx, cy)
    y, x = np.mgrid[0:H, 0:W]
    a = (x + 1 - cx) / fx
    d = (c / (1 - s * a)).astype(np.float32)
    r = oracle.vertex_normals(d, kf.mat3_inverse(K).ravel())
    n = np.array([s, 0, -1]) / math.sqrt(1 + s * s)
    sl = (slice(0, H - 1), slice(0, W - 1))
    for k, v in zip(("nx", "ny", "nz"), n):
        np.testing.assert_allclose(r[k][sl], v, atol=2e-3)


# ---------------------------------------------------------------- A8 integration

def _axis_setup():
    vol = vol_desc((16, 16, 40), (-8, -8, -10), 10.0)
    K = intr(100.0, 32.0, 24.0)
    T = pose(t=(5, 5, 0))  # voxel column x = y = 5 lies on the optical axis
    return vol, K, T, 48, 64


def test_integrate_on_axis_column_is_depth_minus_z(oracle):
    vol, K, T, H, W = _axis_setup()
    prm = band(30, -30, 300)
    D = 200.0
    t, w = oracle.new_volume(vol.dims, "f32")
    oracle.integrate(np.full((H, W), D, np.float32), K, T, vol, prm, True, t, w)
    t3, w3 = t.reshape(vol.dims), w.reshape(vol.dims)
    col_t, col_w = t3[8, 8], w3[8, 8]
    z = (np.arange(40) - 10) * 10.0 + 5.0
    frk = D - z
    inband = (z > 0) & (frk >= -30) & (frk <= 300)
    want = np.where(frk > 0, np.minimum(1, frk / 30), np.maximum(-1, frk / 30))
    assert np.array_equal(col_t[inband], want[inband].astype(np.float32))
    assert np.all(col_w[inband] == 1)
    # behind the camera (z < 0) the front test rejects; outside the band NaN
    assert np.all(np.isnan(col_t[~inband])) and np.all(np.isnan(col_w[~inband]))
    assert np.isnan(col_t[z < 0]).all()


def test_integrate_running_average_and_weight_count(oracle):
    vol, K, T, H, W = _axis_setup()
    prm = band(30, -30, 300)
    t, w = oracle.new_volume(vol.dims, "f32")
    for i, D in enumerate((200.0, 200.0, 210.0)):
        oracle.integrate(np.full((H, W), D, np.float32), K, T, vol, prm, i == 0, t, w)
    t3, w3 = t.reshape(vol.dims), w.reshape(vol.dims)
    z = (np.arange(40) - 10) * 10.0 + 5.0

    def f(D):
        fr = D - z
        return np.where(fr > 0, np.minimum(1, fr / 30), np.maximum(-1, fr / 30)), \
            (z > 0) & (fr >= -30) & (fr <= 300)

    (f1, b1), (f3, b3) = f(200.0), f(210.0)
    both = b1 & b3
    np.testing.assert_allclose(t3[8, 8][both], ((2 * f1 + f3) / 3)[both], rtol=1e-6, atol=1e-7)
    assert np.all(w3[8, 8][both] == 3)
    only3 = b3 & ~b1  # first seen in frame 3: NaN counted as 0 with weight 0
    np.testing.assert_allclose(t3[8, 8][only3], f3[only3], rtol=1e-6)
    assert np.all(w3[8, 8][only3] == 1)


def test_integrate_invariants_and_fp64_twin(kf, oracle):
    rng = np.random.default_rng(6)
    vol = vol_desc((40, 40, 36), (-20, -20, 4), 20.0)
    K = intr(60.0, 31.5, 23.5)
    H, W = 48, 64
    prm = band(60, -60, 600)
    t32, w32 = oracle.new_volume(vol.dims, "f32")
    t64, w64 = oracle.new_volume(vol.dims, "f64")
    n_obs = np.zeros(int(np.prod(vol.dims)))
    for i in range(3):
        T = pose(t=rng.uniform(-30, 30, 3), rx=rng.uniform(-.05, .05), ry=rng.uniform(-.05, .05))
        d = rng.uniform(600, 700, (H, W)).astype(np.float32)
        d[rng.random((H, W)) < 0.1] = 0
        was = np.isfinite(w32).copy()
        oracle.integrate(d, K, T, vol, prm, i == 0, t32, w32)
        oracle.integrate(d, K, T, vol, prm, i == 0, t64, w64)
        n_obs += np.where(np.isfinite(w32), 1, 0) * 0  # weights checked below
        assert np.all(np.isfinite(w32[was]))  # observed voxels never return to NaN
    obs = np.isfinite(w32)
    assert obs.sum() > 10000
    assert np.all(np.abs(t32[obs]) <= 1)
    assert np.all(np.isnan(t32[~obs]))
    assert set(np.unique(w32[obs])) <= {1.0, 2.0, 3.0}
    # FP32 against the FP64 twin: same observation pattern apart from
    # voxels on a rounding boundary; values within 1e-5 of truncation
    agree = np.isfinite(w64) == obs
    assert agree.mean() > 0.999
    both = obs & np.isfinite(w64)
    assert np.quantile(np.abs(t32[both] - t64[both]), 0.999) < 1e-5


def test_integrate_f16_rounds_the_f32_update(oracle):
    vol, K, T, H, W = _axis_setup()
    prm = band(30, -30, 300)
    th, wh = oracle.new_volume(vol.dims, "f16")
    t, w = oracle.new_volume(vol.dims, "f32")
    oracle.integrate(np.full((H, W), 200.0, np.float32), K, T, vol, prm, True, th, wh)
    oracle.integrate(np.full((H, W), 200.0, np.float32), K, T, vol, prm, True, t, w)
    seen = np.isfinite(t)
    assert np.array_equal(th[seen], oracle.f32_to_f16_bits(t[seen]))
    assert np.all(wh[seen] == 1) and np.all(wh[~seen] == 0)


# ---------------------------------------------------------------- A9 raycast

def _plane_volume(dims, Z0, trunc=30.0):
    vol = vol_desc(dims, (-20, -20, 0), 10.0)
    _, _, z = centres(vol)
    t = np.clip((Z0 - z) / trunc, -1, 1).astype(np.float32).ravel()
    return vol, t


def test_raycast_plane_hit_point_normal_and_voxel(kf, oracle):
    Z0, trunc = 203.0, 30.0
    vol, t = _plane_volume((40, 40, 40), Z0, trunc)
    H, W, f, cx, cy = 48, 64, 50.0, 32.5, 24.5
    K = intr(f, cx, cy)
    d = np.full((H, W), Z0, np.float32)
    r = oracle.raycast(d, K, np.eye(4, dtype=np.float32), t, vol, trunc, 0)
    hit = r["hit"] >= 0
    assert hit.mean() > 0.9
    np.testing.assert_allclose(r["z"][hit], Z0, atol=1e-3)
    assert np.all(r["nx"][hit] == 0) and np.all(r["ny"][hit] == 0) and np.all(r["nz"][hit] == -1)
    # the hit voxel is where the ray enters the layer z in [200, 210):
    # brute force in FP64 away from x/y voxel boundaries
    y, x = np.mgrid[0:H, 0:W]
    Ki = kf.mat3_inverse(K).astype(np.float64)
    lx = Ki[0, 0] * (x + 1) + Ki[0, 2]
    ly = Ki[1, 1] * (y + 1) + Ki[1, 2]
    ex, ey = 200.0 * lx + 200.0, 200.0 * ly + 200.0  # entry point, volume-relative
    ix, iy = np.floor(ex / 10).astype(int), np.floor(ey / 10).astype(int)
    clear = (np.abs(ex / 10 - np.round(ex / 10)) > 1e-3) & (np.abs(ey / 10 - np.round(ey / 10)) > 1e-3)
    sel = hit & clear
    assert sel.sum() > 1000
    want = (ix * 40 + iy) * 40 + 20
    assert np.array_equal(r["hit"][sel], want[sel])


def test_raycast_reference_bound_checks(kf, oracle):
    """voxel_traversal.cu:233-281 test ZX-1 and ZY-1 against the z size: in a
    40x40x30 volume a plane hit is rejected where the x or y index at the
    last z step is >= 30, unless fix_bounds is set."""
    Z0, trunc = 203.0, 30.0
    vol, t = _plane_volume((40, 40, 30), Z0, trunc)
    H, W = 48, 64
    K = intr(50.0, 32.5, 24.5)
    d = np.full((H, W), Z0, np.float32)
    r0 = oracle.raycast(d, K, np.eye(4, dtype=np.float32), t, vol, trunc, 0)
    r1 = oracle.raycast(d, K, np.eye(4, dtype=np.float32), t, vol, trunc, 1)
    h0, h1 = r0["hit"], r1["hit"]
    ix, iy = h1 // (40 * 30), (h1 // 30) % 40
    far = (h1 >= 0) & ((ix >= 30) | (iy >= 30))
    assert far.sum() > 50
    assert np.all(h0[far] == -1)
    assert np.array_equal(h0[(h1 >= 0) & ~far], h1[(h1 >= 0) & ~far])


def test_raycast_sphere_surface_and_normals(kf, oracle):
    C, R, trunc = np.array([0.0, 0.0, 250.0]), 100.0, 30.0
    vol = vol_desc((40, 40, 40), (-20, -20, 0), 10.0)
    X, Y, Z = centres(vol)
    dist = np.sqrt((X - C[0]) ** 2 + (Y - C[1]) ** 2 + (Z - C[2]) ** 2)
    t = np.clip((dist - R) / trunc, -1, 1).astype(np.float32).ravel()
    H, W = 48, 64
    K = intr(60.0, 32.5, 24.5)
    y, x = np.mgrid[0:H, 0:W]
    Ki = kf.mat3_inverse(K).astype(np.float64)
    ray = np.stack([Ki[0, 0] * (x + 1) + Ki[0, 2], Ki[1, 1] * (y + 1) + Ki[1, 2],
                    np.ones((H, W))], -1)
    # analytic z-depth of the first intersection
    a = (ray ** 2).sum(-1)
    b = -2 * (ray @ C)
    c = C @ C - R * R
    disc = b * b - 4 * a * c
    s = np.where(disc > 0, (-b - np.sqrt(np.maximum(disc, 0))) / (2 * a), 0)
    d = np.where(disc > 0, s, 0).astype(np.float32)
    r = oracle.raycast(d, K, np.eye(4, dtype=np.float32), t, vol, trunc, 0)
    hit = r["hit"] >= 0
    assert hit.sum() > 0.8 * (d > 0).sum()
    # the hit voxel is the first one inside the surface: its centre lies
    # inside, at most a voxel diagonal deep
    hv = r["hit"][hit]
    ix, iy, iz = hv // 1600, (hv // 40) % 40, hv % 40
    cc = np.stack([(ix - 20) * 10 + 5, (iy - 20) * 10 + 5, iz * 10 + 5], -1)
    depth_in = np.linalg.norm(cc - C, axis=1) - R
    assert np.all(depth_in < 0) and np.all(depth_in > -math.sqrt(3) * vol.voxel_mm)
    # the per-axis interpolation from the previous voxel along each axis
    # (voxel_traversal.cu:318-349) is biased on oblique surfaces, but stays
    # within one voxel of the surface
    p = np.stack([r["x"], r["y"], r["z"]], -1)[hit]
    err = np.abs(np.linalg.norm(p - C, axis=1) - R)
    assert err.max() < vol.voxel_mm and np.median(err) < 0.2 * vol.voxel_mm
    n = np.stack([r["nx"], r["ny"], r["nz"]], -1)[hit]
    true_n = (p - C) / np.linalg.norm(p - C, axis=1, keepdims=True)
    ang = np.degrees(np.arccos(np.clip((n * true_n).sum(1), -1, 1)))
    assert np.quantile(ang, 0.99) < 10.0 and np.median(ang) < 3.0


def test_raycast_miss_cases(kf, oracle):
    Z0, trunc = 203.0, 30.0
    vol, t = _plane_volume((40, 40, 40), Z0, trunc)
    H, W = 8, 8
    K = intr(50.0, 4.5, 4.5)
    I4 = np.eye(4, dtype=np.float32)
    # depth 0: no ray; all-NaN volume: no hit; camera facing away: no hit
    r = oracle.raycast(np.zeros((H, W), np.float32), K, I4, t, vol, trunc, 0)
    assert np.all(r["hit"] == -1) and np.all(r["z"] == 0)
    r = oracle.raycast(np.full((H, W), Z0, np.float32), K, I4, np.full_like(t, np.nan), vol,
                       trunc, 0)
    assert np.all(r["hit"] == -1)
    back = pose(t=(0, 0, 100), ry=math.pi)
    r = oracle.raycast(np.full((H, W), 50.0, np.float32), K, back, t, vol, trunc, 0)
    assert np.all(r["hit"] == -1)


# ---------------------------------------------------------------- A5-A7 ICP

def _frame_from_depth(kf, oracle, d, K):
    fr = oracle.vertex_normals(d, kf.mat3_inverse(K).ravel())
    fr["vz"] = d
    return fr


def test_icp_identity_gives_zero_residual_and_identity_pose(kf, oracle):
    scene = kf.preset_scene(kf.PRESET_ROOM)
    K = intr(131.25, 79.5, 59.5)
    H, W = 120, 160
    d = kf.render(scene, K, np.eye(4, dtype=np.float32), H, W, noise_scale=0.0,
                  hole_frac=0.0).astype(np.float32)
    fr = _frame_from_depth(kf, oracle, d, K)
    model = {"x": fr["vx"], "y": fr["vy"], "z": d, "nx": fr["nx"], "ny": fr["ny"],
             "nz": fr["nz"]}
    I4 = np.eye(4, dtype=np.float32)
    r = oracle.icp(fr, model, K, I4, 3, 100.0, 20.0)
    assert np.all(r["sums"][:, 21:27] == 0)
    valid = (fr["mask"] == 1) & (d != 0) & np.isfinite(fr["nz"]) & (fr["nz"] != 0)
    inl = r["masks"][0].astype(bool)
    assert r["sums"][0][27] == inl.sum() and not np.any(inl & ~valid)
    # a pixel matched to itself has cos = n.n of two re-normalised floats,
    # which rounds above 1 for a share of pixels: acosf gives NaN there and
    # the gate (sensor_pose_estimation.cu:63-69) rejects it
    assert 0.5 * valid.sum() < inl.sum() < valid.sum()
    assert np.array_equal(r["T"], I4)
    # J^T J is symmetric positive semi-definite
    A = np.zeros((6, 6))
    k = 0
    for i in range(6):
        for j in range(i, 6):
            A[i, j] = A[j, i] = r["sums"][0][k]
            k += 1
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > -1e-9 * ev.max()


def _gate_maps(K, H, W, dz, theta_deg):
    """Frame vertices on z = 1000 with normals (0,0,-1); model points moved
    by dz along z and model normals tilted by theta about x."""
    y, x = np.mgrid[0:H, 0:W]
    Ki = np.linalg.inv(K.reshape(3, 3).astype(np.float64))
    vx = (1000.0 * (Ki[0, 0] * (x + 1) + Ki[0, 2])).astype(np.float32)
    vy = (1000.0 * (Ki[1, 1] * (y + 1) + Ki[1, 2])).astype(np.float32)
    vz = np.full((H, W), 1000.0, np.float32)
    fr = {"vx": vx, "vy": vy, "vz": vz, "nx": np.zeros((H, W), np.float32),
          "ny": np.zeros((H, W), np.float32), "nz": np.full((H, W), -1.0, np.float32),
          "mask": np.ones((H, W), np.uint8)}
    th = np.radians(theta_deg)
    model = {"x": vx, "y": vy, "z": (vz + dz).astype(np.float32),
             "nx": np.zeros((H, W), np.float32), "ny": np.sin(th).astype(np.float32),
             "nz": (-np.cos(th)).astype(np.float32)}
    return fr, model


def test_icp_distance_and_angle_gates(kf, oracle):
    K = intr(4.0, 2.5, 2.5)
    H, W = 4, 4
    dz = np.array([[0, 50, 99.9, 100], [100.01, 150, -100, -100.5],
                   [0, 0, 0, 0], [0, 0, 0, 0]], np.float32)
    th = np.array([[0, 0, 0, 0], [0, 0, 0, 0],
                   [10, 19.9, 20.1, 30], [-19.9, -20.1, 0, 45]], np.float64)
    fr, model = _gate_maps(K, H, W, dz, th)
    r = oracle.icp(fr, model, K, np.eye(4, dtype=np.float32), 1, 100.0, 20.0)
    want = np.array([[1, 1, 1, 1], [0, 0, 1, 0], [1, 1, 0, 0], [1, 0, 1, 0]], np.uint8)
    assert np.array_equal(r["masks"][0], want)


def test_icp_sums_exact_for_integer_rows_and_increment_closed_form(kf, oracle):
    """With K = [1 0 0; 0 1 0; 0 0 1] and integer depths, vertices, normals
    and residuals are integers, so the float sums are exact: they must equal
    sum a a^T and sum a b with a = [n x v; n], b = n.(s - v)."""
    rng = np.random.default_rng(7)
    H, W = 5, 6
    K = intr(1.0, 0.0, 0.0)
    y, x = np.mgrid[0:H, 0:W]
    d = rng.integers(2, 20, (H, W)).astype(np.float32)
    V = np.stack([d * (x + 1), d * (y + 1), d], -1).astype(np.float64)
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    n = axes[rng.integers(0, 6, (H, W))].astype(np.float64)
    k = rng.integers(-5, 6, (H, W)).astype(np.float64)
    S = V + k[..., None] * n
    f32 = lambda a: np.ascontiguousarray(a, np.float32)  # noqa: E731
    fr = {"vx": f32(V[..., 0]), "vy": f32(V[..., 1]), "vz": f32(d), "nx": f32(n[..., 0]),
          "ny": f32(n[..., 1]), "nz": f32(n[..., 2]), "mask": np.ones((H, W), np.uint8)}
    model = {"x": f32(S[..., 0]), "y": f32(S[..., 1]), "z": f32(S[..., 2]),
             "nx": f32(n[..., 0]), "ny": f32(n[..., 1]), "nz": f32(n[..., 2])}
    I4 = np.eye(4, dtype=np.float32)
    r = oracle.icp(fr, model, K, I4, 1, 100.0, 20.0)
    a = np.concatenate([np.cross(n, V), n], -1).reshape(-1, 6)
    b = (n * (S - V)).sum(-1).reshape(-1)
    ata = a.T @ a
    want = [ata[i, j] for i in range(6) for j in range(i, 6)] + list(a.T @ b) + [H * W]
    assert np.array_equal(r["sums"][0], np.array(want, np.float64))
    xs = np.linalg.solve(ata, a.T @ b)
    inc = np.array([[1, xs[2], -xs[1], xs[3]], [-xs[2], 1, xs[0], xs[4]],
                    [xs[1], -xs[0], 1, xs[5]], [0, 0, 0, 1]])
    np.testing.assert_allclose(r["T_iters"][0], inc, rtol=1e-6, atol=1e-6)


def test_icp_recovers_a_known_rigid_motion(kf, oracle):
    scene = kf.preset_scene(kf.PRESET_ROOM)
    K = intr(262.5, 159.5, 119.5)
    H, W = 240, 320
    I4 = np.eye(4, dtype=np.float32)
    T1 = pose(t=(8.0, -5.0, 6.0), rx=0.004, ry=-0.006, rz=0.003)
    d0 = kf.render(scene, K, I4, H, W, noise_scale=0.0, hole_frac=0.0).astype(np.float32)
    d1 = kf.render(scene, K, T1, H, W, noise_scale=0.0, hole_frac=0.0).astype(np.float32)
    m = _frame_from_depth(kf, oracle, d0, K)
    model = {"x": m["vx"], "y": m["vy"], "z": d0, "nx": m["nx"], "ny": m["ny"], "nz": m["nz"]}
    fr = _frame_from_depth(kf, oracle, d1, K)
    r = oracle.icp(fr, model, K, I4, 10, 100.0, 20.0)
    T = r["T"].astype(np.float64)
    assert np.abs(T[:3, 3] - T1[:3, 3]).max() < 1.0  # mm (depths are whole mm)
    R = T[:3, :3] @ T1[:3, :3].T.astype(np.float64)
    assert math.degrees(math.acos(min(1.0, (np.trace(R) - 1) / 2))) < 0.05
    assert np.all(r["status"] == 0)


def test_icp_solver_failure_uses_rhs(kf, oracle):
    # one inlier gives a rank-1 J^T J: the solve fails and the right-hand
    # side becomes the increment (cv::Cholesky's result ignored)
    K = intr(4.0, 2.5, 2.5)
    fr, model = _gate_maps(K, 4, 4, np.full((4, 4), 1000.0, np.float32),
                           np.zeros((4, 4)))
    model["z"][1, 2] = fr["vz"][1, 2] + 3.0
    r = oracle.icp(fr, model, K, np.eye(4, dtype=np.float32), 1, 100.0, 20.0)
    assert r["status"][0] != 0 and r["sums"][0][27] == 1
    b = r["sums"][0][21:27]
    inc = np.array([[1, b[2], -b[1], b[3]], [-b[2], 1, b[0], b[4]],
                    [b[1], -b[0], 1, b[5]], [0, 0, 0, 1]])
    np.testing.assert_allclose(r["T_iters"][0], inc, rtol=1e-6)


# ---------------------------------------------------------------- synthetic stream

def test_orbit_pose_is_rigid_and_slow(kf):
    assert np.array_equal(kf.orbit_pose(0), np.eye(4, dtype=np.float32))
    prev = kf.orbit_pose(0).astype(np.float64)
    for f in range(1, 301):
        T = kf.orbit_pose(f).astype(np.float64)
        R = T[:3, :3]
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)
        assert abs(np.linalg.det(R) - 1) < 1e-6
        assert np.linalg.norm(T[:3, 3] - prev[:3, 3]) <= 10.0
        dR = R @ prev[:3, :3].T
        assert math.degrees(math.acos(min(1.0, (np.trace(dR) - 1) / 2))) <= 0.5
        prev = T


def test_render_empty_room_and_one_sphere_closed_form(kf):
    import ctypes as C
    K = intr(100.0, 31.5, 23.5)
    H, W = 48, 64
    rmin = (C.c_float * 3)(-1000, -1000, -500)
    rmax = (C.c_float * 3)(1000, 1000, 2500)
    s = kf.SynthScene()
    kf.lib.kf_synth_build_scene(0, 0, 0, rmin, rmax, C.byref(s))
    d, clean = kf.render(s, K, np.eye(4, dtype=np.float32), H, W, noise_scale=0.0,
                         hole_frac=0.0, clean=True)
    # the far wall z = 2500 fills the central pixels (|x|,|y| < 1000 at z = 2500)
    np.testing.assert_allclose(clean[10:38, 12:52], 2500.0, rtol=1e-6)
    assert np.all(d[10:38, 12:52] == 2500)
    s.n_spheres = 1
    s.spheres[0][:] = [40.0, -30.0, 1200.0, 300.0]
    _, clean = kf.render(s, K, np.eye(4, dtype=np.float32), H, W, noise_scale=0.0,
                         hole_frac=0.0, clean=True)
    y, x = np.mgrid[0:H, 0:W]
    ray = np.stack([(x + 1 - 31.5) / 100.0, (y + 1 - 23.5) / 100.0, np.ones((H, W))], -1)
    Cc = np.array([40.0, -30.0, 1200.0])
    a = (ray ** 2).sum(-1)
    b = -2 * (ray @ Cc)
    c = Cc @ Cc - 300.0 ** 2
    disc = b * b - 4 * a * c
    z = (-b - np.sqrt(np.maximum(disc, 0))) / (2 * a)
    on = disc > 1.0
    np.testing.assert_allclose(clean[on], z[on], rtol=1e-5)


def test_render_noise_and_holes_statistics(kf):
    import ctypes as C
    K = intr(100.0, 31.5, 23.5)
    H, W = 240, 320
    K = intr(200.0, 159.5, 119.5)
    rmin = (C.c_float * 3)(-5000, -5000, -500)
    rmax = (C.c_float * 3)(5000, 5000, 2000)
    s = kf.SynthScene()
    kf.lib.kf_synth_build_scene(0, 0, 0, rmin, rmax, C.byref(s))
    d, clean = kf.render(s, K, np.eye(4, dtype=np.float32), H, W, noise_scale=1.0,
                         hole_frac=0.02, clean=True, seed=1000)
    holes = d == 0
    assert 0.015 < holes.mean() < 0.025
    res = d[~holes].astype(np.float64) - clean[~holes]
    sigma = 1000 * (0.0012 + 0.0019 * (2.0 - 0.4) ** 2)  # mm at z = 2 m
    assert abs(res.std() - math.sqrt(sigma ** 2 + 1 / 12)) < 0.05 * sigma
    assert abs(res.mean()) < 0.1


def test_icp_robust_failed_solve_keeps_the_pose(kf, oracle):
    # the rank-1 case above: with the robust update the failed solve applies
    # no increment (status 1), and the pose stays exactly where it was
    K = intr(4.0, 2.5, 2.5)
    fr, model = _gate_maps(K, 4, 4, np.full((4, 4), 1000.0, np.float32), np.zeros((4, 4)))
    model["z"][1, 2] = fr["vz"][1, 2] + 3.0
    I4 = np.eye(4, dtype=np.float32)
    r = oracle.icp(fr, model, K, I4, 2, 100.0, 20.0, min_inliers=0)
    assert np.all(r["status"] == 1) and np.all(r["sums"][:, 27] == 1)
    assert np.array_equal(r["T"], I4) and np.array_equal(r["T_iters"][0], I4)


def test_icp_robust_inlier_gate(kf, oracle):
    """Known rigid motion: with the gate at 0 the robust update is the
    reference's (every solve succeeds); with the gate above the inlier count
    no iteration is applied (status 2) and the pose is unchanged."""
    scene = kf.preset_scene(kf.PRESET_ROOM)
    K = intr(131.25, 79.5, 59.5)
    H, W = 120, 160
    I4 = np.eye(4, dtype=np.float32)
    T1 = pose(t=(8.0, -5.0, 6.0), rx=0.004, ry=-0.006, rz=0.003)
    d0 = kf.render(scene, K, I4, H, W, noise_scale=0.0, hole_frac=0.0).astype(np.float32)
    d1 = kf.render(scene, K, T1, H, W, noise_scale=0.0, hole_frac=0.0).astype(np.float32)
    m = _frame_from_depth(kf, oracle, d0, K)
    model = {"x": m["vx"], "y": m["vy"], "z": d0, "nx": m["nx"], "ny": m["ny"], "nz": m["nz"]}
    fr = _frame_from_depth(kf, oracle, d1, K)
    ref = oracle.icp(fr, model, K, I4, 5, 100.0, 20.0)
    rob = oracle.icp(fr, model, K, I4, 5, 100.0, 20.0, min_inliers=0)
    assert np.all(ref["status"] == 0)
    assert np.array_equal(ref["T_iters"], rob["T_iters"]) and np.array_equal(rob["status"], ref["status"])
    n0 = int(ref["sums"][0][27])
    gated = oracle.icp(fr, model, K, I4, 5, 100.0, 20.0, min_inliers=n0 + 1)
    assert np.all(gated["status"] == kf.KF_ICP_STATUS_GATED)
    assert np.array_equal(gated["T"], I4)
    # a gate between the counts of the first and the later iterations: the
    # first iteration is applied only if its count reaches the gate
    part = oracle.icp(fr, model, K, I4, 5, 100.0, 20.0, min_inliers=n0)
    assert part["status"][0] == 0 and np.array_equal(part["T_iters"][0], ref["T_iters"][0])


def test_icp_early_stop(kf, oracle):
    """Opt-in early stop (no reference counterpart): with thresholds, the
    iterations end right after the first increment whose rotation parameters
    and translation are all below them, and the pose is that iteration's.
    The increment of iteration i is T_i T_(i-1)^-1 (FP64 here; the small-angle
    matrix [[1, x2, -x1, x3], [-x2, 1, x0, x4], [x1, -x0, 1, x5]]), so the
    stop decision is checked against the recorded poses: a wrong index or
    sign in the test fails the first or the last of these checks."""
    scene = kf.preset_scene(kf.PRESET_ROOM)
    K = intr(131.25, 79.5, 59.5)
    H, W = 120, 160
    I4 = np.eye(4, dtype=np.float32)
    T1 = pose(t=(8.0, -5.0, 6.0), rx=0.004, ry=-0.006, rz=0.003)
    d0 = kf.render(scene, K, I4, H, W, noise_scale=0.0, hole_frac=0.0).astype(np.float32)
    d1 = kf.render(scene, K, T1, H, W, noise_scale=0.0, hole_frac=0.0).astype(np.float32)
    m = _frame_from_depth(kf, oracle, d0, K)
    model = {"x": m["vx"], "y": m["vy"], "z": d0, "nx": m["nx"], "ny": m["ny"], "nz": m["nz"]}
    fr = _frame_from_depth(kf, oracle, d1, K)
    full = oracle.icp(fr, model, K, I4, 10, 100.0, 20.0)
    assert full["iters_done"] == 10
    assert oracle.icp(fr, model, K, I4, 10, 100.0, 20.0, stop=(0.0, 0.0))["iters_done"] == 10

    def increment(Ta, Tb):  # x of the increment that took Tb to Ta
        D = Ta.astype(np.float64) @ np.linalg.inv(Tb.astype(np.float64))
        return np.array([D[1, 2], D[2, 0], D[0, 1]]), D[:3, 3]

    prev = [I4] + list(full["T_iters"])
    incs = [increment(prev[i + 1], prev[i]) for i in range(10)]
    # the motion is recovered in a few iterations, then the increments shrink
    rot_stop, trans_stop = 5e-6, 5e-3
    small = [bool(np.abs(r).max() < rot_stop and np.abs(t).max() < trans_stop) for r, t in incs]
    assert not small[0] and any(small)
    k = small.index(True)  # the first small increment: iteration k is the last
    # far from the thresholds, so the FP64 estimate decides as the float one
    r, t = incs[k]
    assert np.abs(r).max() < 0.5 * rot_stop and np.abs(t).max() < 0.5 * trans_stop
    r, t = incs[k - 1]
    assert np.abs(r).max() > 2 * rot_stop or np.abs(t).max() > 2 * trans_stop
    got = oracle.icp(fr, model, K, I4, 10, 100.0, 20.0, stop=(rot_stop, trans_stop))
    assert got["iters_done"] == k + 1
    assert np.array_equal(got["T_iters"], full["T_iters"][:k + 1])
    assert np.array_equal(got["T"], full["T_iters"][k])
    # recovered motion within the small-angle solve's accuracy
    assert np.abs(got["T"][:3, 3] - T1[:3, 3]).max() < 0.5
