"""Pins for the oracle's trilinear raycast mode and its FP64 adjoint (no GPU).

The mode has no reference counterpart, so it is pinned by what the
mathematics fixes: a planar TSDF is linear, so trilinear interpolation and
the secant are exact; on spheres the surface error must be far below the
reference interpolation's; and the hand-written adjoint must agree with
FP64 central finite differences of the FP64 forward pass, for the se(3)
pose and for the TSDF (pixels whose discrete state -- crossing voxel,
sample index, cells -- changes under the +-h perturbation are excluded).
"""
import math
from types import SimpleNamespace

import numpy as np

from conftest import ROOT  # noqa: F401
from se3 import se3_exp, sphere_tsdf


def vol_desc(dims, start, voxel):
    return SimpleNamespace(dims=list(dims), start=list(start), voxel_mm=float(voxel))


def intr(f, cx, cy):
    return np.array([f, 0, cx, 0, f, cy, 0, 0, 1], np.float32)


def test_trilinear_plane_is_exact(oracle):
    Z0, trunc = 203.0, 30.0
    vol = vol_desc((40, 40, 40), (-20, -20, 0), 10.0)
    z = (np.arange(40) + 0) * 10.0 + 5.0
    t = np.broadcast_to(np.clip((Z0 - z) / trunc, -1, 1), (40, 40, 40)).astype(np.float32).ravel()
    H, W = 48, 64
    K = intr(50.0, 32.5, 24.5)
    d = np.full((H, W), Z0, np.float32)
    for T in (np.eye(4, dtype=np.float32), np.eye(4)):
        r = oracle.raycast_tri(d, K, T, t.astype(np.float64) if T.dtype == np.float64 else t,
                               vol, trunc)
        hit = r["hit"] >= 0
        assert hit.mean() > 0.95
        tol = 1e-9 if T.dtype == np.float64 else 3e-5
        np.testing.assert_allclose(r["depth"][hit], Z0, atol=tol)
        np.testing.assert_allclose(r["z"][hit], Z0, atol=tol)
        assert np.all(r["nz"][hit] == -1) and np.all(r["nx"][hit] == 0)


def test_trilinear_sphere_surface_and_normals(kf, oracle):
    C, R, trunc = np.array([0.0, 0.0, 250.0]), 100.0, 30.0
    vol = vol_desc((40, 40, 40), (-20, -20, 0), 10.0)
    t = sphere_tsdf(vol, [(0.0, 0.0, 250.0, 100.0)], trunc)
    H, W = 48, 64
    K = intr(60.0, 32.5, 24.5)
    y, x = np.mgrid[0:H, 0:W]
    ray = np.stack([(x + 1 - 32.5) / 60.0, (y + 1 - 24.5) / 60.0, np.ones((H, W))], -1)
    a = (ray ** 2).sum(-1)
    b = -2 * (ray @ C)
    disc = b * b - 4 * a * (C @ C - R * R)
    z = np.where(disc > 0, (-b - np.sqrt(np.maximum(disc, 0))) / (2 * a), 0)
    r = oracle.raycast_tri(z.astype(np.float32), K, np.eye(4), t, vol, trunc)
    hit = r["hit"] >= 0
    assert hit.sum() > 0.95 * (z > 0).sum()
    p = np.stack([r["x"], r["y"], r["z"]], -1)[hit]
    err = np.abs(np.linalg.norm(p - C, axis=1) - R)
    # the reference interpolation reaches 0.87 voxel on this sphere
    assert err.max() < 0.05 * vol.voxel_mm
    n = np.stack([r["nx"], r["ny"], r["nz"]], -1)[hit]
    tn = (p - C) / np.linalg.norm(p - C, axis=1, keepdims=True)
    ang = np.degrees(np.arccos(np.clip((n * tn).sum(1), -1, 1)))
    assert ang.max() < 5.0
    # depth = camera z of the vertex
    np.testing.assert_allclose(r["depth"][hit], r["z"][hit], atol=1e-9)


def _scene():
    trunc = 30.0
    vol = vol_desc((40, 40, 40), (-20, -20, 0), 10.0)
    F = sphere_tsdf(vol, [(0, 0, 260.0, 90.0), (70, -40, 200.0, 50.0), (-60, 50, 230.0, 60.0)],
                    trunc)
    H, W = 48, 64
    K = intr(60.0, 32.5, 24.5)
    T0 = np.eye(4)
    T0[:3, :3] = se3_exp([0.02, -0.03, 0.01, 0, 0, 0])[:3, :3]
    T0[:3, 3] = (3.0, -2.0, -10.0)
    return trunc, vol, F, H, W, K, T0


def _guide(oracle, trunc, vol, F, H, W, K, T0, rng):
    r0 = oracle.raycast_tri(np.full((H, W), 150.0, np.float32), K, T0, F, vol, trunc)
    return np.where(r0["hit"] >= 0, r0["depth"] + rng.uniform(-4, 4, (H, W)), 0).astype(
        np.float32)


def test_adjoint_matches_fp64_finite_differences(oracle):
    trunc, vol, F, H, W, K, T0 = _scene()
    rng = np.random.default_rng(11)
    guide = _guide(oracle, trunc, vol, F, H, W, K, T0, rng)
    a, b = rng.normal(size=(H, W)), rng.normal(size=(H, W, 3))
    base = oracle.raycast_tri(guide, K, T0, F, vol, trunc)
    assert (base["hit"] >= 0).sum() > 1500

    def val(T, Fm):
        r = oracle.raycast_tri(guide, K, T, Fm, vol, trunc)
        n = np.stack([r["nx"], r["ny"], r["nz"]], -1)
        same = (r["hit"] >= 0) & np.all(r["disc"] == base["disc"], -1) & (base["hit"] >= 0)
        return a * r["depth"] + (b * n).sum(-1), same

    h = 1e-4
    pairs = []
    stable = base["hit"] >= 0
    for i in range(6):
        e = np.zeros(6)
        e[i] = h
        vp, sp = val(se3_exp(e) @ T0, F)
        vm, sm = val(se3_exp(-e) @ T0, F)
        pairs.append((vp, vm))
        stable &= sp & sm
    u = rng.normal(size=F.shape) * (np.abs(F) < 1)
    vpF, spF = val(T0, F + h * u)
    vmF, smF = val(T0, F - h * u)
    stable &= spF & smF
    assert stable.sum() > 0.9 * (base["hit"] >= 0).sum()
    am, bm = np.where(stable, a, 0), np.where(stable[..., None], b, 0)
    adj = oracle.raycast_tri(guide, K, T0, F, vol, trunc, grad_depth=am, grad_normal=bm)
    fd = np.array([((vp - vm)[stable]).sum() / (2 * h) for vp, vm in pairs])
    assert np.linalg.norm(adj["grad_xi"] - fd) / np.linalg.norm(fd) < 1e-4
    fdF = ((vpF - vmF)[stable]).sum() / (2 * h)
    assert abs((adj["grad_tsdf"] * u).sum() - fdF) / abs(fdF) < 1e-4
    # single voxels next to the surface, one at a time
    gF = adj["grad_tsdf"]
    for v in np.argsort(-np.abs(gF))[:3]:
        e = np.zeros_like(F)
        e[v] = h
        vp, sp = val(T0, F + e)
        vm, sm = val(T0, F - e)
        ok = stable & sp & sm
        fd_v = ((vp - vm)[ok]).sum() / (2 * h)
        a2, b2 = np.where(ok, a, 0), np.where(ok[..., None], b, 0)
        g2 = oracle.raycast_tri(guide, K, T0, F, vol, trunc, grad_depth=a2,
                                grad_normal=b2)["grad_tsdf"][v]
        assert abs(g2 - fd_v) <= 1e-4 * abs(fd_v) + 1e-8


def test_adjoint_is_linear_in_the_seeds(oracle):
    trunc, vol, F, H, W, K, T0 = _scene()
    rng = np.random.default_rng(12)
    guide = _guide(oracle, trunc, vol, F, H, W, K, T0, rng)
    a1, a2 = rng.normal(size=(H, W)), rng.normal(size=(H, W))
    r1 = oracle.raycast_tri(guide, K, T0, F, vol, trunc, grad_depth=a1)
    r2 = oracle.raycast_tri(guide, K, T0, F, vol, trunc, grad_depth=a2)
    r3 = oracle.raycast_tri(guide, K, T0, F, vol, trunc, grad_depth=2 * a1 - a2)
    np.testing.assert_allclose(r3["grad_xi"], 2 * r1["grad_xi"] - r2["grad_xi"], rtol=1e-9,
                               atol=1e-9 * np.abs(r1["grad_xi"]).max())
    np.testing.assert_allclose(r3["grad_tsdf"], 2 * r1["grad_tsdf"] - r2["grad_tsdf"],
                               atol=1e-9 * np.abs(r1["grad_tsdf"]).max())
    assert math.isfinite(float(np.abs(r1["grad_tsdf"]).sum()))
