"""GPU trilinear raycast and its adjoint, through the C ABI.

* forward (KF_RAYCAST_TRILINEAR): vertex, normal, depth and hit index
  bit-exact against the oracle's FP32 transcription (F32 and F16 volumes);
* adjoint: dL/dxi and dL/dTSDF against the oracle's FP64 adjoint of the same
  inputs, to 1e-4 relative (FP32 arithmetic and float atomics on the GPU),
  from a raycast's saved crossings, and without them (the adjoint then runs
  the trilinear raycast into buffers of its own first); the two against each
  other on the room workload, and a changed volume dropping the saved
  crossings;
* the 128^3 three-sphere workload at 320x240 with 16 poses jittered 2 cm /
  2 deg: gradients of an L2 depth loss through a torch.autograd.Function
  against FP64 central finite differences (h = 1e-4) of the oracle, for the
  se(3) pose and for the TSDF on a 64^3 sub-grid, to 1e-3 relative; and the
  adjoint's GPU time against the forward raycast's.
"""
import numpy as np
import pytest
import torch

from conftest import say
from se3 import se3_exp, sphere_tsdf

pytestmark = pytest.mark.gpu


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    ok = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b
    assert ok.all(), f"{what}: {int((~ok).sum())} of {a.size} differ"


def gated(oracle, frame, prm):
    d = oracle.range_gate(frame, prm.depth_min_mm, prm.depth_max_mm)
    return oracle.bilateral(d, prm.bilateral_radius, prm.sigma_s_px, prm.sigma_r_mm)


@pytest.mark.parametrize("precision", ["f32", "f16"])
def test_trilinear_forward_bitexact(kf, room_stream, oracle, precision):
    vol, prm, K, H, W, frames, gt = room_stream
    prec = kf.KF_TSDF_F32 if precision == "f32" else kf.KF_TSDF_F16
    say(f"[gpu] trilinear raycast {precision}")
    with kf.KinfuContext.from_preset(kf.PRESET_ROOM, precision=prec,
                                     raycast_mode=kf.KF_RAYCAST_TRILINEAR) as c:
        t, w = oracle.new_volume(vol.dims, precision)
        for i in (0, 1):
            oracle.integrate(gated(oracle, frames[i], prm), K, gt[i], c.vol, prm, i == 0, t, w)
        if precision == "f32":
            c.upload_volume(t, w)
        else:
            c.upload_volume(oracle.f16_bits_to_f32(t), np.where(w == 0, np.nan, w.astype(np.float32)))
        d = gated(oracle, frames[1], prm)
        c.upload_depth_f32(d)
        c.raycast(K, gt[1])
        got = c.model_maps()
        depth = c.rendered_depth()
    want = oracle.raycast_tri(d, K, gt[1], t, vol, prm.truncation_mm)
    assert_same(got["hit"], want["hit"], "hit")
    for k in ("x", "y", "z", "nx", "ny", "nz"):
        assert_same(got[k], want[k], k)
    assert_same(depth, want["depth"], "depth")
    assert (got["hit"] >= 0).mean() > 0.5


def _small_scene():
    vol = dict(dims=(40, 40, 40), start=(-20, -20, 0), voxel=10.0)
    trunc = 30.0
    K = np.array([60.0, 0, 32.5, 0, 60.0, 24.5, 0, 0, 1], np.float32)
    T0 = np.eye(4)
    T0[:3, :3] = se3_exp([0.02, -0.03, 0.01, 0, 0, 0])[:3, :3]
    T0[:3, 3] = (3.0, -2.0, -10.0)
    return vol, trunc, K, T0.astype(np.float32), 48, 64


@pytest.mark.parametrize("precision", ["f32", "f16"])
@pytest.mark.parametrize("path", ["saved", "no_saved"])
def test_adjoint_matches_fp64_oracle_adjoint(kf, oracle, path, precision):
    """saved: the adjoint right after the raycast at the same inputs, which
    starts from the raycast's saved crossings; no_saved: the depth uploaded again
    in between (same values), which drops them, so the adjoint first runs its
    own trilinear raycast and starts from that one's crossings.

    f16: the volume is stored as half TSDF + u16 weight.  The GPU forward is
    compared bit-exactly with the oracle's FP32 forward over the same half
    values, and the adjoint with the oracle's FP64 adjoint of the
    half-rounded TSDF (the stored values are what the gradient is taken
    of)."""
    geo, trunc, K, T0, H, W = _small_scene()
    vol, prm, _, _, _ = kf.preset(kf.PRESET_SPHERES_128)
    vol.dims[:] = list(geo["dims"])
    vol.start[:] = list(geo["start"])
    vol.voxel_mm = geo["voxel"]
    vol.precision = kf.KF_TSDF_F32 if precision == "f32" else kf.KF_TSDF_F16
    prm.truncation_mm, prm.sdf_min_mm, prm.sdf_max_mm = trunc, -trunc, 10 * trunc
    F = sphere_tsdf(vol, [(0, 0, 260.0, 90.0), (70, -40, 200.0, 50.0), (-60, 50, 230.0, 60.0)],
                    trunc).astype(np.float32)
    # what the forward reads: the float values, or their half bits
    Fs = F if precision == "f32" else oracle.f32_to_f16_bits(F)
    if precision == "f16":
        Fh = oracle.f16_bits_to_f32(Fs)  # the half-rounded values, exact in float
        fin = np.isfinite(F)
        band = fin & (np.abs(F) < 1)  # every voxel inside the band rounds
        assert np.array_equal(fin, np.isfinite(Fh)) and np.all(Fh[band] != F[band])
        F = Fh
    rng = np.random.default_rng(21)
    r0 = oracle.raycast_tri(np.full((H, W), 150.0, np.float32), K, T0, Fs, vol, trunc)
    guide = np.where(r0["hit"] >= 0, r0["depth"] + rng.uniform(-4, 4, (H, W)), 0).astype(
        np.float32)
    a, b = rng.normal(size=(H, W)), rng.normal(size=(H, W, 3))
    # pixels whose FP32 and FP64 discrete states agree
    f32 = oracle.raycast_tri(guide, K, T0, Fs, vol, trunc)
    f64 = oracle.raycast_tri(guide, K, T0.astype(np.float64), F.astype(np.float64), vol, trunc)
    same = (f32["hit"] >= 0) & np.all(f32["disc"] == f64["disc"], -1) & (f64["hit"] >= 0)
    assert same.sum() > 1500
    am, bm = np.where(same, a, 0), np.where(same[..., None], b, 0)
    want = oracle.raycast_tri(guide, K, T0.astype(np.float64), F.astype(np.float64), vol, trunc,
                              grad_depth=am, grad_normal=bm)
    say(f"[gpu] raycast adjoint ({path}, {precision}), 40^3 three-sphere volume")
    with kf.KinfuContext(vol, prm, H, W) as c:
        c.upload_volume(F, np.where(np.isnan(F), np.nan, 1.0).astype(np.float32))
        c.upload_depth_f32(guide)
        c.raycast(K, T0)
        assert_same(c.rendered_depth(), f32["depth"], "forward depth")
        if path == "no_saved":
            c.upload_depth_f32(guide)
        gF, gxi = c.raycast_adjoint(K, T0, am.astype(np.float32), bm.astype(np.float32))
        assert c.adjoint_used_saved() == (path == "saved")
    assert np.linalg.norm(gxi - want["grad_xi"]) / np.linalg.norm(want["grad_xi"]) < 1e-4
    assert np.linalg.norm(gF - want["grad_tsdf"]) / np.linalg.norm(want["grad_tsdf"]) < 1e-4


def test_adjoint_saved_crossings_match_no_saved(kf, room_stream):
    """Room workload after a 4-frame trilinear pipeline (rays finished by the
    second raycast pass included): the adjoint from a raycast's saved
    crossings against the adjoint without them (its own raycast first) at
    the same inputs, with and without dL/dTSDF, to the order of float
    additions; and a volume changed
    after the raycast (TSDF moved 60 mm along z, twice the window tri_hit
    searches around a crossing) must make the adjoint run its own raycast."""
    vol, prm, K, H, W, frames, gt = room_stream
    rng = np.random.default_rng(8)
    a = rng.normal(size=(H, W)).astype(np.float32)
    b = rng.normal(size=(H, W, 3)).astype(np.float32)
    say("[gpu] adjoint from saved crossings vs its own raycast, room 640x480")
    with kf.KinfuContext.from_preset(kf.PRESET_ROOM, raycast_mode=kf.KF_RAYCAST_TRILINEAR,
                                     stage_timing=1) as c:
        c.stream_upload(frames[:4])
        c.pipeline_run(K, 0, 4)
        T = c.pose()
        d = c.depth()

        def adjoint(saved, want_tsdf, reps=1):
            c.reset_stage_times()
            for _ in range(reps):
                if saved:
                    c.raycast(K, T)
                else:
                    c.upload_depth_f32(d)
                out = c.raycast_adjoint(K, T, a, b, want_tsdf=want_tsdf)
                assert c.adjoint_used_saved() == saved
            st = c.stage_times()["raycast_adjoint"]
            return out, st[0] / st[1]

        (gF_s, gxi_s), t_s = adjoint(True, True, 3)
        (gF_w, gxi_w), t_w = adjoint(False, True, 3)
        (_, gxi_sp), _ = adjoint(True, False)
        (_, gxi_wp), _ = adjoint(False, False)
        say(f"[gpu] adjoint with dL/dTSDF: saved {t_s * 1e3:.1f} us, own raycast "
            f"{t_w * 1e3:.1f} us")
        t, w = c.volume()
        n = vol.dims
        t2 = np.roll(t.reshape(n), 6, axis=2).ravel()
        w2 = np.roll(w.reshape(n), 6, axis=2).ravel()
        c.raycast(K, T)
        c.upload_volume(t2, w2)
        gF_x, gxi_x = c.raycast_adjoint(K, T, a, b)
        assert not c.adjoint_used_saved()
        c.raycast(K, T)
        gF_y, gxi_y = c.raycast_adjoint(K, T, a, b)
        assert c.adjoint_used_saved()
        # a different pose has no saved crossings either
        T2 = T.copy()
        T2[0, 3] += 1.0
        c.raycast_adjoint(K, T2, a, b, want_tsdf=False)
        assert not c.adjoint_used_saved()
    assert np.abs(gxi_w).max() > 0 and np.abs(gF_w).max() > 0
    assert np.linalg.norm(gxi_s - gxi_w) <= 1e-6 * np.linalg.norm(gxi_w)
    assert np.linalg.norm(gxi_sp - gxi_wp) <= 1e-6 * np.linalg.norm(gxi_wp)
    assert np.linalg.norm(gxi_sp - gxi_s) <= 1e-6 * np.linalg.norm(gxi_s)
    assert np.abs(gF_s - gF_w).max() <= 1e-5 * np.abs(gF_w).max()
    # changed volume: the adjoint's own raycast (x) and a fresh raycast's
    # crossings (y) agree,
    # and differ from the old volume's sums
    assert np.linalg.norm(gxi_x - gxi_y) <= 1e-6 * np.linalg.norm(gxi_y)
    assert np.abs(gF_x - gF_y).max() <= 1e-5 * np.abs(gF_y).max()
    assert np.linalg.norm(gxi_y - gxi_w) > 1e-2 * np.linalg.norm(gxi_w)


# ---------------------------------------------------------------- 128^3 spheres workload

def _spheres_workload(kf):
    vol, prm, K, H, W = kf.preset(kf.PRESET_SPHERES_128)
    rng = np.random.default_rng(3)
    spheres = [(rng.uniform(-250, 250), rng.uniform(-200, 200), rng.uniform(850, 1150),
                rng.uniform(120, 220)) for _ in range(3)]
    F = sphere_tsdf(vol, spheres, prm.truncation_mm).astype(np.float32)
    base, jit = [], []
    for j in range(16):
        a = 2 * np.pi * j / 16
        # unjittered poses on a 10 cm circle, looking along +z
        b = se3_exp([0.02 * np.sin(a), 0.02 * np.cos(a), 0, 100 * np.cos(a), 100 * np.sin(a), 0])
        w = rng.normal(size=3)
        v = rng.normal(size=3)
        xi = np.concatenate([np.radians(2.0) * w / np.linalg.norm(w), 20.0 * v / np.linalg.norm(v)])
        base.append(b.astype(np.float32))
        jit.append((se3_exp(xi) @ b).astype(np.float32))
    return vol, prm, K, H, W, F, base, jit


def test_spheres_workload_gradients_match_fp64_finite_differences(kf, oracle):
    """TSDF directions: |N(0,1)| weights and the FP64 adjoint's own direction,
    both on the 64^3 sub-grid.  A signed random direction is avoided: its
    directional derivative is a sum of cancelling terms, so the O(h^2) error
    of the few pixels near a kink dominates the relative error of the small
    total."""
    vol, prm, K, H, W, F, base, jit = _spheres_workload(kf)
    trunc = prm.truncation_mm
    Wt = np.where(np.isnan(F), np.nan, 1.0).astype(np.float32)
    F64 = F.astype(np.float64)
    h = 1e-4
    # targets and guides from the unjittered poses (FP32 oracle forward)
    targets = []
    for b in base:
        r = oracle.raycast_tri(np.full((H, W), 900.0, np.float32), K, b, F, vol, trunc)
        targets.append(np.where(r["hit"] >= 0, r["depth"], 0).astype(np.float32))
    sub = np.zeros(vol.dims, bool)
    sub[32:96, 32:96, 32:96] = True
    sub = sub.ravel() & (np.abs(F) < 1)
    rng = np.random.default_rng(5)
    u1 = np.abs(rng.normal(size=F.shape)) * sub

    def fwd64(T, Fm, tgt):
        return oracle.raycast_tri(tgt, K, T, Fm, vol, trunc)

    say("[gpu] 128^3 spheres: 16 jittered poses, finite differences on the CPU oracle")
    T64s = [T.astype(np.float64) for T in jit]
    base_r, masks = [], []
    for T, T64, tgt in zip(jit, T64s, targets):
        r0 = fwd64(T64, F64, tgt)
        r32 = oracle.raycast_tri(tgt, K, T, F, vol, trunc)
        base_r.append(r0)
        masks.append((r0["hit"] >= 0) & (tgt > 0) & (r32["hit"] >= 0) &
                     np.all(r32["disc"] == r0["disc"], -1))

    def perturbed(j, i, sgn, u=None):
        if u is None:
            e = np.zeros(6)
            e[i] = sgn * h
            r = fwd64(se3_exp(e) @ T64s[j], F64, targets[j])
        else:
            r = fwd64(T64s[j], F64 + sgn * h * u, targets[j])
        masks[j] &= (r["hit"] >= 0) & np.all(r["disc"] == base_r[j]["disc"], -1)
        return r["depth"]

    pose_d = [[[perturbed(j, i, s) for s in (1, -1)] for i in range(6)] for j in range(16)]
    u1_d = [[perturbed(j, 0, s, u1) for s in (1, -1)] for j in range(16)]
    # the FP64 adjoint's direction on the sub-grid
    g64 = np.zeros_like(F64)
    for j in range(16):
        gD = np.where(masks[j], base_r[j]["depth"] - targets[j], 0)
        g64 += oracle.raycast_tri(
            targets[j], K, T64s[j], F64, vol, trunc, grad_depth=gD)["grad_tsdf"]
    u2 = g64 * sub
    u2 /= np.abs(u2).max()
    u2_d = [[perturbed(j, 0, s, u2) for s in (1, -1)] for j in range(16)]
    fd_xi, fd_u1, fd_u2 = [], 0.0, 0.0
    for j in range(16):
        ok, tgt = masks[j], targets[j]
        assert ok.sum() > 0.5 * (tgt > 0).sum()
        L = lambda D: 0.5 * ((D - tgt) ** 2)[ok].sum()  # noqa: E731
        fd_xi.append([(L(p) - L(m)) / (2 * h) for p, m in pose_d[j]])
        fd_u1 += (L(u1_d[j][0]) - L(u1_d[j][1])) / (2 * h)
        fd_u2 += (L(u2_d[j][0]) - L(u2_d[j][1])) / (2 * h)
    fd_xi = np.array(fd_xi)

    say("[gpu] 128^3 spheres: autograd through kf_raycast / kf_raycast_adjoint")
    from raycast_autograd import render_depth
    with kf.KinfuContext.from_preset(kf.PRESET_SPHERES_128, stage_timing=1) as c:
        tsdf = torch.from_numpy(F.copy()).requires_grad_(True)
        g_xi = []
        total = 0.0
        for T64, tgt, ok in zip(T64s, targets, masks):
            xi = torch.zeros(6, dtype=torch.float64, requires_grad=True)
            D = render_depth(tsdf, xi, c, K, T64, Wt, tgt)
            loss = 0.5 * ((D - torch.from_numpy(tgt)) ** 2)[torch.from_numpy(ok)].sum()
            loss.backward()
            g_xi.append(xi.grad.numpy().copy())
            total += float(loss)
        g_xi = np.array(g_xi)
        gF = tsdf.grad.numpy().astype(np.float64)
        # cost of the adjoint against the forward raycast (hipEvents, same poses)
        c.reset_stage_times()
        c.upload_volume(F, Wt)
        for _ in range(3):
            for T, tgt in zip(jit, targets):
                c.upload_depth_f32(tgt)
                c.raycast(K, T)
                c.raycast_adjoint(K, T, np.ones((H, W), np.float32), want_tsdf=False)
        st = c.stage_times()
    rel_xi = np.linalg.norm(g_xi - fd_xi) / np.linalg.norm(fd_xi)
    rel_u1 = abs((gF * u1).sum() - fd_u1) / abs(fd_u1)
    rel_u2 = abs((gF * u2).sum() - fd_u2) / abs(fd_u2)
    fwd = st["raycast"][0] / st["raycast"][1]
    adj = st["raycast_adjoint"][0] / st["raycast_adjoint"][1]
    say(f"[gpu] spheres: loss {total:.4g}, rel err xi {rel_xi:.2e}, tsdf |random| "
        f"{rel_u1:.2e}, tsdf grad-direction {rel_u2:.2e}, forward {fwd * 1e3:.1f} us, "
        f"adjoint {adj * 1e3:.1f} us")
    assert rel_xi < 1e-3
    for j in range(16):
        assert np.linalg.norm(g_xi[j] - fd_xi[j]) <= 1e-3 * np.linalg.norm(fd_xi[j])
    assert rel_u1 < 1e-3 and rel_u2 < 1e-3
    assert adj <= 2.0 * fwd
