"""The host/device math shared by the kernels and the oracle, pinned against
numpy/libm and closed forms (no GPU needed)."""
import math

import numpy as np
import pytest

from conftest import ROOT  # noqa: F401  (puts the repo on sys.path)


def ordered(a):
    """float32 -> integers in the same order, one step per ulp, +0 == -0."""
    i = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    return np.where(i < 0, -(i & 0x7fffffff), i)


def ulp_diff(a, b):
    return np.abs(ordered(a) - ordered(b))


def test_expf_within_one_ulp_of_correctly_rounded(kf):
    # the bilateral filter calls exp on -(dx^2+dy^2)/(2 s^2) and -D^2/(2 s^2)
    rng = np.random.default_rng(0)
    xs = np.concatenate([-rng.exponential(4.0, 4000), rng.uniform(-87.0, 88.0, 4000),
                         -np.linspace(0, 1e-6, 50), [0.0, -0.5, -1.0, -20.0, -87.3, -103.0]])
    xs = xs.astype(np.float32)
    got = np.array([kf.lib.kf_expf_abi(float(x)) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64)).astype(np.float32)
    assert ulp_diff(got, want).max() <= 1
    assert kf.lib.kf_expf_abi(0.0) == 1.0
    assert kf.lib.kf_expf_abi(-200.0) == 0.0
    assert kf.lib.kf_expf_abi(float("-inf")) == 0.0


def test_mat4_inverse_of_rigid_pose_is_transpose_form(kf):
    rng = np.random.default_rng(1)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        t = rng.uniform(-2000, 2000, 3)
        T = np.eye(4)
        T[:3, :3], T[:3, 3] = R, t
        T32 = T.astype(np.float32)
        got = kf.mat4_inverse(T32)
        want = np.linalg.inv(T32.astype(np.float64))
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-7 * 2000 + 1e-6)
        # rounded once from FP64: every element within half an ulp of float(want)
        assert ulp_diff(got, want.astype(np.float32)).max() <= 1
        np.testing.assert_allclose(got[:3, :3], R.T, atol=1e-6)


def test_mat4_inverse_rejects_singular(kf):
    from kinect_fusion_slam_amd._abi import fptr
    out = np.zeros(16, np.float32)
    assert kf.lib.kf_mat4_inverse(fptr(np.zeros(16, np.float32)), fptr(out)) != 0


def test_mat3_inverse_of_intrinsics_closed_form(kf):
    for fx, fy, cx, cy in ((525.0, 525.0, 319.5, 239.5), (524.639, 523.503, 316.625, 256.2318),
                           (1000.0, 1000.0, 639.5, 359.5)):
        K = np.array([fx, 0, cx, 0, fy, cy, 0, 0, 1], np.float32)
        k = K.astype(np.float64)
        want = np.array([[1 / k[0], 0, -k[2] / k[0]], [0, 1 / k[4], -k[5] / k[4]], [0, 0, 1]])
        got = kf.mat3_inverse(K)
        assert ulp_diff(got, want.astype(np.float32)).max() <= 1


def test_ldlt6_matches_numpy_and_fails_on_non_pd(kf):
    rng = np.random.default_rng(2)
    for _ in range(50):
        J = rng.normal(size=(40, 6)) * rng.uniform(0.1, 1000, 6)
        A = J.T @ J
        b = rng.normal(size=6) * 100
        st, x = kf.ldlt6(A, b)
        assert st == 0
        want = np.linalg.solve(A, b)
        res = np.linalg.norm(A @ x - b) / (np.linalg.norm(A) * np.linalg.norm(x) + 1e-300)
        assert res < 1e-13
        np.testing.assert_allclose(x, want, rtol=np.linalg.cond(A) * 1e-14, atol=0)
    # zero matrix, indefinite matrix, rank deficiency: status != 0
    assert kf.ldlt6(np.zeros((6, 6)), np.ones(6))[0] != 0
    A = np.eye(6)
    A[3, 3] = -1.0
    assert kf.ldlt6(A, np.ones(6))[0] != 0
    J = rng.normal(size=(3, 6))
    assert kf.ldlt6(J.T @ J, np.ones(6))[0] != 0


def test_f16_conversions_match_ieee_half(oracle):
    # every half pattern widens exactly
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    got = oracle.f16_bits_to_f32(h)
    want = h.view(np.float16).astype(np.float32)
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert same.all()
    # narrowing rounds to nearest even, incl. ties, subnormals, overflow, NaN
    rng = np.random.default_rng(3)
    bits = rng.integers(0, 2 ** 32, 200000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    halfs = np.arange(0, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float64)
    ties = ((halfs[:-1] + halfs[1:]) / 2).astype(np.float32)  # exact midpoints in float
    x = np.concatenate([x, ties, -ties, np.float32([65504, 65519.99, 65520, 1e-8, 5.96e-8,
                                                      2.98e-8, 2.9802322e-08, np.inf,
                                                      -np.inf])])
    got = oracle.f32_to_f16_bits(x)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    nan = np.isnan(x)
    assert np.array_equal(got[~nan], want[~nan])
    assert np.all((got[nan] & 0x7c00) == 0x7c00) and np.all((got[nan] & 0x3ff) != 0)


@pytest.mark.parametrize("gate", [20.0, 30.0, 10.0, 45.0])
def test_gate_thresholds_equal_the_acos_tests(kf, oracle, gate):
    lo, hi = kf.gate_thresholds(gate)
    pif = np.float32(3.141592654)
    k = np.float32(180.0) / pif

    def icp_pass(c):  # acosf(c) * (180/pi_f) <= gate in float, acosf correctly rounded
        a = np.float32(math.acos(float(np.float32(c)))) * k
        return np.float32(0) <= a <= np.float32(gate)

    def flip(c):  # acos((double)c) * (double)(180/pi_f) > 90
        return math.acos(float(np.float32(c))) * float(k) > 90.0

    lo32, hi32 = np.float32(lo), np.float32(hi)
    assert icp_pass(lo32) and not icp_pass(np.nextafter(lo32, np.float32(-1)))
    assert flip(np.nextafter(hi32, np.float32(-1))) and not flip(hi32)
    rng = np.random.default_rng(4)
    cs = np.concatenate([rng.uniform(-1, 1, 3000), np.float32(lo32) + np.arange(-50, 50) *
                         np.float32(6e-8), np.float32(hi32) + np.arange(-50, 50) *
                         np.float32(6e-9), [-1.0, 1.0, 0.0]]).astype(np.float32)
    cs = cs[(cs >= -1) & (cs <= 1)]
    for c in cs:
        assert oracle.icp_angle_gate(c, gate) == int(lo32 <= c <= 1), c
        assert oracle.normal_flip(c) == int(-1 <= c < hi32), c
        assert bool(icp_pass(c)) == bool(lo32 <= c <= 1), c


@pytest.mark.parametrize("gate", [20.0, 100.0, 0.5, 4000.0, 1e-30, 0.0, -1.0, np.inf, np.nan])
def test_dist_gate_threshold_equals_the_sqrt_test(kf, gate):
    """The ICP distance gate, sqrtf(d2) <= gate && sqrtf(d2) >= 0
    (sensor_pose_estimation.cu:47-48), against d2 <= threshold, with numpy's
    float32 sqrt (IEEE, correctly rounded) as the reference side: at the
    threshold, one float either side, random squares of distances around
    the gate, and the special values."""
    hi = np.float32(kf.dist_gate_threshold(gate))
    g = np.float32(gate)

    def passes(d2):
        with np.errstate(invalid="ignore"):
            d = np.sqrt(np.float32(d2))
        return bool(d <= g and d >= np.float32(0))

    if not passes(np.float32(0)):
        assert hi == np.float32(-1)
    else:
        assert passes(hi)
        if np.isfinite(hi):
            assert not passes(np.nextafter(hi, np.float32(np.inf)))
    rng = np.random.default_rng(7)
    scale = np.float32(gate) if np.isfinite(gate) and gate > 0 else np.float32(10.0)
    d = (rng.uniform(0, 2, 4000) * scale).astype(np.float32)
    cand = np.concatenate([d * d, [np.float32(0), np.float32(np.inf), np.float32(np.nan)]])
    if np.isfinite(hi) and hi >= 0:
        cand = np.concatenate([cand, hi + np.arange(-40, 41) * np.spacing(hi)])
    cand = cand.astype(np.float32)
    cand = cand[~(cand < 0)]  # a float sum of squares is >= +0 or NaN
    for s in cand:
        assert passes(s) == bool(s <= hi), (gate, s)
