"""The host/device math shared by the kernels and the oracle, pinned against
numpy/libm and closed forms (no GPU needed)."""
import ctypes as C
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from conftest import ROOT


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


# ---------------------------------------------------------------------------
# Helpers reached only through the shared header (tests/shim/hd_math_shim.cpp
# exposes their host build): pinned against exact rational arithmetic,
# numpy and closed forms.

SHIM = os.path.join(ROOT, "tests", "shim", "libhd_math_shim.so")


@pytest.fixture(scope="module")
def hd():
    if not os.path.exists(SHIM):
        pytest.fail(f"{SHIM} is not built (make -C tests/shim)")
    lib = C.CDLL(SHIM)
    f, d = C.POINTER(C.c_float), C.POINTER(C.c_double)
    lib.hd_round_to_int.argtypes = [C.c_float]
    lib.hd_floor_to_int.argtypes = [C.c_float]
    lib.hd_mat4_mul.argtypes = [f, f, f]
    lib.hd_sym_index.argtypes = [C.c_int, C.c_int]
    lib.hd_icp_increment.argtypes = [d, f]
    lib.hd_icp_increment_robust.argtypes = [d, f, C.c_double]
    lib.hd_icp_update_gated.argtypes = [d, f, C.c_double]
    lib.hd_icp_converged.argtypes = [f, C.c_float, C.c_float]
    lib.hd_inv_norm3.argtypes = [C.c_float, C.c_float, C.c_float]
    lib.hd_inv_norm3.restype = C.c_float
    return lib


def _fp(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _edge_floats():
    halves = [k + 0.5 for k in range(-6, 6)]
    xs = [0.0, -0.0, 1.0, -1.0, 2.5e-8, -2.5e-8, 1e9, -1e9, 2147483520.0, -2147483648.0,
          2147483648.0, -2147483904.0, 3e38, -3e38]
    for h in halves + [639.5, 479.5, -639.5]:
        xs += [h, float(np.nextafter(np.float32(h), np.float32(np.inf))),
               float(np.nextafter(np.float32(h), np.float32(-np.inf)))]
    rng = np.random.default_rng(3)
    xs += list(rng.uniform(-2000, 2000, 400).astype(np.float32))
    return [float(np.float32(x)) for x in xs]


def test_round_and_floor_to_int_match_exact_arithmetic(hd):
    """roundf: the nearest integer, halves away from zero; floorf: the
    largest integer not above; both of the exact float value.  NaN, infinities
    and values outside int32 give INT32_MIN."""
    int_min = -2 ** 31
    for x in _edge_floats():
        q = Fraction(x)
        near = math.floor(abs(q) + Fraction(1, 2))
        want_r = int(near if q >= 0 else -near)
        want_f = math.floor(q)
        want_r = want_r if -2 ** 31 <= want_r < 2 ** 31 and -2 ** 31 < x < 2 ** 31 else int_min
        want_f = want_f if -2 ** 31 <= want_f < 2 ** 31 and -2 ** 31 < x < 2 ** 31 else int_min
        assert hd.hd_round_to_int(x) == want_r, x
        assert hd.hd_floor_to_int(x) == want_f, x
    for x in (float("nan"), float("inf"), -float("inf")):
        assert hd.hd_round_to_int(x) == int_min and hd.hd_floor_to_int(x) == int_min


def test_mat4_mul_within_half_ulp_of_exact_product(hd):
    """Each element is the FP64 sum of four exact products rounded once to
    float: the exact rational product, rounded to float, differs from it by
    at most the FP64 rounding of the sum (far below half a float ulp except
    at float ties), so the two agree to one ulp and almost always exactly; a
    transposed operand or a dropped term misses by far more."""
    rng = np.random.default_rng(5)
    exact_equal = 0
    for trial in range(200):
        A = rng.uniform(-3, 3, 16).astype(np.float32)
        B = rng.uniform(-3, 3, 16).astype(np.float32)
        if trial % 4 == 0:
            A[12:] = [0, 0, 0, 1]
            B[12:] = [0, 0, 0, 1]
        C_ = np.zeros(16, np.float32)
        hd.hd_mat4_mul(_fp(A, C.c_float), _fp(B, C.c_float), _fp(C_, C.c_float))
        for i in range(4):
            for j in range(4):
                s = sum(Fraction(float(A[4 * i + k])) * Fraction(float(B[4 * k + j]))
                        for k in range(4))
                want = np.float32(float(s))  # float(s) is s rounded to FP64
                got = C_[4 * i + j]
                assert ulp_diff(got, want) <= 1, (trial, i, j)
                exact_equal += int(got == want)
    assert exact_equal >= 0.99 * 200 * 16
    # C may alias A
    A = rng.uniform(-1, 1, 16).astype(np.float32)
    I4 = np.eye(4, dtype=np.float32).ravel()
    A0 = A.copy()
    hd.hd_mat4_mul(_fp(A, C.c_float), _fp(I4, C.c_float), _fp(A, C.c_float))
    assert np.array_equal(A, A0)


def test_sym_index_enumerates_the_upper_triangle_row_major(hd):
    order = [(i, j) for i in range(6) for j in range(i, 6)]
    assert [hd.hd_sym_index(i, j) for (i, j) in order] == list(range(21))


def _sums_from_system(A, b, count):
    s = np.zeros(28)
    q = 0
    for i in range(6):
        for j in range(i, 6):
            s[q] = A[i, j]
            q += 1
    s[21:27] = b
    s[27] = count
    return s


def test_icp_increment_is_the_small_angle_matrix_of_the_solution(hd):
    """x = A^-1 b (numpy, FP64) with x = (beta, gamma, alpha, tx, ty, tz);
    the increment is [[1, x2, -x1, x3], [-x2, 1, x0, x4], [x1, -x0, 1, x5],
    [0, 0, 0, 1]] (sensor_pose_estimation.cu:406-425): a wrong sign or
    index of any entry fails."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        J = rng.normal(size=(200, 6)) * np.array([3000, 3000, 3000, 1, 1, 1])
        A = J.T @ J
        b = rng.normal(size=6) * 1e3
        x = np.linalg.solve(A, b)
        inc = np.zeros(16, np.float32)
        assert hd.hd_icp_increment(_fp(_sums_from_system(A, b, 200), C.c_double),
                                   _fp(inc, C.c_float)) == 0
        xf = x.astype(np.float32)
        want = np.array([[1, xf[2], -xf[1], xf[3]], [-xf[2], 1, xf[0], xf[4]],
                         [xf[1], -xf[0], 1, xf[5]], [0, 0, 0, 1]], np.float32).ravel()
        assert ulp_diff(inc, want).max() <= 4
    # a pivot below FLT_EPSILON: the right-hand side itself is the increment
    A = np.zeros((6, 6))
    b = np.array([0.25, -0.5, 0.125, 3.0, -2.0, 1.0])
    inc = np.zeros(16, np.float32)
    assert hd.hd_icp_increment(_fp(_sums_from_system(A, b, 1), C.c_double),
                               _fp(inc, C.c_float)) == 1
    want = np.array([[1, b[2], -b[1], b[3]], [-b[2], 1, b[0], b[4]], [b[1], -b[0], 1, b[5]],
                     [0, 0, 0, 1]], np.float32).ravel()
    assert np.array_equal(inc, want)


def test_robust_increment_and_gated_update(hd):
    """The robust variant gives the identity for a failed solve (status 1)
    and for a solved system with fewer inliers than the gate (status 2), and
    the plain increment otherwise; the gated update composes inc * T, and a
    negative gate selects the plain update."""
    rng = np.random.default_rng(13)
    J = rng.normal(size=(100, 6))
    A, b = J.T @ J, rng.normal(size=6) * 0.01
    s = _sums_from_system(A, b, 100)
    plain, rob = np.zeros(16, np.float32), np.zeros(16, np.float32)
    assert hd.hd_icp_increment(_fp(s, C.c_double), _fp(plain, C.c_float)) == 0
    assert hd.hd_icp_increment_robust(_fp(s, C.c_double), _fp(rob, C.c_float), 100.0) == 0
    assert np.array_equal(plain, rob)
    assert hd.hd_icp_increment_robust(_fp(s, C.c_double), _fp(rob, C.c_float), 101.0) == 2
    assert np.array_equal(rob, np.eye(4, dtype=np.float32).ravel())
    z = _sums_from_system(np.zeros((6, 6)), b, 100)
    assert hd.hd_icp_increment_robust(_fp(z, C.c_double), _fp(rob, C.c_float), 0.0) == 1
    assert np.array_equal(rob, np.eye(4, dtype=np.float32).ravel())
    # gated update: T <- inc * T, rounded once per element
    T = np.array([[0, -1, 0, 100], [1, 0, 0, -50], [0, 0, 1, 2000], [0, 0, 0, 1]],
                 np.float32).ravel()
    got = T.copy()
    assert hd.hd_icp_update_gated(_fp(s, C.c_double), _fp(got, C.c_float), -1.0) == 0
    want = (plain.reshape(4, 4).astype(np.float64) @ T.reshape(4, 4).astype(np.float64))
    assert np.array_equal(got, want.astype(np.float32).ravel())
    got = T.copy()
    assert hd.hd_icp_update_gated(_fp(s, C.c_double), _fp(got, C.c_float), 101.0) == 2
    assert np.array_equal(got, T)


def test_icp_converged_reads_the_six_increment_parameters(hd):
    """The opt-in early stop tests |x0..x2| < stop_rot and |x3..x5| <
    stop_trans, strictly, on the increment matrix's entries x0 = inc[6],
    x1 = inc[8], x2 = inc[1], x3 = inc[3], x4 = inc[7], x5 = inc[11] (as
    test_icp_increment_is_the_small_angle_matrix_of_the_solution lays them
    out); zero thresholds and NaN never stop."""
    def inc_of(x):
        return np.array([1, x[2], -x[1], x[3], -x[2], 1, x[0], x[4], x[1], -x[0], 1, x[5],
                         0, 0, 0, 1], np.float32)

    def conv(x, r=1e-5, t=1e-2):
        a = inc_of(x)
        return hd.hd_icp_converged(_fp(a, C.c_float), r, t)

    assert conv([0, 0, 0, 0, 0, 0]) == 1
    assert conv([0, 0, 0, 0, 0, 0], 0.0, 0.0) == 0
    for k in range(6):
        big = [0.0] * 6
        big[k] = 2e-5 if k < 3 else 2e-2     # above its own threshold
        assert conv(big) == 0, k
        near = [0.0] * 6
        near[k] = 9e-6 if k < 3 else 9e-3    # below it
        assert conv(near) == 1, k
        # a rotation value is not compared with the translation threshold,
        # nor the reverse
        cross = [0.0] * 6
        cross[k] = 5e-3 if k < 3 else 5e-6
        assert conv(cross) == (0 if k < 3 else 1), k
        edge = [0.0] * 6
        edge[k] = 1e-5 if k < 3 else 1e-2    # equal: not below
        assert conv(edge, np.float32(1e-5), np.float32(1e-2)) == 0, k
        nan = [0.0] * 6
        nan[k] = float("nan")
        assert conv(nan) == 0, k


def test_inv_norm3_closed_forms_and_float_steps(hd):
    """kf_inv_norm3 (the factor normalize scales by, cuda_functions.cuh:2-10):
    closed forms (axis vectors 1; (3, 4, 0) and its power-of-two multiples
    2^-k / 5 correctly rounded; the zero vector +inf; NaN in, NaN out), and on
    random vectors equal to the float steps (x x + y y) + z z, square root,
    quotient, each correctly rounded, and within 4 ulp of the exact 1 / |v|
    (the five roundings add at most 3.5 * 2^-24 relative)."""
    f32 = np.float32
    for v in ((1, 0, 0), (0, -1, 0), (0, 0, 1), (-1, 0, 0)):
        assert hd.hd_inv_norm3(*v) == 1.0
    for k in range(-20, 21):
        s = 2.0 ** k
        want = f32(Fraction(1, 5) / Fraction(s))
        for v in ((3 * s, 4 * s, 0.0), (0.0, -4 * s, 3 * s), (4 * s, 0.0, -3 * s)):
            assert hd.hd_inv_norm3(*v) == want, (k, v)
    assert hd.hd_inv_norm3(0.0, 0.0, 0.0) == float("inf")
    assert hd.hd_inv_norm3(-0.0, 0.0, -0.0) == float("inf")
    assert math.isnan(hd.hd_inv_norm3(float("nan"), 1.0, 0.0))
    rng = np.random.default_rng(11)
    vs = rng.normal(size=(2000, 3)).astype(np.float32)
    vs[:500] *= np.float32(1e3)
    vs[500:1000] /= np.linalg.norm(vs[500:1000], axis=1, keepdims=True).astype(np.float32)
    for x, y, z in vs:
        got = hd.hd_inv_norm3(float(x), float(y), float(z))
        s = (f32(x) * f32(x) + f32(y) * f32(y)) + f32(z) * f32(z)
        assert got == f32(1) / np.sqrt(s), (x, y, z)
        exact = 1.0 / math.sqrt(float(Fraction(float(x)) ** 2 + Fraction(float(y)) ** 2 +
                                      Fraction(float(z)) ** 2))
        ulp = float(np.spacing(f32(exact)))
        assert abs(got - exact) <= 4 * ulp, (x, y, z)
