"""Pins the oracle's two-view geometry (oracle/ps_oracle_geometry.c): the classical 9x9
Jacobi against numpy, the Sampson distance closed form, the 8- and 7-point solvers on exact
scenes, fundamental LO-RANSAC under outliers, the adaptive iteration bound, the 3x3 inverse
and the reference's GRIC and 7-point tests (tests/geometry_tests.rs:207-309)."""
import math

import numpy as np
import pytest

import psr
from geometry_scenes import rotation, same_up_to_scale, two_view


def test_jacobi9_matches_numpy():
    rng = np.random.default_rng(4)
    for _ in range(20):
        a = rng.normal(size=(9, 9))
        a = a + a.T
        ev, vec = psr.jacobi9(a)
        w, v = np.linalg.eigh(a)
        assert np.allclose(ev, w, atol=1e-10)
        assert list(ev) == sorted(ev)
        for k in range(9):
            assert abs(abs(np.dot(vec[:, k], v[:, k])) - 1.0) < 1e-8


def test_jacobi9_rejects_non_finite():
    a = np.eye(9)
    a[0, 1] = a[1, 0] = np.inf
    assert psr.jacobi9(a) is None


def test_sampson_closed_form():
    rng = np.random.default_rng(5)
    for _ in range(50):
        F = rng.normal(size=(3, 3))
        x1 = np.array([*rng.uniform(0, 800, 2), 1.0])
        x2 = np.array([*rng.uniform(0, 800, 2), 1.0])
        Fx1, Ftx2 = F @ x1, F.T @ x2
        want = (x2 @ F @ x1) ** 2 / (Fx1[0] ** 2 + Fx1[1] ** 2 + Ftx2[0] ** 2 + Ftx2[1] ** 2)
        got = psr.sampson_sq(F, x1[:2], x2[:2])
        assert abs(got - want) <= 1e-9 * max(1.0, abs(want))
    assert psr.sampson_sq(np.zeros((3, 3)), (1, 2), (3, 4)) == math.inf


def test_8point_recovers_the_true_fundamental():
    m, F = two_view(40, 0, rotation(12, -6, 4), [0.6, -0.15, 0.2], seed=1)
    Fe = psr.fit_fundamental_8point(m)
    assert same_up_to_scale(Fe, F) < 1e-3
    assert abs(np.linalg.det(Fe)) < 1e-9                      # rank 2
    assert abs(np.linalg.norm(Fe) - 1.0) < 1e-12              # Frobenius-normalised
    d = [math.sqrt(psr.sampson_sq(Fe, (p["qx"], p["qy"]), (p["tx"], p["ty"]))) for p in m]
    assert np.mean(d) < 0.05                                  # f32 point rounding only


def test_7point_exact():
    """tests/geometry_tests.rs:207-249: a candidate fits 5 held-out points."""
    m, F = two_view(12, 0, rotation(8, -4, 3), [0.5, 0.1, -0.15], seed=2)
    cands = psr.fit_fundamental_7point(m[:7])
    assert cands
    best = min(np.mean([math.sqrt(psr.sampson_sq(c, (p["qx"], p["qy"]), (p["tx"], p["ty"]))) for p in m[7:]])
               for c in cands)
    assert best < 0.05
    assert min(same_up_to_scale(c, F) for c in cands) < 1e-3


def test_ransac_with_outliers():
    """80 inliers + 26 outliers (the reference's essential-matrix scene size)."""
    m, F = two_view(80, 26, rotation(12, -6, 4), [0.6, -0.15, 0.2], seed=3, noise=0.2)
    inl, Fe, iters = psr.fundamental_ransac(m, 1.5)
    assert Fe is not None and same_up_to_scale(Fe, F) < 2e-2
    assert (inl < 80).sum() >= 76 and (inl >= 80).sum() <= 2
    assert list(inl) == sorted(inl)
    assert 64 <= iters <= 2000


def test_ransac_small_inputs():
    m, _ = two_view(6, 0, rotation(8, -4, 3), [0.5, 0.1, -0.15], seed=6)
    assert psr.fundamental_ransac(m)[1] is None               # < 7 matches: no model
    m7, _ = two_view(7, 0, rotation(8, -4, 3), [0.5, 0.1, -0.15], seed=6)
    inl, F7, iters = psr.fundamental_ransac(m7)
    assert F7 is not None and iters == 0 and len(inl) == 7    # the 7-point path


def test_adaptive_epipolar_iters_closed_form():
    assert psr.adaptive_epipolar_iters(40, 80, 8, 0.999, 64, 2000) == min(
        2000, math.ceil(math.log(1 - 0.999) / math.log(1 - 0.5 ** 8)))
    assert psr.adaptive_epipolar_iters(5, 8, 8, 0.999, 64, 2000) == 2000      # total <= sample size
    assert psr.adaptive_epipolar_iters(0, 80, 8, 0.999, 64, 2000) == 2000
    assert psr.adaptive_epipolar_iters(9999, 10000, 8, 0.999, 64, 2000) == 64  # w >= 0.9999
    assert psr.adaptive_epipolar_iters(1, 1000, 8, 0.999, 64, 2000) == 2000   # w^8 <= 1e-12
    assert psr.adaptive_epipolar_iters(90, 100, 8, 0.99, 64, 2000) == 64


def test_invert3_matches_numpy():
    rng = np.random.default_rng(7)
    h = rng.normal(size=(3, 3)).astype(np.float32)
    assert np.allclose(psr.invert3_f32(h), np.linalg.inv(h.astype(np.float64)), atol=1e-9)
    assert psr.invert3_f32(np.zeros((3, 3), np.float32)) is None


def test_gric_planar_and_3d():
    """tests/geometry_tests.rs:251-309: homography on a planar scene, fundamental with parallax."""
    K = np.array([[700.0, 0, 350], [0, 700, 280], [0, 0, 1]])
    xs, ys = np.meshgrid(-2.0 + np.arange(8) * 0.55, -1.5 + np.arange(5) * 0.7)
    P = np.stack([xs.ravel(), ys.ravel(), np.full(40, 6.0)], 1)
    from geometry_scenes import project
    p1 = project(K, np.eye(3), np.zeros(3), P).astype(np.float32)
    p2 = project(K, rotation(5, 14, -3), np.array([0.35, -0.1, 0.05]), P).astype(np.float32)
    planar = psr.make_matches(p1, p2, ratio=0.4, distance=0.1)
    _, H = psr.homography_ransac_ref(planar, 3.0)
    _, Fp, _ = psr.fundamental_ransac(planar, 1.5)
    sel, hg, fg = psr.select_model_gric(planar, H, Fp, 1.0)
    assert sel == 0 and hg <= fg
    m3, _ = two_view(60, 0, rotation(10, -8, 2), [0.9, 0.25, 0.35], K=K, seed=8)
    _, H3 = psr.homography_ransac_ref(m3, 3.0)
    _, F3, _ = psr.fundamental_ransac(m3, 1.5)
    sel, hg, fg = psr.select_model_gric(m3, H3, F3, 1.0)
    assert sel == 1 and fg < hg
    assert psr.select_model_gric(m3, None, None, 1.0) is None
    assert psr.select_model_gric(m3[:0], H3, F3, 1.0) is None


def test_gric_closed_form_terms():
    """Exact model: every data term is 0, so GRIC = lambda1 * d * n + lambda2 * k."""
    m, F = two_view(30, 0, rotation(0, 0, 0), [0.0, 0.0, 0.0], seed=9)   # identical views
    sel, hg, fg = psr.select_model_gric(m, np.eye(3, dtype=np.float32), None, 1.0)
    assert sel == 0 and hg == 2.0 * 2.0 * 30 + 2.0 * 8.0 and fg == math.inf


def _corrected_hypot(x, y):
    """The device's hypot (pure-sift_amd/csrc/fundamental.hip glibc_hypot) in numpy f64, each
    operation rounded on its own: sqrt(ax^2 + ay^2) then one correction step (Borges' corrected
    algorithm without fused multiply-add)."""
    ax, ay = np.maximum(np.abs(x), np.abs(y)), np.minimum(np.abs(x), np.abs(y))
    h = np.sqrt(ax * ax + ay * ay)
    small = h <= 2.0 * ay
    d1 = h - ay
    t1a, t2a = ax * (2.0 * d1 - ax), (d1 - 2.0 * (ax - ay)) * d1
    d2 = h - ax
    t1b, t2b = 2.0 * d2 * (ax - 2.0 * ay), (4.0 * d2 - ay) * ay + d2 * d2
    t1, t2 = np.where(small, t1a, t1b), np.where(small, t2a, t2b)
    out = h - (t1 + t2) / (2.0 * h)
    tiny = ay <= ax * 2.0 ** -54
    return np.where(ay == 0.0, ax, np.where(tiny, ax + ay, out))


def test_device_hypot_algorithm_equals_libm():
    """The oracle's Hartley normalisation calls the C library's hypot, which is what the
    reference's f64::hypot binds on Linux; the device restates it. Over centred-coordinate
    magnitudes (1e-3 .. 1e5, both signs, equal and very unequal arguments) the restatement must
    equal libm's hypot bit for bit, and differ from the correctly rounded value somewhere
    (otherwise the test would not tell the two apart)."""
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.hypot.restype = ctypes.c_double
    libm.hypot.argtypes = [ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(7)
    n = 40000
    mag = 10.0 ** rng.uniform(-3, 5, size=(n, 2))
    xy = mag * rng.choice([-1.0, 1.0], size=(n, 2))
    xy[:200, 1] = xy[:200, 0]                      # equal magnitudes
    xy[200:400, 1] = xy[200:400, 0] * 1e-17        # ratio beyond 2^54
    xy[400:500, 1] = 0.0
    got = _corrected_hypot(xy[:, 0], xy[:, 1])
    want = np.array([libm.hypot(float(a), float(b)) for a, b in xy])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    naive = np.sqrt(xy[:, 0] * xy[:, 0] + xy[:, 1] * xy[:, 1])
    assert np.count_nonzero(naive != want) > 100
