"""Pins the oracle's RANSAC (pipeline.rs:1860-2755) with the reference's known-answer
tests, a published SplitMix64 vector, SVD-based DLT and closed-form iteration counts."""
import math

import numpy as np
import pytest

import psr


def project(H, x, y):
    """Homography2D::project in f32 (pipeline.rs:1122-1132)."""
    H = np.asarray(H, np.float32)
    x, y = np.float32(x), np.float32(y)
    w = H[2, 0] * x + H[2, 1] * y + H[2, 2]
    iw = np.float32(1.0) / w
    return (H[0, 0] * x + H[0, 1] * y + H[0, 2]) * iw, (H[1, 0] * x + H[1, 1] * y + H[1, 2]) * iw


def test_splitmix64_published_vector():
    # SplitMix64 from state 0: first outputs (Vigna's reference implementation)
    assert psr.splitmix64(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_adaptive_iters_closed_form():
    for inl, tot, ss in [(10, 14, 2), (12, 60, 4), (5, 100, 4), (100, 100, 4)]:
        w = min(max(inl / tot, 1e-3), 1 - 1e-6)
        p = min(max(w ** ss, 1e-12), 1 - 1e-9)
        est = math.ceil(math.log(1 - 0.999) / math.log(1 - p))
        assert psr.adaptive_ransac_iters(inl, tot, ss, 0.999, 64, 2000) == min(max(est, 64), 2000)
    assert psr.adaptive_ransac_iters(0, 10, 4, 0.999, 64, 2000) == 2000


def test_similarity_known_answer():
    """pipeline_tests.rs:343-445: 10 inliers (offset -341) + 4 outliers."""
    q, t, r = [], [], []
    for i in range(10):
        qx, qy = 400.0 + i * 85.0, 50.0 + ((i * 73) % 600)
        q.append((qx, qy)); t.append((qx - 341.0, qy)); r.append(0.10)
    for k, (ox, oy) in enumerate([(220.0, 180.0), (-90.0, -310.0), (410.0, -95.0), (150.0, 260.0)]):
        qx, qy = 500.0 + k * 60.0, 200.0 + k * 50.0
        q.append((qx, qy)); t.append((qx + ox, qy + oy)); r.append(0.71)
    m = psr.make_matches(q, t)
    m["ratio"] = r
    inl, mod = psr.filter_matches_ransac(m, 3.0)
    assert sorted(inl.tolist()) == list(range(10))
    assert abs(mod["tx"] + 341.0) < 1e-3 and abs(mod["ty"]) < 1e-3
    assert abs(mod["scale"] - 1.0) < 1e-3 and abs(mod["rotation_deg"]) < 1e-3


def test_homography_known_answer_10_plus_3():
    """pipeline_tests.rs:447-524"""
    H = [[1.08, -0.04, 35.0], [0.03, 0.96, -18.0], [0.00012, -0.00008, 1.0]]
    pts = [(50, 60), (320, 45), (480, 110), (95, 240), (260, 275), (440, 310), (70, 460), (290, 490),
           (495, 470), (180, 170)]
    q = [(float(a), float(b)) for a, b in pts]
    t = [project(H, a, b) for a, b in q]
    r = [0.08] * 10
    for (qx, qy, tx, ty) in [(120, 130, 490, 20), (350, 220, 15, 480), (200, 400, 420, 90)]:
        q.append((qx, qy)); t.append((tx, ty)); r.append(0.72)
    m = psr.make_matches(q, t)
    m["ratio"] = r
    inl, h = psr.homography_ransac_ref(m, 3.0)
    assert sorted(inl.tolist()) == list(range(10))
    for (a, b) in pts:
        e = np.hypot(*(np.array(project(H, a, b)) - np.array(project(h, a, b))))
        assert e < 0.05


def _eighty_percent_outliers():
    H = [[1.12, -0.06, 42.0], [0.05, 0.94, -24.0], [0.00015, -0.0001, 1.0]]
    pts = [(40, 50), (200, 40), (420, 75), (560, 110), (80, 210), (270, 240), (490, 280), (60, 410),
           (230, 450), (440, 430), (580, 470), (330, 340)]
    q = [(float(a), float(b)) for a, b in pts]
    t = [project(H, a, b) for a, b in q]
    r = [0.25] * len(q)
    for k in range(48):
        qx = 100.0 + k * 30.0 if k < 8 else float((k * 97) % 600)
        qy = 100.0 if k < 8 else float((k * 131) % 480)
        q.append((qx, qy)); t.append((float((k * 173 + 50) % 620), float((k * 211 + 80) % 490))); r.append(0.72)
    m = psr.make_matches(q, t)
    m["ratio"] = r
    return H, pts, m


def test_homography_80_percent_outliers():
    """tests/homography_properties.rs:70-149"""
    H, pts, m = _eighty_percent_outliers()
    inl, h = psr.homography_ransac_ref(m, 3.0)
    assert sorted(inl.tolist()) == list(range(12))
    for (a, b) in pts:
        assert np.hypot(*(np.array(project(H, a, b)) - np.array(project(h, a, b)))) < 0.05


def test_seed_determinism():
    """tests/homography_properties.rs:151-237"""
    _, _, m = _eighty_percent_outliers()
    a = psr.homography_ransac_ref(m, 3.0, seed=0x123456789ABCDEF0)
    b = psr.homography_ransac_ref(m, 3.0, seed=0x123456789ABCDEF0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    c = psr.homography_ransac_ref(m, 3.0)
    d = psr.homography_ransac(m, 3.0, 0xD1B54A32D192ED03, 0x517CC1B727220A95, sampler=0, confidence=0.999,
                              min_iters=64, max_iters=2000, local_opt=True, sim_seed_model=True)
    assert np.array_equal(c[0], d[0]) and np.array_equal(c[1], d[1])


def test_dlt_vs_svd():
    rng = np.random.default_rng(0)
    H = np.array([[0.9, 0.1, 12.0], [-0.05, 1.1, -7.0], [1e-4, -2e-4, 1.0]])
    for n in (4, 10):
        q = rng.uniform(0, 600, (n, 2)).astype(np.float32)
        ph = np.c_[q.astype(np.float64), np.ones(n)] @ H.T
        t = (ph[:, :2] / ph[:, 2:]).astype(np.float32)
        h = psr.fit_homography_dlt(psr.make_matches(q, t))
        # reference: unnormalised DLT via numpy SVD in float64
        A = []
        for (x, y), (u, v) in zip(q.astype(np.float64), t.astype(np.float64)):
            A.append([-x, -y, -1, 0, 0, 0, u * x, u * y, u])
            A.append([0, 0, 0, -x, -y, -1, v * x, v * y, v])
        _, _, Vt = np.linalg.svd(np.array(A))
        Hs = Vt[-1].reshape(3, 3)
        Hs /= Hs[2, 2]
        assert np.allclose(h, Hs, rtol=1e-6, atol=1e-8)


def test_degenerate_inputs():
    m = psr.make_matches([(0.0, 5.0), (1.0, 5.0), (2.0, 5.0), (3.0, 5.0)], [(7.0, 0.0), (7.0, 1.0), (7.0, 2.0), (7.0, 3.0)])
    inl, h = psr.homography_ransac_ref(m, 3.0)
    assert len(inl) <= 4
    rep = psr.make_matches([(10.0, 10.0)] * 4, [(20.0, 30.0)] * 4)
    inl, h = psr.homography_ransac_ref(rep, 3.0)
    assert len(inl) <= 4
    assert len(psr.homography_ransac_ref(psr.make_matches(np.zeros((0, 2)), np.zeros((0, 2))), 3.0)[0]) == 0


@pytest.mark.parametrize("sampler", [1, 2])
def test_variant_samplers_recover_model(sampler):
    H, pts, m = _eighty_percent_outliers()
    w = np.where(np.arange(len(m)) < 12, 0.9, 0.1).astype(np.float32) if sampler == 2 else None
    inl, h, hyps = psr.homography_ransac(m, 3.0, 0, 0x5EED, sampler=sampler, weights=w, confidence=0.99,
                                         min_iters=1, max_iters=100000, adaptive=True)
    assert sorted(inl.tolist()) == list(range(12))
    assert hyps >= 1
    if sampler == 2:
        # guided mass of inliers = 12*0.9 / (12*0.9 + 48*0.1) -> far fewer hypotheses than uniform
        _, _, hu = psr.homography_ransac(m, 3.0, 0, 0x5EED, sampler=1, confidence=0.99, min_iters=1,
                                         max_iters=100000, adaptive=True)
        assert hyps * 3 <= hu


@pytest.mark.parametrize("bad", [0.0, np.nan, -1.0, np.inf])
def test_weighted_sampler_fallback_is_uniform(bad):
    """Weights without usable mass (all zero, or one NaN / negative / infinite weight) sample
    uniformly with the uniform stop rule: same inliers, model and hypothesis count as sampler 1."""
    H, pts, m = _eighty_percent_outliers()
    w = np.zeros(len(m), np.float32) if bad == 0.0 else np.full(len(m), 0.5, np.float32)
    w[3] = bad
    kw = dict(confidence=0.99, min_iters=1, max_iters=100000, adaptive=True)
    got = psr.homography_ransac(m, 3.0, 0, 0x5EED, sampler=2, weights=w, **kw)
    uni = psr.homography_ransac(m, 3.0, 0, 0x5EED, sampler=1, **kw)
    assert np.array_equal(got[0], uni[0]) and got[2] == uni[2]
    assert np.array_equal(got[1], uni[1])
    # a usable weight vector does change the sampling
    w2 = np.full(len(m), 0.5, np.float32)
    w2[:12] = 5.0
    assert psr.homography_ransac(m, 3.0, 0, 0x5EED, sampler=2, weights=w2, **kw)[2] < uni[2]


def test_mlp_gradient_finite_differences():
    """The oracle's MLP gradient against central differences of its own loss (float64
    arithmetic on f32 parameters; small steps stay off the ReLU kinks)."""
    rng = np.random.default_rng(0)
    p = (rng.standard_normal(psr.MLP_PARAMS) * 0.2).astype(np.float32)
    x = rng.standard_normal((40, 8)).astype(np.float32)
    y = (rng.random(40) < 0.5).astype(np.float32)
    _, g = psr.mlp_loss_grad(p, x, y, bf16=False)
    for i in [0, 77, 511, 512, 575, 600, 4000, 4671, 4672, 4700, 4736, 4799, 4800]:
        lo, hi = p.copy(), p.copy()
        hi[i] = np.float32(p[i] + 3e-5)
        lo[i] = np.float32(p[i] - 3e-5)
        fd = (psr.mlp_loss_grad(hi, x, y, bf16=False)[0] - psr.mlp_loss_grad(lo, x, y, bf16=False)[0]) / (
            float(hi[i]) - float(lo[i]))
        assert abs(fd - g[i]) <= 1e-4 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])


def test_mlp_forward_known_answer():
    """All-zero weights except b3 -> logit b3; one active path gives a closed form."""
    p = np.zeros(psr.MLP_PARAMS, np.float32)
    p[4800] = 0.75
    x = np.ones((3, 8), np.float32)
    assert np.all(psr.mlp_forward(p, x) == np.float32(0.75))
    p[0] = 2.0          # W1[0][0]
    p[576] = 3.0        # W2[0][0]
    p[4736] = 0.5       # w3[0]
    x = np.array([[1.0, 0, 0, 0, 0, 0, 0, 0], [-1.0, 0, 0, 0, 0, 0, 0, 0]], np.float32)
    assert psr.mlp_forward(p, x).tolist() == [0.75 + 0.5 * 3.0 * 2.0, 0.75]
