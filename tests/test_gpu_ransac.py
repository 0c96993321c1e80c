"""GPU homography RANSAC (pipeline.rs:2055-2191 similarity seed, :2663-2755 LO-RANSAC)
through the C ABI against the oracle's sequential algorithm.

Bar: bit-exact. Inlier index lists, the hypothesis count, and the f32 homography and
similarity model equal the oracle's for every problem: the device reproduces the
SplitMix64 stream draw for draw and every f32 / f64 expression in the same order.
"""
import numpy as np
import pytest

import psr
import synth

pytestmark = pytest.mark.gpu


def _problem(n_in, n_out, seed, noise=0.5, w=1024, h=768):
    rng = np.random.default_rng(seed)
    H = synth.random_homography(w, h, rng)
    q = np.stack([rng.uniform(0, w, n_in + n_out), rng.uniform(0, h, n_in + n_out)], 1)
    t = np.empty_like(q)
    x, y = q[:n_in, 0], q[:n_in, 1]
    den = H[2, 0] * x + H[2, 1] * y + H[2, 2]
    t[:n_in, 0] = (H[0, 0] * x + H[0, 1] * y + H[0, 2]) / den + rng.normal(0, noise, n_in)
    t[:n_in, 1] = (H[1, 0] * x + H[1, 1] * y + H[1, 2]) / den + rng.normal(0, noise, n_in)
    t[n_in:] = np.stack([rng.uniform(0, w, n_out), rng.uniform(0, h, n_out)], 1)
    ratio = np.concatenate([rng.uniform(0.2, 0.8, n_in), rng.uniform(0.5, 0.98, n_out)])
    perm = rng.permutation(n_in + n_out)
    m = psr.make_matches(q[perm], t[perm])
    m["ratio"] = ratio[perm]
    m["distance"] = rng.uniform(0.1, 0.6, n_in + n_out).astype(np.float32)
    return m


def _same_h(a, b):
    if a is None or b is None:
        return a is None and b is None
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def test_known_answer_10_plus_3(gpu):
    """pipeline_tests.rs:447-524 scene: the 10 inliers and the oracle's model."""
    H = np.array([[1.08, -0.04, 35.0], [0.03, 0.96, -18.0], [0.00012, -0.00008, 1.0]])
    pts = [(50, 60), (320, 45), (480, 110), (95, 240), (260, 275), (440, 310), (70, 460), (290, 490), (495, 470),
           (180, 170)]
    q = [(float(a), float(b)) for a, b in pts]
    t = []
    for a, b in q:
        w = H[2, 0] * a + H[2, 1] * b + H[2, 2]
        t.append(((H[0, 0] * a + H[0, 1] * b + H[0, 2]) / w, (H[1, 0] * a + H[1, 1] * b + H[1, 2]) / w))
    r = [0.08] * 10
    for (qx, qy, tx, ty) in [(120, 130, 490, 20), (350, 220, 15, 480), (200, 400, 420, 90)]:
        q.append((qx, qy))
        t.append((tx, ty))
        r.append(0.72)
    m = psr.make_matches(q, t)
    m["ratio"] = r
    inl, Hg = gpu.homography_estimate(m, gpu.RansacOptions(3.0))
    rin, Hr = psr.homography_ransac_ref(m, 3.0)
    assert sorted(inl.tolist()) == list(range(10))
    assert np.array_equal(inl, rin) and _same_h(Hg, Hr)


@pytest.mark.parametrize("n_in,n_out", [(0, 0), (1, 0), (2, 0), (3, 1), (4, 0), (5, 3), (20, 10), (60, 140),
                                        (300, 300), (700, 1300), (2500, 500)])
@pytest.mark.parametrize("seed", [None, 12345])
def test_homography_estimate_matches_oracle(gpu, n_in, n_out, seed):
    m = _problem(n_in, n_out, seed=n_in * 31 + n_out)
    inl, Hg = gpu.homography_estimate(m, gpu.RansacOptions(3.0, seed))
    rin, Hr = psr.homography_ransac_ref(m, 3.0, seed)
    assert np.array_equal(inl, rin)
    assert _same_h(Hg, Hr)


def test_degenerate_collinear(gpu):
    """All query points on a line: every 4-sample is degenerate (pipeline.rs:2357-2404)."""
    x = np.linspace(10, 900, 50)
    q = np.stack([x, 0.5 * x + 7], 1)
    m = psr.make_matches(q, q + np.array([5.0, -3.0]))
    inl, Hg = gpu.homography_estimate(m, gpu.RansacOptions(2.0))
    rin, Hr = psr.homography_ransac_ref(m, 2.0)
    assert np.array_equal(inl, rin) and _same_h(Hg, Hr)


CONFIGS = [
    dict(sampler=0, confidence=0.999, min_iters=64, max_iters=2000, local_opt=True, sim_seed_model=True, adaptive=True),
    dict(sampler=1, confidence=0.99, min_iters=1, max_iters=3000, local_opt=False, sim_seed_model=False, adaptive=True),
    dict(sampler=2, confidence=0.99, min_iters=1, max_iters=3000, local_opt=False, sim_seed_model=False, adaptive=True),
    dict(sampler=1, confidence=0.99, min_iters=1, max_iters=700, local_opt=False, sim_seed_model=False,
         adaptive=False),
    dict(sampler=2, confidence=0.95, min_iters=10, max_iters=1500, local_opt=True, sim_seed_model=True,
         adaptive=True),
]


@pytest.mark.parametrize("cfg", range(len(CONFIGS)))
def test_batch_matches_oracle(gpu, cfg):
    c = CONFIGS[cfg]
    rng = np.random.default_rng(cfg)
    sizes = [(0, 0), (3, 0), (4, 0), (12, 30), (40, 60), (150, 350), (500, 200), (90, 910), (35, 15)]
    ms, probs, off = [], [], 0
    for k, (a, b) in enumerate(sizes):
        m = _problem(a, b, seed=100 * cfg + k)
        ms.append(m)
        probs.append((off, len(m), int(rng.integers(0, 2 ** 63)), int(rng.integers(0, 2 ** 63))))
        off += len(m)
    allm = np.concatenate(ms)
    # weights: a noisy inlier score (higher for low ratio)
    wts = (1.0 - allm["ratio"]).astype(np.float32) + np.float32(0.01)
    pr = np.array(probs, gpu.RANSAC_PROBLEM_DTYPE)
    prm = gpu.ransac_params(reproj_threshold=3.0, **c)
    res, inl = gpu.ransac_batch(allm, pr, prm, weights=wts)
    for k, m in enumerate(ms):
        o, n = probs[k][0], probs[k][1]
        rin, Hr, hyps = psr.homography_ransac(m, 3.0, sim_seed=probs[k][2], hom_seed=probs[k][3],
                                              sampler=c["sampler"], weights=wts[o:o + n],
                                              confidence=c["confidence"], min_iters=c["min_iters"],
                                              max_iters=c["max_iters"], local_opt=c["local_opt"],
                                              sim_seed_model=c["sim_seed_model"], adaptive=c["adaptive"])
        r = res[k]
        assert r["n_inliers"] == len(rin), (k, r["n_inliers"], len(rin))
        assert np.array_equal(inl[o:o + r["n_inliers"]], rin), k
        assert bool(r["has_model"]) == (Hr is not None)
        if Hr is not None:
            assert _same_h(r["h"].reshape(3, 3), Hr), k
        if n >= 4:
            assert r["hyps"] == hyps, (k, r["hyps"], hyps)


def _same_result(a, b):
    """Equal RANSAC result records (the similarity fields only count when has_sim is set)."""
    for f in ("h", "has_model", "n_inliers", "hyps", "has_sim", "n_sim"):
        if not np.array_equal(a[f], b[f]):
            return False
    return not a["has_sim"] or np.array_equal(a["sim"], b["sim"])


def _fallback_batch(gpu, weight_fn, host=True):
    """Three problems sampled by weights from weight_fn(m), next to the same problems sampled
    uniformly; returns both result sets."""
    ms, probs, off = [], [], 0
    for k, (a, b) in enumerate([(40, 60), (150, 350), (12, 30)]):
        m = _problem(a, b, seed=700 + k)
        ms.append(m)
        probs.append((off, len(m), 11 + k, 1000 + 7 * k))
        off += len(m)
    allm = np.concatenate(ms)
    pr = np.array(probs, gpu.RANSAC_PROBLEM_DTYPE)
    kw = dict(confidence=0.99, min_iters=1, max_iters=3000, local_opt=True, sim_seed_model=False, adaptive=True)
    wts = weight_fn(allm).astype(np.float32)
    if host:
        res_w, inl_w = gpu.ransac_batch(allm, pr, gpu.ransac_params(reproj_threshold=3.0, sampler=2, **kw),
                                        weights=wts)
    else:
        import torch
        dev = torch.device("cuda:0")
        eng = gpu.Engine(0)
        try:
            dm = torch.from_numpy(allm.view(np.uint8).copy()).to(dev)
            dw = torch.from_numpy(wts).to(dev)
            dp = torch.from_numpy(pr.view(np.uint8).copy()).to(dev)
            di = torch.zeros(len(allm), dtype=torch.int32, device=dev)
            dr = torch.zeros(len(pr) * gpu.RANSAC_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
            prm = gpu.ransac_params(reproj_threshold=3.0, sampler=2, **kw)
            eng.ransac_batch(dm.data_ptr(), dw.data_ptr(), dp.data_ptr(), len(pr), len(allm), prm, di.data_ptr(),
                             dr.data_ptr())
            eng.sync()
            eng.check()
            res_w = dr.cpu().numpy().view(gpu.RANSAC_RESULT_DTYPE).copy()
            inl_w = di.cpu().numpy().view(np.uint32).copy()
        finally:
            eng.close()
    res_u, inl_u = gpu.ransac_batch(allm, pr, gpu.ransac_params(reproj_threshold=3.0, sampler=1, **kw))
    return ms, probs, wts, (res_w, inl_w), (res_u, inl_u), kw


def test_weighted_sampler_without_mass_falls_back_to_uniform(gpu):
    """All-zero weights (no sampling mass): each problem gets exactly the uniform sampler's result
    (inliers, model, hypothesis count), which also equals the oracle's run."""
    ms, probs, wts, (rw, iw), (ru, iu), kw = _fallback_batch(gpu, lambda m: np.zeros(len(m)))
    assert all(_same_result(a, b) for a, b in zip(rw, ru)) and np.array_equal(iw, iu)
    for k, m in enumerate(ms):
        o, n = probs[k][0], probs[k][1]
        rin, Hr, hyps = psr.homography_ransac(m, 3.0, sim_seed=probs[k][2], hom_seed=probs[k][3], sampler=2,
                                              weights=wts[o:o + n], **kw)
        assert np.array_equal(iw[o:o + rw[k]["n_inliers"]], rin) and rw[k]["hyps"] == hyps
        assert rw[k]["has_model"] == 1


@pytest.mark.parametrize("bad", [np.nan, -1.0, np.inf])
def test_weighted_sampler_bad_weights(gpu, bad):
    """A NaN, negative or infinite weight: the host-pointer entry point rejects the call; the
    device-resident form samples that problem uniformly (the kernel cannot return an error per
    problem)."""
    def fn(m):
        w = 1.0 - m["ratio"].astype(np.float64) + 0.01
        w[5] = bad
        return w
    with pytest.raises(gpu.DeviceError):
        _fallback_batch(gpu, fn, host=True)
    ms, probs, wts, (rw, iw), (ru, iu), kw = _fallback_batch(gpu, fn, host=False)
    # problem 0 holds the bad weight and falls back; the others keep their weighted sampling
    assert _same_result(rw[0], ru[0])
    o, n = probs[0][0], probs[0][1]
    assert np.array_equal(iw[o:o + rw[0]["n_inliers"]], iu[o:o + ru[0]["n_inliers"]])
    for k in (1, 2):
        o, n = probs[k][0], probs[k][1]
        rin, Hr, hyps = psr.homography_ransac(ms[k], 3.0, sim_seed=probs[k][2], hom_seed=probs[k][3], sampler=2,
                                              weights=wts[o:o + n], **kw)
        assert np.array_equal(iw[o:o + rw[k]["n_inliers"]], rin) and rw[k]["hyps"] == hyps
