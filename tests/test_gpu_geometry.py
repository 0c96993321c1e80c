"""GPU parity for the per-pair two-view geometry (cli.rs:290-318 post-processing) through the
C ABI, against the oracle: similarity RANSAC (Homography::estimate_similarity), fundamental
LO-RANSAC (geometry.rs:857-1016) and GRIC (geometry.rs:1233-1341).

Bar: bit-exact inlier lists, iteration counts, models (f32 similarity, f64 fundamental) and
GRIC scores; the 7-point path (exactly 7 matches, device cbrt/acos/cos before the Newton
polish) is held to 1e-9 on the model with equal inliers.
"""
import numpy as np
import pytest

import psr
from geometry_scenes import rotation, two_view

pytestmark = pytest.mark.gpu

SCENES = [  # (inliers, outliers, yaw, pitch, roll, t, noise, seed, planar depth)
    (80, 26, 12, -6, 4, [0.6, -0.15, 0.2], 0.2, 1, None),
    (60, 0, 10, -8, 2, [0.9, 0.25, 0.35], 0.0, 2, None),
    (200, 150, 5, 3, -2, [0.3, 0.05, 0.1], 0.5, 3, None),
    (40, 40, 4, 14, -3, [0.35, -0.1, 0.05], 0.3, 4, 6.0),
    (9, 3, 8, -4, 3, [0.5, 0.1, -0.15], 0.1, 5, None),
    (8, 0, 8, -4, 3, [0.5, 0.1, -0.15], 0.0, 6, None),
    (30, 90, 20, 5, 1, [0.7, 0.3, 0.1], 1.0, 7, None),
]


def scene(i):
    ni, no, y, p, r, t, noise, seed, depth = SCENES[i]
    m, _ = two_view(ni, no, rotation(y, p, r), t, seed=seed, noise=noise, planar_depth=depth)
    print(f"scene {i}: {ni} inliers, {no} outliers", flush=True)
    return m


@pytest.mark.parametrize("i", range(len(SCENES)))
@pytest.mark.parametrize("thr", [1.5, 0.1, 3.0])
def test_fundamental_equals_oracle(gpu, i, thr):
    m = scene(i)
    want_inl, want_f, want_it = psr.fundamental_ransac(m, thr)
    got_inl, got_f = gpu.estimate_fundamental_ransac(m, thr)
    assert (got_f is None) == (want_f is None)
    assert np.array_equal(got_inl, want_inl)
    if want_f is not None:
        assert np.array_equal(got_f, want_f), np.abs(got_f - want_f).max()


def test_fundamental_seven_point_path(gpu):
    m = scene(0)[:7]
    want_inl, want_f, _ = psr.fundamental_ransac(m, 1.5)
    got_inl, got_f = gpu.estimate_fundamental_ransac(m, 1.5)
    assert want_f is not None and got_f is not None
    assert np.array_equal(got_inl, want_inl)
    assert np.abs(got_f - want_f).max() <= 1e-9


def test_fundamental_batch_and_iterations(gpu):
    ms = [scene(i) for i in range(len(SCENES))]
    off = np.cumsum([0] + [len(x) for x in ms])
    probs = np.zeros(len(ms), gpu.FUNDAMENTAL_PROBLEM_DTYPE)
    probs["offset"], probs["n"] = off[:-1], [len(x) for x in ms]
    probs["seed"] = [psr.FUNDAMENTAL_DEFAULT_SEED + 17 * k for k in range(len(ms))]
    res, inl = gpu.fundamental_batch(np.concatenate(ms), probs, 1.5)
    for k, x in enumerate(ms):
        wi, wf, wit = psr.fundamental_ransac(x, 1.5, seed=int(probs["seed"][k]))
        assert bool(res[k]["has_model"]) == (wf is not None)
        assert res[k]["n_inliers"] == len(wi)
        assert np.array_equal(inl[off[k]:off[k] + len(wi)], wi)
        if wf is not None:
            assert np.array_equal(res[k]["f"].reshape(3, 3), wf)
            assert res[k]["iters"] == wit


@pytest.mark.parametrize("i", range(len(SCENES)))
def test_gric_equals_oracle(gpu, i):
    m = scene(i)
    _, H = psr.homography_ransac_ref(m, 3.0)
    _, F, _ = psr.fundamental_ransac(m, 1.5)
    for h, f in ((H, F), (H, None), (None, F)):
        if h is None and f is None:
            continue
        want = psr.select_model_gric(m, h, f, 1.0)
        got = gpu.select_model_gric(m, h, f, 1.0)
        assert (got is None) == (want is None)
        if want is not None:
            assert got[0] == ("Fundamental" if want[0] else "Homography")
            assert got[1] == want[1] and got[2] == want[2]


@pytest.mark.parametrize("i", range(len(SCENES)))
@pytest.mark.parametrize("seed", [None, 12345])
def test_similarity_equals_oracle(gpu, i, seed):
    m = scene(i)
    want_inl, want = psr.filter_matches_ransac(m, 3.0, *( [] if seed is None else [seed]))
    got_inl, got = gpu.estimate_similarity(m, gpu.RansacOptions(3.0, seed))
    assert np.array_equal(got_inl, want_inl)
    assert (got is None) == (want is None)
    if want is not None:
        assert got == tuple(np.float32(want[k]) for k in ("a", "b", "tx", "ty", "scale", "rotation_deg"))
