"""Pins the oracle's panorama pairwise stage (oracle/ps_oracle_panorama.c, psr.panorama_*):
the reference's own safe-alignment tests (pipeline_tests.rs:1035-1061,
coverage_gaps.rs:108-146), closed forms of the sanity test (area bounds, orientation,
horizon), the patch sampler against hand-computed patches (clamping at borders, Rust's
round-half-away-from-zero, non-finite points), and the reference's 4-tile panorama scene
(tests/panorama_tests.rs:29-79) rebuilt from a synthetic image: every neighbouring tile pair
gets an edge whose alignment is the known horizontal shift and whose mean-intensity ratio
follows the applied exposure factors."""
import numpy as np
import pytest

import psr
import synth


def test_reference_degenerate_warp_falls_back_to_similarity():
    # pipeline_tests.rs:1035-1061
    deg = np.array([[-0.7689, -0.2804, 1340.6], [-0.6761, -0.2462, 1178.4], [-0.000574, -0.000209, 1.0]],
                   np.float32)
    assert not psr.is_sane_homography(deg, 1024, 1024)
    sim = dict(a=0.99, b=-0.05, tx=12.0, ty=-8.0)
    h = psr.resolve_safe_alignment(deg, 8, sim, 1024, 1024)
    assert h is not None
    assert abs(h[2, 0]) < 1e-6 and abs(h[0, 2] - 12.0) < 1e-4
    # visualize.rs:296: [[a, -b, tx], [b, a, ty], [0, 0, 1]]
    assert h[0, 1] == np.float32(0.05) and h[1, 0] == np.float32(-0.05) and h[1, 1] == np.float32(0.99)
    assert h[1, 2] == np.float32(-8.0) and h[2, 2] == 1.0


def test_reference_resolve_fallbacks():
    # coverage_gaps.rs:108-146
    sane = np.array([[1, 0, 12], [0, 1, -8], [0, 0, 1]], np.float32)
    sim = dict(a=1.0, b=0.0, tx=5.0, ty=-3.0)
    assert np.array_equal(psr.resolve_safe_alignment(sane, 6, sim, 200, 200), sane)
    low = psr.resolve_safe_alignment(sane, 5, sim, 200, 200)
    assert low[0, 2] == 5.0 and low[1, 2] == -3.0
    deg = np.array([[1, 0, 0], [0, 1, 0], [0.05, 0.05, 1]], np.float32)
    assert psr.resolve_safe_alignment(deg, 20, sim, 200, 200)[0, 2] == 5.0
    assert psr.resolve_safe_alignment(None, 0, None, 200, 200) is None
    assert psr.resolve_safe_alignment(deg, 20, None, 200, 200) is None


@pytest.mark.parametrize("s,ok", [(0.19, False), (0.21, True), (1.0, True), (4.9, True), (5.1, False)])
def test_sane_area_bounds(s, ok):
    # a uniform scale s maps the source rectangle to s^2 times its area: sane iff 0.04 <= s^2 <= 25
    h = np.array([[s, 0, 3], [0, s, -2], [0, 0, 1]], np.float32)
    assert psr.is_sane_homography(h, 640, 480) == ok


def test_sane_orientation_and_horizon():
    mirror = np.array([[-1, 0, 639], [0, 1, 0], [0, 0, 1]], np.float32)
    assert psr.is_sane_homography(mirror, 640, 480)            # all four turns negative: still convex
    rot = np.array([[0, -1, 480], [1, 0, 0], [0, 0, 1]], np.float32)
    assert psr.is_sane_homography(rot, 640, 480)
    horizon = np.array([[1, 0, 0], [0, 1, 0], [-0.004, 0, 1]], np.float32)   # w < 0 at the right edge
    assert not psr.is_sane_homography(horizon, 640, 480)
    squash = np.array([[1, 0, 0], [0, 0, 0], [0, 0, 1]], np.float32)        # collinear corners
    assert not psr.is_sane_homography(squash, 640, 480)
    assert not psr.is_sane_homography(np.eye(3, dtype=np.float32), 1, 480)


def test_patch_intensity_hand_computed():
    u8 = (np.arange(5 * 6).reshape(5, 6) * 7 % 256).astype(np.uint8)
    g = psr.luma_to_f32(u8)
    v = g.astype(np.float64)

    def patch(y0, y1, x0, x1):
        s = 0.0
        for yy in range(y0, y1 + 1):
            for xx in range(x0, x1 + 1):
                s += v[yy, xx]
        return s / ((y1 - y0 + 1) * (x1 - x0 + 1)) * 255.0

    assert psr.sample_patch_intensity(g, 2.0, 1.0) == patch(0, 2, 1, 3)      # interior 3x3
    assert psr.sample_patch_intensity(g, -3.0, -7.0) == patch(0, 1, 0, 1)    # clamped corner 2x2
    assert psr.sample_patch_intensity(g, 99.0, 2.2) == patch(1, 3, 4, 5)     # right edge 3x2
    assert psr.sample_patch_intensity(g, 0.5, 3.5) == patch(3, 4, 0, 2)      # halves round up: (1, 4)
    assert psr.sample_patch_intensity(g, 2.49, 1.51) == patch(1, 3, 1, 3)    # (2, 2)
    assert psr.sample_patch_intensity(g, -0.5, 0.0) == patch(0, 1, 0, 1)     # -0.5 -> -1 -> 0
    assert psr.sample_patch_intensity(g, float("nan"), 1.0) == 0.0
    assert psr.sample_patch_intensity(g, 1.0, float("inf")) == 0.0
    c = psr.luma_to_f32(np.full((9, 9), 200, np.uint8))
    assert abs(psr.sample_patch_intensity(c, 4.0, 4.0) - 200.0) < 1e-4


def test_four_tile_panorama_edges():
    """tests/panorama_tests.rs:29-79 on a synthetic base image: 4 tiles 42% wide, exposures
    0.85 / 1.0 / 1.18 / 0.92. Neighbouring tiles overlap by 45%: each such pair has an edge,
    its alignment is the horizontal shift by the tile step, and mean_i_from / mean_i_to follows
    the exposure ratio."""
    f = [0.85, 1.0, 1.18, 0.92]
    tiles, step = synth.panorama_tiles(1200, 400, 4, 300, seed=3, exposure=f)
    res = psr.panorama_pairs(tiles, 3, 3, 0.015, 0.75, 3.0)
    assert [(i, j) for i, j, _, _ in res] == [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    by = {(i, j): (n, e) for i, j, n, e in res}
    for i in range(3):
        n, e = by[(i, i + 1)]
        assert e is not None and e["inliers"] >= 100 and n >= e["inliers"]
        h = e["h"] / e["h"][2, 2]
        assert abs(h[0, 2] + step) < 1.5 and abs(h[1, 2]) < 1.5
        assert np.allclose(h[:2, :2], np.eye(2), atol=0.01)
        assert np.allclose(e["h_inv"] @ e["h"], np.eye(3), atol=1e-9)
        assert abs(e["mean_i"] / e["mean_j"] - f[i] / f[i + 1]) < 0.03 * f[i] / f[i + 1]
    # fewer than 6 raw matches make no edge
    for (i, j), (n, e) in by.items():
        if n < 6:
            assert e is None
