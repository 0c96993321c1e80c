"""Pins the float64 restatements (fp64_errors.py) used to state the f32 results' errors, and
states the oracle's own errors against them on the 640x480 pair.

Pins: the vectorised float64 descriptor and orientation histogram against the explicit-loop
float64 restatement of pipeline.rs:1233-1480 kept with the oracle tests; the float64 Taylor
step against an analytic quadratic whose extremum is known; the float64 pyramid against the
f32 oracle pyramid to f32 rounding."""
import math

import numpy as np
import pytest

import fp64_errors as fe
import psr
import synth
from test_oracle_describe_match import _f64_descriptor, _f64_orientation_hist

THR = 0.015


@pytest.fixture(scope="module")
def pair():
    a, b, _ = synth.blob_pair(640, 480, 200, seed=0)
    img = psr.luma_to_f32(a)
    return a, b, img, fe.pyramid64(img, 4, 3, True), psr.build_pyramid(img, 4, 3, True)


def test_pyramid64_against_f32_oracle(pair):
    _, _, _, p64, p32 = pair
    assert len(p64) == len(p32)
    for (g, m), r in zip(p64, p32):
        assert g.shape == r["gauss"].shape and m == r["scale_factor"]
        assert np.abs(g - r["gauss"]).max() < 2e-6


def test_refine64_analytic_quadratic():
    """D(o) = c + g.o + o^T H o / 2 sampled on the 3x3x3 grid: one Newton step lands exactly on
    the extremum -H^-1 g, and interp equals c + g.o / 2 there."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        A = rng.normal(size=(3, 3))
        Hm = -(A @ A.T + 3 * np.eye(3))          # negative definite: a maximum
        off = rng.uniform(-0.45, 0.45, 3)
        g = -Hm @ off
        c = 1.0
        cube = np.zeros((3, 3, 3))
        for s in range(3):
            for y in range(3):
                for x in range(3):
                    o = np.array([x - 1, y - 1, s - 1], float)
                    cube[s, y, x] = c + g @ o + 0.5 * o @ Hm @ o
        dx, dy, ds, interp, used3 = fe._refine64(cube)
        assert used3
        assert np.allclose([dx, dy, ds], off, atol=1e-12)
        assert math.isclose(interp, c + 0.5 * g @ off, rel_tol=1e-12)


def test_vectorised_descriptor_equals_loop_restatement(pair):
    _, _, img, p64, _ = pair
    kps = psr.detect_keypoints(img, 4, 3, THR, True)
    for kp in kps[:10]:
        o, layer = int(kp["octave"]), int(kp["layer"])
        g, m = p64[o]
        mag, ang = fe.gradient_planes64(g[layer])
        loc = kp.copy()
        loc["x"], loc["y"] = kp["x"] / m, kp["y"] / m
        loc["scale"] = max(kp["scale"] / m, 0.8)
        x, y, sc = float(loc["x"]), float(loc["y"]), float(loc["scale"])
        loc = {"x": x, "y": y, "scale": sc}       # python floats: the loop restatement stays in float64
        sm = _f64_orientation_hist(mag, ang, loc)
        oris = fe.orientations64(mag, ang, x, y, sc)
        assert abs(sm.argmax() - int(oris[0] / (2 * math.pi / 36))) <= 1
        for t in oris:
            assert np.abs(fe.descriptor64(mag, ang, x, y, sc, t) - _f64_descriptor(mag, ang, loc, t)).max() < 1e-12


def test_oracle_errors_against_fp64(pair):
    """The f32 oracle (and so the bit-exact GPU path) against float64: keypoint position and
    scale, descriptors, match distances and the homography, each within the stated bound."""
    a, b, img, p64, p32 = pair
    kps = psr.detect_keypoints(img, 4, 3, THR, True)
    ke = fe.keypoint_errors(kps, [(r["gauss"], r["dog"]) for r in p32], p64, 3)
    assert ke["unlocated"] == 0
    assert ke["solve"]["pos_px_max"] < 1e-4 and ke["solve"]["scale_oct_max"] < 1e-6
    assert ke["total"]["pos_px_p99"] < 1e-3 and ke["total"]["scale_oct_p99"] < 1e-4
    fh, fd = psr.extract_features(img, 4, 3, THR, True)
    de = fe.descriptor_errors(fh, fd, kps, p64, 4, sample=64)
    assert de["l2_max"] < 1e-5 and de["orientation_rad_median"] < 1e-5
    th, td = psr.extract_features(psr.luma_to_f32(b), 4, 3, THR, True)
    m = psr.match(fh, fd, th, td, 0.8)
    me = fe.match_errors(fd, td, m)
    assert me["distance_abs_max"] < 1e-6 and me["ratio_abs_max"] < 1e-6 and me["nearest_index_differs"] == 0
    inl, H = psr.homography_ransac_ref(m, 3.0)
    assert fe.homography_errors(m, inl, H)["reproj_px_max"] < 1e-3
