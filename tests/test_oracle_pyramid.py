"""Pins the oracle's scale-space pyramid (pipeline.rs:275-659) with closed forms,
float64 re-derivations and the reference's own invariants."""
import math

import numpy as np
import pytest

import psr


def checkerboard_512():
    """pipeline_tests.rs:31-51"""
    img = np.zeros((512, 512), np.float32)
    for y in range(64, 448):
        for x in range(64, 448):
            cy, cx = (y - 64) // 32, (x - 64) // 32
            if (cx + cy) % 2 == 0:
                ly, lx = (y - 64) % 32, (x - 64) % 32
                if 4 <= ly < 28 and 4 <= lx < 28:
                    img[y, x] = 1.0
    return img


@pytest.mark.parametrize("sigma", [0.5, 1.0, 1.2490, 1.6, 2.4525, 3.09, 4.0])
def test_kernel_closed_form(sigma):
    k = psr.gaussian_kernel_1d(sigma)
    r = max(math.ceil(3 * sigma), 1)
    assert len(k) == 2 * r + 1 and len(k) % 2 == 1
    x = np.arange(-r, r + 1, dtype=np.float64)
    ref = np.exp(-x * x / (2 * sigma * sigma))
    ref /= ref.sum()
    assert np.allclose(k, ref, rtol=2e-6, atol=1e-8)
    assert abs(float(k.astype(np.float64).sum()) - 1.0) < 1e-5        # pipeline_tests.rs:196-200
    assert np.all(np.abs(k - k[::-1]) < 1e-6)                         # :202-208


def test_kernel_sigma4_len25():
    assert len(psr.gaussian_kernel_1d(4.0)) == 25                     # pipeline_tests.rs:1067-1068


def test_blur_dc_invariance():
    out = psr.gaussian_blur(np.full((64, 64), 0.75, np.float32), 1.6)  # pipeline_tests.rs:217-225
    assert np.all(np.abs(out - 0.75) < 1e-4)


def _direct_blur_f64(img, sigma):
    """Non-separable 2-D convolution in float64 with clamped (edge-replicated) borders."""
    k = psr.gaussian_kernel_1d(sigma).astype(np.float64)
    r = len(k) // 2
    h, w = img.shape
    pad = np.pad(img.astype(np.float64), r, mode="edge")
    out = np.zeros((h, w))
    for dy in range(2 * r + 1):
        for dx in range(2 * r + 1):
            out += k[dy] * k[dx] * pad[dy:dy + h, dx:dx + w]
    return out


@pytest.mark.parametrize("shape,sigma", [((40, 50), 1.6), ((23, 31), 3.09), ((5, 7), 3.09), ((1, 9), 1.2), ((12, 3), 2.0)])
def test_blur_vs_direct_f64(shape, sigma):
    rng = np.random.default_rng(shape[0] * 100 + shape[1])
    img = rng.random(shape).astype(np.float32)
    out = psr.gaussian_blur(img, sigma)
    assert np.max(np.abs(out - _direct_blur_f64(img, sigma))) < 1e-6


def test_upsample_bilinear_on_ramp():
    h, w = 9, 13
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = (0.01 * xx + 0.02 * yy + 0.1).astype(np.float32)
    up = psr.upsample_2x(img).astype(np.float64)
    uy, ux = np.mgrid[0:2 * h, 0:2 * w].astype(np.float64)
    sy = np.minimum(uy / 2.0, h - 1)
    sx = np.minimum(ux / 2.0, w - 1)
    # interior: exact bilinear midpoints; last row/col replicate (clamped x1/y1)
    ref = 0.01 * sx + 0.02 * sy + 0.1
    ref[:, -1] = 0.01 * (w - 1) + 0.02 * sy[:, -1] + 0.1
    ref[-1, :] = 0.01 * sx[-1, :] + 0.02 * (h - 1) + 0.1
    ref[-1, -1] = 0.01 * (w - 1) + 0.02 * (h - 1) + 0.1
    # pipeline.rs:477: x1 = min(x0 + (ux&1), w-1) -> odd last column averages col w-1 with itself
    assert up.shape == (2 * h, 2 * w)
    assert np.max(np.abs(up - ref)) < 1e-6


def test_downsample_is_stride2():
    img = np.random.default_rng(1).random((11, 8)).astype(np.float32)
    assert np.array_equal(psr.downsample_2x(img), img[0:10:2, 0:8:2])


def test_luma_conversion():
    u8 = np.arange(256, dtype=np.uint8).reshape(16, 16)
    f = psr.luma_to_f32(u8)
    assert np.array_equal(f, u8.astype(np.float32) * np.float32(1.0 / 255.0))
    assert f[-1, -1] <= 1.0 and f[0, 0] == 0.0


def test_octave_dimensions():
    """pipeline_tests.rs:167-189"""
    pyr = psr.build_pyramid(checkerboard_512(), 3, 3, False)
    assert [p["dog"].shape for p in pyr] == [(5, 512, 512), (5, 256, 256), (5, 128, 128)]
    assert [p["gauss"].shape[0] for p in pyr] == [6, 6, 6]
    assert [p["scale_factor"] for p in pyr] == [1.0, 2.0, 4.0]
    up = psr.build_pyramid(np.zeros((10, 6), np.float32), 8, 3, True)
    # 20x12 -> 10x6 -> 5x3 -> stop at 2x1 (< 3)
    assert [p["dog"].shape[1:] for p in up] == [(20, 12), (10, 6), (5, 3)]
    assert [p["scale_factor"] for p in up] == [0.5, 1.0, 2.0]


@pytest.mark.parametrize("scales", [3, 5])
def test_pyramid_sigma_schedule_and_dog(scales):
    """Each Gaussian level equals a float64 blur of the previous one with
    delta = sqrt(target_s^2 - target_{s-1}^2), target_s = 1.6 * 2^(s/S); DoG = difference;
    the next octave starts from level S subsampled by 2."""
    rng = np.random.default_rng(scales)
    img = rng.random((48, 40)).astype(np.float32)
    pyr = psr.build_pyramid(img, 2, scales, False)
    for o, p in enumerate(pyr):
        g, d = p["gauss"], p["dog"]
        for s in range(1, scales + 3):
            t1, t0 = 1.6 * 2 ** (s / scales), 1.6 * 2 ** ((s - 1) / scales)
            delta = math.sqrt(t1 * t1 - t0 * t0)
            assert np.max(np.abs(g[s] - _direct_blur_f64(g[s - 1], delta))) < 2e-6
            assert np.array_equal(d[s - 1], g[s] - g[s - 1])
        if o + 1 < len(pyr):
            assert np.array_equal(pyr[o + 1]["gauss"][0], g[scales][::2, ::2][:pyr[o + 1]["gauss"].shape[1],
                                                                                :pyr[o + 1]["gauss"].shape[2]])
    # first level: init sigma sqrt(1.6^2 - 0.5^2) applied to the input (native start)
    init = math.sqrt(1.6 ** 2 - 0.25)
    assert np.max(np.abs(pyr[0]["gauss"][0] - _direct_blur_f64(img, init))) < 2e-6


def test_pyramid_upsampled_first_octave():
    rng = np.random.default_rng(7)
    img = rng.random((20, 30)).astype(np.float32)
    pyr = psr.build_pyramid(img, 1, 3, True)
    init = math.sqrt(1.6 ** 2 - 1.0)
    up = psr.upsample_2x(img)
    assert np.max(np.abs(pyr[0]["gauss"][0] - _direct_blur_f64(up, init))) < 2e-6


def test_threads_do_not_change_results():
    """tests/invariance_sweeps.rs:219-244: identical output across thread counts."""
    img = checkerboard_512()
    psr.set_threads(1)
    a = psr.build_pyramid(img, 3, 3, True)
    psr.set_threads(4)
    b = psr.build_pyramid(img, 3, 3, True)
    psr.set_threads(0)
    for x, y in zip(a, b):
        assert np.array_equal(x["gauss"], y["gauss"]) and np.array_equal(x["dog"], y["dog"])
