"""Pins the oracle's orientation / descriptor (pipeline.rs:1158-1695), matcher
(pipeline.rs:1697-1858) and Hamming primitive (catalog.rs:37-45, :707-728)."""
import math

import numpy as np
import pytest

import psr
import synth
from test_oracle_pyramid import checkerboard_512

THR = 0.015
TWO_PI = 2 * math.pi


def test_descriptor_invariants_checkerboard():
    """pipeline_tests.rs:254-279"""
    f, d = psr.extract_features(checkerboard_512(), 3, 3, THR, True)
    assert len(f) > 0
    assert np.all((f["orientation"] >= 0) & (f["orientation"] < TWO_PI))
    assert np.all(np.abs(np.linalg.norm(d, axis=1) - 1) < 1e-4)
    assert np.all((d >= 0) & (d <= 1))


def test_secondary_orientation_split():
    """pipeline_tests.rs:989-1033: the symmetric cross emits more features than keypoints,
    and at least one keypoint carries two orientations more than 0.1 rad apart."""
    img = np.zeros((96, 96), np.float32)
    img[24:72, 42:54] = 1
    img[42:54, 24:72] = 1
    kps = psr.detect_keypoints(img, 2, 3, THR, True)
    f, _ = psr.extract_features(img, 2, 3, THR, True)
    assert len(kps) > 0 and len(f) > len(kps)
    split = False
    for k in kps:
        m = f[(np.abs(f["x"] - k["x"]) < 1e-5) & (np.abs(f["y"] - k["y"]) < 1e-5) & (np.abs(f["scale"] - k["scale"]) < 1e-5)]
        if len(m) >= 2 and abs(m["orientation"][0] - m["orientation"][1]) > 0.1:
            split = True
    assert split


def test_rootsift_and_u8_quantisation():
    """pipeline_tests.rs:1282-1332 (RootSIFT unit norm; u8 round-trip L2 < 0.03, pipeline.rs:1024-1044)."""
    img = checkerboard_512()
    fs, ds = psr.extract_features(img, 3, 3, THR, False, 0)
    fr, dr = psr.extract_features(img, 3, 3, THR, False, 1)
    assert len(fs) == len(fr) > 0
    assert np.all(np.abs(np.linalg.norm(dr, axis=1) - 1) < 1e-4)
    # RootSIFT closed form from the SIFT descriptor: sqrt(d / sum(d))
    ref = np.sqrt(ds.astype(np.float64) / ds.astype(np.float64).sum(axis=1, keepdims=True))
    assert np.max(np.abs(dr - ref)) < 1e-6
    q = np.clip(np.rint(dr * 512.0), 0, 255).astype(np.uint8)
    back = q.astype(np.float32) / 512.0
    back /= np.linalg.norm(back, axis=1, keepdims=True)
    assert np.max(np.linalg.norm(back - dr, axis=1)) < 0.03


def test_descriptor_scale_invariance():
    """pipeline_tests.rs:1064-1136: 1x vs 0.5x descriptor L2 < 0.25."""
    blobs = [(128.0, 128.0, 3.2, 1.0), (140.0, 122.0, 2.4, -0.65), (118.0, 136.0, 2.8, 0.55)]
    def render(n, scale):
        yy, xx = np.mgrid[0:n, 0:n].astype(np.float32)
        v = np.zeros((n, n), np.float32)
        for cx, cy, s, a in blobs:
            cx, cy, s = np.float32(cx * scale), np.float32(cy * scale), np.float32(s * scale)
            v += np.float32(a) * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (np.float32(2) * s * s))
        return np.clip(v, 0, 1).astype(np.float32)
    f1, d1 = psr.extract_features(render(256, 1.0), 4, 3, THR, True)
    fh, dh = psr.extract_features(render(128, 0.5), 4, 3, THR, True)
    best = np.inf
    for i in range(len(f1)):
        for j in range(len(fh)):
            pe = math.hypot(0.5 * f1["x"][i] - fh["x"][j], 0.5 * f1["y"][i] - fh["y"][j])
            sr = 0.5 * f1["scale"][i] / fh["scale"][j]
            if pe < 1.5 and 0.75 <= sr <= 1.35:
                best = min(best, float(np.linalg.norm(d1[i] - dh[j])))
    assert best < 0.25


def _f64_descriptor(mag, ang, kp, theta):
    """float64 restatement of pipeline.rs:1352-1480 for one keypoint (window in numpy)."""
    h, w = mag.shape
    c, s = math.cos(theta), math.sin(theta)
    hw = max(3 * kp["scale"], 1.5)
    R = max(math.ceil(hw * math.sqrt(2) * 5 * 0.5), 4)
    cx, cy = int(round(kp["x"])), int(round(kp["y"]))
    desc = np.zeros(128)
    for y in range(max(cy - R, 1), max(min(cy + R, h - 2), 1) + 1):
        for x in range(max(cx - R, 1), max(min(cx + R, w - 2), 1) + 1):
            dx, dy = x - kp["x"], y - kp["y"]
            rr = (dx * c + dy * s) / hw
            cc = (-dx * s + dy * c) / hw
            rb, cb = rr + 1.5, cc + 1.5
            if rb <= -1 or rb >= 4 or cb <= -1 or cb >= 4 or mag[y, x] <= 1e-7:
                continue
            wm = mag[y, x] * math.exp(-(rr * rr + cc * cc) / 8)
            ob = ((ang[y, x] - theta) % TWO_PI) * 8 / TWO_PI
            r0, c0, o0 = math.floor(rb), math.floor(cb), math.floor(ob)
            for ir in (0, 1):
                for ic in (0, 1):
                    for io in (0, 1):
                        if 0 <= r0 + ir < 4 and 0 <= c0 + ic < 4:
                            wgt = (rb - r0 if ir else 1 - (rb - r0)) * (cb - c0 if ic else 1 - (cb - c0)) * \
                                  (ob - o0 if io else 1 - (ob - o0))
                            desc[((r0 + ir) * 4 + c0 + ic) * 8 + (o0 + io) % 8] += wm * wgt
    desc /= np.linalg.norm(desc)
    desc = np.minimum(desc, 0.2)
    return desc / np.linalg.norm(desc)


def _f64_orientation_hist(mag, ang, kp):
    h, w = mag.shape
    so = max(1.5 * kp["scale"], 1.0)
    R = max(math.ceil(3 * so), 2)
    cx, cy = int(round(kp["x"])), int(round(kp["y"]))
    hist = np.zeros(36)
    for y in range(max(cy - R, 1), max(min(cy + R, h - 2), 1) + 1):
        for x in range(max(cx - R, 1), max(min(cx + R, w - 2), 1) + 1):
            r2 = (x - cx) ** 2 + (y - cy) ** 2
            if r2 > R * R or mag[y, x] <= 1e-7:
                continue
            hist[int(math.floor(ang[y, x] * 36 / TWO_PI)) % 36] += mag[y, x] * math.exp(-r2 / (2 * so * so))
    sm = (np.roll(hist, 1) + hist + np.roll(hist, -1)) / 3
    return sm


def test_orientation_and_descriptor_vs_float64_restatement():
    img = psr.luma_to_f32(synth.blob_image(160, 120, 25, seed=9))
    pyr = psr.build_pyramid(img, 2, 3, False)
    kps = psr.detect_keypoints(img, 2, 3, THR, False)
    assert len(kps) > 3
    checked = 0
    for kp in kps[:12]:
        o = int(kp["octave"])
        g = pyr[o]["gauss"][int(kp["layer"])]
        mag, ang = psr.gradient_planes(g)
        # gradient planes: closed form with zero border (pipeline.rs:1168-1203)
        gx = np.zeros_like(g); gy = np.zeros_like(g)
        gx[1:-1, 1:-1] = g[1:-1, 2:] - g[1:-1, :-2]
        gy[1:-1, 1:-1] = g[2:, 1:-1] - g[:-2, 1:-1]
        assert np.array_equal(mag, np.sqrt(gx * gx + gy * gy))
        assert np.max(np.abs(ang - np.mod(np.arctan2(gy, gx), TWO_PI)) *
                      (np.abs(np.abs(ang - np.mod(np.arctan2(gy, gx), TWO_PI)) - TWO_PI) > 1e-3)) < 1e-5
        loc = kp.copy()
        inv = np.float32(1.0) / np.float32(pyr[o]["scale_factor"])
        loc["x"], loc["y"] = kp["x"] * inv, kp["y"] * inv
        loc["scale"] = max(kp["scale"] * inv, np.float32(0.8))
        oris = psr.keypoint_orientations(mag, ang, loc)
        sm = _f64_orientation_hist(mag.astype(np.float64), ang.astype(np.float64), loc)
        best = sm.max()
        # every returned orientation sits on a >=80% local peak of the float64 histogram
        for t in oris:
            b = int(math.floor(t / (TWO_PI / 36) - 0.5 + 1e-9)) % 36
            cand = [(b + d) % 36 for d in (-1, 0, 1)]
            assert any(sm[c] >= 0.8 * best - 1e-6 for c in cand)
        for t in oris:
            d = psr.descriptor_128(mag, ang, loc, t)
            ref = _f64_descriptor(mag.astype(np.float64), ang.astype(np.float64), loc, float(t))
            assert np.linalg.norm(d - ref) < 1e-4
            checked += 1
    assert checked >= 5


def test_rotation_covariance_90deg():
    """Rotating the image by 90 degrees rotates every keypoint's orientation by pi/2 and
    leaves its descriptor unchanged (the gradient field rotates exactly on the pixel grid).
    One octave only: stride-2 subsampling is not covariant under a 90-degree rotation."""
    img = psr.luma_to_f32(synth.blob_image(128, 128, 20, seed=21))
    f0, d0 = psr.extract_features(img, 1, 3, THR, False)
    rot = np.ascontiguousarray(np.rot90(img))        # new[i, j] = old[j, W-1-i]
    f1, d1 = psr.extract_features(rot, 1, 3, THR, False)
    W = img.shape[1]
    matched = 0
    for i in range(len(f0)):
        nx, ny = f0["y"][i], (W - 1) - f0["x"][i]
        m = np.nonzero((np.abs(f1["x"] - nx) < 1e-3) & (np.abs(f1["y"] - ny) < 1e-3))[0]
        for j in m:
            dt = (f1["orientation"][j] - f0["orientation"][i]) % TWO_PI
            if min(abs(dt - 1.5 * math.pi), abs(dt - 1.5 * math.pi + TWO_PI)) < 1e-3 or \
               min(abs(dt - 0.5 * math.pi), abs(dt - 0.5 * math.pi - TWO_PI)) < 1e-3:
                if np.linalg.norm(d1[j] - d0[i]) < 1e-3:
                    matched += 1
                    break
    assert len(f0) >= 5 and matched >= 0.8 * len(f0), (matched, len(f0))


# ------------------------------------------------------------------ matching
def _unit(rng, n):
    x = np.maximum(rng.standard_normal((n, 128)), 0).astype(np.float32)
    return x / np.linalg.norm(x, axis=1, keepdims=True)


def test_nn2_vs_float64_brute_force():
    rng = np.random.default_rng(0)
    q, t = _unit(rng, 300), _unit(rng, 500)
    bi, b, s = psr.nn2(q, t)
    d = ((q[:, None, :].astype(np.float64) - t[None, :, :].astype(np.float64)) ** 2).sum(-1)
    order = np.argsort(d, axis=1, kind="stable")
    rows = np.arange(len(q))
    gap = d[rows, order[:, 1]] - d[rows, order[:, 0]]
    robust = gap > 1e-5
    assert robust.mean() > 0.95
    assert np.array_equal(bi[robust], order[robust, 0])
    assert np.allclose(b, d[rows, order[:, 0]], rtol=1e-5, atol=1e-6)
    assert np.allclose(s, d[rows, order[:, 1]], rtol=1e-5, atol=1e-6)


def test_nn2_ties_take_first_index():
    rng = np.random.default_rng(1)
    t = _unit(rng, 10)
    t[7] = t[2]
    q = t[[2, 7]].copy()
    bi, b, s = psr.nn2(q, t)
    assert list(bi) == [2, 2] and np.all(b == 0) and np.all(s == 0)


def test_match_ratio_and_fields():
    rng = np.random.default_rng(2)
    q, t = _unit(rng, 200), _unit(rng, 300)
    t[:100] = q[:100] + 0.01 * rng.standard_normal((100, 128)).astype(np.float32)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    qh = np.zeros(len(q), psr.FEAT_DTYPE); th = np.zeros(len(t), psr.FEAT_DTYPE)
    qh["x"] = np.arange(len(q)); th["x"] = np.arange(len(t)) * 2.0; th["y"] = 5.0
    m = psr.match(qh, q, th, t, 0.8)
    d = np.sqrt(((q[:, None, :].astype(np.float64) - t[None].astype(np.float64)) ** 2).sum(-1))
    srt = np.sort(d, axis=1)
    ratio = srt[:, 0] / srt[:, 1]
    exp = np.nonzero(ratio < 0.8)[0]
    assert np.array_equal(m["query_index"], exp)
    assert np.allclose(m["ratio"], ratio[exp], rtol=1e-5)
    assert np.array_equal(m["dx"], th["x"][m["target_index"]] - qh["x"][m["query_index"]])
    assert np.all(m["target_index"][:50] == np.arange(50)[: len(m["target_index"][:50])])


def test_match_degenerate_inputs():
    rng = np.random.default_rng(3)
    q, t = _unit(rng, 4), _unit(rng, 1)
    h4, h1 = np.zeros(4, psr.FEAT_DTYPE), np.zeros(1, psr.FEAT_DTYPE)
    assert len(psr.match(h4, q, h1, t, 0.8)) == 0           # target.len() < 2 (pipeline.rs:1749)
    assert len(psr.match(h4[:0], q[:0], h4, q, 0.8)) == 0   # empty query


def test_mutual_known_answer():
    """pipeline_tests.rs:1221-1280"""
    def desc(v0, v1):
        d = np.zeros(128, np.float32)
        n = np.float32(np.sqrt(np.float32(v0 * v0 + v1 * v1)))
        d[0], d[1] = np.float32(v0) / n, np.float32(v1) / n
        return d
    qh = np.zeros(2, psr.FEAT_DTYPE); th = np.zeros(2, psr.FEAT_DTYPE)
    qd = np.stack([desc(1.0, 0.0), desc(0.95, 0.31)]); td = np.stack([desc(1.0, 0.0), desc(0.0, 1.0)])
    assert len(psr.match(qh, qd, th, td, 0.75, False)) == 2
    m = psr.match(qh, qd, th, td, 0.75, True)
    assert len(m) == 1 and m[0]["query_index"] == 0 and m[0]["target_index"] == 0


# ------------------------------------------------------------------ Hamming
def test_quantize_u128_closed_form():
    rng = np.random.default_rng(4)
    d = _unit(rng, 50)
    bits = d > np.float32(0.065)
    w = np.zeros((50, 2), np.uint64)
    for i in range(128):
        w[:, i // 64] |= bits[:, i].astype(np.uint64) << np.uint64(i % 64)
    assert np.array_equal(psr.quantize_u128(d), w)


def test_hamming_topk_vs_numpy():
    cat, qs, src, flips = synth.hamming_catalog(20000, 40, seed=5)
    idx, dist = psr.hamming_topk(cat, qs, k=100)
    for i in range(len(qs)):
        x = np.bitwise_count(cat ^ qs[i]).sum(axis=1).astype(np.int64)
        order = np.lexsort((np.arange(len(cat)), x))[:100]
        assert np.array_equal(idx[i], order.astype(np.uint32))
        assert np.array_equal(dist[i], x[order].astype(np.uint8))
        assert dist[i, 0] <= flips[i]


def test_coarse_votes():
    cat, qs, src, flips = synth.hamming_catalog(500, 30, seed=6)
    v = psr.count_coarse_hamming_votes(qs, cat)
    exp = sum(1 for i in range(len(qs)) if np.bitwise_count(cat ^ qs[i]).sum(axis=1).min() <= 24)
    assert v == exp


def test_rerank_known_answer():
    """Rows at exactly representable offsets from the query: squared distances are exact,
    equal distances order by index, invalid slots go last."""
    q = np.zeros((1, 128), np.float32)
    q[0, :] = 0.25
    deltas = [0.5, 0.25, 1.0, 0.25, 0.125]
    cd = np.tile(q, (len(deltas), 1))
    for i, dl in enumerate(deltas):
        cd[i, i] += dl
    sl = np.array([[0, 1, 2, 3, 4, 0xFFFFFFFF, 9]], np.uint32)
    idx, d2 = psr.rerank_l2(q, cd, sl)
    assert idx[0].tolist() == [4, 1, 3, 0, 2, 0xFFFFFFFF, 0xFFFFFFFF]
    assert d2[0, :5].tolist() == [0.015625, 0.0625, 0.0625, 0.25, 1.0]
    assert np.all(np.isinf(d2[0, 5:]))


def test_rerank_vs_float64():
    q, db, _ = synth.unit_descriptors(20, 3000, seed=4)
    rng = np.random.default_rng(4)
    sl = rng.integers(0, 3000, (20, 64)).astype(np.uint32)
    idx, d2 = psr.rerank_l2(q, db, sl)
    ref = ((db[sl.astype(np.int64)].astype(np.float64) - q[:, None, :].astype(np.float64)) ** 2).sum(-1)
    for i in range(20):
        got = ((db[idx[i].astype(np.int64)].astype(np.float64) - q[i].astype(np.float64)) ** 2).sum(-1)
        assert np.allclose(got, d2[i], rtol=1e-5, atol=1e-6)
        assert np.all(np.diff(d2[i]) >= 0)
        assert sorted(idx[i].tolist()) == sorted(sl[i].tolist())
        assert np.allclose(np.sort(ref[i]), d2[i], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("upsample", [True, False])
def test_compute_descriptors_infers_the_detection_parameters(upsample):
    """pipeline.rs:1591-1617: for keypoints detected with (4 octaves, 3 scales, upsample) the
    inferred pyramid is that one (scales from the largest layer, upsampling from the first
    keypoint's scale / (1.6 * 2^octave * k^layer) < 0.707, octaves from the largest octave), so
    the features equal those computed on the explicit pyramid."""
    img = psr.luma_to_f32(synth.blob_image(320, 240, 60, 3))
    kps = psr.detect_keypoints(img, 4, 3, 0.015, upsample)
    assert len(kps) > 20
    f1, d1 = psr.compute_descriptors(img, kps, 0)
    f2, d2 = psr.descriptors_for_keypoints(img, kps, int(kps["octave"].max()) + 1, 3, upsample, 0)
    assert np.array_equal(f1, f2) and np.array_equal(d1, d2)


def test_compute_descriptors_layer0_resolution():
    """resolve_octave_and_layer (pipeline.rs:1206-1227): a layer-0 keypoint of scale
    1.6 * 2^o * 2^(l/3) takes octave o and layer l; the explicit (o, l) keypoint of the same
    scale (no upsampling: scale / expected = 1) gives the same feature."""
    img = psr.luma_to_f32(synth.blob_image(320, 240, 60, 4))
    for o, l in [(0, 1), (1, 2), (2, 1), (1, 1)]:
        s = np.float32(1.6 * 2 ** o * 2 ** (l / 3))
        k0 = np.zeros(2, psr.KP_DTYPE)
        k0["x"], k0["y"], k0["scale"], k0["response"] = [120, 200], [100, 150], s, 0.02
        k0[1]["scale"] = np.float32(1.6 * 2 ** 2)   # same octave count for both calls
        k1 = k0.copy()
        k1[0]["octave"], k1[0]["layer"] = o, l
        k1[1]["octave"], k1[1]["layer"] = 2, 1   # scale 6.4: octave floor(log2 4) = 2, local sigma 1 -> layer 1
        f0, d0 = psr.compute_descriptors(img, k0, 0)
        f1, d1 = psr.compute_descriptors(img, k1, 0)
        assert np.array_equal(d0, d1), (o, l)
