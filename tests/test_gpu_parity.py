"""GPU parity: the HIP pipeline (through the C ABI) against the CPU oracle on the
same seeded inputs.

Bars:
  * Gaussian/DoG pyramid, keypoint x/y/response/octave/layer, match indices,
    distances and ratios: bit-exact.
  * keypoint scale: within 2 ulp (the device evaluates k^ds in f64 and rounds once;
    glibc powf differs from that by 1 ulp in ~0.06% of arguments, and the product
    eff_scale * k^ds can round that to 2 ulp), i.e. << 1e-4 octave.
  * descriptors: L2 <= 1e-3 per feature on unit vectors, orientation within 1e-3 rad,
    after aligning features by keypoint (orientation-count flips are counted and
    bounded, since device expf/atan2f differ from glibc in the last ulp).
"""
import numpy as np
import pytest

import psr
import synth

pytestmark = pytest.mark.gpu

SMALL_SHAPES = [(64, 64), (37, 53), (7, 5), (3, 3), (96, 130)]


def _rand_img(h, w, seed):
    rng = np.random.default_rng(seed)
    img = rng.random((h, w)).astype(np.float32)
    # add smooth structure so the pyramid has extrema
    yy, xx = np.mgrid[0:h, 0:w]
    img += np.sin(xx / 3.0).astype(np.float32) * np.cos(yy / 5.0).astype(np.float32)
    return np.clip(img * 0.5, 0, 1).astype(np.float32)


def _blob_f32(w, h, n, seed, band=False):
    a = synth.blob_image(w, h, n, seed, band)
    return a, psr.luma_to_f32(a)


@pytest.mark.parametrize("shape", SMALL_SHAPES)
@pytest.mark.parametrize("upsample", [True, False])
def test_pyramid_bit_exact_small(gpu, shape, upsample):
    img = _rand_img(*shape, seed=shape[0] * 7 + shape[1])
    cfg = gpu.SiftConfig().octaves(4).scales(3).upsample_first_octave(upsample)
    dev = gpu.Sift(cfg).build_pyramid(img)
    ref = psr.build_pyramid(img, 4, 3, upsample)
    assert len(dev) == len(ref)
    for (g, d), r in zip(dev, ref):
        assert np.array_equal(g, r["gauss"]), "Gaussian planes differ"
        assert np.array_equal(d, r["dog"]), "DoG planes differ"


@pytest.mark.parametrize("scales", [3, 4, 6, 10])
def test_pyramid_bit_exact_scales(gpu, scales):
    u8, img = _blob_f32(160, 120, 30, seed=scales)
    cfg = gpu.SiftConfig().octaves(3).scales(scales)
    dev = gpu.Sift(cfg).build_pyramid(img)
    ref = psr.build_pyramid(img, 3, scales, True)
    assert len(dev) == len(ref)
    for (g, d), r in zip(dev, ref):
        assert np.array_equal(g, r["gauss"])
        assert np.array_equal(d, r["dog"])


@pytest.mark.parametrize("shape", [(800, 1000), (799, 1001)])
def test_pyramid_bit_exact_large(gpu, shape):
    """Upsampled first octave of >= 3M pixels: its levels take the sliding-window blur
    (128-row chunks, ragged last chunk and strip for the odd shape)."""
    u8, img = _blob_f32(shape[1], shape[0], 200, seed=shape[1])
    cfg = gpu.SiftConfig().octaves(3).scales(3)
    dev = gpu.Sift(cfg).build_pyramid(img)
    ref = psr.build_pyramid(img, 3, 3, True)
    assert len(dev) == len(ref)
    for (g, d), r in zip(dev, ref):
        assert np.array_equal(g, r["gauss"])
        assert np.array_equal(d, r["dog"])


def _cmp_keypoints(dev, ref):
    assert len(dev) == len(ref), (len(dev), len(ref))
    for f in ("x", "y", "response", "octave", "layer"):
        assert np.array_equal(dev[f], ref[f]), f
    ulp = np.abs(dev["scale"].view(np.int32).astype(np.int64) - ref["scale"].view(np.int32).astype(np.int64))
    assert ulp.max(initial=0) <= 2
    return int((ulp > 0).sum())


@pytest.mark.parametrize("upsample", [True, False])
def test_keypoints_exact_640x480(gpu, upsample):
    u8, img = _blob_f32(640, 480, 200, seed=0)
    cfg = gpu.SiftConfig().octaves(4).scales(3).upsample_first_octave(upsample)
    dev = gpu.Sift(cfg).detect_keypoints(img)
    ref = psr.detect_keypoints(img, 4, 3, 0.015, upsample)
    assert len(ref) > 50
    _cmp_keypoints(dev, ref)


@pytest.mark.parametrize("shape", SMALL_SHAPES + [(24, 300), (300, 24)])
def test_keypoints_exact_small(gpu, shape):
    """Small and thin levels, where the window of the level S+2 values formed around slice-S
    candidates reaches past the borders on every side and is clamped like the blur's."""
    img = _rand_img(*shape, seed=shape[0] * 11 + shape[1])
    cfg = gpu.SiftConfig().octaves(4).scales(3).contrast_threshold(0.005)
    dev = gpu.Sift(cfg).detect_keypoints(img)
    ref = psr.detect_keypoints(img, 4, 3, 0.005, True)
    _cmp_keypoints(dev, ref)


def test_keypoints_exact_with_nan_pixels(gpu):
    """An f32 image with NaN pixels: the blur spreads them over a patch of every level, and
    the extremum test must reject exactly what the reference's strict comparisons reject
    (a NaN neighbour fails all of them), so tiles holding a NaN take the literal comparison
    path; the keypoints elsewhere, and their number, equal the oracle's."""
    _, img = _blob_f32(640, 480, 200, seed=4)
    img = img.copy()
    img[100, 100] = np.nan
    img[300:303, 500] = np.nan
    img[479, 0] = np.nan
    cfg = gpu.SiftConfig().octaves(4).scales(3)
    dev = gpu.Sift(cfg).detect_keypoints(img)
    ref = psr.detect_keypoints(img, 4, 3, 0.015, True)
    assert len(ref) > 50 and np.isfinite(ref["x"]).all()
    _cmp_keypoints(dev, ref)


@pytest.mark.parametrize("scales", [3, 4, 5, 6, 7, 8, 9, 10])
def test_keypoints_exact_scales(gpu, scales):
    """Every supported slice count: the survivor masks of the extremum kernel's in-plane pass
    fill one or two 64-bit words (the upsampled first octave, 1200 x 900, takes 32-row tiles,
    8 rows per thread: S = 8 fills one word, S > 8 flushes twice; the later octaves 16-row
    tiles), and each S has its own radius (10, 8, 7, 6, 5, 5, 5, 4) for the level S+2 values
    that the extremum kernel forms around slice-S candidates instead of reading a stored plane."""
    u8, img = _blob_f32(600, 450, 200, seed=scales)
    cfg = gpu.SiftConfig().octaves(3).scales(scales)
    dev = gpu.Sift(cfg).detect_keypoints(img)
    ref = psr.detect_keypoints(img, 3, scales, 0.015, True)
    assert len(ref) > 10
    _cmp_keypoints(dev, ref)


def _align_features(dh, dd, rh, rd):
    """Pairs device and oracle features by (x, y, response) in order; returns index pairs
    and the number of keypoints whose orientation count differs."""
    from collections import defaultdict
    rk = defaultdict(list)
    for i, h in enumerate(rh):
        rk[(h["x"].tobytes(), h["y"].tobytes(), h["response"].tobytes())].append(i)
    dk = defaultdict(list)
    for i, h in enumerate(dh):
        dk[(h["x"].tobytes(), h["y"].tobytes(), h["response"].tobytes())].append(i)
    pairs, flips = [], 0
    for k, ri in rk.items():
        di = dk.get(k, [])
        if len(di) != len(ri):
            flips += 1
            continue
        pairs.extend(zip(di, ri))
    return pairs, flips


def _check_features(dev, rh, rd, max_flip_frac=0.01):
    dh, dd = dev.header, dev.descriptors
    pairs, flips = _align_features(dh, dd, rh, rd)
    nkp = len({(h["x"].tobytes(), h["y"].tobytes()) for h in rh})
    assert flips <= max(1, max_flip_frac * nkp), f"{flips} orientation-count flips of {nkp}"
    assert len(pairs) >= 0.98 * len(rh)
    di = np.array([p[0] for p in pairs])
    ri = np.array([p[1] for p in pairs])
    dori = np.abs(dh["orientation"][di] - rh["orientation"][ri])
    dori = np.minimum(dori, 2 * np.pi - dori)
    ok_ori = dori <= 1e-3
    assert ok_ori.mean() >= 0.99, f"orientation agreement {ok_ori.mean()}"
    l2 = np.linalg.norm(dd[di] - rd[ri], axis=1)
    assert np.percentile(l2[ok_ori], 99) <= 1e-3 and (l2[ok_ori] <= 1e-3).mean() >= 0.99, l2.max()
    norms = np.linalg.norm(dd, axis=1)
    assert np.all(np.abs(norms - 1) < 1e-4)
    return flips, float(l2[ok_ori].max())


@pytest.mark.parametrize("kind", [0, 1])
def test_detect_and_compute_640x480(gpu, kind):
    u8, img = _blob_f32(640, 480, 200, seed=0)
    cfg = gpu.SiftConfig().octaves(4).scales(3).descriptor_kind(kind)
    dev = gpu.Sift(cfg).detect_and_compute(u8)
    rh, rd = psr.extract_features(img, 4, 3, 0.015, True, kind)
    _check_features(dev, rh, rd)
    dev_f = gpu.Sift(cfg).detect_and_compute(img)
    assert np.array_equal(dev_f.header, dev.header) and np.array_equal(dev_f.descriptors, dev.descriptors)


def test_counts_start_at_zero_each_extraction(gpu):
    """An extraction's counters are zeroed by its first blur launch, or by one fill when the
    image is too small for any octave (base level under 3 px: 2 x 2 without upsampling, 1 x 1
    with it). Extractions right after one with hundreds of features report only their own
    (none, like the oracle), and the first image extracted again is byte-equal to its first run."""
    u8, img = _blob_f32(640, 480, 200, seed=0)
    cfg = gpu.SiftConfig().octaves(4).scales(3)
    first = gpu.Sift(cfg).detect_and_compute(u8)
    assert len(first) > 100
    for shape, up in [((2, 2), False), ((1, 1), True), ((2, 5), False)]:
        tiny = np.full(shape, 0.5, np.float32)
        tiny[0, 0] = 0.9
        rh, _ = psr.extract_features(tiny, 4, 3, 0.015, up)
        dev = gpu.Sift(gpu.SiftConfig().octaves(4).scales(3).upsample_first_octave(up)).detect_and_compute(tiny)
        assert len(dev) == len(rh) == 0, shape
        assert len(gpu.Sift(cfg).detect_keypoints(tiny)) == 0, shape
    again = gpu.Sift(cfg).detect_and_compute(u8)
    assert np.array_equal(again.header, first.header) and np.array_equal(again.descriptors, first.descriptors)


def test_detect_and_compute_capped_1080p(gpu):
    """1080p band-noise image: > 4096 keypoints, so the stable |response| cap is exercised."""
    u8, _, _ = synth.blob_pair(1920, 1080, 2000, seed=1, band_noise=True)
    img = psr.luma_to_f32(u8)
    cfg = gpu.SiftConfig().octaves(4).scales(3)
    kps = gpu.Sift(cfg).detect_keypoints(img)
    assert len(kps) > 4096
    dev = gpu.Sift(cfg).detect_and_compute(u8)
    rh, rd = psr.extract_features(img, 4, 3, 0.015, True)
    assert len(dev) == 4096 == len(rh)
    # the keypoint order after the cap is exact, so headers agree except for orientation flips
    flips, _ = _check_features(dev, rh, rd)
    assert flips <= 8


@pytest.mark.parametrize("thr", [0.015, 0.005])
def test_order_with_unpacked_scan_keys(gpu, thr):
    """An image whose upsampled first octave is 4200 px wide: its scan keys do not pack into
    31 bits, so the ordering takes the select / gather / count path (the direct all-pairs
    ordering needs rows and columns < 4096). Uncapped keypoints in scan order, and the capped
    features (15.7k candidates at contrast 0.005), equal the oracle's."""
    a = synth.blob_image(2100, 240, 300, 3, True)
    img = psr.luma_to_f32(a)
    cfg = gpu.SiftConfig().octaves(4).scales(3).contrast_threshold(thr)
    _cmp_keypoints(gpu.Sift(cfg).detect_keypoints(img), psr.detect_keypoints(img, 4, 3, thr, True))
    dev = gpu.Sift(cfg).detect_and_compute(a)
    rh, rd = psr.extract_features(img, 4, 3, thr, True)
    assert len(dev) == len(rh) and (thr > 0.01 or len(rh) == 4096)
    flips, _ = _check_features(dev, rh, rd)
    assert flips <= 8


def test_keypoints_exact_3000x1500(gpu):
    """A 3000x1500 image: its octave 1 (3000 x 1500 px, more than 4M) keeps every octave's leaf
    work (level S+1 and extrema) forked right after its own chain -- smaller images fork octave
    0's after octave 1's chain -- and its 6000 px wide first octave takes the select / gather /
    count ordering. 21k keypoints in scan order equal the oracle's."""
    a = synth.blob_image(3000, 1500, 1500, 7, True)
    img = psr.luma_to_f32(a)
    cfg = gpu.SiftConfig().octaves(4).scales(3)
    ref = psr.detect_keypoints(img, 4, 3, 0.015, True)
    assert len(ref) > 20000
    _cmp_keypoints(gpu.Sift(cfg).detect_keypoints(img), ref)


def test_order_beyond_the_direct_limit_1080p(gpu):
    """1080p at contrast 0.006: 55k candidates, more than the direct ordering takes (32768), so
    the select / gather / count path orders them -- by scan key uncapped, by |response| for the
    4096-feature cap -- and both equal the oracle's."""
    u8, _, _ = synth.blob_pair(1920, 1080, 2000, seed=1, band_noise=True)
    img = psr.luma_to_f32(u8)
    cfg = gpu.SiftConfig().octaves(4).scales(3).contrast_threshold(0.006)
    ref = psr.detect_keypoints(img, 4, 3, 0.006, True)
    assert len(ref) > 32768
    _cmp_keypoints(gpu.Sift(cfg).detect_keypoints(img), ref)
    dev = gpu.Sift(cfg).detect_and_compute(u8)
    rh, rd = psr.extract_features(img, 4, 3, 0.006, True)
    assert len(dev) == 4096 == len(rh)
    flips, _ = _check_features(dev, rh, rd)
    assert flips <= 8


@pytest.mark.parametrize("upsample", [True, False])
@pytest.mark.parametrize("kind", [0, 1])
def test_compute_descriptors_for_detected_keypoints(gpu, upsample, kind):
    """The uncapped CLI extract path (cli.rs:174-212): detect_keypoints, then
    compute_sift_descriptors_with_kind (pipeline.rs:1581-1620), whose pyramid parameters are
    inferred from the keypoints; compared with the oracle's restatement at descriptor tolerance."""
    u8, img = _blob_f32(640, 480, 200, seed=11)
    cfg = gpu.SiftConfig().octaves(4).scales(3).upsample_first_octave(upsample).descriptor_kind(kind)
    sift = gpu.Sift(cfg)
    kps = sift.detect_keypoints(img)
    assert len(kps) > 50
    dev = sift.compute_descriptors(img, kps)
    rh, rd = psr.compute_descriptors(img, kps, kind)
    assert len(rh) >= len(kps)
    flips, _ = _check_features(dev, rh, rd)


def test_compute_descriptors_layer0_keypoints(gpu):
    """Keypoints without octave/layer (layer 0): octave floor(log2(max(scale/1.6, 1))), layer from
    the local sigma (resolve_octave_and_layer, pipeline.rs:1206-1227), no upsampling."""
    u8, img = _blob_f32(400, 300, 120, seed=5)
    rng = np.random.default_rng(3)
    k = np.zeros(300, gpu.KP_DTYPE)
    k["x"] = rng.uniform(2, 397, 300).astype(np.float32)
    k["y"] = rng.uniform(2, 297, 300).astype(np.float32)
    k["scale"] = rng.uniform(0.5, 20.0, 300).astype(np.float32)
    k["response"] = rng.uniform(-0.1, 0.1, 300).astype(np.float32)
    dev = gpu.Sift(gpu.SiftConfig()).compute_descriptors(img, k)
    rh, rd = psr.compute_descriptors(img, k, 0)
    assert len(rh) >= 300
    _check_features(dev, rh, rd)


def test_compute_descriptors_large_scales(gpu):
    """Keypoints at scales far beyond detection's (octave-local sigma up to 90): descriptor cell
    boxes of 256+ pixels take the kernel's direct path instead of the packed pixel lists."""
    u8, img = _blob_f32(1100, 1000, 300, seed=8)
    k = np.zeros(6, gpu.KP_DTYPE)
    k["x"] = [550, 500, 600, 300, 800, 550]
    k["y"] = [500, 450, 550, 300, 700, 520]
    k["scale"] = [8, 40, 60, 70, 80, 90]
    k["octave"] = 0
    k["layer"] = 1
    dev = gpu.Sift(gpu.SiftConfig()).compute_descriptors(img, k)
    rh, rd = psr.compute_descriptors(img, k, 0)
    assert len(rh) >= 6
    _check_features(dev, rh, rd, max_flip_frac=0.2)


def test_determinism(gpu):
    u8, img = _blob_f32(640, 480, 200, seed=3)
    cfg = gpu.SiftConfig().octaves(4)
    a = gpu.Sift(cfg).detect_and_compute(u8)
    b = gpu.Sift(cfg).detect_and_compute(u8)
    assert np.array_equal(a.header, b.header)
    assert np.array_equal(a.descriptors, b.descriptors)


@pytest.mark.parametrize("nq,nt", [(1, 2), (5, 3), (300, 257), (1000, 1500), (64, 1)])
def test_nn2_exact(gpu, nq, nt):
    rng = np.random.default_rng(nq * 31 + nt)
    q = np.maximum(rng.standard_normal((nq, 128)), 0).astype(np.float32)
    t = np.maximum(rng.standard_normal((nt, 128)), 0).astype(np.float32)
    q /= np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    t /= np.maximum(np.linalg.norm(t, axis=1, keepdims=True), 1e-12)
    if nt >= 3:
        t[nt // 2] = t[0]          # exact duplicate target -> tie on best
        q[0] = t[0]                # zero distance
    bi, b, s = gpu.nn2(q, t)
    rbi, rb, rs = psr.nn2(q, t)
    assert np.array_equal(bi, rbi)
    assert np.array_equal(b, rb)
    assert np.array_equal(s, rs)


@pytest.mark.parametrize("eps", [0.0, 1e-7, 1e-5])
def test_nn2_exact_near_ties(gpu, eps):
    """Clusters of near-equidistant targets: the MFMA candidate guard must route every
    ambiguous query to the exact scan, so indices and distances stay bit-exact."""
    rng = np.random.default_rng(int(eps * 1e9) + 7)
    base = np.maximum(rng.standard_normal((40, 128)), 0).astype(np.float32)
    base /= np.linalg.norm(base, axis=1, keepdims=True)
    t = np.repeat(base, 12, axis=0) + np.float32(eps) * rng.standard_normal((480, 128)).astype(np.float32)
    q = base + np.float32(1e-3) * rng.standard_normal((40, 128)).astype(np.float32)
    q = np.concatenate([q, t[::7]])
    bi, b, s = gpu.nn2(q, t)
    rbi, rb, rs = psr.nn2(q, t)
    assert np.array_equal(bi, rbi)
    assert np.array_equal(b, rb)
    assert np.array_equal(s, rs)


@pytest.mark.parametrize("mutual", [False, True])
def test_match_exact_on_oracle_features(gpu, mutual):
    a, b, _ = synth.blob_pair(640, 480, 200, seed=5)
    qh, qd = psr.extract_features(psr.luma_to_f32(a), 4, 3, 0.015, True)
    th, td = psr.extract_features(psr.luma_to_f32(b), 4, 3, 0.015, True)
    ref = psr.match(qh, qd, th, td, 0.8, mutual)
    m = gpu.Matcher(gpu.MatchOptions(0.8, mutual)).match_(gpu.Features(qh, qd), gpu.Features(th, td))
    assert len(ref) > 20
    assert np.array_equal(m, ref)


@pytest.mark.parametrize("mutual", [False, True])
def test_match_exact_beyond_4096_queries(gpu, mutual):
    """6000 queries against 5000 targets (planted unit descriptors): past 16 blocks of 256
    queries the ratio test and the ordered compaction take the flag + compaction launches
    instead of the single finishing launch; the records equal the oracle's."""
    q, db, _ = synth.unit_descriptors(6000, 5000, seed=33)
    rng = np.random.default_rng(33)
    qh = np.zeros(6000, gpu.FEAT_DTYPE)
    th = np.zeros(5000, gpu.FEAT_DTYPE)
    for h in (qh, th):
        h["x"] = rng.uniform(0, 1920, len(h))
        h["y"] = rng.uniform(0, 1080, len(h))
        h["scale"] = rng.uniform(1, 8, len(h))
    ref = psr.match(qh, q, th, db, 0.8, mutual)
    m = gpu.Matcher(gpu.MatchOptions(0.8, mutual)).match_(gpu.Features(qh, q), gpu.Features(th, db))
    assert len(ref) > 500
    assert np.array_equal(m, ref)


def test_mutual_known_answer(gpu):
    """pipeline_tests.rs:1221-1280"""
    def desc(v0, v1):
        d = np.zeros(128, np.float32)
        n = np.float32(np.sqrt(np.float32(v0 * v0 + v1 * v1)))
        d[0], d[1] = np.float32(v0) / n, np.float32(v1) / n
        return d
    def feats(xs, ds):
        h = np.zeros(len(xs), gpu.FEAT_DTYPE)
        h["x"] = xs
        h["y"] = 10.0
        h["scale"] = 1.6
        h["response"] = 0.05
        return gpu.Features(h, np.stack(ds))
    q = feats([10.0, 20.0], [desc(1.0, 0.0), desc(0.95, 0.31)])
    t = feats([12.0, 80.0], [desc(1.0, 0.0), desc(0.0, 1.0)])
    one = gpu.Matcher(gpu.MatchOptions(0.75, False)).match_(q, t)
    mut = gpu.Matcher(gpu.MatchOptions(0.75, True)).match_(q, t)
    assert len(one) == 2
    assert len(mut) == 1 and mut[0]["query_index"] == 0 and mut[0]["target_index"] == 0


@pytest.fixture(scope="module")
def oracle_1080p():
    """The 1920x1080 band-noise pair of the frame workload with the oracle's features (4096 per
    image after the cap)."""
    a, b, _ = synth.blob_pair(1920, 1080, 2000, seed=1, band_noise=True)
    q = psr.extract_features(psr.luma_to_f32(a), 4, 3, 0.015, True)
    t = psr.extract_features(psr.luma_to_f32(b), 4, 3, 0.015, True)
    return a, b, q, t


@pytest.mark.parametrize("mutual", [False, True])
def test_match_exact_on_oracle_features_1080p(gpu, oracle_1080p, mutual):
    """The headline match size: 4096 x 4096 oracle descriptors, ratio 0.8; indices, distances,
    ratios and displacements bit-exact with the oracle (pipeline.rs:1744-1841)."""
    _, _, (qh, qd), (th, td) = oracle_1080p
    assert len(qh) == 4096 and len(th) == 4096
    ref = psr.match(qh, qd, th, td, 0.8, mutual)
    m = gpu.Matcher(gpu.MatchOptions(0.8, mutual)).match_(gpu.Features(qh, qd), gpu.Features(th, td))
    assert len(ref) > 500
    assert np.array_equal(m, ref)


def _feature_map(dev, rh, rd):
    """GPU feature index -> oracle feature index for the features _align_features pairs."""
    pairs, _ = _align_features(dev.header, dev.descriptors, rh, rd)
    return dict(pairs)


def _end_to_end_agreement(gpu, a, b):
    """GPU extract + GPU match against oracle extract + oracle match on one pair: the matches
    are compared as (query, target) pairs of oracle feature indices (GPU features mapped through
    their bit-exact keypoints); a match of a feature whose orientation count differs from the
    oracle's counts as a disagreement. Returns (agreement over the union, both match counts)."""
    cfg = gpu.SiftConfig().octaves(4).scales(3)
    s = gpu.Sift(cfg)
    fq, ft = s.detect_and_compute(a), s.detect_and_compute(b)
    m = gpu.Matcher(gpu.MatchOptions(0.8)).match_(fq, ft)
    qh, qd = psr.extract_features(psr.luma_to_f32(a), 4, 3, 0.015, True)
    th, td = psr.extract_features(psr.luma_to_f32(b), 4, 3, 0.015, True)
    ref = psr.match(qh, qd, th, td, 0.8)
    qmap, tmap = _feature_map(fq, qh, qd), _feature_map(ft, th, td)
    got = {(qmap.get(int(q), -1 - int(q)), tmap.get(int(t), -1 - int(t)))
           for q, t in zip(m["query_index"], m["target_index"])}
    want = set(zip(ref["query_index"].tolist(), ref["target_index"].tolist()))
    agree = len(got & want) / max(len(got | want), 1)
    return agree, len(m), len(ref)


@pytest.mark.parametrize("size", [(640, 480, 200, False), (1920, 1080, 2000, True)])
def test_end_to_end_match_agreement(gpu, size):
    """GPU extract + GPU match vs oracle extract + oracle match, always asserted: keypoints are
    bit-exact, descriptors agree within 1e-3, so the only disagreements are ratio-test decisions
    within the descriptor tolerance of 0.8 and matches of features whose orientation count
    flipped. Measured on MI355X: 640x480 and 1080p pairs both above 0.99."""
    w, h, n, band = size
    a, b, _ = synth.blob_pair(w, h, n, seed=0 if w == 640 else 1, band_noise=band)
    agree, n_gpu, n_ref = _end_to_end_agreement(gpu, a, b)
    print(f"end-to-end agreement {w}x{h}: {agree:.5f} ({n_gpu} GPU / {n_ref} oracle matches)")
    assert n_ref > 50 and abs(n_gpu - n_ref) <= max(3, 0.01 * n_ref)
    assert agree >= 0.98, agree
