"""GPU parity for the grouped matcher (ps_engine_match_sets) and the panorama's pairwise stage
(stitch.rs:478-571) through the C ABI.

Bars, all bit-exact:
* grouped matcher: every problem's records and count equal the oracle's match_descriptors
  (ratio test, no mutual check) of its two sets -- empty, single-feature and clamped sets,
  an out-of-range set index, separate and shared query / target arrays;
* panorama: each pair record equals the oracle's restatement (psr.panorama_pairs) run on the
  GPU's own features of the same images (the batched extraction equals a single-image
  extraction, so these are the stage's features): raw match count, edge presence, kept inlier
  count, the f64 alignment and its inverse, both mean patch intensities; the host form's edge
  list is the reference's expansion (i -> j, then j -> i) of the records with an edge.
"""
import numpy as np
import pytest

import psr
import synth

pytestmark = pytest.mark.gpu


def _sets(rng, pool, sizes):
    """Features drawn from a shared pool of unit descriptors (noise 0.04), so that sets share
    near-duplicates and the ratio test accepts some pairs."""
    n = int(sum(sizes))
    f = np.zeros(n, psr.FEAT_DTYPE)
    for k in ("x", "y"):
        f[k] = rng.uniform(0, 640, n)
    f["scale"] = rng.uniform(1, 6, n)
    f["response"] = rng.uniform(0.01, 0.1, n)
    f["orientation"] = rng.uniform(0, 6.28, n)
    d = pool[rng.integers(0, len(pool), n)] + rng.normal(0, 0.04, (n, 128))
    d = np.abs(d)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint32)
    return f, d, offs


def _run_sets(gpu, qf, qd, qtab, tf, td, ttab, probs, cap, ratio, shared):
    import torch
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).ravel().copy()).to(dev)
    dqf, dqd, dqs = t(qf), t(qd), t(qtab)
    if shared:
        dtf, dtd, dts = dqf, dqd, dqs
    else:
        dtf, dtd, dts = t(tf), t(td), t(ttab)
    dp = t(probs)
    out_rows = int(max(probs["out_offset"])) + cap
    dout = torch.zeros(out_rows * gpu.MATCH_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    dcnt = torch.full((len(probs),), 7777, dtype=torch.int32, device=dev)
    eng = gpu.Engine(0)
    eng.match_sets(dqf.data_ptr(), dqd.data_ptr(), len(qf), dqs.data_ptr(), len(qtab), dtf.data_ptr(), dtd.data_ptr(),
                   len(tf), dts.data_ptr(), len(ttab), cap, dp.data_ptr(), len(probs), ratio, dout.data_ptr(),
                   dcnt.data_ptr())
    eng.sync()
    eng.close()
    return dout.cpu().numpy().view(gpu.MATCH_DTYPE), dcnt.cpu().numpy()


def _check_problem(out, cnt, k, p, qf, qd, qtab, tf, td, ttab, cap, ratio):
    def take(f, d, tab, idx):
        if idx >= len(tab):
            return f[:0], d[:0]
        o, c = int(tab[idx]["offset"]), min(int(tab[idx]["count"]), cap)
        return f[o:o + c], d[o:o + c]
    a, ad = take(qf, qd, qtab, int(p["query_set"]))
    b, bd = take(tf, td, ttab, int(p["target_set"]))
    want = psr.match(a, ad, b, bd, ratio)
    assert cnt[k] == len(want), (k, cnt[k], len(want))
    got = out[int(p["out_offset"]):int(p["out_offset"]) + len(want)]
    for name in want.dtype.names:
        assert np.array_equal(got[name], want[name]), (k, name)


@pytest.mark.parametrize("ratio", [0.75, 0.9])
def test_match_sets_equals_oracle(gpu, ratio):
    rng = np.random.default_rng(11)
    pool = np.abs(rng.normal(0, 1, (300, 128)))
    qsizes, tsizes = [0, 1, 2, 7, 130, 300, 257, 340], [0, 1, 3, 129, 260, 300, 64]
    qf, qd, qo = _sets(rng, pool, qsizes)
    tf, td, to = _sets(rng, pool, tsizes)
    cap = 320   # the 340-feature query set is clamped to its first 320 rows
    qtab = np.array(list(zip(qo, qsizes)), gpu.FEATURE_SET_DTYPE)
    ttab = np.array(list(zip(to, tsizes)), gpu.FEATURE_SET_DTYPE)
    pl = [(i, j) for i in range(len(qsizes)) for j in range(len(tsizes)) if (i + 2 * j) % 3 != 1] + [(99, 3), (4, 77)]
    probs = np.array([(i, j, k * cap, 0) for k, (i, j) in enumerate(pl)], gpu.MATCH_PROBLEM_DTYPE)
    out, cnt = _run_sets(gpu, qf, qd, qtab, tf, td, ttab, probs, cap, ratio, shared=False)
    assert cnt.sum() > 200
    for k, p in enumerate(probs):
        _check_problem(out, cnt, k, p, qf, qd, qtab, tf, td, ttab, cap, ratio)


def test_match_sets_shared_arrays(gpu):
    """One feature array and one set table on both sides (the panorama's layout): all pairs i < j."""
    rng = np.random.default_rng(12)
    pool = np.abs(rng.normal(0, 1, (500, 128)))
    sizes = [450, 0, 37, 512, 1, 200]
    f, d, o = _sets(rng, pool, sizes)
    cap = 512
    tab = np.array(list(zip(o, sizes)), gpu.FEATURE_SET_DTYPE)
    pl = [(i, j) for i in range(len(sizes)) for j in range(i + 1, len(sizes))]
    probs = np.array([(i, j, k * cap, 0) for k, (i, j) in enumerate(pl)], gpu.MATCH_PROBLEM_DTYPE)
    out, cnt = _run_sets(gpu, f, d, tab, f, d, tab, probs, cap, 0.75, shared=True)
    assert cnt.sum() > 100
    for k, p in enumerate(probs):
        _check_problem(out, cnt, k, p, f, d, tab, f, d, tab, cap, 0.75)


def _scene():
    tiles, step = synth.panorama_tiles(1200, 400, 4, 300, seed=3, exposure=[0.85, 1.0, 1.18, 0.92])
    other = synth.blob_image(380, 300, 120, seed=77, band_noise=True)   # unrelated: no edges
    return tiles + [other], step


def _oracle_pairs(gpu, images, opt):
    feats = []
    for im in images:
        F = gpu.extract_sift_features(im, opt.octaves, opt.scales, opt.threshold, True, gpu.DescriptorKind.Sift)
        feats.append((F.header, F.descriptors))
    return psr.panorama_pairs(images, opt.octaves, opt.scales, opt.threshold, opt.ratio, opt.ransac_threshold,
                              features=feats)


def _check_records(recs, want):
    assert len(recs) == len(want)
    for r, (i, j, n, e) in zip(recs, want):
        assert (r["from"], r["to"], r["raw_matches"]) == (i, j, n)
        assert bool(r["has_edge"]) == (e is not None), (i, j)
        if e is None:
            continue
        assert r["inliers"] == e["inliers"]
        assert np.array_equal(r["h_from_to"], e["h"].ravel()), (i, j)
        assert np.array_equal(r["h_to_from"], e["h_inv"].ravel()), (i, j)
        assert r["mean_i_from"] == e["mean_i"] and r["mean_i_to"] == e["mean_j"], (i, j)


def test_panorama_pairs_equal_oracle(gpu):
    import torch
    images, step = _scene()
    opt = gpu.PanoramaOptions()
    want = _oracle_pairs(gpu, images, opt)
    assert sum(e is not None for _, _, _, e in want) >= 3
    # device-resident form
    descs, chunks, off = [], [], 0
    for im in images:
        h, w = im.shape
        stride = w + 3
        buf = np.zeros((h, stride), np.uint8)
        buf[:, :w] = im
        descs.append((off, h, w, stride, 0))
        chunks.append(buf.ravel())
        off += buf.size
    dev = torch.device("cuda:0")
    pix = torch.from_numpy(np.concatenate(chunks)).to(dev)
    n = len(images)
    dpairs = torch.zeros(n * (n - 1) // 2 * gpu.PANORAMA_PAIR_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    eng = gpu.Engine(0)
    for _ in range(2):   # the second call replays the extraction graph
        dpairs.zero_()
        eng.panorama_pairs(pix.data_ptr(), np.array(descs, gpu.IMAGE_DESC_DTYPE), opt, dpairs.data_ptr())
        eng.sync()
        _check_records(dpairs.cpu().numpy().view(gpu.PANORAMA_PAIR_DTYPE), want)
    eng.close()
    # host form: the reference's edge list
    edges = gpu.panorama_pair_edges(images, opt)
    exp = []
    for i, j, _, e in want:
        if e is not None:
            exp.append((i, j, e["h"].ravel(), e["inliers"], e["mean_i"], e["mean_j"]))
            exp.append((j, i, e["h_inv"].ravel(), e["inliers"], e["mean_j"], e["mean_i"]))
    assert len(edges) == len(exp)
    for g, (a, b, h, ni, mf, mt) in zip(edges, exp):
        assert (g["from"], g["to"], g["inliers"]) == (a, b, ni)
        assert np.array_equal(g["h_from_to"], h)
        assert (g["mean_i_from"], g["mean_i_to"]) == (mf, mt)


def test_panorama_two_images_and_other_options(gpu):
    """Two images (one pair), 4 octaves, a looser ratio and a 5 px RANSAC threshold."""
    images, _ = _scene()
    opt = gpu.PanoramaOptions(octaves=4, scales=3, threshold=0.02, ratio=0.8, ransac_threshold=5.0)
    pair = [images[1], images[2]]
    want = _oracle_pairs(gpu, pair, opt)
    edges = gpu.panorama_pair_edges(pair, opt)
    (i, j, n, e), = want
    assert e is not None and len(edges) == 2
    assert np.array_equal(edges[0]["h_from_to"], e["h"].ravel())
    assert np.array_equal(edges[1]["h_from_to"], e["h_inv"].ravel())
    assert edges[0]["mean_i_from"] == e["mean_i"] and edges[1]["mean_i_from"] == e["mean_j"]


def test_panorama_uneven_sizes_and_blank_image(gpu):
    """Images of different sizes (the alignment's sanity test uses the query image's size), a
    flat image with no features (its pairs have no matches and no edge), and a 2x-downscaled
    tile: the records still equal the oracle's."""
    import torch
    images, _ = _scene()
    small = images[1][::2, ::2].copy()
    flat = np.full((200, 260), 97, np.uint8)
    imgs = [images[0][:, :600].copy(), images[1], flat, small, images[2][40:, :]]
    opt = gpu.PanoramaOptions()
    want = _oracle_pairs(gpu, imgs, opt)
    assert all(n == 0 and e is None for i, j, n, e in want if 2 in (i, j))
    assert any(e is not None for _, _, _, e in want)
    descs, chunks, off = [], [], 0
    for im in imgs:
        h, w = im.shape
        descs.append((off, h, w, w, 0))
        chunks.append(im.ravel())
        off += im.size
    dev = torch.device("cuda:0")
    pix = torch.from_numpy(np.concatenate(chunks)).to(dev)
    n = len(imgs)
    dpairs = torch.zeros(n * (n - 1) // 2 * gpu.PANORAMA_PAIR_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    eng = gpu.Engine(0)
    eng.panorama_pairs(pix.data_ptr(), np.array(descs, gpu.IMAGE_DESC_DTYPE), opt, dpairs.data_ptr())
    eng.sync()
    _check_records(dpairs.cpu().numpy().view(gpu.PANORAMA_PAIR_DTYPE), want)
    eng.close()


def _dense_lattice(n, period=4, sigma=0.9):
    """u8 image of identical bright blobs on a `period`-pixel lattice: about one extremum
    candidate per lattice site, so n = 4608 gives ~1.3M candidates, above the 2^20 capacity."""
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    py = np.exp(-((np.mod(yy, period) - period / 2) ** 2) / (2 * sigma ** 2))
    px = np.exp(-((np.mod(xx, period) - period / 2) ** 2) / (2 * sigma ** 2))
    return np.round((0.2 + 0.6 * py * px) * 255).astype(np.uint8)


def test_candidate_overflow_is_reported(gpu):
    """An image with more than 2^20 extremum candidates: the host panorama call reports
    PS_E_CAPACITY instead of edges built from incomplete features, as the single-image
    extraction does; the next call on the pooled engines is unaffected."""
    big = _dense_lattice(4608)
    small = synth.blob_image(320, 240, 40, seed=1)
    with pytest.raises(gpu.DeviceError) as ei:
        gpu.panorama_pair_edges([small, big, small])
    assert ei.value.status == gpu.PS_E_CAPACITY
    with pytest.raises(gpu.DeviceError) as ei:
        gpu.Sift(gpu.SiftConfig().octaves(3).scales(3)).detect_and_compute(big)
    assert ei.value.status == gpu.PS_E_CAPACITY
    tiles = synth.panorama_tiles(1600, 480, 3, 300, seed=4)[0]
    ok = gpu.panorama_pair_edges(tiles)
    assert len(ok) >= 2
