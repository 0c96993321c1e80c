"""GPU parity for the device-resident feature catalog (catalog.rs) through the C ABI, against
the oracle: PSIFTCAT load and save, upsert / remove, both coarse-vote paths and the full
search_features.

Bar: bit-exact for everything integer or index (entries, signatures, votes, shortlist, hit
order, match and inlier counts) and for the f32 values whose arithmetic is reproduced
operation by operation (dequantised descriptors, inlier ratio, mean / median displacement,
mean scale ratio, similarity and homography models); mean_rotation_deg sums device sinf /
cosf values, so it is held to 1e-3 degrees.
"""
import ctypes as C
import os

import numpy as np
import pytest

import psr
from catalog_scenes import (SCENE, brute_votes, random_code_catalog, scene_catalog, scene_features, through_file)

pytestmark = pytest.mark.gpu

EXACT = ("entry", "coarse_votes", "raw_match_count", "match_count", "inlier_ratio", "mean_dx", "mean_dy",
         "median_dx", "median_dy", "mean_scale_ratio", "has_transform", "has_homography")


def load(gpu, cat):
    """The oracle catalog written as a PSIFTCAT file, loaded on the device; and the oracle's
    own parse of the same bytes."""
    data = psr.catalog_write(cat)
    print(f"catalog: {len(cat.ids)} entries, {len(cat.feat)} features, {len(data)} bytes", flush=True)
    return gpu.FeatureCatalog.from_binary_bytes(data), psr.catalog_parse(data)


def features(gpu, f, d):
    return gpu.Features.from_arrays(f, d)


def compare_hits(got, want, ids):
    assert len(got) == len(want), (len(got), len(want))
    for g, w in zip(got, want):
        for k in EXACT:
            assert g[k] == w[k] or (k in ("median_dx", "median_dy") and g[k] == 0 == w[k]), \
                (ids[int(w["entry"])], k, g[k], w[k])
        if w["has_transform"]:
            assert np.array_equal(g["transform"], w["transform"]), (g["transform"], w["transform"])
        if w["has_homography"]:
            assert np.array_equal(g["homography"], w["homography"]), (g["homography"], w["homography"])
        assert abs(float(g["mean_rotation_deg"]) - float(w["mean_rotation_deg"])) <= 1e-3


def test_load_equals_oracle_parse(gpu):
    cat = scene_catalog([("a.png", 11), ("b.png", 12)])
    cat.ids[1] = "bé.png"
    g, ref = load(gpu, cat)
    assert len(g) == 2 and g.n_features == len(ref.feat)
    for e in range(2):
        s = g.summary(e)
        assert (s.id, s.width, s.height) == (ref.ids[e], ref.width[e], ref.height[e])
        f, d, sg = g.entry_features(e)
        a, b = int(ref.off[e]), int(ref.off[e + 1])
        assert np.array_equal(f, ref.feat[a:b])
        assert np.array_equal(sg, ref.sig[a:b])
        assert np.array_equal(d, ref.desc[a:b]), "dequantised descriptors differ"


@pytest.mark.parametrize("n_entries,per_entry,mih", [(20, (0, 300), False), (32, (1, 80), False),
                                                    (33, (0, 80), True), (90, (0, 40), True)])
def test_coarse_votes_both_paths(gpu, n_entries, per_entry, mih):
    """catalog.rs:595-633: <= 32 entries exhaustive, more through the inverted index."""
    cat, q, _ = random_code_catalog(n_entries, per_entry, 400, seed=n_entries)
    g, ref = load(gpu, cat)
    from catalog_scenes import bits_to_desc
    votes = g.coarse_votes(bits_to_desc(q))
    want = psr.catalog_votes(ref, q, mih)
    assert np.array_equal(want, brute_votes(ref, q, mih))
    assert np.array_equal(votes, want)


@pytest.mark.parametrize("n_entries,per_entry", [(2000, (100, 900)), (32, (15000, 25000))])
def test_coarse_votes_large(gpu, n_entries, per_entry):
    """A million catalog signatures: 2000 entries (inverted-index path) and 32 large ones
    (exhaustive path), 1000 planted queries."""
    cat, q, _ = random_code_catalog(n_entries, per_entry, 1000, seed=7)
    g, ref = load(gpu, cat)
    from catalog_scenes import bits_to_desc
    votes = g.coarse_votes(bits_to_desc(q))
    assert np.array_equal(votes, psr.catalog_votes(ref, q, n_entries > 32))


def scene_case(name):
    if name == "shifted":    # tests/pipeline_tests.rs:839-877
        return [(f"scene_{i}.png", 100 + i) for i in range(3)], (100, 16, 14), 3, 2
    if name == "mih":        # tests/pipeline_tests.rs:1438-1483
        spec = [(f"item_{i:02d}.png", 200 if i == 17 else (201 if i % 2 == 0 else 202)) for i in range(40)]
        return spec, (200, 0, 0), 3, 8
    if name == "exact":      # tests/coverage_gaps.rs:29-58
        return [("alpha", 301), ("beta", 307), ("gamma", 319)], (307, 0, 0), 3, 16
    if name == "many":       # every entry shortlisted, all hits returned
        return [(f"s{i}", 500 + i % 5) for i in range(12)], (502, 9, 4), 12, 12
    raise KeyError(name)


@pytest.mark.parametrize("name", ["shifted", "mih", "exact", "many"])
def test_search_equals_oracle(gpu, name):
    spec, (qs, dx, dy), top_k, shortlist = scene_case(name)
    g, ref = load(gpu, scene_catalog(spec))
    qf, qd = scene_features(qs, dx, dy)
    got = g.search_raw(features(gpu, qf, qd), top_k, shortlist, 0.75, 3.0)
    want = psr.catalog_search(ref, qf, qd, top_k, shortlist, 0.75, 3.0)
    assert len(want) >= 1
    compare_hits(got, want, ref.ids)
    hits = g.search_features(features(gpu, qf, qd), top_k, shortlist)
    assert hits[0].id == ref.ids[int(want[0]["entry"])]
    assert hits[0].tracking.match_count == int(want[0]["match_count"])


@pytest.mark.parametrize("thr", [1.0, 5.0, 3.0])
@pytest.mark.parametrize("ratio", [0.6, 0.8])
def test_search_thresholds(gpu, thr, ratio):
    """Another RANSAC threshold: summarize_tracking keeps its own 3 px similarity run."""
    g, ref = load(gpu, scene_catalog([(f"t{i}", 600 + i % 3) for i in range(5)]))
    qf, qd = scene_features(601, 5, 11)
    got = g.search_raw(features(gpu, qf, qd), 5, 16, ratio, thr)
    want = psr.catalog_search(ref, qf, qd, 5, 16, ratio, thr)
    compare_hits(got, want, ref.ids)


def test_device_query_form(gpu):
    import torch
    g, ref = load(gpu, scene_catalog([(f"scene_{i}.png", 100 + i) for i in range(3)]))
    qf, qd = scene_features(100, 16, 14)
    dev = torch.device("cuda:0")
    t_f = torch.from_numpy(qf.view(np.uint8).copy()).to(dev)
    t_d = torch.from_numpy(qd).to(dev)
    e = gpu.Engine(0)
    torch.cuda.synchronize()
    got = g.search_raw_device(e, t_f.data_ptr(), t_d.data_ptr(), len(qf), 3, 2)
    want = g.search_raw(features(gpu, qf, qd), 3, 2)
    assert np.array_equal(got, want)


def test_upsert_and_remove(gpu):
    """catalog.rs:309-339: upsert appends or replaces in place; remove swaps the last entry in."""
    fa, da = scene_features(701)
    fb, db = scene_features(702)
    fc, dc = scene_features(703)
    g = gpu.FeatureCatalog(3, 3, 0.015)
    g.upsert_entry("a", 1, 2, features(gpu, fa, da))
    g.upsert_entry("b", 3, 4, features(gpu, fb, db))
    g.upsert_entry("c", 5, 6, features(gpu, fc, dc))
    g.upsert_entry("b", 7, 8, features(gpu, fc[:7], dc[:7]))     # replaced in place
    assert [(s.id, s.width, s.feature_count) for s in g.list_summaries()] == [("a", 1, len(fa)), ("b", 7, 7),
                                                                                ("c", 5, len(fc))]
    f, d, sg = g.entry_features(1)
    assert np.array_equal(f, fc[:7]) and np.array_equal(d, dc[:7])
    assert np.array_equal(sg, psr.quantize_u128(dc[:7]))
    assert g.remove("a") and not g.remove("a")
    assert [s.id for s in g.list_summaries()] == ["c", "b"]       # Vec::swap_remove
    ref = psr.Catalog.from_entries([("c", 5, 6, fc, dc), ("b", 7, 8, fc[:7], dc[:7])])
    qf, qd = scene_features(703, 6, 3)
    compare_hits(g.search_raw(features(gpu, qf, qd), 5, 16), psr.catalog_search(ref, qf, qd, 5, 16), ref.ids)
    g.upsert_entry("d", 1, 1, features(gpu, fb, db))
    assert [s.id for s in g.list_summaries()] == ["c", "b", "d"]
    assert g.remove("d") and [s.id for s in g.list_summaries()] == ["c", "b"]


def test_save_binary_equals_oracle_writer(gpu, tmp_path):
    cat = scene_catalog([("x", 801), ("y", 802), ("z", 803)])
    g, ref = load(gpu, cat)
    path = os.path.join(str(tmp_path), "sub", "cat.psift")
    g.save_binary(path)
    data = open(path, "rb").read()
    assert data == psr.catalog_write(ref)
    assert not [p for p in os.listdir(os.path.dirname(path)) if ".tmp" in p]
    g2 = gpu.FeatureCatalog.load_auto(path)
    assert [s.id for s in g2.list_summaries()] == ["x", "y", "z"]
    jpath = os.path.join(str(tmp_path), "cat.json")
    g.save_auto(jpath)
    g3 = gpu.FeatureCatalog.load_auto(jpath)
    for e in range(3):
        a, b = g.entry_features(e), g3.entry_features(e)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        # load_json recomputes signatures from the descriptors (catalog.rs:386-395), which here
        # are the dequantised ones, so they need not equal the stored signatures
        assert np.array_equal(b[2], psr.quantize_u128(a[1]))


def test_trivial_searches(gpu):
    fa, da = scene_features(901)
    q = features(gpu, fa, da)
    empty = gpu.FeatureCatalog(3, 3, 0.015)
    assert len(empty.search_raw(q)) == 0
    g = gpu.FeatureCatalog(3, 3, 0.015)
    g.upsert_entry("one", 1, 1, features(gpu, fa[:1], da[:1]))     # < 2 features: never a hit
    g.upsert_entry("none", 1, 1, features(gpu, fa[:0], da[:0]))
    assert len(g.search_raw(q)) == 0
    g.upsert_entry("full", 1, 1, q)
    assert len(g.search_raw(q, top_k=0)) == 0
    assert len(g.search_raw(features(gpu, fa[:0], da[:0]))) == 0
    hits = g.search_raw(q, top_k=5)
    assert len(hits) == 1 and g.summary(int(hits[0]["entry"])).id == "full"
    n = C.c_size_t(0)
    out = np.zeros(1, gpu.CATALOG_HIT_DTYPE)
    hdr = np.ascontiguousarray(fa)
    st = gpu.lib().ps_catalog_search(g.h, hdr.ctypes.data, da.ctypes.data, len(fa), 5, 16, 0.75, 3.0,
                                     out.ctypes.data, 0, C.byref(n))
    assert st == gpu.PS_E_CAPACITY and n.value == 1
