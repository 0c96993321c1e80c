"""Pins the oracle's feature-catalog restatement (oracle/ps_oracle_catalog.c): the PSIFTCAT
byte layout, u8 descriptor quantisation, both coarse-vote paths, the shortlist and hit order
rules, summarize_tracking's closed forms, and the reference's own catalog tests."""
import math
import struct

import numpy as np
import pytest

import psr
from catalog_scenes import (SCENE, bits_to_desc, brute_votes, random_code_catalog, scene_catalog, scene_features,
                            through_file)


def f32_seq_norm(v):
    acc = np.zeros(len(v), np.float32)
    for i in range(v.shape[1]):
        acc = (acc + v[:, i] * v[:, i]).astype(np.float32)
    return np.sqrt(acc).astype(np.float32)


def test_u8_quantisation_closed_form():
    """pipeline.rs:1024-1030: round half away from zero, clamp to [0, 255], NaN -> 0."""
    rng = np.random.default_rng(1)
    d = rng.uniform(-0.2, 0.7, (64, 128)).astype(np.float32)
    d[0, :6] = np.array([0.5 / 512, 1.5 / 512, 2.5 / 512, -0.5 / 512, 255.5 / 512, np.nan], np.float32)
    d[1, :2] = np.array([np.nextafter(np.float32(0.5), np.float32(0)) / 512, 3.4e38], np.float32)
    x = (d * np.float32(512)).astype(np.float32).astype(np.float64)
    want = np.where(np.isnan(x), 0, np.clip(np.sign(x) * np.floor(np.abs(x) + 0.5), 0, 255)).astype(np.uint8)
    got = psr.quantize_u8(d)
    assert np.array_equal(got, want)
    assert list(got[0, :6]) == [1, 2, 3, 0, 255, 0]
    assert list(got[1, :2]) == [0, 255]                 # 0.49999997 rounds down; inf clamps


def test_u8_dequantisation_closed_form():
    """pipeline.rs:1034-1044: b / 512, divided by the sequential f32 norm when it exceeds 1e-7."""
    rng = np.random.default_rng(2)
    q = rng.integers(0, 256, (50, 128)).astype(np.uint8)
    q[3] = 0
    v = (q.astype(np.float32) * np.float32(1 / 512)).astype(np.float32)
    norm = f32_seq_norm(v)
    inv = (np.float32(1) / np.where(norm > 1e-7, norm, 1)).astype(np.float32)
    want = np.where((norm > 1e-7)[:, None], (v * inv[:, None]).astype(np.float32), v)
    got = psr.dequantize_u8(q)
    assert np.array_equal(got, want)
    assert not got[3].any()
    nz = np.arange(50) != 3
    assert np.allclose(np.linalg.norm(got[nz].astype(np.float64), axis=1), 1.0, atol=1e-6)


def small_catalog():
    rng = np.random.default_rng(3)
    entries = []
    for ident, n in (("a", 3), ("bé.png", 0), ("scene_2", 5)):
        f = np.zeros(n, psr.FEAT_DTYPE)
        for k in f.dtype.names:
            f[k] = rng.uniform(0, 50, n)
        d = rng.random((n, 128)).astype(np.float32)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        entries.append((ident, 100 + n, 80 + n, f, d))
    return psr.Catalog.from_entries(entries, octaves=4, scales=5, threshold=0.02, kind=1)


def test_format_layout_closed_form():
    """catalog.rs:400-445: 33-byte header, 20 + id bytes per entry, 164 bytes per feature."""
    cat = small_catalog()
    b = psr.catalog_write(cat)
    ids = [i.encode() for i in cat.ids]
    assert len(b) == 33 + sum(20 + len(i) for i in ids) + 164 * len(cat.feat)
    assert b[:8] == b"PSIFTCAT"
    version, octv, scales, thr, kind, ne = struct.unpack_from("<IIIfBQ", b, 8)
    assert (version, octv, scales, kind, ne) == (1, 4, 5, 1, 3)
    assert thr == np.float32(0.02)
    pos = 33
    f = 0
    for e, ident in enumerate(ids):
        (idl,) = struct.unpack_from("<I", b, pos)
        assert b[pos + 4:pos + 4 + idl] == ident
        w, h, fc = struct.unpack_from("<IIQ", b, pos + 4 + idl)
        assert (w, h, fc) == (cat.width[e], cat.height[e], cat.off[e + 1] - cat.off[e])
        pos += 20 + idl
        for _ in range(fc):
            rec = struct.unpack_from("<5f", b, pos)
            assert rec == tuple(float(cat.feat[f][k]) for k in psr.FEAT_DTYPE.names)
            lo, hi = struct.unpack_from("<QQ", b, pos + 20)
            assert (lo, hi) == (int(cat.sig[f, 0]), int(cat.sig[f, 1]))
            assert b[pos + 36:pos + 164] == psr.quantize_u8(cat.desc[f]).tobytes()
            pos += 164
            f += 1
    assert pos == len(b)


def hand_file(entries, octaves=3, scales=3, thr=0.015, kind=0, version=1, magic=b"PSIFTCAT"):
    out = [magic, struct.pack("<IIIfBQ", version, octaves, scales, thr, kind, len(entries))]
    for ident, w, h, feats in entries:
        out.append(struct.pack("<I", len(ident)) + ident + struct.pack("<IIQ", w, h, len(feats)))
        for rec, sig, q in feats:
            out.append(struct.pack("<5f", *rec) + struct.pack("<QQ", *sig) + bytes(q))
    return b"".join(out)


def test_parse_of_hand_assembled_file():
    """catalog.rs:447-540 field by field on a file assembled with struct.pack."""
    q0 = list(range(128))
    q1 = [255] * 64 + [0] * 64
    data = hand_file([("é".encode(), 7, 9, [((1, 2, 3, 4, 5), (11, 1 << 63), q0), ((6, 7, 8, 9, 10), (0, 5), q1)]),
                      (b"x", 1, 2, [])], octaves=2, scales=4, thr=0.5, kind=1) + b"trailing bytes are ignored"
    c = psr.catalog_parse(data)
    assert c.ids == ["é", "x"]
    assert list(c.width) == [7, 1] and list(c.height) == [9, 2]
    assert list(c.off) == [0, 2, 2]
    assert (c.octaves, c.scales, c.kind) == (2, 4, 1) and c.threshold == np.float32(0.5)
    assert tuple(c.feat[1]) == (6, 7, 8, 9, 10)
    assert [int(v) for v in c.sig[0]] == [11, 1 << 63]
    assert np.array_equal(c.desc, psr.dequantize_u8(np.array([q0, q1], np.uint8)))


@pytest.mark.parametrize("data,msg", [
    (b"PSIFT", "Unexpected EOF while reading binary catalog"),
    (b"NOTACAT!" + b"\0" * 40, "Invalid binary catalog magic header"),
    (hand_file([], version=2), "Unsupported binary catalog version 2 (expected 1)"),
    (hand_file([], kind=2), "Unknown descriptor kind 2 in binary catalog"),
    (hand_file([(b"abc", 1, 1, [])])[:-18], "Truncated entry id in binary catalog"),
    (hand_file([(b"\xff", 1, 1, [])]), "Invalid UTF-8 entry id in binary catalog"),
    (hand_file([(b"\xed\xa0\x80", 1, 1, [])]), "Invalid UTF-8 entry id in binary catalog"),
    (hand_file([(b"a", 1, 1, [((0,) * 5, (0, 0), [0] * 128)])])[:-1], "Unexpected EOF while reading binary catalog"),
])
def test_parse_errors(data, msg):
    with pytest.raises(ValueError, match="^" + msg.replace("(", r"\(").replace(")", r"\)") + "$"):
        psr.catalog_parse(data)


def test_write_parse_round_trip():
    cat = small_catalog()
    c2 = through_file(cat)
    assert c2.ids == cat.ids
    assert np.array_equal(c2.sig, cat.sig) and np.array_equal(c2.off, cat.off)
    assert np.array_equal(c2.feat, cat.feat)
    assert np.abs(c2.desc - cat.desc).max() < 2.0 / 512


def test_vote_paths_equal_their_predicates():
    """The inverted-index path (catalog.rs:597-620) votes exactly for a signature within 24
    bits sharing a byte with the query; the exhaustive path (:707-728) for one within 24 bits."""
    cat, q, kinds = random_code_catalog(40, (0, 50), 240, seed=5)
    mih = psr.catalog_votes(cat, q, True)
    exh = psr.catalog_votes(cat, q, False)
    assert np.array_equal(mih, brute_votes(cat, q, True))
    assert np.array_equal(exh, brute_votes(cat, q, False))
    assert (exh >= mih).all() and exh.sum() > mih.sum()   # the all-bytes-differ plants
    cat.build_index()                                      # a kept index votes the same
    assert np.array_equal(psr.catalog_votes(cat, q, True), mih)
    # each query votes for an entry at most once
    assert exh.max() <= len(q)


def test_vote_known_answers():
    base = np.array([0x0123456789ABCDEF, 0xFEDCBA9876543210], np.uint64)
    from catalog_scenes import flip_bits
    entries = [base.copy(), flip_bits(base, range(0, 128, 5))]   # 26 flips away
    sig = np.array(entries, np.uint64)
    cat = psr.Catalog(["p", "q"], [1, 1], [1, 1], [0, 1, 2], np.zeros(2, psr.FEAT_DTYPE), sig, bits_to_desc(sig))
    q = np.array([flip_bits(base, range(24)),                      # 24 bits: voted by both paths
                  flip_bits(base, range(25)),                      # 25 bits: by neither
                  flip_bits(base, [8 * b for b in range(16)])],    # 16 bits, every byte differs
                 np.uint64)
    assert list(psr.catalog_votes(cat, q, False)) == [2, 0]
    assert list(psr.catalog_votes(cat, q, True)) == [1, 0]


def test_shortlist_order_rules():
    """catalog.rs:594-638: cut to max(k, 1) by (votes desc, id asc) only when longer; the
    inverted-index path drops entries without votes first."""
    ids = ["d", "b", "a", "c", "e"]
    cat = psr.Catalog(ids, [1] * 5, [1] * 5, [0, 1, 2, 3, 4, 5], np.zeros(5, psr.FEAT_DTYPE),
                      np.zeros((5, 2), np.uint64), np.zeros((5, 128), np.float32))
    votes = np.array([3, 5, 3, 0, 5], np.uint32)
    e, v = psr.catalog_shortlist(cat, votes, False, 3)
    assert [ids[i] for i in e] == ["b", "e", "a"] and list(v) == [5, 5, 3]
    e, v = psr.catalog_shortlist(cat, votes, False, 5)
    assert list(e) == [0, 1, 2, 3, 4]                # not longer than k: entry order kept
    e, v = psr.catalog_shortlist(cat, votes, True, 4)
    assert list(e) == [0, 1, 2, 4]                   # zero votes dropped, then 4 <= k
    e, v = psr.catalog_shortlist(cat, votes, True, 0)
    assert [ids[i] for i in e] == ["b"]              # k = 0 behaves as 1


def tracking_case(n, dx, dy, scale_q, scale_t, dtheta, jitter=0.0, seed=0):
    rng = np.random.default_rng(seed)
    qp = rng.uniform(0, 200, (n, 2)).astype(np.float32)
    tp = (qp + np.array([dx, dy], np.float32) + rng.uniform(-jitter, jitter, (n, 2))).astype(np.float32)
    m = psr.make_matches(qp, tp, ratio=0.5)
    qf = np.zeros(n, psr.FEAT_DTYPE)
    tf = np.zeros(n, psr.FEAT_DTYPE)
    qf["scale"], tf["scale"] = scale_q, scale_t
    qf["orientation"] = rng.uniform(0, 6, n)
    tf["orientation"] = qf["orientation"] + np.float32(dtheta)
    return qf, tf, m


def test_tracking_closed_forms():
    """pipeline.rs:2210-2272 on a pure translation with constant scale and rotation."""
    qf, tf, m = tracking_case(41, 7.0, -3.0, 2.0, 3.0, 0.3)
    t = psr.summarize_tracking(qf, tf, m)
    assert t["raw_match_count"] == 41 and t["match_count"] == 41 and t["inlier_ratio"] == 1.0
    assert abs(t["mean_dx"] - 7.0) < 1e-4 and abs(t["mean_dy"] + 3.0) < 1e-4
    assert abs(t["median_dx"] - 7.0) < 1e-4 and abs(t["median_dy"] + 3.0) < 1e-4
    assert t["mean_scale_ratio"] == np.float32(1.5)
    assert abs(t["mean_rotation_deg"] - math.degrees(0.3)) < 1e-3


@pytest.mark.parametrize("n", [6, 7])
def test_tracking_median_rule(n):
    """median_f32 (pipeline.rs:2194-2206): middle element, or the mean of the middle two."""
    qf, tf, m = tracking_case(n, 4.0, 1.0, 1.0, 1.0, 0.0, jitter=1.0, seed=n)
    t = psr.summarize_tracking(qf, tf, m)
    assert t["match_count"] == n                      # jitter within 3 px: every match is an inlier
    for comp, key in (("dx", "median_dx"), ("dy", "median_dy")):
        s = np.sort(m[comp])
        want = s[n // 2] if n % 2 else np.float32(0.5) * (s[n // 2 - 1] + s[n // 2])
        assert t[key] == np.float32(want)
    acc = np.float32(0)
    for v in m["dx"]:                                 # sequential f32 sum in match order
        acc = np.float32(acc + v)
    assert t["mean_dx"] == np.float32(acc * (np.float32(1) / np.float32(n)))


def test_tracking_empty_defaults():
    t = psr.summarize_tracking(np.zeros(0, psr.FEAT_DTYPE), np.zeros(0, psr.FEAT_DTYPE), np.zeros(0, psr.MATCH_DTYPE))
    assert (t["raw_match_count"], t["match_count"], t["inlier_ratio"]) == (0, 0, 0.0)
    assert t["mean_scale_ratio"] == 1.0 and t["mean_rotation_deg"] == 0.0


def hit_ids(cat, hits):
    return [cat.ids[h["entry"]] for h in hits]


def assert_hit_order(cat, hits):
    keys = [(-int(h["match_count"]), -float(h["inlier_ratio"]), -int(h["coarse_votes"]), cat.ids[h["entry"]])
            for h in hits]
    assert keys == sorted(keys)


def test_search_shifted_scene_ranks_first():
    """tests/pipeline_tests.rs:839-877: shortlist_k = 2 forces the Hamming prefilter."""
    cat = scene_catalog([(f"scene_{i}.png", 100 + i) for i in range(3)])
    qf, qd = scene_features(100, 16, 14)
    hits = psr.catalog_search(cat, qf, qd, 3, 2, 0.75, 3.0)
    assert len(hits) >= 1 and hit_ids(cat, hits)[0] == "scene_0.png"
    assert hits[0]["match_count"] > 0
    assert_hit_order(cat, hits)
    # the query window is the entry moved by (16, 14): target = query + (16, 14)
    assert abs(hits[0]["median_dx"] - 16) < 0.5 and abs(hits[0]["median_dy"] - 14) < 0.5
    assert hits[0]["has_transform"] and abs(hits[0]["transform"][2] - 16) < 1.0
    # the same search through the binary format (tests/pipeline_tests.rs:1384-1436)
    hb = psr.catalog_search(through_file(cat), qf, qd, 3, 2, 0.75, 3.0)
    assert hit_ids(cat, hb)[0] == "scene_0.png"


def test_search_inverted_index_path():
    """tests/pipeline_tests.rs:1438-1483: 40 entries (> 32) take the inverted-index path."""
    spec = [(f"item_{i:02d}.png", 200 if i == 17 else (201 if i % 2 == 0 else 202)) for i in range(40)]
    cat = scene_catalog(spec)
    qf, qd = scene_features(200)
    hits = psr.catalog_search(cat, qf, qd, 3, 8, 0.75, 3.0)
    assert hit_ids(cat, hits)[0] == "item_17.png"
    assert_hit_order(cat, hits)


def test_search_exact_entry_ranks_first():
    """tests/coverage_gaps.rs:29-58: the entry holding the query's own features wins."""
    cat = scene_catalog([("alpha", 301), ("beta", 307), ("gamma", 319)])
    qf, qd = scene_features(307)
    hits = psr.catalog_search(cat, qf, qd, 3, 16, 0.75, 3.0)
    assert hit_ids(cat, hits)[0] == "beta"
    assert hits[0]["match_count"] > 0 and hits[0]["inlier_ratio"] > 0.9


def test_search_trivial_inputs():
    cat = scene_catalog([("a", 401)])
    qf, qd = scene_features(401)
    assert len(psr.catalog_search(cat, qf, qd, 0, 16)) == 0
    assert len(psr.catalog_search(cat, qf[:0], qd[:0], 5, 16)) == 0
    empty = psr.Catalog([], [], [], [0], np.zeros(0, psr.FEAT_DTYPE), np.zeros((0, 2), np.uint64),
                        np.zeros((0, 128), np.float32))
    assert len(psr.catalog_search(empty, qf, qd, 5, 16)) == 0


def test_similarity_rotation_uses_the_f32_degree_constant():
    """pipeline.rs:1964 b.atan2(a).to_degrees(): Rust's f32::to_degrees multiplies by the f32
    constant 57.29578 (57.2957795... rounded), not by 180 / (f32 pi) = 57.295776."""
    a, b = np.float32(0.9999016), np.float32(-4.3980202e-05)
    qp = np.array([[0, 0], [100, 0], [0, 100], [100, 100]], np.float32)
    rot = np.array([[a, -b], [b, a]], np.float32)
    tp = (qp @ rot.T).astype(np.float32)
    _, model = psr.filter_matches_ransac(psr.make_matches(qp, tp), 3.0)
    ang = math.atan2(float(np.float32(model["b"])), float(np.float32(model["a"])))
    deg = np.float32(np.float32(ang) * np.float32(57.2957795130823208767981548141051703))
    assert model["rotation_deg"] == deg or abs(model["rotation_deg"] - deg) <= abs(deg) * 2e-7
    assert np.float32(57.2957795130823208767981548141051703) != np.float32(np.float32(180) / np.float32(math.pi))
