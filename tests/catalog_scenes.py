"""Seeded catalog workloads shared by the catalog tests: small blob scenes (the shape of the
reference's catalog tests, tests/pipeline_tests.rs:839-877 and :1438-1483) with features from
the oracle's extraction, and random-code catalogs with planted query codes."""
from __future__ import annotations

import numpy as np

import psr
import synth

SCENE = 220


def scene_u8(seed: int, dx: int = 0, dy: int = 0, size: int = SCENE) -> np.ndarray:
    """A size x size window at (dx, dy) of a seeded blob image (content moves by (-dx, -dy))."""
    base = synth.blob_image(size + 48, size + 48, 70, seed)
    return np.ascontiguousarray(base[dy:dy + size, dx:dx + size])


def scene_features(seed: int, dx: int = 0, dy: int = 0, octaves: int = 3):
    img = psr.luma_to_f32(scene_u8(seed, dx, dy))
    return psr.extract_features(img, octaves, 3, 0.015, True)


def scene_catalog(ids_seeds, octaves: int = 3) -> psr.Catalog:
    cache = {}
    entries = []
    for ident, seed in ids_seeds:
        if seed not in cache:
            cache[seed] = scene_features(seed, octaves=octaves)
        f, d = cache[seed]
        entries.append((ident, SCENE, SCENE, f, d))
    return psr.Catalog.from_entries(entries, octaves=octaves)


def through_file(cat: psr.Catalog) -> psr.Catalog:
    """The catalog as a PSIFTCAT file reads back (u8 descriptors, stored signatures)."""
    return psr.catalog_parse(psr.catalog_write(cat))


def bits_to_desc(codes: np.ndarray) -> np.ndarray:
    """Descriptors whose 128-bit quantisation (bit i = d[i] > 0.065) is exactly `codes`."""
    codes = np.asarray(codes, np.uint64).reshape(-1, 2)
    bits = np.unpackbits(codes.view(np.uint8).reshape(-1, 16), axis=1, bitorder="little")
    return (bits.astype(np.float32) * 0.125).reshape(-1, 128)


def flip_bits(code: np.ndarray, positions) -> np.ndarray:
    c = np.asarray(code, np.uint64).reshape(2).copy()
    for p in positions:
        c[p // 64] ^= np.uint64(1) << np.uint64(p % 64)
    return c


def random_code_catalog(n_entries: int, per_entry, n_queries: int, seed: int = 0):
    """Entries of random 128-bit codes (Bernoulli(0.5) bits) and query codes planted near
    catalog codes: 0-24 flips at random positions, exactly 25 flips, and 16-24 flips placed so
    that every byte differs (within 24 bits yet sharing no byte with their source)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = (rng.integers(per_entry[0], per_entry[1] + 1, n_entries) if isinstance(per_entry, tuple)
              else np.full(n_entries, per_entry))
    n = int(counts.sum())
    sig = rng.integers(0, 2**63, (n, 2), dtype=np.int64).astype(np.uint64) ^ (
        rng.integers(0, 2, (n, 2)).astype(np.uint64) << np.uint64(63))
    off = np.zeros(n_entries + 1, np.uint64)
    off[1:] = np.cumsum(counts)
    q = []
    kinds = []
    for i in range(n_queries):
        src = sig[rng.integers(0, n)] if n else np.zeros(2, np.uint64)
        mode = i % 4
        if mode == 0:
            k = int(rng.integers(0, 25))
            pos = rng.choice(128, k, replace=False)
        elif mode == 1:
            pos = rng.choice(128, 25, replace=False)
        elif mode == 2:
            extra = int(rng.integers(0, 9))
            pos = [8 * b + int(rng.integers(0, 8)) for b in range(16)]
            rest = [p for p in range(128) if p not in pos]
            pos = pos + list(rng.choice(rest, extra, replace=False))
        else:
            pos = rng.choice(128, 64, replace=False)
        q.append(flip_bits(src, pos))
        kinds.append(mode)
    q = np.array(q, np.uint64).reshape(-1, 2)
    feat = np.zeros(n, psr.FEAT_DTYPE)
    feat["x"] = rng.uniform(0, 100, n)
    feat["y"] = rng.uniform(0, 100, n)
    feat["scale"] = rng.uniform(1, 4, n)
    desc = bits_to_desc(sig)
    ids = [f"e{i:05d}" for i in range(n_entries)]
    cat = psr.Catalog(ids, np.full(n_entries, 64), np.full(n_entries, 48), off, feat, sig, desc)
    return cat, q, np.array(kinds)


def brute_votes(cat: psr.Catalog, q_codes: np.ndarray, mih: bool) -> np.ndarray:
    """Direct predicate form of catalog.rs:595-633: an entry's vote counts the queries with a
    signature of the entry within 24 bits that (on the inverted-index path) also equals the
    query in one of its 16 bytes."""
    qb = np.ascontiguousarray(q_codes, np.uint64).reshape(-1, 2).view(np.uint8).reshape(-1, 16)
    sb = cat.sig.view(np.uint8).reshape(-1, 16)
    votes = np.zeros(len(cat.ids), np.uint32)
    for e in range(len(cat.ids)):
        a, b = int(cat.off[e]), int(cat.off[e + 1])
        if a == b:
            continue
        x = qb[:, None, :] ^ sb[None, a:b, :]
        d = np.unpackbits(x, axis=2).sum(axis=2)
        ok = d <= 24
        if mih:
            ok &= (x == 0).any(axis=2)
        votes[e] = ok.any(axis=1).sum()
    return votes
