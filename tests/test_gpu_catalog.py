"""GPU parity for the 128-bit code path (catalog.rs:37-45, :707-728), the exact top-k
Hamming shortlist and the FP32 rerank, through the C ABI, against the oracle.

Bar: bit-exact (codes, vote counts, shortlist indices and distances, rerank order and
squared distances). The shortlist cases cover the fast bounded scan, the exact
fallback (short candidate lists, overflowing lists, ties at the k-th distance), k > n,
n at and next to the sample size, and the full 1M-code catalog.
"""
import numpy as np
import pytest

import psr
import synth

pytestmark = pytest.mark.gpu


def _flip(code, bits):
    c = code.copy()
    for b in bits:
        c[b // 64] ^= np.uint64(1) << np.uint64(b % 64)
    return c


def _check_topk(gpu, cat, qs, k):
    idx, dist = gpu.hamming_topk(cat, qs, k)
    ridx, rdist = psr.hamming_topk(cat, qs, k)
    assert np.array_equal(dist, rdist), "shortlist distances differ"
    assert np.array_equal(idx, ridx), "shortlist indices differ"
    return idx, dist


def test_quantize_exact(gpu):
    rng = np.random.default_rng(3)
    d = rng.normal(0, 0.08, (1000, 128)).astype(np.float32)
    d[0, :] = np.float32(0.065)                      # not strictly greater
    d[1, :] = np.nextafter(np.float32(0.065), np.float32(1))
    d[2, ::3] = np.nan
    d[3, :] = -1.0
    assert np.array_equal(gpu.quantize_descriptor_u128(d), psr.quantize_u128(d))


@pytest.mark.parametrize("nq,ne", [(1, 1), (300, 777), (4096, 4096), (37, 20000)])
def test_coarse_votes_exact(gpu, nq, ne):
    cat, _, _, _ = synth.hamming_catalog(ne, 1, seed=nq + ne)
    rng = np.random.default_rng(nq)
    src = rng.integers(0, ne, nq)
    qs = np.stack([_flip(cat[s], rng.choice(128, rng.integers(0, 40), replace=False)) for s in src])
    assert gpu.count_coarse_hamming_votes(qs, cat) == psr.count_coarse_hamming_votes(qs, cat)


def test_coarse_votes_empty(gpu):
    cat, qs, _, _ = synth.hamming_catalog(10, 5, seed=1)
    assert gpu.count_coarse_hamming_votes(qs, cat[:0]) == 0
    assert gpu.count_coarse_hamming_votes(qs[:0], cat) == 0


@pytest.mark.parametrize("n,nq,k", [(200000, 300, 100), (16384, 50, 100), (16385, 50, 100),
                                    (300000, 20, 1024), (5000, 64, 1), (100, 10, 100), (50, 7, 100), (1, 3, 1)])
def test_topk_random_catalog(gpu, n, nq, k):
    cat, qs, src, flips = synth.hamming_catalog(n, nq, seed=n + k)
    idx, dist = _check_topk(gpu, cat, qs, k)
    kk = min(k, n)
    assert np.all(idx[:, kk:] == 0xFFFFFFFF) and np.all(dist[:, kk:] == 255)


def test_topk_duplicates_overflow(gpu):
    """All codes identical: every code ties at the k-th distance (candidate list overflows)."""
    cat = np.tile(np.array([[0x0123456789ABCDEF, 0xFEDCBA9876543210]], np.uint64), (6000, 1))
    qs = np.stack([_flip(cat[0], range(b)) for b in (0, 3, 64, 128)])
    _check_topk(gpu, cat, qs, 100)


def test_topk_few_distinct_codes(gpu):
    """16 distinct codes repeated: ties at the k-th distance resolved by index."""
    base, qs, _, _ = synth.hamming_catalog(16, 30, seed=9)
    rng = np.random.default_rng(9)
    cat = base[rng.integers(0, 16, 70000)]
    _check_topk(gpu, cat, qs, 100)
    _check_topk(gpu, cat, qs, 1024)


def test_topk_short_candidate_list(gpu):
    """The sampled prefix (first 16384 codes) holds 70 near codes and otherwise the complement
    of the query, the rest of the catalog is random: the sampled bound admits fewer than k
    codes, so the exact fallback runs (70 near codes + the 30 nearest random ones)."""
    rng = np.random.default_rng(11)
    q = synth.hamming_catalog(1, 1, seed=11)[0][0]
    n = 400000
    cat = synth.hamming_catalog(n, 1, seed=12)[0]
    cat[:16384] = ~q
    for j in rng.choice(16384, 70, replace=False):
        cat[j] = _flip(q, rng.choice(128, rng.integers(0, 6), replace=False))
    _check_topk(gpu, cat, np.stack([q, q, _flip(q, [5])]), 100)


def test_topk_full_catalog_1m(gpu):
    """1M random codes, planted queries with 0..20 flipped bits (200 queries vs the oracle;
    all 1000 through size-independent properties)."""
    cat, qs, src, flips = synth.hamming_catalog(1_000_000, 1000, seed=2024)
    idx, dist = gpu.hamming_topk(cat, qs, 100)
    ridx, rdist = psr.hamming_topk(cat, qs[:200], 100)
    assert np.array_equal(idx[:200], ridx) and np.array_equal(dist[:200], rdist)
    # distances recomputed, sorted by (distance, index), planted source present
    x = cat[idx.astype(np.int64)] ^ qs[:, None, :]
    d = np.bitwise_count(x).sum(-1)
    assert np.array_equal(d, dist)
    key = dist.astype(np.int64) * (1 << 32) + idx.astype(np.int64)
    assert np.all(np.diff(key, axis=1) > 0)
    for i in range(len(qs)):
        if dist[i, -1] > flips[i]:
            assert src[i] in idx[i]


def test_rerank_exact(gpu):
    rng = np.random.default_rng(5)
    qd, cd, _ = synth.unit_descriptors(64, 20000, seed=5)
    sl = rng.integers(0, 20000, (64, 100)).astype(np.uint32)
    sl[0, :] = 7                                   # duplicated index: ties by index
    sl[1, 50:] = 0xFFFFFFFF                        # invalid slots
    sl[2, 10] = 20000                              # out of range
    idx, d2 = gpu.rerank_l2(qd, cd, sl)
    ridx, rd2 = psr.rerank_l2(qd, cd, sl)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(d2.view(np.uint32), rd2.view(np.uint32))


def test_shortlist_then_rerank_engine(gpu):
    """Device-resident quantize -> top-k -> rerank on one engine stream equals the host calls."""
    import torch
    qd, cd, _ = synth.unit_descriptors(100, 50000, seed=8)
    codes = gpu.quantize_descriptor_u128(cd)
    qcodes = gpu.quantize_descriptor_u128(qd)
    idx, _ = gpu.hamming_topk(codes, qcodes, 100)
    ridx, rd2 = gpu.rerank_l2(qd, cd, idx)
    dev = torch.device("cuda:0")
    t_cd = torch.from_numpy(cd).to(dev)
    t_qd = torch.from_numpy(qd).to(dev)
    t_codes = torch.empty((len(cd), 2), dtype=torch.int64, device=dev)
    t_q = torch.empty((len(qd), 2), dtype=torch.int64, device=dev)
    t_idx = torch.empty((len(qd), 100), dtype=torch.int32, device=dev)
    t_dist = torch.empty((len(qd), 100), dtype=torch.uint8, device=dev)
    t_ridx = torch.empty((len(qd), 100), dtype=torch.int32, device=dev)
    t_rd2 = torch.empty((len(qd), 100), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    eng = gpu.Engine(0)
    eng.quantize_u128(t_cd.data_ptr(), len(cd), t_codes.data_ptr())
    eng.quantize_u128(t_qd.data_ptr(), len(qd), t_q.data_ptr())
    eng.hamming_topk(t_codes.data_ptr(), len(cd), t_q.data_ptr(), len(qd), 100, t_idx.data_ptr(), t_dist.data_ptr())
    eng.rerank_l2(t_qd.data_ptr(), len(qd), t_cd.data_ptr(), len(cd), t_idx.data_ptr(), 100, t_ridx.data_ptr(),
                  t_rd2.data_ptr())
    eng.sync()
    eng.close()
    assert np.array_equal(t_codes.cpu().numpy().view(np.uint64), codes)
    assert np.array_equal(t_idx.cpu().numpy().view(np.uint32), idx)
    assert np.array_equal(t_ridx.cpu().numpy().view(np.uint32), ridx)
    assert np.array_equal(t_rd2.cpu().numpy(), rd2)
