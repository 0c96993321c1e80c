"""GPU tests of the precision matcher (BF16 / FP8-E4M3 MFMA distances, top-8 rescored in
exact f32) through the C ABI.

Bars:
  * the returned best / second squared distances are the exact f32 distances of the
    returned targets (bit-exact with the oracle's sequential sum);
  * the rescored candidates are a subset of the targets, so best_d2 >= the exact best
    and second_d2 >= the exact second always, with equality whenever the exact best and
    second are among the 8 (every query with a planted match; at least 99% (BF16) /
    95% (FP8) of all queries on these dense random sets, whose second-nearest distances
    are often closer together than the format's rounding); the ratio-0.8 decision at the
    full 50k x 200k size disagrees with the FP32 path on at most 0.1% (BF16) / 0.5% (FP8);
  * the low-precision squared distance of the best target is within the rounding bound
    of the format versus float64: BF16 (8-bit significand) |err| <= 2 * 2^-8 * |q||t|,
    FP8 E4M3 (4-bit significand) |err| <= 2 * (2^-4 + 2^-8) * |q||t| (unit rows).
"""
import numpy as np
import pytest

import psr
import synth

pytestmark = pytest.mark.gpu

BOUND = {1: 2 * 2.0 ** -8, 2: 2 * (2.0 ** -4 + 2.0 ** -8)}


def _f64(q, t):
    return ((q.astype(np.float64) - t.astype(np.float64)) ** 2).sum(-1)


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("nq,nt", [(1, 1), (3, 5), (100, 7), (300, 1000), (2000, 8000), (257, 65), (513, 4097)])
def test_precision_nn2_matches_fp32(gpu, prec, nq, nt):
    q, db, planted = synth.unit_descriptors(nq, nt, seed=nq * 7 + nt)
    bi, b, s = psr.nn2(q, db)
    gi, gb, gs, ga = gpu.nn2_precision(q, db, prec, with_approx=True)
    assert np.all(gb >= b) and np.all(gs >= s)
    same_best = (gi == bi) & (gb.view(np.uint32) == b.view(np.uint32))
    same_second = gs.view(np.uint32) == s.view(np.uint32)
    assert np.all(same_best[planted >= 0]) and np.all(same_second[planted >= 0])
    assert same_best.mean() >= (0.999 if prec == 1 else 0.99)
    assert same_second.mean() >= (0.99 if prec == 1 else 0.95)
    err = np.abs(ga.astype(np.float64) - _f64(q, db[gi]))
    assert err.max() <= BOUND[prec], err.max()


@pytest.mark.parametrize("prec", [1, 2])
def test_precision_distances_are_exact_f32(gpu, prec):
    """Whatever targets are returned, their distances are the exact sequential f32 sums."""
    rng = np.random.default_rng(4)
    q = rng.standard_normal((500, 128)).astype(np.float32)
    db = rng.standard_normal((3000, 128)).astype(np.float32)   # not unit norm, wider range
    gi, gb, gs = gpu.nn2_precision(q, db, prec)
    for i in range(0, 500, 25):
        _, d, _ = psr.nn2(q[i:i + 1], db[gi[i]:gi[i] + 1])
        assert gb[i].view(np.uint32) == d[0].view(np.uint32)
    assert np.all(gs >= gb)


@pytest.mark.parametrize("prec", [1, 2])
def test_precision_match_and_mutual(gpu, prec):
    """ps_match with the precision option accepts the same matches as the FP32 ps_match
    (at ratio 0.8 the accepted queries are planted ones, whose best and second are
    always among the rescored 8)."""
    q, db, _ = synth.unit_descriptors(1500, 3000, seed=21)
    rng = np.random.default_rng(21)
    qh = np.zeros(1500, gpu.FEAT_DTYPE)
    th = np.zeros(3000, gpu.FEAT_DTYPE)
    for h in (qh, th):
        h["x"] = rng.uniform(0, 640, len(h))
        h["y"] = rng.uniform(0, 480, len(h))
    qf = gpu.Features(qh, q)
    tf = gpu.Features(th, db)
    for mutual in (False, True):
        ref = gpu.Matcher(gpu.MatchOptions(0.8, mutual)).match_(qf, tf)
        got = gpu.Matcher(gpu.MatchOptions(0.8, mutual, prec)).match_(qf, tf)
        assert np.array_equal(ref, got)


def test_precision_full_size_decisions(gpu):
    """50k x 200k: ratio decisions versus the exact FP32 search on the GPU."""
    import torch
    nq, nt = 50_000, 200_000
    q, db, planted = synth.unit_descriptors(nq, nt, seed=1)
    dev = torch.device("cuda:0")
    tq = torch.from_numpy(q).to(dev)
    tt = torch.from_numpy(db).to(dev)
    cnt = torch.tensor([nq, nt], dtype=torch.int32, device=dev)
    eng = gpu.Engine(0)
    out = {}
    for prec in (0, 1, 2):
        bi = torch.empty(nq, dtype=torch.int32, device=dev)
        b = torch.empty(nq, dtype=torch.float32, device=dev)
        s = torch.empty(nq, dtype=torch.float32, device=dev)
        eng.nn2(tq.data_ptr(), cnt.data_ptr(), nq, tt.data_ptr(), cnt.data_ptr() + 4, nt, prec, bi.data_ptr(),
                b.data_ptr(), s.data_ptr())
        eng.sync()
        out[prec] = (bi.cpu().numpy(), b.cpu().numpy(), s.cpu().numpy())
    eng.close()
    # the FP32 path against the oracle on a sample of queries
    sel = np.random.default_rng(3).choice(nq, 300, replace=False)
    ri, rb, rs = psr.nn2(q[sel], db)
    assert np.array_equal(out[0][0][sel].view(np.uint32), ri)
    assert np.array_equal(out[0][1][sel], rb) and np.array_equal(out[0][2][sel], rs)

    def decide(r):
        return (r[2] > 1e-12) & (np.sqrt(r[1]) / np.sqrt(r[2]) < 0.8)

    ref = decide(out[0])
    for prec, bound in ((1, 1e-3), (2, 5e-3)):
        acc = decide(out[prec])
        dis = np.mean((acc != ref) | (acc & (out[prec][0] != out[0][0])))
        assert dis <= bound, (prec, dis)
        assert np.mean(out[prec][0][planted >= 0] == planted[planted >= 0]) >= 0.999
