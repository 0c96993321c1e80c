"""GPU parity for the per-pair match post-processing of run_match_from_bytes (cli.rs:290-318)
through the C ABI, against the oracle's restatement (psr.pair_geometry): the tracking summary,
similarity, homography, fundamental matrix, GRIC and the kept inliers, for one pair (host
form) and for many pairs in one device-resident batch.

Bar: bit-exact counts, inlier lists, models and GRIC scores, mean / median displacement and
scale ratio; the mean rotation (device sinf / cosf sums) within 1e-3 degrees and the
similarity's rotation_deg (device atan2f) within 1e-4 degrees.
"""
import numpy as np
import pytest

import psr
import synth
from geometry_scenes import rotation, two_view

pytestmark = pytest.mark.gpu


def real_pair(seed):
    a, b, _ = synth.blob_pair(640, 480, 200, seed=seed)
    qf, qd = psr.extract_features(psr.luma_to_f32(a), 4, 3, 0.015, True)
    tf, td = psr.extract_features(psr.luma_to_f32(b), 4, 3, 0.015, True)
    return qf, qd, tf, td, psr.match(qf, qd, tf, td, 0.8)


def scene_pair(i):
    n_in, n_out = [(80, 26), (40, 40), (9, 3), (8, 0), (7, 0), (5, 0), (1, 0), (0, 0)][i]
    m, _ = two_view(n_in, n_out, rotation(12, -6, 4), [0.6, -0.15, 0.2], seed=20 + i, noise=0.3)
    rng = np.random.default_rng(i)
    qf = np.zeros(len(m), psr.FEAT_DTYPE)
    tf = np.zeros(len(m), psr.FEAT_DTYPE)
    qf["scale"], tf["scale"] = rng.uniform(1, 5, len(m)), rng.uniform(1, 5, len(m))
    qf["orientation"], tf["orientation"] = rng.uniform(0, 6.2, len(m)), rng.uniform(0, 6.2, len(m))
    return qf, tf, m


def check(rec, best, want, n):
    tr = want["tracking"]
    assert rec["raw_match_count"] == n
    assert rec["track_match_count"] == tr["match_count"]
    for k in ("inlier_ratio", "mean_dx", "mean_dy", "median_dx", "median_dy", "mean_scale_ratio"):
        assert rec[k] == tr[k] or (rec[k] == 0 == tr[k]), (k, rec[k], tr[k])
    assert abs(float(rec["mean_rotation_deg"]) - float(tr["mean_rotation_deg"])) <= 1e-3
    sim = want["similarity"]
    assert bool(rec["has_transform"]) == (sim is not None)
    assert rec["n_similarity"] == len(want["sim_inliers"])
    if sim is not None:
        for j, k in enumerate(("a", "b", "tx", "ty", "scale")):
            assert rec["transform"][j] == np.float32(sim[k]), k
        assert abs(float(rec["transform"][5]) - sim["rotation_deg"]) <= 1e-4
    H = want["homography"]
    assert bool(rec["has_homography"]) == (H is not None)
    assert rec["n_homography"] == len(want["hom_inliers"])
    if H is not None:
        assert np.array_equal(rec["homography"].reshape(3, 3), H)
    F = want["fundamental"]
    assert bool(rec["has_fundamental"]) == (F is not None)
    if F is not None:
        assert np.array_equal(rec["fundamental"].reshape(3, 3), F)
        assert rec["n_fundamental"] == want["n_fundamental"]
    g = want["gric"]
    assert bool(rec["has_gric"]) == (g is not None)
    if g is not None:
        assert (rec["gric_selected"], rec["homography_gric"], rec["fundamental_gric"]) == g
    assert bool(rec["best_is_homography"]) == want["best_is_homography"]
    assert np.array_equal(best, want["best"])


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("thr,rseed", [(3.0, None), (3.0, 777), (1.0, None), (5.0, 99)])
def test_real_pair_equals_oracle(gpu, seed, thr, rseed):
    qf, qd, tf, td, m = real_pair(seed)
    print(f"pair seed {seed}: {len(qf)} / {len(tf)} features, {len(m)} matches", flush=True)
    rec, best = gpu.pair_geometry(gpu.Features.from_arrays(qf, qd), gpu.Features.from_arrays(tf, td), m,
                                  gpu.RansacOptions(thr, rseed))
    check(rec, best, psr.pair_geometry(qf, tf, m, thr, rseed), len(m))


@pytest.mark.parametrize("i", range(8))
def test_scene_pairs_equal_oracle(gpu, i):
    qf, tf, m = scene_pair(i)
    rec, best = gpu.pair_geometry(gpu.Features(qf, np.zeros((len(qf), 128), np.float32)),
                                  gpu.Features(tf, np.zeros((len(tf), 128), np.float32)), m, gpu.RansacOptions(3.0))
    check(rec, best, psr.pair_geometry(qf, tf, m, 3.0), len(m))


def test_batch_equals_single_pairs(gpu):
    import torch
    pairs = [scene_pair(i) for i in range(8)] + [real_pair(3)[0::2] + (real_pair(3)[4],)]
    pairs = [(p[0], p[1], p[2]) for p in pairs]
    offs, qo, to = np.cumsum([0] + [len(p[2]) for p in pairs]), np.cumsum([0] + [len(p[0]) for p in pairs]), \
        np.cumsum([0] + [len(p[1]) for p in pairs])
    probs = np.zeros(len(pairs), gpu.PAIR_PROBLEM_DTYPE)
    probs["offset"], probs["n"] = offs[:-1], [len(p[2]) for p in pairs]
    probs["qfeat_offset"], probs["tfeat_offset"] = qo[:-1], to[:-1]
    dev = torch.device("cuda:0")
    tens = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to(dev)
    d_m = tens(np.concatenate([p[2] for p in pairs]))
    d_q = tens(np.concatenate([p[0] for p in pairs]))
    d_t = tens(np.concatenate([p[1] for p in pairs]))
    d_p = tens(probs)
    d_best = torch.zeros(max(int(offs[-1]), 1), dtype=torch.int32, device=dev)
    d_out = torch.zeros(len(pairs) * gpu.PAIR_GEOMETRY_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    e = gpu.Engine(0)
    torch.cuda.synchronize()
    e.pair_geometry_batch(d_m.data_ptr(), d_q.data_ptr(), d_t.data_ptr(), d_p.data_ptr(), len(pairs), int(offs[-1]),
                          gpu.RansacOptions(3.0), d_best.data_ptr(), d_out.data_ptr())
    e.sync()
    recs = d_out.cpu().numpy().view(gpu.PAIR_GEOMETRY_DTYPE)
    best = d_best.cpu().numpy().view(np.uint32)
    for k, (qf, tf, m) in enumerate(pairs):
        want = psr.pair_geometry(qf, tf, m, 3.0)
        check(recs[k], best[offs[k]:offs[k] + recs[k]["n_best"]], want, len(m))
