"""Learned proposal ordering for homography RANSAC, end to end on one GPU.

Workload: synthetic 1024x768 image pairs (blob images, each partner warped by a random
homography: rotation U[-30, 30] deg, scale U[0.8, 1.25] about the centre, translation
U[-40, 40] px, perspective U[-1e-4, 1e-4]); SIFT extraction and ratio-0.8 matching of every
pair with the library's kernels; the putative matches mixed with 30-80% outliers drawn as
random point pairs (random query and target features, their true descriptor distance and
the distance over the query's second-nearest distance as ratio).

Learning: inlier labels come from a plain RANSAC homography (uniform sampler, 10,000
hypotheses, no local optimisation); the 8-64-64-1 match-scoring MLP is trained on the
labels of the training pairs with the library's forward / backward / Adam kernels; on the
held-out pairs, RANSAC sampling proportionally to the MLP's inlier probability (weighted
sampler, adaptive stop on the sampled inlier mass) is compared with uniform sampling
(adaptive stop on the inlier ratio) and with the reference's 1 / ratio sampler, all at the
same confidence: hypotheses drawn until the stop, final inlier count, inlier recall against
the 10k-hypothesis labels, and reprojection error against the true homography.
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import pure_sift as ps  # noqa: E402
import synth  # noqa: E402

W, H = 1024, 768
MAX_FEAT = ps.MAX_FEATURES_PER_IMAGE


def render_blobs(n, seed, device, width=W, height=H, n_blobs=200):
    """n u8 blob images (background 128, 200 isotropic Gaussian blobs with sigma U[2, 12]
    and signed amplitude U[40, 200], i.i.d. N(0, 3) noise), rendered as separable sums on
    the GPU: image = 128 + Gy^T diag(a) Gx."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    xs = torch.arange(width, device=device, dtype=torch.float32)
    ys = torch.arange(height, device=device, dtype=torch.float32)
    out = torch.empty((n, height, width), dtype=torch.uint8, device=device)
    for i in range(n):
        cx = torch.rand(n_blobs, device=device, generator=g) * width
        cy = torch.rand(n_blobs, device=device, generator=g) * height
        sg = 2.0 + 10.0 * torch.rand(n_blobs, device=device, generator=g)
        amp = (40.0 + 160.0 * torch.rand(n_blobs, device=device, generator=g)) * torch.where(
            torch.rand(n_blobs, device=device, generator=g) < 0.5, -1.0, 1.0)
        gx = torch.exp(-((xs[None, :] - cx[:, None]) ** 2) / (2 * sg[:, None] ** 2))
        gy = torch.exp(-((ys[None, :] - cy[:, None]) ** 2) / (2 * sg[:, None] ** 2))
        img = 128.0 + (gy * amp[:, None]).t().contiguous() @ gx
        img = img + 3.0 * torch.randn((height, width), device=device, generator=g)
        out[i] = torch.clamp(torch.round(img), 0, 255).to(torch.uint8)
    return out


def warp(img, Hm, fill=128.0):
    """Target(x') = bilinear source(H^-1 x'), out-of-bounds = fill (one u8 image on the GPU)."""
    import torch
    import torch.nn.functional as F
    h, w = img.shape
    dev = img.device
    Hi = torch.tensor(np.linalg.inv(Hm), dtype=torch.float32, device=dev)
    yy, xx = torch.meshgrid(torch.arange(h, device=dev, dtype=torch.float32),
                            torch.arange(w, device=dev, dtype=torch.float32), indexing="ij")
    den = Hi[2, 0] * xx + Hi[2, 1] * yy + Hi[2, 2]
    sx = (Hi[0, 0] * xx + Hi[0, 1] * yy + Hi[0, 2]) / den
    sy = (Hi[1, 0] * xx + Hi[1, 1] * yy + Hi[1, 2]) / den
    grid = torch.stack([sx / (w - 1) * 2 - 1, sy / (h - 1) * 2 - 1], -1)[None]
    src = img.to(torch.float32)[None, None] - fill
    out = F.grid_sample(src, grid, mode="bilinear", padding_mode="zeros", align_corners=True)[0, 0] + fill
    return torch.clamp(torch.round(out), 0, 255).to(torch.uint8)


class Dataset:
    """Device-resident pairs: features, putative + outlier matches, pair tables."""


def build_dataset(eng, n_pairs, n_base, seed=0, ratio=0.8, outlier_attrs="true", log=None):
    """All device work runs on the engine's stream (torch ops included), so buffers are
    reused in stream order. outlier_attrs: "true" gives each injected outlier its own
    descriptor distance and distance / second-nearest ratio; "resampled" instead draws its
    distance and ratio from the pair's putative matches, so that only geometric cues
    (scale change, orientation change, position) tell outliers apart."""
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    fsz = ps.FEAT_DTYPE.itemsize
    n_img = n_base + n_pairs
    feat = torch.empty((n_img, MAX_FEAT * fsz), dtype=torch.uint8, device=dev)
    desc = torch.empty((n_img, MAX_FEAT, 128), dtype=torch.float32, device=dev)
    cnt = torch.zeros(n_img, dtype=torch.int32, device=dev)
    cfg = ps.SiftConfig().octaves(4).scales(3).contrast_threshold(0.015)
    base = render_blobs(n_base, seed, dev)
    Hs = [synth.random_homography(W, H, rng) for _ in range(n_pairs)]
    for b in range(n_base):
        eng.extract_u8(base[b].data_ptr(), H, W, W, cfg, feat[b].data_ptr(), desc[b].data_ptr(), MAX_FEAT,
                       cnt.data_ptr() + 4 * b)
    for k in range(n_pairs):
        tgt = warp(base[k % n_base], Hs[k])
        i = n_base + k
        eng.extract_u8(tgt.data_ptr(), H, W, W, cfg, feat[i].data_ptr(), desc[i].data_ptr(), MAX_FEAT,
                       cnt.data_ptr() + 4 * i)
    eng.sync()
    counts = cnt.cpu().numpy()
    if log:
        log(f"extracted {n_img} images, mean {counts.mean():.0f} features")
    # putative matches (ratio test) and nearest / second-nearest of every query feature
    msz = ps.MATCH_DTYPE.itemsize
    mout = torch.empty((n_pairs, MAX_FEAT * msz), dtype=torch.uint8, device=dev)
    mcnt = torch.zeros(n_pairs, dtype=torch.int32, device=dev)
    sec = torch.empty((n_pairs, MAX_FEAT), dtype=torch.float32, device=dev)
    bidx = torch.empty((n_pairs, MAX_FEAT), dtype=torch.int32, device=dev)
    bd = torch.empty((n_pairs, MAX_FEAT), dtype=torch.float32, device=dev)
    opt = ps.MatchOptions(ratio, False)
    for k in range(n_pairs):
        q, t = k % n_base, n_base + k
        eng.match(feat[q].data_ptr(), desc[q].data_ptr(), cnt.data_ptr() + 4 * q, MAX_FEAT, feat[t].data_ptr(),
                  desc[t].data_ptr(), cnt.data_ptr() + 4 * t, MAX_FEAT, opt, mout[k].data_ptr(),
                  mcnt.data_ptr() + 4 * k)
        eng.nn2(desc[q].data_ptr(), cnt.data_ptr() + 4 * q, MAX_FEAT, desc[t].data_ptr(), cnt.data_ptr() + 4 * t,
                MAX_FEAT, ps.PRECISION_FP32, bidx[k].data_ptr(), bd[k].data_ptr(), sec[k].data_ptr())
    eng.sync()
    mc = mcnt.cpu().numpy()
    mall = mout.cpu().numpy()
    feats_h = feat.cpu().numpy().view(ps.FEAT_DTYPE).reshape(n_img, MAX_FEAT)
    # outliers: random point pairs with their true descriptor distance
    parts, pairs, off = [], [], 0
    qsel, tsel, owner = [], [], []
    for k in range(n_pairs):
        q, t = k % n_base, n_base + k
        put = mall[k].view(ps.MATCH_DTYPE)[:mc[k]].copy()
        f = rng.uniform(0.3, 0.8)
        n_out = int(round(len(put) * f / (1.0 - f)))
        qi = rng.integers(0, max(counts[q], 1), n_out)
        ti = rng.integers(0, max(counts[t], 1), n_out)
        out = np.zeros(n_out, ps.MATCH_DTYPE)
        out["query_index"], out["target_index"] = qi, ti
        out["qx"], out["qy"] = feats_h[q]["x"][qi], feats_h[q]["y"][qi]
        out["tx"], out["ty"] = feats_h[t]["x"][ti], feats_h[t]["y"][ti]
        out["dx"], out["dy"] = out["tx"] - out["qx"], out["ty"] - out["qy"]
        allk = np.concatenate([put, out])
        perm = rng.permutation(len(allk))
        allk = allk[perm]
        is_out = np.concatenate([np.zeros(len(put), bool), np.ones(n_out, bool)])[perm]
        parts.append(allk)
        pairs.append((off, len(allk), q * MAX_FEAT, t * MAX_FEAT, float(W), float(H)))
        idx = np.nonzero(is_out)[0]
        qsel.append(q * MAX_FEAT + allk["query_index"][idx].astype(np.int64))
        tsel.append(t * MAX_FEAT + allk["target_index"][idx].astype(np.int64))
        owner.append(off + idx)
        off += len(allk)
    matches = np.concatenate(parts)
    qsel, tsel, owner = np.concatenate(qsel), np.concatenate(tsel), np.concatenate(owner)
    dflat = desc.view(-1, 128)
    sflat = sec.view(-1)
    tq = torch.from_numpy(qsel).to(dev)
    tt = torch.from_numpy(tsel).to(dev)
    dist = (dflat[tq] - dflat[tt]).norm(dim=1)
    # second-nearest of the query feature within its own pair: row k of `sec`
    pair_of = np.searchsorted(np.array([p[0] for p in pairs] + [off]), owner, side="right") - 1
    qloc = matches["query_index"][owner].astype(np.int64)
    s2 = sflat[torch.from_numpy(pair_of * MAX_FEAT + qloc).to(dev)]
    ratio_o = dist / torch.sqrt(torch.clamp(s2, min=1e-12))
    matches["distance"][owner] = dist.cpu().numpy()
    matches["ratio"][owner] = ratio_o.cpu().numpy()
    if outlier_attrs == "resampled":
        is_out = np.zeros(len(matches), bool)
        is_out[owner] = True
        for o, n in ((p[0], p[1]) for p in pairs):
            put = o + np.nonzero(~is_out[o:o + n])[0]
            dst = o + np.nonzero(is_out[o:o + n])[0]
            if len(put) == 0 or len(dst) == 0:
                continue
            src = put[rng.integers(0, len(put), len(dst))]
            matches["distance"][dst] = matches["distance"][src]
            matches["ratio"][dst] = matches["ratio"][src]
    ds = Dataset()
    ds.n_pairs, ds.n_base = n_pairs, n_base
    ds.Hs = np.stack(Hs)
    ds.matches = matches
    ds.pairs = np.array(pairs, ps.MLP_PAIR_DTYPE)
    ds.feat = feat
    ds.d_matches = torch.from_numpy(matches.view(np.uint8)).to(dev)
    ds.d_pairs = torch.from_numpy(ds.pairs.view(np.uint8)).to(dev)
    ds.x = torch.empty((len(matches), 8), dtype=torch.float32, device=dev)
    eng.mlp_features(ds.d_matches.data_ptr(), feat.data_ptr(), feat.data_ptr(), ds.d_pairs.data_ptr(), n_pairs,
                     ds.x.data_ptr())
    eng.sync()
    if log:
        log(f"{n_pairs} pairs, {len(matches)} matches ({len(owner)} injected outliers)")
    return ds


def run_ransac(eng, ds, pair_ids, sampler, weights=None, confidence=0.99, max_iters=10000, min_iters=1,
               local_opt=True, adaptive=True, seed=7):
    """Batched RANSAC over the given pairs; returns (results, inlier index array, problem table)."""
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    probs = np.zeros(len(pair_ids), ps.RANSAC_PROBLEM_DTYPE)
    probs["offset"] = ds.pairs["match_offset"][pair_ids]
    probs["n"] = ds.pairs["n"][pair_ids]
    probs["sim_seed"] = rng.integers(0, 2 ** 63, len(pair_ids), dtype=np.int64).astype(np.uint64)
    probs["hom_seed"] = rng.integers(0, 2 ** 63, len(pair_ids), dtype=np.int64).astype(np.uint64)
    prm = ps.ransac_params(sampler=sampler, reproj_threshold=3.0, confidence=confidence, min_iters=min_iters,
                           max_iters=max_iters, local_opt=local_opt, sim_seed_model=False, adaptive=adaptive)
    tprob = torch.from_numpy(probs.view(np.uint8)).to(dev)
    inl = torch.empty(len(ds.matches), dtype=torch.int32, device=dev)
    res = torch.empty(len(pair_ids) * ps.RANSAC_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    w = 0 if weights is None else weights.data_ptr()
    eng.ransac_batch(ds.d_matches.data_ptr(), w, tprob.data_ptr(), len(pair_ids), len(ds.matches), prm,
                     inl.data_ptr(), res.data_ptr())
    eng.sync()
    return res.cpu().numpy().view(ps.RANSAC_RESULT_DTYPE), inl.cpu().numpy().view(np.uint32), probs


def labels_from_ransac(eng, ds, iters=10000):
    """Inlier labels (1 / 0 per match) of a plain uniform RANSAC with `iters` hypotheses."""
    ids = np.arange(ds.n_pairs)
    res, inl, probs = run_ransac(eng, ds, ids, ps.SAMPLER_UNIFORM, max_iters=iters, local_opt=False, adaptive=False)
    y = np.zeros(len(ds.matches), np.float32)
    for k in ids:
        o, n = probs["offset"][k], res["n_inliers"][k]
        y[o + inl[o:o + n].astype(np.int64)] = 1.0
    return y, res


def pair_rows(ds, pair_ids):
    return np.concatenate([np.arange(ds.pairs["match_offset"][k], ds.pairs["match_offset"][k] + ds.pairs["n"][k])
                           for k in pair_ids])


def train(eng, ds, y, train_ids, steps=600, batch=65536, lr=3e-3, seed=0, log=None, resume=None, checkpoint=None):
    """Adam training of the match-scoring MLP on the GPU kernels. Each step's minibatch is drawn
    from a generator seeded by (seed, step), so a run resumed from a checkpoint (`resume`: the
    dict of pure_sift.load_mlp_checkpoint, with the Adam moments) continues exactly as the
    uninterrupted run would. `checkpoint`: a path that gets the final masters, moments and step.
    Returns (params, loss curve)."""
    import torch
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(pair_rows(ds, train_ids)).to(dev)
    ty = torch.from_numpy(y).to(dev)
    first = 1
    if resume is not None:
        if resume["m"] is None:
            raise ValueError("resuming needs a checkpoint with the Adam moments")
        tp = torch.from_numpy(resume["params"]).to(dev)
        tm, tv = torch.from_numpy(resume["m"]).to(dev), torch.from_numpy(resume["v"]).to(dev)
        first = resume["step"] + 1
    else:
        tp = torch.from_numpy(ps.mlp_init(seed)).to(dev)
        tm, tv = torch.zeros_like(tp), torch.zeros_like(tp)
    tg = torch.zeros_like(tp)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    g = torch.Generator(device=dev)
    xb = torch.empty((batch, 8), dtype=torch.float32, device=dev)
    yb = torch.empty(batch, dtype=torch.float32, device=dev)
    curve = []
    for step in range(first, first + steps):
        g.manual_seed(seed * 1000003 + step)
        sel = rows[torch.randint(0, len(rows), (batch,), device=dev, generator=g)]
        xb.copy_(ds.x[sel])
        yb.copy_(ty[sel])
        eng.mlp_grad(tp.data_ptr(), xb.data_ptr(), yb.data_ptr(), batch, tg.data_ptr(), loss.data_ptr())
        eng.mlp_adam(tp.data_ptr(), tm.data_ptr(), tv.data_ptr(), tg.data_ptr(), step, lr)
        if step % 100 == 0 or step == first + steps - 1:
            eng.sync()
            curve.append((step, float(loss.item())))
            if log:
                log(f"step {step}: loss {curve[-1][1]:.4f}")
    eng.sync()
    params = tp.cpu().numpy()
    if checkpoint is not None:
        ps.save_mlp_checkpoint(checkpoint, params, tm.cpu().numpy(), tv.cpu().numpy(), first + steps - 1)
    return params, curve


def reproj_error(Hest, Htrue):
    """Mean |H_est(p) - H_true(p)| (px) over a 32 x 24 grid of the query image."""
    xs, ys = np.meshgrid(np.linspace(0, W - 1, 32), np.linspace(0, H - 1, 24))
    p = np.stack([xs.ravel(), ys.ravel(), np.ones(xs.size)])

    def proj(M):
        q = M @ p
        return q[:2] / q[2]

    return float(np.mean(np.linalg.norm(proj(Hest.astype(np.float64)) - proj(Htrue), axis=0)))


def evaluate(eng, ds, y, params, test_ids, confidence=0.99):
    import torch
    dev = torch.device("cuda:0")
    tp = torch.from_numpy(params).to(dev)
    logit = torch.empty(len(ds.matches), dtype=torch.float32, device=dev)
    prob = torch.empty(len(ds.matches), dtype=torch.float32, device=dev)
    eng.mlp_forward(tp.data_ptr(), ds.x.data_ptr(), len(ds.matches), logit.data_ptr(), prob.data_ptr())
    eng.sync()
    out = {}
    for name, sampler, w in (("uniform", ps.SAMPLER_UNIFORM, None), ("guided", ps.SAMPLER_WEIGHTS, prob),
                             ("reference_sampler", ps.SAMPLER_REFERENCE, None)):
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        res, inl, probs = run_ransac(eng, ds, test_ids, sampler, weights=w, confidence=confidence)
        t1.record()
        t1.synchronize()
        hyps = res["hyps"].astype(np.float64)
        recall, err = [], []
        for j, k in enumerate(test_ids):
            o, n = probs["offset"][j], res["n_inliers"][j]
            lab = np.nonzero(y[o:o + probs["n"][j]] > 0.5)[0]
            got = inl[o:o + n].astype(np.int64)
            recall.append(len(np.intersect1d(lab, got)) / max(len(lab), 1))
            if res["has_model"][j]:
                err.append(reproj_error(res["h"][j].reshape(3, 3), ds.Hs[k]))
        out[name] = {"hyps_mean": float(hyps.mean()), "hyps_median": float(np.median(hyps)),
                     "inliers_mean": float(res["n_inliers"].mean()), "inlier_recall_mean": float(np.mean(recall)),
                     "reproj_err_px_median": float(np.median(err)) if err else None,
                     "has_model_frac": float(res["has_model"].mean()), "ms_batch": t0.elapsed_time(t1),
                     "n_inliers": res["n_inliers"].copy()}
    u, g = out["uniform"], out["guided"]
    ni_u, ni_g = u.pop("n_inliers"), g.pop("n_inliers")
    out["reference_sampler"].pop("n_inliers")
    out["hyps_ratio_uniform_over_guided"] = u["hyps_mean"] / max(g["hyps_mean"], 1e-9)
    out["equal_inlier_count_frac"] = float(np.mean(np.abs(ni_u.astype(np.int64) - ni_g.astype(np.int64)) <=
                                                   np.maximum(1, 0.01 * ni_u)))
    pr = prob.cpu().numpy()
    test_rows = pair_rows(ds, test_ids)
    out["mlp_test_accuracy"] = float(np.mean((pr[test_rows] > 0.5) == (y[test_rows] > 0.5)))
    return out


def split_by_base(n_pairs, n_base, train_frac):
    """Train / test pair ids split by base image (pair k uses base k % n_base): the first
    round(n_base * train_frac) bases train, the others test."""
    ids = np.arange(n_pairs)
    n_train_base = int(round(n_base * train_frac))
    if not 0 < n_train_base < n_base:
        raise ValueError("the split needs at least one training and one test base image")
    base = ids % n_base
    return ids[base < n_train_base], ids[base >= n_train_base]


def run(n_pairs=4096, n_base=256, train_frac=0.75, steps=600, seed=0, outlier_attrs="true", log=None):
    import torch
    eng = ps.Engine(0)
    try:
        with torch.cuda.stream(torch.cuda.ExternalStream(eng.stream())):
            return _run(eng, n_pairs, n_base, train_frac, steps, seed, outlier_attrs, log)
    finally:
        eng.close()


def _run(eng, n_pairs, n_base, train_frac, steps, seed, outlier_attrs, log):
    import time
    t = time.perf_counter()
    ds = build_dataset(eng, n_pairs, n_base, seed=seed, outlier_attrs=outlier_attrs, log=log)
    t_data = time.perf_counter() - t
    t = time.perf_counter()
    y, lab_res = labels_from_ransac(eng, ds)
    t_lab = time.perf_counter() - t
    if log:
        log(f"labels: {y.mean():.3f} inlier fraction, {t_lab:.1f} s")
    # split by base image: pair k warps base image k % n_base, so the held-out pairs' query
    # images (and their features) never appear in a training pair
    train_ids, test_ids = split_by_base(n_pairs, n_base, train_frac)
    t = time.perf_counter()
    params, curve = train(eng, ds, y, train_ids, steps=steps, seed=seed, log=log)
    t_train = time.perf_counter() - t
    ev = evaluate(eng, ds, y, params, test_ids)
    ev.update({"pairs": n_pairs, "train_pairs": int(len(train_ids)), "test_pairs": int(len(test_ids)),
               "split": f"by base image: bases [0, {int(round(n_base * train_frac))}) train, the rest test",
               "matches": int(len(ds.matches)), "label_inlier_frac": float(y.mean()), "loss_curve": curve,
               "seconds": {"data": t_data, "labels": t_lab, "train": t_train}})
    return ev, params
