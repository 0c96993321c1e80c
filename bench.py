#!/usr/bin/env python3
"""1080p SIFT extract + match throughput on one MI355X.

One step ("frame") = extract one 1920x1080 u8 image (4 octaves x 3 scales, first
octave upsampled, contrast 0.015, capped at 4096 features) and ratio-0.8 match it
against the pre-extracted warped partner. Inputs are resident in HBM before the
timed region; the whole step runs on one engine stream through the C ABI.

Prints one JSON line. `roofline` covers the scale-space stage (pyramid + extrema): algorithmic
bytes per image W*H + 4*P*(5 + 4 + 4) = 574,905,600 B at 1080p -- the u8 input, the 5 Gaussian
planes written per octave, the 4 source planes the blur levels read and the 4 Gaussian planes
the extrema pass reads (P = pixels summed over octaves) -- over the stage time measured with
HIP events on the engine's stream; beside it the stage's algorithmic VALU floor (the same
arithmetic at the 32-bit VALU issue rate, no FMA), which binds slightly above HBM. `cpu_baseline`
is the CPU oracle
(the C restatement of the reference's Rayon path, parallelised at the same sites)
timed on a bounded sample on this host.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "pure-sift_amd"))

HBM_PEAK_GBS = 8000.0
# 32-bit integer VALU issue rate: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_INT_PEAK_TOPS = 39.32
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "pyramid_extrema_1080p_hbm_traffic.md")


def alg_bytes(h, w, octaves=4, upsample=True):
    """Algorithmic HBM bytes of the scale-space stage (S = 3) per image: the u8 input, then per
    octave of P pixels 5 Gaussian planes written (levels 0..S+1; level S+2 is never stored on
    the extraction path), the 4 source planes the blur levels 1..S+1 read, and the 4 planes
    (levels 1..S+1) the extrema pass reads: W*H + 4*P*(5 + 4 + 4)."""
    hh, ww = (2 * h, 2 * w) if upsample else (h, w)
    P = 0
    for _ in range(octaves):
        if hh < 3 or ww < 3:
            break
        P += hh * ww
        hh //= 2
        ww //= 2
    return h * w + 4 * P * (5 + 4 + 4)


ALG_BYTES_1080P = alg_bytes(1080, 1920)


def alg_valu_ops(h, w, octaves=4, upsample=True):
    """Algorithmic 32-bit VALU operations of the scale-space stage (S = 3) per image, counting the
    reference's own arithmetic with no FMA (each tap one multiply and one add): the u8 -> f32
    conversion (2 per input pixel); per octave-0 pixel the 2x bilinear upsampling (4) and the
    first blur (radius 4: 2 passes x 9 taps x 2); per pixel of every octave the blur levels
    1..S+1 (radii 4, 5, 6, 8: 4 per tap) and the extremum scan (3 DoG differences, then per
    searchable slice the 8-neighbour comparison and the threshold, ~10). Level S+2, the later
    octaves' 2x decimation and the refinement of the few candidates are left out."""
    hh, ww = (2 * h, 2 * w) if upsample else (h, w)
    ops, first = 2 * h * w, True
    for _ in range(octaves):
        if hh < 3 or ww < 3:
            break
        px = hh * ww
        if first:
            ops += px * ((4 if upsample else 0) + 4 * 9)
            first = False
        ops += px * (4 * (9 + 11 + 13 + 17) + 3 + 3 * 10)
        hh //= 2
        ww //= 2
    return ops


ALG_VALU_OPS_1080P = alg_valu_ops(1080, 1920)


def traffic_bytes_1080p():
    """HBM bytes per 1080p frame moved by the scale-space kernels (blur + extrema), as measured
    by rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by a 256 MiB copy; read from the
    committed profile summary (its `traffic_bytes_per_frame:` line), None when absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_frame:"):
                    return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def _host_cores() -> int:
    """Threads for the CPU baseline: every core this process may run on, within the host CPU
    share granted per GPU (OMP_NUM_THREADS where the launcher sets it; 16 on the MI355X boxes
    used here, one GPU each)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        cores = min(cores, int(share))
    return max(1, cores)


def _cpu_info() -> dict:
    """nproc, the CPU model and the affinity of this process, for the baseline's record."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = None
    return {"nproc": os.cpu_count(), "affinity": aff, "cpu_model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def _median_runs(fn, runs: int, warmup: int):
    """Median wall time (s) of `runs` calls of fn after `warmup` untimed calls."""
    import statistics
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), ts


def _stream_events(eng):
    """(start, end) HIP events on the engine's own stream."""
    import torch
    st = torch.cuda.ExternalStream(eng.stream())
    return st, torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def bench_hamming(steps: int = 20, cpu_queries: int = 1000) -> dict:
    """Top-100 Hamming shortlist of 1000 planted queries over 1M random 128-bit codes,
    then the FP32 128-D rerank of each shortlist (catalog descriptors resident in HBM)."""
    import numpy as np
    import torch
    import pure_sift as ps
    import synth
    n, nq, k = 1_000_000, 1000, 100
    cat, qs, src, flips = synth.hamming_catalog(n, nq, seed=2024)
    dev = torch.device("cuda:0")
    tc = torch.from_numpy(cat.view(np.int64)).to(dev)
    tq = torch.from_numpy(qs.view(np.int64)).to(dev)
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    cd = torch.relu(torch.randn((n, 128), device=dev, generator=g))
    cd /= cd.norm(dim=1, keepdim=True).clamp_min(1e-12)
    qd = torch.relu(torch.randn((nq, 128), device=dev, generator=g))
    qd /= qd.norm(dim=1, keepdim=True).clamp_min(1e-12)
    ti = torch.empty((nq, k), dtype=torch.int32, device=dev)
    tdist = torch.empty((nq, k), dtype=torch.uint8, device=dev)
    ri = torch.empty((nq, k), dtype=torch.int32, device=dev)
    rd = torch.empty((nq, k), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    eng = ps.Engine(0)
    st, e0, e1 = _stream_events(eng)

    def topk():
        eng.hamming_topk(tc.data_ptr(), n, tq.data_ptr(), nq, k, ti.data_ptr(), tdist.data_ptr())

    def rerank():
        eng.rerank_l2(qd.data_ptr(), nq, cd.data_ptr(), n, ti.data_ptr(), k, ri.data_ptr(), rd.data_ptr())

    for _ in range(3):
        topk()
        rerank()
    eng.sync()
    e0.record(st)
    for _ in range(steps):
        topk()
    e1.record(st)
    eng.sync()
    t_topk = e0.elapsed_time(e1) / steps
    e0.record(st)
    for _ in range(steps):
        rerank()
    e1.record(st)
    eng.sync()
    t_rr = e0.elapsed_time(e1) / steps
    idx = ti.cpu().numpy().view(np.uint32)
    planted_found = float(np.mean([src[i] in idx[i] for i in range(nq)]))
    eng.close()
    ops = 8.0 * n * nq      # 4 XOR + 4 POPCNT (32-bit) per (query, code)
    achieved = ops / (t_topk * 1e-3) / 1e12
    out = {"workload": "top-100 Hamming shortlist, 1000 planted queries (0-20 flipped bits) over 1M random "
                       "128-bit codes, then FP32 128-D L2 rerank of each shortlist",
           "value": nq / ((t_topk + t_rr) * 1e-3), "unit": "queries/s",
           "ms_topk": t_topk, "ms_rerank": t_rr, "planted_in_shortlist": planted_found,
           "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_INT_PEAK_TOPS, "unit": "Tops/s",
                        "frac": achieved / VALU_INT_PEAK_TOPS,
                        "hbm_frac_per_pass": (16.0 * n / (t_topk * 1e-3)) / (HBM_PEAK_GBS * 1e9)}}
    if cpu_queries > 0:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import psr
        cores = _host_cores()
        psr.set_threads(cores)
        dt, ts = _median_runs(lambda: psr.hamming_topk(cat, qs[:cpu_queries], k), runs=5, warmup=1)
        out["cpu_baseline"] = {"value": cpu_queries / dt, "unit": "queries/s (top-k only)", "cores": cores,
                               "kind": "port", "sample": f"all {cpu_queries} queries over the full 1M catalog, median "
                                                         f"of 5 runs after 1 warm-up ({sum(ts):.1f} s timed)"}
    return out


def bench_precision(steps: int = 5, cpu_queries: int = 200) -> dict:
    """Precision form of matching: 50k query x 200k database unit descriptors (ReLU(N(0,1)),
    L2-normalised, 20% of queries planted), ratio 0.8; BF16 and FP8-E4M3 MFMA distances with
    the 8 best rescored in FP32, against the exact FP32 search on the same GPU."""
    import numpy as np
    import torch
    import pure_sift as ps
    import synth
    nq, nt = 50_000, 200_000
    q, db, planted = synth.unit_descriptors(nq, nt, seed=1)
    dev = torch.device("cuda:0")
    tq = torch.from_numpy(q).to(dev)
    tt = torch.from_numpy(db).to(dev)
    cnt = torch.tensor([nq, nt], dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    eng = ps.Engine(0)
    st, e0, e1 = _stream_events(eng)
    res, ms = {}, {}
    for prec, name in ((ps.PRECISION_FP32, "fp32"), (ps.PRECISION_BF16, "bf16"), (ps.PRECISION_FP8, "fp8")):
        bi = torch.empty(nq, dtype=torch.int32, device=dev)
        b = torch.empty(nq, dtype=torch.float32, device=dev)
        s2 = torch.empty(nq, dtype=torch.float32, device=dev)
        ap = torch.empty(nq, dtype=torch.float32, device=dev)

        def run():
            eng.nn2(tq.data_ptr(), cnt.data_ptr(), nq, tt.data_ptr(), cnt.data_ptr() + 4, nt, prec, bi.data_ptr(),
                    b.data_ptr(), s2.data_ptr(), ap.data_ptr() if prec else 0)

        run()
        eng.sync()
        k = max(1, steps if prec else 2)
        e0.record(st)
        for _ in range(k):
            run()
        e1.record(st)
        eng.sync()
        ms[name] = e0.elapsed_time(e1) / k
        res[name] = (bi.cpu().numpy().view(np.uint32), b.cpu().numpy(), s2.cpu().numpy(),
                     ap.cpu().numpy() if prec else None)
    eng.close()

    def decide(r):
        return (r[2] > 1e-12) & (np.sqrt(r[1]) / np.sqrt(r[2]) < 0.8)

    ref_acc = decide(res["fp32"])
    flop = 2.0 * nq * nt * 128
    out = {"workload": "precision matcher: 50k query x 200k database 128-D unit descriptors (20% planted), "
                       "ratio 0.8, top-8 rescored in FP32",
           "ms_fp32_exact": ms["fp32"], "unit": "solves/s"}
    sel = np.random.default_rng(0).choice(nq, 2000, replace=False)
    for name, peak in (("bf16", 2500.0), ("fp8", 5000.0)):
        r = res[name]
        acc = decide(r)
        dis = float(np.mean((acc != ref_acc) | (acc & (r[0] != res["fp32"][0]))))
        d64 = ((q[sel].astype(np.float64) - db[r[0][sel]].astype(np.float64)) ** 2).sum(-1)
        err = np.abs(r[3][sel].astype(np.float64) - d64)
        tf = flop / (ms[name] * 1e-3) / 1e12
        out[name] = {"ms": ms[name], "value": 1e3 / ms[name], "ratio_decision_disagreement_vs_fp32": dis,
                     "best_index_differs_vs_fp32": float(np.mean(r[0] != res["fp32"][0])),
                     "planted_recovered": float(np.mean(r[0][planted >= 0] == planted[planted >= 0])),
                     "approx_d2_abs_err_vs_fp64": {"max": float(err.max()), "mean": float(err.mean()),
                                                   "sample": len(sel)},
                     "roofline": {"bound": "mfma", "achieved": tf, "peak": peak, "unit": "TFLOP/s",
                                  "frac": tf / peak}}
    out["value"] = out["bf16"]["value"]
    if cpu_queries > 0:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import psr
        cores = _host_cores()
        psr.set_threads(cores)
        dt, ts = _median_runs(lambda: psr.nn2(q[:cpu_queries], db), runs=3, warmup=1)
        out["cpu_baseline"] = {"value": cpu_queries / nq / dt, "unit": "solves/s (exact f32)", "cores": cores,
                               "kind": "port", "sample": f"{cpu_queries} of the 50k queries against all 200k "
                                                         f"targets, median of 3 runs after 1 warm-up "
                                                         f"({sum(ts):.1f} s timed), scaled to a full solve"}
    return out


def bench_ransac(n_pairs: int = 4096, n_base: int = 256, steps: int = 600, cpu_pairs: int = 64) -> dict:
    """Learned-proposal RANSAC: synthetic 1024x768 warped pairs, GPU extraction + matching,
    30-80% injected outliers, labels from a plain 10k-hypothesis RANSAC, MLP trained on the
    GPU, then hypotheses to 99% confidence: MLP-guided vs uniform (and the reference's
    1 / ratio sampler) on the held-out pairs; also with the outliers' distance and ratio
    resampled from the putative matches (geometric cues only)."""
    import numpy as np
    import learned_ransac as lr
    t0 = time.perf_counter()
    log = lambda m: print(f"[ransac {time.perf_counter() - t0:6.1f}s] {m}", file=sys.stderr, flush=True)
    ev, params = lr.run(n_pairs=n_pairs, n_base=n_base, steps=steps, log=log)
    ev_hard, _ = lr.run(n_pairs=max(n_pairs // 4, 64), n_base=max(n_base // 4, 16), steps=steps,
                        outlier_attrs="resampled", seed=1, log=log)
    wall = time.perf_counter() - t0
    out = {"workload": f"{n_pairs} synthetic 1024x768 pairs (random homographies), GPU SIFT + ratio-0.8 matches "
                       f"+ 30-80% random point-pair outliers; labels from 10k-hypothesis uniform RANSAC; "
                       f"MLP 8-64-64-1 trained on 75% of the pairs; confidence 0.99 on the rest",
           "unit": "hypotheses/solve", "higher_is_better": False,
           "hyps_guided": ev["guided"]["hyps_mean"], "hyps_uniform": ev["uniform"]["hyps_mean"],
           "hyps_reference_sampler": ev["reference_sampler"]["hyps_mean"],
           "uniform_over_guided": ev["hyps_ratio_uniform_over_guided"], "detail": ev,
           "geometric_cues_only": {k: ev_hard[k] for k in ("guided", "uniform", "reference_sampler",
                                                            "hyps_ratio_uniform_over_guided",
                                                            "equal_inlier_count_frac", "mlp_test_accuracy")},
           "wall_s": wall}
    out["value"] = out["hyps_guided"]
    return out


def bench_batch(n_images: int = 256, reps: int = 5, cpu_images: int = 4) -> dict:
    """Batched extraction of 256 seeded blob/noise u8 images with widths U[320, 2048] and
    heights U[240, 1536]: grouped persistent launches (one per pyramid level / stage for all
    images) captured in one HIP graph, launch shapes picked by the batch autotuner; compared
    with the per-image loop and the eager batch."""
    import numpy as np
    import torch
    import pure_sift as ps
    import synth
    import learned_ransac as lr
    dev = torch.device("cuda:0")
    sizes = synth.batch_sizes(n_images, seed=3)          # (w, h)
    descs, off = [], 0
    for (w, h) in sizes:
        descs.append((off, h, w, w, 0))
        off += w * h
    pix = torch.empty(off, dtype=torch.uint8, device=dev)
    for i, (w, h) in enumerate(sizes):
        nb = max(20, int(200 * w * h / (640 * 480)))
        pix[descs[i][0]:descs[i][0] + w * h] = lr.render_blobs(1, 1000 + i, dev, width=w, height=h, n_blobs=nb)[0].reshape(-1)
    torch.cuda.synchronize()
    im = np.array(descs, ps.IMAGE_DESC_DTYPE)
    cap = ps.MAX_FEATURES_PER_IMAGE
    fsz = ps.FEAT_DTYPE.itemsize
    feat = torch.empty((n_images, cap * fsz), dtype=torch.uint8, device=dev)
    desc = torch.empty((n_images, cap, 128), dtype=torch.float32, device=dev)
    cnt = torch.zeros(n_images, dtype=torch.int32, device=dev)
    cfg = ps.SiftConfig().octaves(4).scales(3).contrast_threshold(0.015)
    eng = ps.Engine(0)
    st, e0, e1 = _stream_events(eng)

    def timed(fn, k):
        fn()
        eng.sync()
        e0.record(st)
        for _ in range(k):
            fn()
        e1.record(st)
        eng.sync()
        return e0.elapsed_time(e1) / k

    def loop():
        for i, (o, h, w, s, _) in enumerate(descs):
            eng.extract_u8(pix.data_ptr() + o, h, w, s, cfg, feat[i].data_ptr(), desc[i].data_ptr(), cap,
                           cnt.data_ptr() + 4 * i)

    def batch(flags):
        return lambda: eng.extract_batch_u8(pix.data_ptr(), im, cfg, feat.data_ptr(), desc.data_ptr(), cap,
                                            cnt.data_ptr(), flags)

    ps.lib().ps_autotune_reset()
    ms_loop = timed(loop, reps)
    ms_eager = timed(batch(ps.BATCH_NO_GRAPH), reps)
    ms_graph = timed(batch(0), reps)
    n_tuned = eng.autotune_batch(pix.data_ptr(), im, cfg)
    ms_tuned = timed(batch(0), reps)
    counts = cnt.cpu().numpy()
    eng.check()
    eng.close()
    alg = sum(alg_bytes(h, w) for (w, h) in sizes)
    best = min(ms_graph, ms_tuned)
    achieved = alg / (best * 1e-3) / 1e9
    out = {"workload": f"{n_images} blob/noise u8 images, widths U[320,2048], heights U[240,1536] "
                       f"({off / 1e6:.1f} MP), 4 octaves x 3 scales, upsampled, cap 4096",
           "value": n_images / (best * 1e-3), "unit": "images/s",
           "ms_per_batch": {"per_image_loop": ms_loop, "batch_eager": ms_eager, "batch_graph": ms_graph,
                            "batch_graph_autotuned": ms_tuned},
           "autotuned_launch_classes": n_tuned, "features_mean": float(counts.mean()),
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "alg_bytes_per_batch": alg,
                        "note": "whole extraction (pyramid + extrema + describe) timed; bytes are the "
                                "pyramid+extrema unit"}}
    if cpu_images > 0:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import psr
        cores = _host_cores()
        psr.set_threads(cores)
        sel = list(range(cpu_images))
        host = pix.cpu().numpy()
        t0 = time.perf_counter()
        mp = 0.0
        for i in sel:
            o, h, w, s, _ = descs[i]
            psr.extract_features(psr.luma_to_f32(host[o:o + h * w].reshape(h, w)), 4, 3, 0.015, True)
            mp += h * w / 1e6
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": len(sel) / dt, "unit": "images/s", "cores": cores, "kind": "port",
                               "sample": f"the first {len(sel)} images ({mp:.1f} MP), {dt:.1f} s wall"}
    return out


def bench_catalog(n_entries: int = 1000, steps: int = 20, cpu_searches: int = 1) -> dict:
    """GPU-resident catalog search (catalog.rs:576-702): a PSIFTCAT catalog of 1000 entries,
    each the features of a seeded 640x480 blob image (4 octaves x 3 scales, upsampled, cap
    4096), searched with the features of a homography-warped view of entry 517; top_k 5,
    shortlist 16, ratio 0.75, RANSAC 3 px (the reference defaults). Query features are
    device-resident before the timed region; each search returns its hits to the host."""
    import numpy as np
    import torch
    import pure_sift as ps
    import learned_ransac as lr
    import synth
    dev = torch.device("cuda:0")
    W, H = 640, 480
    cfg = ps.SiftConfig().octaves(4).scales(3).contrast_threshold(0.015)
    cap = ps.MAX_FEATURES_PER_IMAGE
    fsz = ps.FEAT_DTYPE.itemsize
    eng = ps.Engine(0)
    chunk = 200
    feat = torch.empty((chunk, cap * fsz), dtype=torch.uint8, device=dev)
    desc = torch.empty((chunk, cap, 128), dtype=torch.float32, device=dev)
    cnt = torch.zeros(chunk, dtype=torch.int32, device=dev)
    entries = []
    for c0 in range(0, n_entries, chunk):
        imgs = lr.render_blobs(chunk, 9000 + c0, dev, W, H, 200)
        im = np.array([(i * W * H, H, W, W, 0) for i in range(chunk)], ps.IMAGE_DESC_DTYPE)
        eng.extract_batch_u8(imgs.data_ptr(), im, cfg, feat.data_ptr(), desc.data_ptr(), cap, cnt.data_ptr())
        eng.sync()
        counts = cnt.cpu().numpy()
        fh, dh = feat.cpu().numpy(), desc.cpu().numpy()
        for i in range(chunk):
            n = int(counts[i])
            entries.append((f"img_{c0 + i:04d}.png", W, H,
                            ps.Features(fh[i, :n * fsz].view(ps.FEAT_DTYPE).copy(), dh[i, :n].copy())))
        if c0 <= 517 < c0 + chunk:
            target = imgs[517 - c0].clone()
    data = ps.FeatureCatalog.binary_bytes(entries, 4, 3, 0.015)
    t0 = time.perf_counter()
    cat = ps.FeatureCatalog.from_binary_bytes(data)
    load_s = time.perf_counter() - t0
    n_feat = cat.n_features
    rng = np.random.default_rng(5)
    Hm = synth.random_homography(W, H, rng)
    query = lr.warp(target, Hm)
    qf = torch.empty(cap * fsz, dtype=torch.uint8, device=dev)
    qd = torch.empty((cap, 128), dtype=torch.float32, device=dev)
    qc = torch.zeros(1, dtype=torch.int32, device=dev)
    eng.extract_u8(query.data_ptr(), H, W, W, cfg, qf.data_ptr(), qd.data_ptr(), cap, qc.data_ptr())
    eng.sync()
    nq = int(qc.item())

    def search():
        return cat.search_raw_device(eng, qf.data_ptr(), qd.data_ptr(), nq)

    for _ in range(2):
        hits = search()
    t0 = time.perf_counter()
    for _ in range(steps):
        hits = search()
    ms = (time.perf_counter() - t0) / steps * 1e3
    qdh = qd[:nq].cpu().numpy()
    cat.coarse_votes(qdh)
    t0 = time.perf_counter()
    for _ in range(steps):
        cat.coarse_votes(qdh)
    vote_ms = (time.perf_counter() - t0) / steps * 1e3
    top = cat.summary(int(hits[0]["entry"])).id if len(hits) else None
    pairs = nq * n_feat
    ops = 10 * pairs   # per (query, signature): 4 xor, 4 popcount, add3, compare (the byte test only on near codes)
    out = {"workload": f"{n_entries}-entry PSIFTCAT catalog ({n_feat} features, 640x480 blob images), "
                       f"query {nq} features of a warped view of entry 517; top 5, shortlist 16, ratio 0.75, 3 px",
           "value": 1e3 / ms, "unit": "searches/s", "ms_per_search": ms, "load_s": load_s,
           "vote_pass_ms_host_timed": vote_ms, "top_hit": top, "top_hit_correct": top == "img_0517.png",
           "top_hit_match_count": int(hits[0]["match_count"]) if len(hits) else 0,
           "roofline": {"bound": "valu", "kernel": "catalog_vote_kernel (inverted-index predicate over every "
                        "catalog signature)", "achieved": ops / (vote_ms * 1e-3) / 1e12,
                        "peak": VALU_INT_PEAK_TOPS, "unit": "TOP/s", "frac": ops / (vote_ms * 1e-3) / 1e12 / VALU_INT_PEAK_TOPS,
                        "pairs": pairs, "note": "host-timed vote call (includes the query upload, launch "
                        "and readback); 10 integer VALU ops per (query, signature) pair"}}
    if cpu_searches > 0:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import psr
        cores = _host_cores()
        psr.set_threads(cores)
        ref = psr.catalog_parse(data)
        ref.build_index()   # built once at load, as the reference's rebuild_indices (catalog.rs:284-297)
        qfh = qf[:nq * fsz].cpu().numpy().view(psr.FEAT_DTYPE).copy()
        box = {}

        def search():
            box["rh"] = psr.catalog_search(ref, qfh, qdh, 5, 16, 0.75, 3.0)
        dt, ts = _median_runs(search, runs=cpu_searches, warmup=2)
        rh = box["rh"]
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "searches/s", "cores": cores, "kind": "port",
                               "sample": f"median of {cpu_searches} searches of the same query and catalog after 2 "
                                         f"warm-ups ({sum(ts):.1f} s timed)"}
        out["hits_equal_oracle"] = bool(len(rh) == len(hits) and all(
            int(a["entry"]) == int(b["entry"]) and int(a["match_count"]) == int(b["match_count"])
            for a, b in zip(rh, hits)))
    eng.close()
    return out


def bench_pair(steps: int = 20, n_batch: int = 64, cpu_pairs: int = 2) -> dict:
    """The per-pair match path of run_match_from_bytes (cli.rs:270-318) on 640x480 blob pairs
    (200 blobs, warped partner): extract both images (4 octaves x 3 scales, upsampled), ratio-0.8
    match, then the post-processing -- tracking summary, similarity, homography, fundamental
    matrix, GRIC, kept inliers -- all on one engine stream with the images resident in HBM.
    Reported for one pair at a time (latency) and for 64 pairs per batch (throughput: one batched
    extraction of the 128 images, one grouped match of the 64 pairs, one post-processing batch;
    the same 64 pairs as a loop of per-pair calls beside it, with a check that both give the
    same records)."""
    import numpy as np
    import torch
    import pure_sift as ps
    import synth
    dev = torch.device("cuda:0")
    W, H = 640, 480
    cap = ps.MAX_FEATURES_PER_IMAGE
    fsz, msz = ps.FEAT_DTYPE.itemsize, ps.MATCH_DTYPE.itemsize
    cfg = ps.SiftConfig().octaves(4).scales(3).contrast_threshold(0.015)
    opt = ps.MatchOptions(ratio=0.8, mutual=False)
    ro = ps.RansacOptions(3.0)
    pairs = [synth.blob_pair(W, H, 200, seed=100 + k)[:2] for k in range(n_batch)]
    imgs = torch.from_numpy(np.stack([np.stack([a, b]) for a, b in pairs])).to(dev)   # [k][2][H][W]
    qf = torch.empty((n_batch, cap * fsz), dtype=torch.uint8, device=dev)
    tf = torch.empty((n_batch, cap * fsz), dtype=torch.uint8, device=dev)
    qd = torch.empty((n_batch, cap, 128), dtype=torch.float32, device=dev)
    td = torch.empty((n_batch, cap, 128), dtype=torch.float32, device=dev)
    cnt = torch.zeros((n_batch, 4), dtype=torch.int32, device=dev)
    mout = torch.empty((n_batch, cap * msz), dtype=torch.uint8, device=dev)
    probs = np.zeros(n_batch, ps.PAIR_PROBLEM_DTYPE)
    probs["offset"] = np.arange(n_batch) * cap
    probs["qfeat_offset"] = np.arange(n_batch) * cap
    probs["tfeat_offset"] = np.arange(n_batch) * cap
    d_probs = torch.from_numpy(probs.view(np.uint8).copy()).to(dev)
    d_best = torch.empty(n_batch * cap, dtype=torch.int32, device=dev)
    d_out = torch.empty(n_batch * ps.PAIR_GEOMETRY_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    eng = ps.Engine(0)
    c = cnt.data_ptr()
    st = torch.cuda.ExternalStream(eng.stream())

    def run(n):
        """the first n pairs; pair k's matches and features sit in slab k of the buffers"""
        for k in range(n):
            base = imgs[k].data_ptr()
            ck = c + 16 * k
            eng.extract_u8(base, H, W, W, cfg, qf[k].data_ptr(), qd[k].data_ptr(), cap, ck)
            eng.extract_u8(base + W * H, H, W, W, cfg, tf[k].data_ptr(), td[k].data_ptr(), cap, ck + 4)
            eng.match(qf[k].data_ptr(), qd[k].data_ptr(), ck, cap, tf[k].data_ptr(), td[k].data_ptr(), ck + 4, cap, opt,
                      mout[k].data_ptr(), ck + 8)
            with torch.cuda.stream(st):   # the raw match count into the pair's record, stream-ordered
                d_probs.view(-1, 16)[k, 4:8].copy_(cnt[k, 2:3].view(torch.uint8), non_blocking=True)
        eng.pair_geometry_batch(mout.data_ptr(), qf.data_ptr(), tf.data_ptr(), d_probs.data_ptr(), n, n * cap, ro,
                                d_best.data_ptr(), d_out.data_ptr())

    # the batch path: all 2n images in one batched extraction (captured graph), one grouped match
    # of every pair, one post-processing batch; features of pair k at rows 2k (query) and 2k+1
    # (target) of one array
    descs = np.array([((k * 2 + side) * H * W, H, W, W, 0) for k in range(n_batch) for side in (0, 1)],
                     ps.IMAGE_DESC_DTYPE)
    bf = torch.empty((2 * n_batch, cap * fsz), dtype=torch.uint8, device=dev)
    bd = torch.empty((2 * n_batch, cap, 128), dtype=torch.float32, device=dev)
    bcnt = torch.zeros(2 * n_batch, dtype=torch.int32, device=dev)
    sets = torch.zeros((2 * n_batch, 2), dtype=torch.int32, device=dev)
    sets[:, 0] = torch.arange(2 * n_batch, device=dev, dtype=torch.int32) * cap
    mprob = np.array([(2 * k, 2 * k + 1, k * cap, 0) for k in range(n_batch)], ps.MATCH_PROBLEM_DTYPE)
    d_mprob = torch.from_numpy(mprob.view(np.uint8).copy()).to(dev)
    bprobs = np.zeros(n_batch, ps.PAIR_PROBLEM_DTYPE)
    bprobs["offset"] = np.arange(n_batch) * cap
    bprobs["qfeat_offset"] = np.arange(n_batch) * 2 * cap
    bprobs["tfeat_offset"] = (np.arange(n_batch) * 2 + 1) * cap
    d_bprobs = torch.from_numpy(bprobs.view(np.int32).copy()).to(dev).view(-1, 4)
    mcnt = torch.zeros(n_batch, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    def run_batch():
        eng.extract_batch_u8(imgs.data_ptr(), descs, cfg, bf.data_ptr(), bd.data_ptr(), cap, bcnt.data_ptr())
        with torch.cuda.stream(st):   # set sizes = the extracted counts, stream-ordered
            sets[:, 1].copy_(bcnt, non_blocking=True)
        eng.match_sets(bf.data_ptr(), bd.data_ptr(), 2 * n_batch * cap, sets.data_ptr(), 2 * n_batch, bf.data_ptr(),
                       bd.data_ptr(), 2 * n_batch * cap, sets.data_ptr(), 2 * n_batch, cap, d_mprob.data_ptr(),
                       n_batch, opt.ratio, mout.data_ptr(), mcnt.data_ptr())
        with torch.cuda.stream(st):
            d_bprobs[:, 1].copy_(mcnt, non_blocking=True)
        eng.pair_geometry_batch(mout.data_ptr(), bf.data_ptr(), bf.data_ptr(), d_bprobs.data_ptr(), n_batch,
                                n_batch * cap, ro, d_best.data_ptr(), d_out.data_ptr())

    def timed(fn, reps):
        fn()
        eng.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        eng.sync()
        return (time.perf_counter() - t0) / reps * 1e3
    ms_single = timed(lambda: run(1), steps)
    ms_loop = timed(lambda: run(n_batch), max(2, steps // 10))
    recs_loop = d_out.cpu().numpy().view(ps.PAIR_GEOMETRY_DTYPE).copy()
    ms_batch = timed(run_batch, max(2, steps // 10))
    # the batch path gives every pair the record of the per-pair path
    same_as_loop = bool(np.array_equal(d_out.cpu().numpy().view(ps.PAIR_GEOMETRY_DTYPE), recs_loop))
    recs = d_out.cpu().numpy().view(ps.PAIR_GEOMETRY_DTYPE)
    eng.check()
    out = {"workload": f"640x480 blob pairs (200 blobs, warped partner): extract both + ratio-0.8 match + tracking, "
                       f"similarity, homography, fundamental, GRIC (RANSAC 3 px)",
           "value": n_batch / (ms_batch * 1e-3), "unit": "pairs/s", "ms_per_pair_single": ms_single,
           "ms_per_batch": ms_batch, "batch": n_batch, "ms_per_batch_pair_loop": ms_loop,
           "batch_equals_pair_loop": same_as_loop,
           "raw_matches_mean": float(recs["raw_match_count"].mean()),
           "homography_inliers_mean": float(recs["n_homography"].mean()),
           "fundamental_frac": float(recs["has_fundamental"].mean()),
           "gric_selects_homography_frac": float((recs["gric_selected"] == 0).mean())}
    if cpu_pairs > 0:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import psr
        cores = _host_cores()
        psr.set_threads(cores)
        t0 = time.perf_counter()
        for k in range(cpu_pairs):
            a, b = pairs[k]
            qh, qdh = psr.extract_features(psr.luma_to_f32(a), 4, 3, 0.015, True)
            th, tdh = psr.extract_features(psr.luma_to_f32(b), 4, 3, 0.015, True)
            mm = psr.match(qh, qdh, th, tdh, 0.8)
            psr.pair_geometry(qh, th, mm, 3.0)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_pairs / dt, "unit": "pairs/s", "cores": cores, "kind": "port",
                               "sample": f"the first {cpu_pairs} pairs, {dt:.2f} s wall"}
    eng.close()
    return out


def bench_panorama(n_tiles: int = 16, steps: int = 10, cpu_sample: bool = True) -> dict:
    """The panorama's pairwise stage (stitch.rs:478-571) on 16 overlapping 652x480 tiles of one
    synthetic strip (1700 blobs + band-limited noise, step 319 px, exposures 0.85-1.18):
    extract every tile (3 octaves x 3 scales, SIFT descriptors, cap 4096), ratio-0.75 match all
    120 pairs, similarity + homography RANSAC (3 px), safe alignment, inverse, patch intensities
    -- one engine call with the tiles resident in HBM. The grouped matcher is also timed alone
    (HIP events on the engine stream) on the same features, with its split-BF16 MFMA work
    (3 products of 2 * nq * nt * 128 flops per pair) against the dense BF16 peak."""
    import numpy as np
    import torch
    import pure_sift as ps
    import synth
    dev = torch.device("cuda:0")
    width = 5440
    rng = np.random.default_rng(5)
    exposure = rng.uniform(0.85, 1.18, n_tiles)
    tiles, step = synth.panorama_tiles(width, 480, n_tiles, 1700, seed=9, tile_frac=0.12, exposure=exposure)
    h, w = tiles[0].shape
    n = len(tiles)
    npairs = n * (n - 1) // 2
    pix = torch.from_numpy(np.concatenate([t.ravel() for t in tiles])).to(dev)
    descs = np.array([(k * h * w, h, w, w, 0) for k in range(n)], ps.IMAGE_DESC_DTYPE)
    opt = ps.PanoramaOptions()
    d_pairs = torch.zeros(npairs * ps.PANORAMA_PAIR_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    eng = ps.Engine(0)
    run = lambda: eng.panorama_pairs(pix.data_ptr(), descs, opt, d_pairs.data_ptr())
    run()
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    eng.sync()
    ms = (time.perf_counter() - t0) / steps * 1e3
    recs = d_pairs.cpu().numpy().view(ps.PANORAMA_PAIR_DTYPE)
    # the grouped matcher alone, on the same features (batched extraction into slabs of 4096)
    cap = ps.MAX_FEATURES_PER_IMAGE
    fsz = ps.FEAT_DTYPE.itemsize
    feat = torch.empty(n * cap * fsz, dtype=torch.uint8, device=dev)
    desc = torch.empty((n * cap, 128), dtype=torch.float32, device=dev)
    cnt = torch.zeros(n, dtype=torch.int32, device=dev)
    cfg = ps.SiftConfig().octaves(3).scales(3).contrast_threshold(0.015).descriptor_kind(ps.DescriptorKind.Sift)
    eng.extract_batch_u8(pix.data_ptr(), descs, cfg, feat.data_ptr(), desc.data_ptr(), cap, cnt.data_ptr())
    eng.sync()
    counts = cnt.cpu().numpy().astype(np.int64)
    sets = np.array([(k * cap, counts[k]) for k in range(n)], ps.FEATURE_SET_DTYPE)
    probs = np.array([(i, j, k * cap, 0) for k, (i, j) in enumerate((i, j) for i in range(n) for j in range(i + 1, n))],
                     ps.MATCH_PROBLEM_DTYPE)
    d_sets = torch.from_numpy(sets.view(np.uint8).copy()).to(dev)
    d_probs = torch.from_numpy(probs.view(np.uint8).copy()).to(dev)
    mout = torch.empty(npairs * cap * ps.MATCH_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    mcnt = torch.zeros(npairs, dtype=torch.int32, device=dev)
    st = torch.cuda.ExternalStream(eng.stream())
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms_match = []
    for r in range(steps + 1):
        with torch.cuda.stream(st):
            ev0.record(st)
        eng.match_sets(feat.data_ptr(), desc.data_ptr(), n * cap, d_sets.data_ptr(), n, feat.data_ptr(),
                       desc.data_ptr(), n * cap, d_sets.data_ptr(), n, cap, d_probs.data_ptr(), npairs, opt.ratio,
                       mout.data_ptr(), mcnt.data_ptr())
        with torch.cuda.stream(st):
            ev1.record(st)
        eng.sync()
        if r:
            ms_match.append(ev0.elapsed_time(ev1))
    ms_m = float(np.median(ms_match))
    pq = sum(int(counts[i]) * int(counts[j]) for i, j in zip(probs["query_set"], probs["target_set"]))
    flops = 3 * 2.0 * pq * 128
    assert np.array_equal(mcnt.cpu().numpy(), recs["raw_matches"].astype(np.int32))
    eng.check()
    eng.close()
    out = {"workload": f"{n} overlapping {w}x{h} tiles of a synthetic strip (step {step} px): extract all, "
                       f"ratio-0.75 match all {npairs} pairs, similarity + homography RANSAC 3 px, pair edges",
           "value": npairs / (ms * 1e-3), "unit": "image pairs/s", "ms_per_stage": ms, "images": n, "pairs": npairs,
           "features_mean": float(counts.mean()), "raw_matches_mean": float(recs["raw_matches"].mean()),
           "edges": int(recs["has_edge"].sum()),
           "grouped_match": {"ms": ms_m, "pairs_nq_nt": pq, "tflops": flops / (ms_m * 1e-3) / 1e12,
                             "peak_tflops": 2500.0, "frac": flops / (ms_m * 1e-3) / 2.5e15,
                             "note": "chain of prep, MFMA distances, rescore, compaction for all pairs; "
                                     "3 BF16 products (hi*hi, hi*lo, lo*hi) of 2*nq*nt*128 flops per pair"}}
    if cpu_sample:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import psr
        cores = _host_cores()
        psr.set_threads(cores)
        ns = 6   # 6 tiles and their 15 pairs: neighbours and unrelated pairs, as in the full stage
        t0 = time.perf_counter()
        fs = [psr.extract_features(psr.luma_to_f32(t), 3, 3, 0.015, True, kind=0) for t in tiles[:ns]]
        t_ext = (time.perf_counter() - t0) / ns
        grays = [psr.luma_to_f32(t) for t in tiles[:ns]]
        sp = [(i, j) for i in range(ns) for j in range(i + 1, ns)]
        t0 = time.perf_counter()
        for i, j in sp:
            m = psr.match(fs[i][0], fs[i][1], fs[j][0], fs[j][1], opt.ratio)
            psr.panorama_pair(grays[i], grays[j], m, opt.ransac_threshold)
        t_pair = (time.perf_counter() - t0) / len(sp)
        t_stage = n * t_ext + npairs * t_pair
        out["cpu_baseline"] = {"value": npairs / t_stage, "unit": "image pairs/s", "cores": cores, "kind": "port",
                               "sample": f"{ns} tiles extracted ({t_ext:.3f} s each) and their {len(sp)} pairs "
                                         f"matched + RANSAC + edges ({t_pair:.3f} s each), scaled to the "
                                         f"{n}-tile stage"}
    return out


def fp64_statement(a, b) -> dict:
    """Errors of the GPU results on the frame pair against the float64 evaluation of the same
    algorithm (fp64_errors.py): keypoints (all, uncapped), a sample of 256 descriptors, the
    ratio matches and the homography of the matches."""
    import numpy as np
    import fp64_errors as fe
    import pure_sift as ps
    t0 = time.perf_counter()
    img = a.astype(np.float32) * np.float32(1.0 / 255.0)      # luma_to_array2, pipeline.rs:957-966
    sift = ps.Sift(ps.SiftConfig().octaves(4).scales(3).contrast_threshold(0.015))
    pyr32 = sift.build_pyramid(img)
    p64 = fe.pyramid64(img, 4, 3, True)
    kps = sift.detect_keypoints(img)
    fq, ft = sift.detect_and_compute(a), sift.detect_and_compute(b)
    m = ps.Matcher(ps.MatchOptions(0.8)).match_(fq, ft)
    inl, H = ps.homography_estimate(m, ps.RansacOptions(3.0))
    out = {"gaussian_planes_abs_max": max(float(np.abs(g - np.asarray(g32, np.float64)).max())
                                          for (g, _), (g32, _) in zip(p64, pyr32)),
           "keypoints": fe.keypoint_errors(kps, pyr32, p64, 3),
           "descriptors": fe.descriptor_errors(fq.header, fq.descriptors, kps, p64, 4, sample=256),
           "matches": fe.match_errors(fq.descriptors, ft.descriptors, m),
           "homography": fe.homography_errors(m, inl, H) if H is not None else None}
    out["seconds"] = time.perf_counter() - t0
    return out


def cpu_baseline(a, b, runs: int = 10, warmup: int = 2, runs_1: int = 10):
    """The oracle (C restatement of the reference's CPU path, OpenMP where the reference uses
    Rayon) on the same 1080p frame: median of `runs` frames after `warmup` warm-ups with all
    granted cores, and the same with one thread; extract and match timed separately."""
    import statistics
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import psr
    cores = _host_cores()
    fb = psr.extract_features(psr.luma_to_f32(b), 4, 3, 0.015, True)

    def measure(threads, n, w):
        psr.set_threads(threads)
        te, tm = [], []
        for i in range(w + n):
            t0 = time.perf_counter()
            qh, qd = psr.extract_features(psr.luma_to_f32(a), 4, 3, 0.015, True)
            t1 = time.perf_counter()
            psr.match(qh, qd, fb[0], fb[1], 0.8)
            t2 = time.perf_counter()
            if i >= w:
                te.append(t1 - t0)
                tm.append(t2 - t1)
        frame = [x + y for x, y in zip(te, tm)]
        return statistics.median(frame), statistics.median(te), statistics.median(tm), sum(frame)

    med, med_e, med_m, tot = measure(cores, runs, warmup)
    med1, med1_e, med1_m, tot1 = measure(1, runs_1, warmup)
    psr.set_threads(cores)
    return {"value": 1.0 / med, "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"median of {runs} frames after {warmup} warm-ups of the same 1080p workload (oracle "
                      f"extract + ratio match, partner pre-extracted), {tot:.1f} s timed; 1-thread: median of "
                      f"{runs_1} after {warmup} warm-ups, {tot1:.1f} s timed",
            "t_extract_ms": med_e * 1e3, "t_match_ms": med_m * 1e3,
            "value_1_thread": 1.0 / med1, "t_extract_ms_1_thread": med1_e * 1e3, "t_match_ms_1_thread": med1_m * 1e3,
            "host": _cpu_info()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-runs", type=int, default=10,
                    help="CPU baseline: timed frames (median) after 2 warm-ups, all granted cores and 1 thread")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["frame", "hamming", "precision", "ransac", "batch", "catalog", "pair",
                                           "panorama"],
                    default="frame",
                    help="frame: the 1080p extract+match headline (default); hamming: the catalog shortlist")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device-to-device copy first (a known-size reference for HBM counters)")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the errors of the frame's results against the float64 evaluation")
    ap.add_argument("--no-secondary", action="store_true",
                    help="frame workload only, without the secondary workloads in the same JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(images shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "panorama":
        print(json.dumps(bench_panorama(cpu_sample=not args.no_cpu_baseline), default=float), flush=True)
        return
    if args.workload == "pair":
        print(json.dumps(bench_pair(cpu_pairs=0 if args.no_cpu_baseline else 64), default=float), flush=True)
        return
    if args.workload == "catalog":
        print(json.dumps(bench_catalog(steps=min(args.steps, 20), cpu_searches=0 if args.no_cpu_baseline else 10),
                         default=float), flush=True)
        return
    if args.workload == "batch":
        print(json.dumps(bench_batch(cpu_images=0 if args.no_cpu_baseline else 16), default=float), flush=True)
        return
    if args.workload == "ransac":
        print(json.dumps(bench_ransac(), default=float), flush=True)
        return
    if args.workload == "precision":
        print(json.dumps(bench_precision(steps=min(args.steps, 10), cpu_queries=0 if args.no_cpu_baseline else 2000)),
              flush=True)
        return
    if args.workload == "hamming":
        print(json.dumps(bench_hamming(steps=args.steps, cpu_queries=0 if args.no_cpu_baseline else 1000)), flush=True)
        return

    import numpy as np
    import torch
    import pure_sift as ps
    import synth

    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst

    W, H = 1920, 1080
    a, b, _ = synth.blob_pair(W, H, 2000, seed=1, band_noise=True)
    dev = torch.device("cuda:0")
    da = torch.from_numpy(a).to(dev)
    db = torch.from_numpy(b).to(dev)
    cap = ps.MAX_FEATURES_PER_IMAGE
    fsz = ps.FEAT_DTYPE.itemsize
    qf = torch.empty(cap * fsz, dtype=torch.uint8, device=dev)
    qd = torch.empty((cap, 128), dtype=torch.float32, device=dev)
    tf = torch.empty(cap * fsz, dtype=torch.uint8, device=dev)
    td = torch.empty((cap, 128), dtype=torch.float32, device=dev)
    cnt = torch.zeros(4, dtype=torch.int32, device=dev)
    mout = torch.empty(cap * ps.MATCH_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    cfg = ps.SiftConfig().octaves(4).scales(3).contrast_threshold(0.015)
    opt = ps.MatchOptions(ratio=0.8, mutual=False)
    eng = ps.Engine(0)
    c = cnt.data_ptr()
    eng.extract_u8(db.data_ptr(), H, W, W, cfg, tf.data_ptr(), td.data_ptr(), cap, c + 4)
    eng.sync()

    def step():
        eng.extract_u8(da.data_ptr(), H, W, W, cfg, qf.data_ptr(), qd.data_ptr(), cap, c)
        eng.match(qf.data_ptr(), qd.data_ptr(), c, cap, tf.data_ptr(), td.data_ptr(), c + 4, cap, opt,
                  mout.data_ptr(), c + 8)

    for _ in range(args.warmup):
        step()
    eng.sync()
    eng.check()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    eng.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    eng.check()
    # stage times come from a second pass: the HIP timing events between stages cost a few
    # microseconds each on the device, so they stay out of the timed loop
    prof_steps = min(args.steps, 20)
    eng.profile_begin(prof_steps)
    for _ in range(prof_steps):
        step()
    stages, calls = eng.profile_end()
    eng.check()
    counts = cnt.cpu().tolist()
    n_q, n_t, n_m = counts[0], counts[1], counts[2]
    if n_q == 0 or n_m == 0:
        raise SystemExit(f"bench produced no features/matches: {counts}")

    K = args.steps
    ms = (t1 - t0) / K * 1e3
    stage_ms = {k: v / max(calls, 1) for k, v in stages.items()}
    pyr_ms = stage_ms["octave0"] + stage_ms["octaves_extrema"]
    achieved = ALG_BYTES_1080P / (pyr_ms * 1e-3) / 1e9
    out = {
        "metric": "1080p frames/s (extract+match) & RANSAC hyps/solve, guided vs uniform",
        "value": K / (t1 - t0),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded blob images + homography-warped partner)",
        "config": {"workload": "1920x1080 u8 pair, 2000 blobs + band-limited noise; extract query "
                               "(4 octaves x 3 scales, upsampled, contrast 0.015, cap 4096) + FP32 ratio-0.8 "
                               "match against the pre-extracted partner",
                   "features": [n_q, n_t], "matches": n_m, "parallelism": "single GPU"},
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        # one frame = t_extract + t_match (a pair-based definition is 2 * t_extract + t_match)
        "t_extract_ms": round(sum(v for k, v in stage_ms.items() if k != "match"), 4),
        "t_match_ms": round(stage_ms["match"], 4),
        "roofline": {"bound": "hbm", "kernel": "scale-space stage: 17 blur launches (tile / sliding-window kernels) + 4 extrema launches "
                               "(which form level S+2 only around slice-S candidates), octaves 1-3 overlapped on "
                               "side streams",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_bytes_1080p(),
                     "alg_bytes_per_image": ALG_BYTES_1080P,
                     "stage_us": pyr_ms * 1e3,
                     "hbm_floor_us": ALG_BYTES_1080P / (HBM_PEAK_GBS * 1e9) * 1e6,
                     "valu_floor_us": ALG_VALU_OPS_1080P / (VALU_INT_PEAK_TOPS * 1e12) * 1e6,
                     "alg_valu_ops_per_image": ALG_VALU_OPS_1080P,
                     "frac_of_binding_floor": max(ALG_BYTES_1080P / (HBM_PEAK_GBS * 1e9),
                                                  ALG_VALU_OPS_1080P / (VALU_INT_PEAK_TOPS * 1e12)) / (pyr_ms * 1e-3)},
    }
    eng.close()
    if not args.no_fp64:
        out["fp64_errors"] = fp64_statement(a, b)
    if not args.no_cpu_baseline and args.cpu_runs > 0:
        out["cpu_baseline"] = cpu_baseline(a, b, runs=args.cpu_runs, runs_1=args.cpu_runs)
    if not args.no_secondary:
        sec = {}
        for name, fn in (("hamming", lambda: bench_hamming(cpu_queries=0 if args.no_cpu_baseline else 1000)),
                         ("precision", lambda: bench_precision(cpu_queries=0 if args.no_cpu_baseline else 2000)),
                         ("ransac", bench_ransac),
                         ("batch", lambda: bench_batch(cpu_images=0 if args.no_cpu_baseline else 16)),
                         ("catalog", lambda: bench_catalog(cpu_searches=0 if args.no_cpu_baseline else 10)),
                         ("pair", lambda: bench_pair(cpu_pairs=0 if args.no_cpu_baseline else 64)),
                         ("panorama", lambda: bench_panorama(cpu_sample=not args.no_cpu_baseline))):
            try:
                sec[name] = fn()
            except Exception as ex:   # a secondary workload never hides the headline line
                sec[name] = {"error": f"{type(ex).__name__}: {ex}"}
        out["secondary"] = sec
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
