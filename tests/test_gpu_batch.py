"""Batched extraction (grouped persistent launches, captured HIP graph) and the blur-shape
autotuners, through the C ABI.

Bars: every image of a batch gets exactly the features, descriptors and count of a
single-image extraction of the same pixels (graph capture, graph replay and eager launch
alike, with default and with autotuned grouped launch shapes); after single-level
autotuning, the pyramid is still bit-exact with the oracle."""
import ctypes

import numpy as np
import pytest

import psr
import synth

pytestmark = pytest.mark.gpu

# (1, 40) has no octave at all (its upsampled height is 2 < 3): it must come out empty
SIZES = [(240, 320), (241, 333), (480, 640), (77, 1500), (1000, 300), (768, 1024), (240, 320), (513, 515),
         (900, 1700), (333, 777), (16, 16), (1, 40), (1536, 2048)]


def _pack(images, pad):
    descs, chunks, off = [], [], 0
    for img, p in zip(images, pad):
        h, w = img.shape
        stride = w + p
        buf = np.zeros((h, stride), np.uint8)
        buf[:, :w] = img
        descs.append((off, h, w, stride, 0))
        chunks.append(buf.ravel())
        off += buf.size
    return np.concatenate(chunks), descs


def test_batch_equals_single(gpu):
    import torch
    dev = torch.device("cuda:0")
    images = [synth.blob_image(w, h, max(8, (w * h) // 8000), seed=i) for i, (h, w) in enumerate(SIZES)]
    pix, descs = _pack(images, [i % 3 * 5 for i in range(len(images))])
    im = np.array(descs, gpu.IMAGE_DESC_DTYPE)
    n, cap = len(images), gpu.MAX_FEATURES_PER_IMAGE
    fsz = gpu.FEAT_DTYPE.itemsize
    tpix = torch.from_numpy(pix).to(dev)
    cfg = gpu.SiftConfig().octaves(4).scales(3).contrast_threshold(0.015)
    # single-image reference on the same device pixels
    eng = gpu.Engine(0)
    ref_f = torch.zeros((n, cap * fsz), dtype=torch.uint8, device=dev)
    ref_d = torch.zeros((n, cap, 128), dtype=torch.float32, device=dev)
    ref_c = torch.zeros(n, dtype=torch.int32, device=dev)
    for i, (off, h, w, stride, _) in enumerate(descs):
        eng.extract_u8(tpix.data_ptr() + off, h, w, stride, cfg, ref_f[i].data_ptr(), ref_d[i].data_ptr(), cap,
                       ref_c.data_ptr() + 4 * i)
    eng.sync()
    rc = ref_c.cpu().numpy()
    assert rc.min() >= 0 and rc.max() > 100 and rc[SIZES.index((1, 40))] == 0
    bf = torch.zeros_like(ref_f)
    bd = torch.zeros_like(ref_d)
    bc = torch.zeros_like(ref_c)
    runs = [0, 0, gpu.BATCH_NO_GRAPH, "tune", 0, 0, gpu.BATCH_NO_GRAPH]   # capture, replay, eager; tuned
    for flags in runs:
        if flags == "tune":
            assert eng.autotune_batch(tpix.data_ptr(), im, cfg) > 0
            continue
        bf.zero_()
        bd.zero_()
        bc.fill_(-1)
        torch.cuda.synchronize()
        eng.extract_batch_u8(tpix.data_ptr(), im, cfg, bf.data_ptr(), bd.data_ptr(), cap, bc.data_ptr(), flags)
        eng.sync()
        eng.check()
        c = bc.cpu().numpy()
        assert np.array_equal(c, rc)
        F, D = bf.cpu().numpy(), bd.cpu().numpy()
        RF, RD = ref_f.cpu().numpy(), ref_d.cpu().numpy()
        for i in range(n):
            k = rc[i]
            assert np.array_equal(F[i, :k * fsz], RF[i, :k * fsz]), (flags, i)
            assert np.array_equal(D[i, :k], RD[i, :k]), (flags, i)
    eng.close()
    gpu.lib().ps_autotune_reset()


def test_autotuned_pyramid_still_bit_exact(gpu):
    cfg = gpu.SiftConfig().octaves(4).scales(3)
    eng = gpu.Engine(0)
    try:
        n = eng.autotune_blur([(120, 160), (97, 131), (480, 640)], cfg)
        L = gpu.lib()
        assert n > 0 and L.ps_autotune_entries() > 0
        out = (ctypes.c_int32 * 5)()
        for i in range(L.ps_autotune_entries()):
            assert L.ps_autotune_entry(i, out) == 0
            r, mode, hc, wc, shape = list(out)
            assert r >= 1 and 0 <= mode <= 4 and hc > 0 and wc > 0
            if shape & gpu.BLUR_SLIDE:
                assert (shape & 0xffff) in (64, 128, 192)
            else:
                assert shape in (16, 32, 64)
        assert L.ps_autotune_entry(L.ps_autotune_entries(), out) != 0
        for (h, w) in [(120, 160), (97, 131)]:
            img = psr.luma_to_f32(synth.blob_image(w, h, 20, seed=h))
            dev = gpu.Sift(cfg).build_pyramid(img)
            ref = psr.build_pyramid(img, 4, 3, True)
            for (g, d), r in zip(dev, ref):
                assert np.array_equal(g, r["gauss"]) and np.array_equal(d, r["dog"])
    finally:
        eng.close()
        gpu.lib().ps_autotune_reset()


def test_full_256_image_batch(gpu):
    """The 256-image batching workload at full size: seeded blob/noise u8 images with widths
    U[320, 2048] and heights U[240, 1536] (1000 + i seeds, about 200 blobs per 640x480 of
    area), blur shapes autotuned over all 256 sizes, one captured graph then its replay.
    Every image must equal a single-image extraction of the same device pixels bit for bit
    (count, feature records, descriptors); a sample of four images (the two smallest, a
    mid-size and the largest under 1 Mpx) is also checked against the oracle."""
    import torch
    import learned_ransac as lr
    from test_gpu_parity import _check_features
    dev = torch.device("cuda:0")
    sizes = synth.batch_sizes(256, seed=3)          # (w, h)
    descs, off = [], 0
    for (w, h) in sizes:
        descs.append((off, h, w, w, 0))
        off += w * h
    pix = torch.empty(off, dtype=torch.uint8, device=dev)
    for i, (w, h) in enumerate(sizes):
        nb = max(20, int(200 * w * h / (640 * 480)))
        pix[descs[i][0]:descs[i][0] + w * h] = lr.render_blobs(1, 1000 + i, dev, width=w, height=h,
                                                                 n_blobs=nb)[0].reshape(-1)
    torch.cuda.synchronize()
    n, cap, fsz = len(sizes), gpu.MAX_FEATURES_PER_IMAGE, gpu.FEAT_DTYPE.itemsize
    im = np.array(descs, gpu.IMAGE_DESC_DTYPE)
    cfg = gpu.SiftConfig().octaves(4).scales(3).contrast_threshold(0.015)
    eng = gpu.Engine(0)
    try:
        assert eng.autotune_batch(pix.data_ptr(), np.array(descs, gpu.IMAGE_DESC_DTYPE), cfg) > 0
        ref_f = torch.zeros((n, cap * fsz), dtype=torch.uint8, device=dev)
        ref_d = torch.zeros((n, cap, 128), dtype=torch.float32, device=dev)
        ref_c = torch.zeros(n, dtype=torch.int32, device=dev)
        for i, (o, h, w, s, _) in enumerate(descs):
            eng.extract_u8(pix.data_ptr() + o, h, w, s, cfg, ref_f[i].data_ptr(), ref_d[i].data_ptr(), cap,
                           ref_c.data_ptr() + 4 * i)
        eng.sync()
        eng.check()
        rc = ref_c.cpu().numpy()
        RF, RD = ref_f.cpu().numpy(), ref_d.cpu().numpy()
        del ref_f, ref_d
        assert rc.min() > 50 and rc.max() > 1000
        bf = torch.zeros((n, cap * fsz), dtype=torch.uint8, device=dev)
        bd = torch.zeros((n, cap, 128), dtype=torch.float32, device=dev)
        bc = torch.zeros(n, dtype=torch.int32, device=dev)
        for _ in range(2):                         # capture, then replay
            bc.fill_(-1)
            bd.zero_()
            torch.cuda.synchronize()
            eng.extract_batch_u8(pix.data_ptr(), im, cfg, bf.data_ptr(), bd.data_ptr(), cap, bc.data_ptr(), 0)
            eng.sync()
            eng.check()
            assert np.array_equal(bc.cpu().numpy(), rc)
            F, D = bf.cpu().numpy(), bd.cpu().numpy()
            for i in range(n):
                k = rc[i]
                assert np.array_equal(F[i, :k * fsz], RF[i, :k * fsz]), i
                assert np.array_equal(D[i, :k], RD[i, :k]), i
    finally:
        eng.close()
        gpu.lib().ps_autotune_reset()
    area = np.array([w * h for (w, h) in sizes])
    order = np.argsort(area, kind="stable")
    small = [int(order[0]), int(order[1])]
    mid = int(order[len(order) // 4])
    big = int(order[np.searchsorted(area[order], 1 << 20) - 1])
    host = pix.cpu().numpy()
    for i in small + [mid, big]:
        o, h, w, s, _ = descs[i]
        img = psr.luma_to_f32(host[o:o + w * h].reshape(h, w))
        rh, rd = psr.extract_features(img, 4, 3, 0.015, True)
        k = rc[i]
        got = gpu.Features.from_arrays(np.frombuffer(F[i, :k * fsz].tobytes(), gpu.FEAT_DTYPE), D[i, :k])
        assert len(got) == len(rh), (i, w, h)
        _check_features(got, rh, rd)


def test_counters_clean_when_the_batch_grows(gpu):
    """A batch of more images than the last one on the same engine: the counter block grows, and
    its new memory can have the old block's address. Every counter word of the new batch,
    error words included, must start at zero: check() reports nothing and the counts equal a
    fresh engine's, ten times over."""
    import torch
    dev = torch.device("cuda:0")
    sizes = [(240, 320), (300, 280), (200, 360), (256, 256), (180, 400)]
    imgs = [synth.blob_image(w, h, 40, seed=50 + i) for i, (h, w) in enumerate(sizes)]
    pix, descs = _pack(imgs, [0] * len(imgs))
    tpix = torch.from_numpy(pix).to(dev)
    cap, fsz = gpu.MAX_FEATURES_PER_IMAGE, gpu.FEAT_DTYPE.itemsize
    cfg = gpu.SiftConfig().octaves(3).scales(3).contrast_threshold(0.015)
    f = torch.zeros((5, cap * fsz), dtype=torch.uint8, device=dev)
    d = torch.zeros((5, cap, 128), dtype=torch.float32, device=dev)
    c = torch.zeros(5, dtype=torch.int32, device=dev)
    want = None
    for _ in range(10):
        eng = gpu.Engine(0)
        for n in (3, 5):
            im = np.array(descs[:n], gpu.IMAGE_DESC_DTYPE)
            eng.extract_batch_u8(tpix.data_ptr(), im, cfg, f.data_ptr(), d.data_ptr(), cap, c.data_ptr())
            eng.sync()
            eng.check()
        got = c.cpu().numpy().copy()
        if want is None:
            want = got
        assert np.array_equal(got, want) and got.min() > 20
        eng.close()
