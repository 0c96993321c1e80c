"""ps_engine_match_static (the matcher without device allocation: scratch in a static device
array) against ps_engine_match on the same device features, through the C ABI.

Bar: bit-exact -- the same match records and count -- on two calls in a row, the second with
a smaller input (the static scratch keeps the first call's contents), with and without the
mutual check; sizes beyond 4096 features and the non-FP32 precisions are rejected."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def _extract(gpu, eng, dev, img):
    import torch
    cap = gpu.MAX_FEATURES_PER_IMAGE
    f = torch.zeros(cap * gpu.FEAT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    d = torch.zeros((cap, 128), dtype=torch.float32, device=dev)
    c = torch.zeros(1, dtype=torch.int32, device=dev)
    t = torch.from_numpy(np.ascontiguousarray(img)).to(dev)
    h, w = img.shape
    eng.extract_u8(t.data_ptr(), h, w, w, gpu.SiftConfig().octaves(4).scales(3), f.data_ptr(), d.data_ptr(), cap,
                   c.data_ptr())
    eng.sync()
    eng.check()
    return f, d, c


def _run(gpu, eng, dev, q, t, mutual, static):
    import torch
    cap = gpu.MAX_FEATURES_PER_IMAGE
    out = torch.full((cap * gpu.MATCH_DTYPE.itemsize,), 0xAB, dtype=torch.uint8, device=dev)
    cnt = torch.full((1,), -1, dtype=torch.int32, device=dev)
    fn = eng.match_static if static else eng.match
    fn(q[0].data_ptr(), q[1].data_ptr(), q[2].data_ptr(), cap, t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(),
       cap, gpu.MatchOptions(0.8, mutual), out.data_ptr(), cnt.data_ptr())
    eng.sync()
    n = int(cnt.cpu().item())
    return n, out.cpu().numpy().view(gpu.MATCH_DTYPE)[:n].copy()


@pytest.mark.parametrize("mutual", [False, True])
def test_static_matcher_equals_engine_matcher(gpu, mutual):
    import torch
    dev = torch.device("cuda:0")
    eng = gpu.Engine(0)
    try:
        big = synth.blob_pair(1920, 1080, 2000, seed=1, band_noise=True)[:2]
        small = synth.blob_pair(640, 480, 200, seed=5)[:2]
        fb = [_extract(gpu, eng, dev, im) for im in big]
        fs = [_extract(gpu, eng, dev, im) for im in small]
        counts = [int(x[2].cpu().item()) for x in fb + fs]
        assert counts[0] == counts[1] == 4096 and 0 < counts[2] < 4096 and 0 < counts[3] < 4096
        for q, t in ((fb[0], fb[1]), (fs[0], fs[1])):        # second call smaller than the first
            n_ref, ref = _run(gpu, eng, dev, q, t, mutual, static=False)
            n, got = _run(gpu, eng, dev, q, t, mutual, static=True)
            assert n == n_ref and n > 20
            assert np.array_equal(got, ref)
        # and once more at full size after the small call
        n_ref, ref = _run(gpu, eng, dev, fb[0], fb[1], mutual, static=False)
        n, got = _run(gpu, eng, dev, fb[0], fb[1], mutual, static=True)
        assert n == n_ref and np.array_equal(got, ref)
    finally:
        eng.close()


def test_static_matcher_limits(gpu):
    import torch
    dev = torch.device("cuda:0")
    eng = gpu.Engine(0)
    try:
        buf = torch.zeros(1 << 20, dtype=torch.uint8, device=dev)
        p = buf.data_ptr()
        with pytest.raises(gpu.DeviceError) as ei:
            eng.match_static(p, p, p, 4097, p, p, p, 16, gpu.MatchOptions(0.8), p, p)
        assert ei.value.status == gpu.PS_E_INVALID_ARG
        with pytest.raises(gpu.DeviceError) as ei:
            eng.match_static(p, p, p, 16, p, p, p, 4097, gpu.MatchOptions(0.8), p, p)
        assert ei.value.status == gpu.PS_E_INVALID_ARG
        opt = gpu.MatchOptions(0.8)
        opt.precision = 1
        with pytest.raises(gpu.DeviceError) as ei:
            eng.match_static(p, p, p, 16, p, p, p, 16, opt, p, p)
        assert ei.value.status == gpu.PS_E_INVALID_ARG
    finally:
        eng.close()
