"""GPU kernels of the match-scoring MLP (forward, backward, Adam, input features) through
the C ABI, against the oracle.

Bars: logits and gradients within f32-accumulation tolerance of the oracle's float64
evaluation at the same BF16 weights (|diff| <= 1e-4 * scale + 1e-6); features within 1e-5
of their float64 definition; one Adam step equal to its closed form to 1e-6; training on
a separable labelling reaches >= 95% accuracy.
"""
import numpy as np
import pytest

import psr

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.uniform(0, 1.5, (n, 1)), rng.normal(0, 0.3, (n, 1)), rng.uniform(-1, 1, (n, 2)),
                        rng.uniform(0, 1.2, (n, 1)), rng.uniform(0, 1, (n, 3))], 1).astype(np.float32)
    y = (rng.random(n) < 0.4).astype(np.float32)
    return x, y


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 40000])
def test_forward_matches_oracle(gpu, n):
    p = gpu.mlp_init(seed=n)
    x, _ = _data(n, seed=n)
    got = gpu.mlp_forward(p, x)
    ref = psr.mlp_forward(p, x, bf16=True)
    assert np.allclose(got, ref, rtol=1e-4, atol=1e-5), np.abs(got - ref).max()


@pytest.mark.parametrize("n", [1, 77, 4096, 50000])
def test_gradient_matches_oracle(gpu, n):
    p = gpu.mlp_init(seed=3)
    x, y = _data(n, seed=n + 1)
    loss, g = gpu.mlp_loss_grad(p, x, y)
    rl, rg = psr.mlp_loss_grad(p, x, y, bf16=True)
    assert abs(loss - rl) <= 1e-4 * abs(rl) + 1e-6
    scale = np.abs(rg).max()
    assert np.abs(g - rg).max() <= 1e-4 * scale + 1e-6, (np.abs(g - rg).max(), scale)


def test_adam_step_closed_form(gpu):
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(gpu.MLP_PARAMS).astype(np.float32)
    g = rng.standard_normal(gpu.MLP_PARAMS).astype(np.float32)
    tp = torch.from_numpy(p0.copy()).to(dev)
    tm = torch.zeros_like(tp)
    tv = torch.zeros_like(tp)
    tg = torch.from_numpy(g).to(dev)
    eng = gpu.Engine(0)
    for step in (1, 2):
        eng.mlp_adam(tp.data_ptr(), tm.data_ptr(), tv.data_ptr(), tg.data_ptr(), step, 0.01)
    eng.sync()
    eng.close()
    m = np.zeros_like(p0, np.float64)
    v = np.zeros_like(p0, np.float64)
    p = p0.astype(np.float64)
    for step in (1, 2):
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g.astype(np.float64) ** 2
        p -= 0.01 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    assert np.allclose(tp.cpu().numpy(), p, rtol=1e-6, atol=1e-6)


def test_features_match_definition(gpu):
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(8)
    pairs, ms, qfs, tfs, mo, qo, to = [], [], [], [], 0, 0, 0
    for k, (nm, nqf, ntf) in enumerate([(50, 80, 70), (1, 3, 3), (300, 400, 500), (0, 5, 5)]):
        qf = np.zeros(nqf, gpu.FEAT_DTYPE)
        tf = np.zeros(ntf, gpu.FEAT_DTYPE)
        for f in (qf, tf):
            f["scale"] = rng.uniform(0.5, 20, len(f))
            f["orientation"] = rng.uniform(0, 2 * np.pi, len(f))
        m = psr.make_matches(rng.uniform(0, 1000, (nm, 2)), rng.uniform(0, 1000, (nm, 2)))
        m["query_index"] = rng.integers(0, nqf, nm)
        m["target_index"] = rng.integers(0, ntf, nm)
        m["ratio"] = np.round(rng.uniform(0.1, 1.4, nm), 1)   # ties exercise the (ratio, index) rank
        if nm:
            m["ratio"][0] = np.nan
        m["distance"] = rng.uniform(0, 1.4, nm)
        pairs.append((mo, nm, qo, to, 1024.0, 768.0))
        ms.append(m)
        qfs.append(qf)
        tfs.append(tf)
        mo, qo, to = mo + nm, qo + nqf, to + ntf
    allm = np.concatenate(ms)
    td = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.uint8)).to(dev)
    tm, tq, tt = td(allm), td(np.concatenate(qfs)), td(np.concatenate(tfs))
    tp = td(np.array(pairs, gpu.MLP_PAIR_DTYPE))
    tx = torch.zeros((mo, 8), dtype=torch.float32, device=dev)
    eng = gpu.Engine(0)
    eng.mlp_features(tm.data_ptr(), tq.data_ptr(), tt.data_ptr(), tp.data_ptr(), len(pairs), tx.data_ptr())
    eng.sync()
    eng.close()
    got = tx.cpu().numpy()
    for k, (o, nm, _, _, w, h) in enumerate(pairs):
        if nm == 0:
            continue
        ref = psr.mlp_features(ms[k], qfs[k], tfs[k], w, h)
        assert np.allclose(got[o:o + nm], ref, rtol=1e-5, atol=1e-5), k


def test_training_learns_separable_labels(gpu):
    """Adam on the GPU kernels: label = (ratio < 0.6) XOR (x5 > 0.7) is learned to >= 95%."""
    import torch
    dev = torch.device("cuda:0")
    x, _ = _data(60000, seed=11)
    y = ((x[:, 0] < 0.6) ^ (x[:, 5] > 0.7)).astype(np.float32)
    tx = torch.from_numpy(x).to(dev)
    ty = torch.from_numpy(y).to(dev)
    tp = torch.from_numpy(gpu.mlp_init(seed=1)).to(dev)
    tm, tv, tg = torch.zeros_like(tp), torch.zeros_like(tp), torch.zeros_like(tp)
    eng = gpu.Engine(0)
    for step in range(1, 401):
        eng.mlp_grad(tp.data_ptr(), tx.data_ptr(), ty.data_ptr(), len(x), tg.data_ptr())
        eng.mlp_adam(tp.data_ptr(), tm.data_ptr(), tv.data_ptr(), tg.data_ptr(), step, 0.01)
    logit = torch.empty(len(x), dtype=torch.float32, device=dev)
    eng.mlp_forward(tp.data_ptr(), tx.data_ptr(), len(x), logit.data_ptr())
    eng.sync()
    eng.close()
    acc = np.mean((logit.cpu().numpy() > 0) == (y > 0.5))
    assert acc >= 0.95, acc
