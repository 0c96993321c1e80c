"""The learned RANSAC proposal end to end on the synthetic warped-pair workload (GPU
extraction, matching, injected outliers, 10k-hypothesis labels, MLP training with the
library's kernels, guided vs uniform sampling), with the train / test split by base image so
that no held-out pair's query image is seen in training.

Bars: at confidence 0.99 the MLP-guided sampler stops after at least 3x fewer hypotheses
than uniform sampling, with the final inlier count equal (within 1%) on at least 90% of
the held-out pairs and inlier recall against the labels of at least 0.95; the same holds
when the outliers' distance and ratio are resampled from the putative matches (only
geometric cues left). The full-size case is the 4096-pair, 256-base workload."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(ev):
    g, u = ev["guided"], ev["uniform"]
    assert u["hyps_mean"] >= 3.0 * g["hyps_mean"], (u["hyps_mean"], g["hyps_mean"])
    assert ev["equal_inlier_count_frac"] >= 0.9
    assert g["inlier_recall_mean"] >= 0.95 and u["inlier_recall_mean"] >= 0.95
    assert g["has_model_frac"] == 1.0
    assert ev["mlp_test_accuracy"] >= 0.9


@pytest.mark.parametrize("attrs", ["true", "resampled"])
def test_guided_beats_uniform(gpu, attrs):
    import learned_ransac as lr
    ev, params = lr.run(n_pairs=256, n_base=32, steps=300, outlier_attrs=attrs, seed=3)
    assert ev["test_pairs"] == 64 and ev["train_pairs"] == 192
    _check(ev)


def test_guided_beats_uniform_4096_pairs(gpu):
    """The full learned-proposal workload: 4096 pairs of 1024x768 images over 256 base images,
    600 training steps, 1024 held-out pairs of 64 unseen base images."""
    import learned_ransac as lr
    ev, params = lr.run(n_pairs=4096, n_base=256, steps=600, seed=0)
    assert ev["test_pairs"] == 1024 and ev["train_pairs"] == 3072
    print(f"4096 pairs: guided {ev['guided']['hyps_mean']:.2f} vs uniform {ev['uniform']['hyps_mean']:.1f} "
          f"vs reference {ev['reference_sampler']['hyps_mean']:.1f} hypotheses; equal inliers "
          f"{ev['equal_inlier_count_frac']:.3f}")
    _check(ev)
    assert np.isfinite(ev["guided"]["reproj_err_px_median"])


def test_training_resumes_from_checkpoint(gpu, tmp_path):
    """Checkpoint / resume of the MLP training: 40 Adam steps in one run equal 25 steps, a
    checkpoint (masters, moments, step), a load and 15 more steps, bit for bit (the minibatch of
    a step depends only on the seed and the step; the gradient reduction has a fixed order)."""
    import torch
    import learned_ransac as lr
    import pure_sift as ps
    eng = ps.Engine(0)
    try:
        with torch.cuda.stream(torch.cuda.ExternalStream(eng.stream())):
            ds = lr.build_dataset(eng, 48, 8, seed=5)
            y, _ = lr.labels_from_ransac(eng, ds, iters=2000)
            ids = np.arange(48)
            whole, _ = lr.train(eng, ds, y, ids, steps=40, batch=4096, seed=2)
            path = str(tmp_path / "mlp.ckpt")
            lr.train(eng, ds, y, ids, steps=25, batch=4096, seed=2, checkpoint=path)
            ck = ps.load_mlp_checkpoint(path)
            assert ck["step"] == 25 and ck["m"] is not None
            resumed, curve = lr.train(eng, ds, y, ids, steps=15, batch=4096, seed=2, resume=ck)
            assert curve[-1][0] == 40
            assert np.array_equal(resumed.view(np.uint32), whole.view(np.uint32))
            assert not np.array_equal(whole, ck["params"])
    finally:
        eng.close()
