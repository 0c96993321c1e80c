"""The learned RANSAC proposal end to end on a small version of the synthetic warped-pair
workload (GPU extraction, matching, injected outliers, 10k-hypothesis labels, MLP training
with the library's kernels, guided vs uniform sampling).

Bars: at confidence 0.99 the MLP-guided sampler stops after at least 3x fewer hypotheses
than uniform sampling, with the final inlier count equal (within 1%) on at least 90% of
the held-out pairs and inlier recall against the labels of at least 0.95; the same holds
when the outliers' distance and ratio are resampled from the putative matches (only
geometric cues left)."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("attrs", ["true", "resampled"])
def test_guided_beats_uniform(gpu, attrs):
    import learned_ransac as lr
    ev, params = lr.run(n_pairs=256, n_base=32, steps=300, outlier_attrs=attrs, seed=3)
    g, u = ev["guided"], ev["uniform"]
    assert u["hyps_mean"] >= 3.0 * g["hyps_mean"], (u["hyps_mean"], g["hyps_mean"])
    assert ev["equal_inlier_count_frac"] >= 0.9
    assert g["inlier_recall_mean"] >= 0.95 and u["inlier_recall_mean"] >= 0.95
    assert g["has_model_frac"] == 1.0
    assert ev["mlp_test_accuracy"] >= 0.9
