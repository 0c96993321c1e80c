"""Pins the oracle's feature cap and output order (pipeline.rs:1633-1658 with :1566-1571):

* more than 4096 keypoints -> a STABLE sort by |response| descending (ties keep the scan order
  octave, slice, y, x), truncated to 4096 keypoints (pipeline.rs:1643-1651);
* features come in keypoint order, each keypoint's orientations in peak order (the flat_map of
  pipeline.rs:1566-1571), and the feature list is truncated to 4096 (pipeline.rs:1654-1656).

The expected keypoint sequence is rebuilt here with numpy's stable argsort over the oracle's
uncapped scan-order keypoints (a different sort implementation than the oracle's merge sort),
and the per-keypoint orientations with the oracle's orientation routine on the keypoint's own
Gaussian level, so a wrong comparator direction, an unstable sort, a missing truncation or
reordered orientations each fail one of these tests.
"""
import numpy as np
import pytest

import psr
import synth

CAP = 4096
THR = 0.015


def _runs(feats):
    """Splits the feature headers into runs of consecutive features of one keypoint."""
    key = np.stack([feats[f].view(np.uint32) for f in ("x", "y", "scale", "response")], axis=1)
    starts = [0] + [i for i in range(1, len(feats)) if not np.array_equal(key[i], key[i - 1])]
    ends = starts[1:] + [len(feats)]
    return list(zip(starts, ends))


def _same_kp(f, k):
    return all(np.float32(f[n]).view(np.uint32) == np.float32(k[n]).view(np.uint32)
               for n in ("x", "y", "scale", "response"))


def _check_capped_order(img, octaves, scales, upsample):
    kps = psr.detect_keypoints(img, octaves, scales, THR, upsample)
    assert len(kps) > CAP, len(kps)
    order = np.argsort(-np.abs(kps["response"]).astype(np.float64), kind="stable")[:CAP]
    expect = kps[order]
    feats, _ = psr.extract_features(img, octaves, scales, THR, upsample)
    assert len(feats) == CAP
    runs = _runs(feats)
    # features walk the capped keypoint list in order, one run per keypoint
    for j, (a, b) in enumerate(runs):
        assert _same_kp(feats[a], expect[j]), j
        assert 1 <= b - a <= 4
    # every run but the last is complete; the list stops inside or at the end of run len(runs)-1
    assert len(runs) <= CAP
    return kps, expect, feats, runs


def test_cap_stable_order_1080p():
    """The 1920x1080 band-noise query image (about 10k keypoints before the cap)."""
    u8, _, _ = synth.blob_pair(1920, 1080, 2000, seed=1, band_noise=True)
    img = psr.luma_to_f32(u8)
    kps, expect, feats, runs = _check_capped_order(img, 4, 3, True)
    # the 4096th kept |response| is >= every dropped one
    kept = np.abs(expect["response"])
    dropped = np.setdiff1d(np.arange(len(kps)), np.argsort(-np.abs(kps["response"]).astype(np.float64),
                                                             kind="stable")[:CAP])
    assert kept.min() >= np.abs(kps["response"][dropped]).max()


def test_cap_ties_keep_scan_order():
    """A lattice of identical blobs on a 16-pixel period gives thousands of keypoints with
    bit-identical |response| (the pyramid is shift-covariant by 16 / 2^o pixels in octave o);
    a stable sort keeps them in scan order, an unstable one or a reversed tie rule does not."""
    h = w = 1024
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    # a periodic bump: identical blob profile at every lattice site
    prof_y = np.exp(-((np.mod(yy, 16) - 8.0) ** 2) / (2 * 2.2 ** 2))
    prof_x = np.exp(-((np.mod(xx, 16) - 8.0) ** 2) / (2 * 2.2 ** 2))
    img = (0.3 + 0.5 * prof_y * prof_x).astype(np.float32)
    kps, expect, feats, runs = _check_capped_order(img, 3, 3, True)
    # there really are exact ties among the kept keypoints, and they span different rows
    r = np.abs(expect["response"])
    vals, counts = np.unique(r, return_counts=True)
    assert counts.max() >= 100
    tied = expect[r == vals[np.argmax(counts)]]
    # inside a tie class (one octave and slice: shifted copies of one blob) the scan order
    # y, x is kept
    key = [(int(k["octave"]), float(k["y"]), float(k["x"])) for k in tied]
    assert len({k[0] for k in key}) == 1 and len({k[1] for k in key}) > 1
    assert key == sorted(key)


def test_orientations_in_peak_order_per_keypoint():
    """Each keypoint's features carry the orientation routine's angles in its peak order
    (highest smoothed bin first, pipeline.rs:1323-1327), computed on the keypoint's octave and
    layer with octave-local coordinates (pipeline.rs:1519-1528)."""
    u8 = synth.blob_image(480, 360, 150, seed=21)
    img = psr.luma_to_f32(u8)
    pyr = psr.build_pyramid(img, 4, 3, True)
    kps = psr.detect_keypoints(img, 4, 3, THR, True)
    feats, _ = psr.extract_features(img, 4, 3, THR, True)
    assert 0 < len(kps) <= CAP
    runs = _runs(feats)
    assert len(runs) == len(kps)
    planes = {}
    multi = 0
    for j, (a, b) in enumerate(runs):
        kp = kps[j]
        assert _same_kp(feats[a], kp)
        o, l = int(kp["octave"]), int(kp["layer"])
        if (o, l) not in planes:
            planes[(o, l)] = psr.gradient_planes(pyr[o]["gauss"][l])
        mag, ang = planes[(o, l)]
        inv = np.float32(1.0) / np.float32(pyr[o]["scale_factor"])
        loc = kp.copy()
        loc["x"], loc["y"] = kp["x"] * inv, kp["y"] * inv
        loc["scale"] = max(np.float32(kp["scale"] * inv), np.float32(0.8))
        oris = psr.keypoint_orientations(mag, ang, loc)
        assert np.array_equal(feats["orientation"][a:b], oris), j
        multi += (b - a) > 1
    assert multi > 0


@pytest.mark.parametrize("seed", [3, 4])
def test_no_cap_at_or_below_the_limit(seed):
    """At most 4096 keypoints: no sort, features in scan order (pipeline.rs:1643 only sorts
    when the count exceeds the limit)."""
    u8 = synth.blob_image(320, 240, 40, seed=seed)
    img = psr.luma_to_f32(u8)
    kps = psr.detect_keypoints(img, 3, 3, THR, True)
    feats, _ = psr.extract_features(img, 3, 3, THR, True)
    runs = _runs(feats)
    assert len(runs) == len(kps) <= CAP
    for j, (a, _) in enumerate(runs):
        assert _same_kp(feats[a], kps[j])
