"""SHA-256 pins of the seeded synthetic inputs the benchmark and the parity tests use, so a
change of generator (or of the numpy random streams it draws from) is seen before it changes
a measured figure. Generators: synth.py (numpy PCG64 streams, float64 blob sums,
rounded and clipped to u8)."""
import numpy as np
import pytest

import synth

DIGESTS = {
    # 640x480, 200 blobs, N(0,3) noise, seed 0, and its homography-warped partner
    "pair_640x480": ("35d59b977898ee6bfcf5592e92ae258581ac202f1f45a083ecba35e09c1eda82",
                     "8a1e48cc90aa469bd92cac55d53ecf0ebc320bf337d259adbdad21538aaa27d6"),
    # 1920x1080, 2000 blobs, band-limited noise, seed 1 (the frame workload)
    "pair_1080p": ("1b050db619fda36610dad7402f11f16070434d1cb983794855a2e6c255e4f87d",
                   "71c2603f502927c0f3204af9aa63e4718d7082f37adff8c17638f4f7918139cc"),
    # 50k queries x 200k database unit descriptors, 20% planted, seed 1: queries, database, planted rows
    "descriptors_50k_200k": ("f5fe5a5393c4614245a2154932f915fa331f4e5f6e2290e831933944a0331d56",
                             "51a0b2f65321595906d7f2528b38dfe71d63ff8f21894bc05914287f4840ca83",
                             "5a9907ed354acba768327e92a3d904f19682107381936d278e1191fdab70800b"),
    # 1M random 128-bit codes and 1000 planted queries, seed 2024
    "codes_1m": ("ec7ec1e8486dbd3eec5e0ddffdf2383eb9ff783a50649a2e785049b823a44a16",
                 "73cef22bfb4a70bdc52b3ee87f9c76d6183ac3ae830d5c17a201b84566eda861"),
    # the 256 (width, height) sizes of the batch workload, seed 3
    "batch_sizes_256": ("b3c1b0ea04c97875e89b0e546a53dbe487060676bc69e11e252a50861306a341",),
}


def _generate(name):
    if name == "pair_640x480":
        a, b, _ = synth.blob_pair(640, 480, 200, seed=0)
        return a, b
    if name == "pair_1080p":
        a, b, _ = synth.blob_pair(1920, 1080, 2000, seed=1, band_noise=True)
        return a, b
    if name == "descriptors_50k_200k":
        return synth.unit_descriptors(50000, 200000, seed=1)
    if name == "codes_1m":
        cat, q, _, _ = synth.hamming_catalog(1000000, 1000, seed=2024)
        return cat, q
    if name == "batch_sizes_256":
        return (np.array(synth.batch_sizes(256, seed=3), np.int64),)
    raise KeyError(name)


@pytest.mark.parametrize("name", sorted(DIGESTS))
def test_synthetic_input_digest(name):
    got = tuple(synth.sha256(a) for a in _generate(name))
    assert got == DIGESTS[name]


def test_batch_size_ranges():
    sz = np.array(synth.batch_sizes(256, seed=3))
    assert sz[:, 0].min() >= 320 and sz[:, 0].max() <= 2048
    assert sz[:, 1].min() >= 240 and sz[:, 1].max() <= 1536
