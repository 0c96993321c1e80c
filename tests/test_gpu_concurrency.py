"""The C ABI under concurrent host threads, as the reference's callers use the hot path: Rayon
workers in catalog builds, stitching and catalog verification (cli.rs:2093-2113,
stitch.rs:493-499, catalog.rs:644-684) and up to four tokio blocking threads (main.rs:71, :279).

Eight threads call Sift.detect_and_compute, Matcher.match_ and FeatureCatalog.search_features
at once (ctypes drops the GIL for the duration of each foreign call, so the calls overlap inside
the library); every result must equal the same call made serially."""
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import psr
import synth
from catalog_scenes import scene_catalog, scene_u8

pytestmark = pytest.mark.gpu


def _hits_key(hits):
    return [(h.id, h.coarse_votes, h.tracking.match_count, h.tracking.raw_match_count,
             None if h.homography is None else h.homography.tobytes()) for h in hits]


def test_concurrent_calls_equal_serial(gpu):
    cfg = gpu.SiftConfig().octaves(4).scales(3)
    sift = gpu.Sift(cfg)
    matcher = gpu.Matcher(gpu.MatchOptions(0.8))
    pairs = [synth.blob_pair(480 + 32 * k, 360 + 16 * k, 120, seed=40 + k)[:2] for k in range(6)]
    cat_src = scene_catalog([(f"img_{i}.png", 300 + i) for i in range(12)])
    cat = gpu.FeatureCatalog.from_binary_bytes(psr.catalog_write(cat_src))
    q_imgs = [scene_u8(300 + i, 9, 5) for i in (1, 4, 7, 10)]
    q_sift = gpu.Sift(gpu.SiftConfig().octaves(3).scales(3))

    def extract(i):
        f = sift.detect_and_compute(pairs[i][0])
        return f.header.tobytes() + f.descriptors.tobytes()

    feats = [(sift.detect_and_compute(a), sift.detect_and_compute(b)) for a, b in pairs]

    def match(i):
        return matcher.match_(*feats[i]).tobytes()

    q_feats = [q_sift.detect_and_compute(im) for im in q_imgs]

    def search(i):
        return _hits_key(cat.search_features(q_feats[i]))

    def search_extract(i):
        return _hits_key(cat.search_features(q_sift.detect_and_compute(q_imgs[i])))

    jobs = ([(extract, i) for i in range(6)] + [(match, i) for i in range(6)] + [(search, i) for i in range(4)] +
            [(search_extract, i) for i in range(4)])
    serial = [fn(i) for fn, i in jobs]
    assert all(len(s) > 0 for s in serial)
    threads = set()

    def run(job):
        threads.add(threading.get_ident())
        fn, i = job
        return fn(i)

    for rep in range(3):
        order = np.random.default_rng(rep).permutation(len(jobs))
        with ThreadPoolExecutor(max_workers=8) as ex:
            got = list(ex.map(run, [jobs[k] for k in order]))
        for k, g in zip(order, got):
            assert g == serial[k], (rep, jobs[k][0].__name__, jobs[k][1])
    assert len(threads) > 1
