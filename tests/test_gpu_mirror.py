"""The C++ host mirror (pure-sift_amd/pure_sift.hpp) end to end: tests/cpp/mirror_demo (built
with the library) runs validation, detect_keypoints, detect_and_compute (f32 and u8),
compute_descriptors, match_ (plain and mutual), the feature catalog (upsert, search, save /
load, remove) and the panorama pair edges through the C ABI in its own process; its counts must
equal the Python mirror's on the same images."""
import os
import re
import subprocess

import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(REPO, "tests", "cpp", "mirror_demo")


def test_cpp_mirror_demo(gpu, tmp_path):
    assert os.path.exists(DEMO), "tests/cpp/mirror_demo is built by the library Makefile (build())"
    a, b, _ = synth.blob_pair(640, 480, 200, seed=7)
    qa, ta = tmp_path / "q.u8", tmp_path / "t.u8"
    qa.write_bytes(np.ascontiguousarray(a).tobytes())
    ta.write_bytes(np.ascontiguousarray(b).tobytes())
    env = dict(os.environ, MIRROR_DEMO_TMP=str(tmp_path))
    r = subprocess.run([DEMO, str(qa), str(ta), "480", "640"], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    got = dict((k, v) for k, v in re.findall(r"(\w+)=([\d,]+)", r.stdout))
    assert got["bad"] == "0"
    sift = gpu.Sift(gpu.SiftConfig().octaves(4).scales(3))
    img = a.astype(np.float32) * np.float32(1.0 / 255.0)
    kps = sift.detect_keypoints(img)
    fa, fb = sift.detect_and_compute(a), sift.detect_and_compute(b)
    assert int(got["keypoints"]) == len(kps)
    assert got["features"] == f"{len(fa)},{len(fb)}"
    assert int(got["described"]) == len(sift.compute_descriptors(img, kps))
    assert int(got["matches"]) == len(gpu.Matcher(gpu.MatchOptions(0.8)).match_(fa, fb))
    assert int(got["mutual"]) == len(gpu.Matcher(gpu.MatchOptions(0.8, True)).match_(fa, fb))
    assert int(got["catalog_hits"]) >= 1
    assert int(got["pano_edges"]) == len(gpu.panorama_pair_edges([a, b, a])) >= 2
    assert (tmp_path / "demo.psift").exists()
