"""C ABI checks that need no GPU: the library loads, exports every entry point the
header declares, validates configs with the reference's messages, and fails loudly
(no CPU fallback) when no device is usable."""
import ctypes as C
import re

import numpy as np
import pytest

import pure_sift as ps
from conftest import has_gpu


def declared_symbols():
    txt = open(ps.HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ps_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    names = declared_symbols()
    assert len(names) >= 20
    L = ps.lib()
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/ps_sift.h but not exported"


def test_struct_layouts_match_header():
    assert C.sizeof(ps._SiftConfigC) == 16
    assert C.sizeof(ps._MatchOptionsC) == 8
    assert ps.FEAT_DTYPE.itemsize == 20
    assert ps.KP_DTYPE.itemsize == 24
    assert ps.MATCH_DTYPE.itemsize == 48


@pytest.mark.parametrize("cfg,msg", [
    (ps.SiftConfig().octaves(0), "Invalid `octaves`=0: must be between 1 and 8"),
    (ps.SiftConfig().octaves(9), "Invalid `octaves`=9: must be between 1 and 8"),
    (ps.SiftConfig().scales(2), "Invalid `scales`=2: must be between 3 and 10 for 3D extrema extraction"),
    (ps.SiftConfig().scales(11), "Invalid `scales`=11: must be between 3 and 10 for 3D extrema extraction"),
    (ps.SiftConfig().contrast_threshold(0.0), "Invalid `contrast_threshold`=0: must be finite and in (0.0, 1.0]"),
    (ps.SiftConfig().contrast_threshold(float("nan")), "must be finite and in (0.0, 1.0]"),
    (ps.SiftConfig().contrast_threshold(1.5), "must be finite and in (0.0, 1.0]"),
])
def test_sift_config_validation(cfg, msg):
    """pipeline.rs:80-103, pipeline_tests.rs:1334-1338"""
    with pytest.raises(ValueError) as ei:
        ps.Sift.new(cfg)
    assert msg in str(ei.value)


def test_valid_configs_accepted():
    for o in (1, 8):
        for s in (3, 10):
            ps.SiftConfig().octaves(o).scales(s).validate()
    ps.SiftConfig().contrast_threshold(1.0).validate()


@pytest.mark.parametrize("ratio", [0.0, -0.1, 1.5, float("inf"), float("nan")])
def test_matcher_validation(ratio):
    """pipeline.rs:156-164"""
    with pytest.raises(ValueError, match="Invalid `ratio`"):
        ps.Matcher.new(ps.MatchOptions(ratio=ratio))
    ps.Matcher.new(ps.MatchOptions(ratio=1.0))


def test_injected_failure_is_reported():
    """The host-settable fault switch forces the error path without touching a device."""
    ps.lib().ps_debug_fail_next()
    with pytest.raises(ps.DeviceError) as ei:
        ps.Sift(ps.SiftConfig()).detect_and_compute(np.zeros((16, 16), np.float32))
    assert ei.value.status == ps.PS_E_INJECTED


@pytest.mark.skipif(has_gpu(), reason="checks the no-device path")
def test_no_device_fails_loudly():
    with pytest.raises(ps.DeviceError) as ei:
        ps.Sift(ps.SiftConfig()).detect_and_compute(np.zeros((16, 16), np.float32))
    assert ei.value.status == ps.PS_E_NO_DEVICE
    with pytest.raises(ps.DeviceError):
        ps.nn2(np.zeros((2, 128), np.float32), np.zeros((3, 128), np.float32))


def test_empty_match_needs_no_device():
    """pipeline.rs:1749-1751: empty query or < 2 targets -> empty result."""
    f = ps.Features(np.zeros(0, ps.FEAT_DTYPE), np.zeros((0, 128), np.float32))
    g = ps.Features(np.zeros(1, ps.FEAT_DTYPE), np.zeros((1, 128), np.float32))
    m = ps.Matcher(ps.MatchOptions(0.8))
    assert len(m.match_(f, g)) == 0
    assert len(m.match_(g, g)) == 0


def test_autotune_table_inspection_needs_no_device():
    txt = open(ps.HEADER_PATH).read()
    assert int(re.search(r"#define PS_BLUR_SLIDE (\d+)", txt).group(1)) == ps.BLUR_SLIDE
    L = ps.lib()
    L.ps_autotune_reset()
    assert L.ps_autotune_entries() == 0
    out = (C.c_int32 * 5)()
    assert L.ps_autotune_entry(0, out) == ps.PS_E_INVALID_ARG
    assert L.ps_autotune_entry(0, None) == ps.PS_E_INVALID_ARG


def test_status_strings():
    L = ps.lib()
    for s in (ps.PS_OK, ps.PS_E_INVALID_ARG, ps.PS_E_CONFIG, ps.PS_E_DEVICE, ps.PS_E_CAPACITY, ps.PS_E_NO_DEVICE,
              ps.PS_E_OUT_OF_MEMORY, ps.PS_E_INJECTED):
        assert L.ps_status_string(s).decode() not in ("", "unknown status")


@pytest.mark.parametrize("k", [0, 1025])
def test_hamming_topk_rejects_k(k):
    cat = np.zeros((4, 2), np.uint64)
    with pytest.raises(ps.DeviceError) as ei:
        ps.hamming_topk(cat, cat[:1], k)
    assert ei.value.status == ps.PS_E_INVALID_ARG
    with pytest.raises(ps.DeviceError) as ei:
        ps.rerank_l2(np.zeros((1, 128), np.float32), np.zeros((4, 128), np.float32), np.zeros((1, max(k, 1)), np.uint32)
                     if k else np.zeros((1, 0), np.uint32))
    assert ei.value.status == ps.PS_E_INVALID_ARG


def test_hamming_empty_inputs_need_no_device():
    """catalog.rs:708-710: no entries -> no votes; an empty catalog gives an all-empty shortlist."""
    q = np.ones((3, 2), np.uint64)
    assert ps.count_coarse_hamming_votes(q, np.zeros((0, 2), np.uint64)) == 0
    idx, dist = ps.hamming_topk(np.zeros((0, 2), np.uint64), q, 5)
    assert np.all(idx == 0xFFFFFFFF) and np.all(dist == 255)
    idx, dist = ps.hamming_topk(q, np.zeros((0, 2), np.uint64), 5)
    assert idx.shape == (0, 5)


def test_precision_option_validation():
    """The precision byte sits in MatchOptions' padding; values past FP8 are rejected."""
    with pytest.raises(ValueError, match="precision"):
        ps.Matcher.new(ps.MatchOptions(0.8, False, 3))
    for p in (ps.PRECISION_FP32, ps.PRECISION_BF16, ps.PRECISION_FP8):
        ps.Matcher.new(ps.MatchOptions(0.8, False, p))


def test_compute_descriptors_trivial_inputs_need_no_device():
    """pipeline.rs:1587-1589: an image under 3x3 or no keypoints gives no features; an unknown
    descriptor kind is rejected before any device work."""
    L = ps.lib()
    n = C.c_size_t(7)
    kp = np.zeros(1, ps.KP_DTYPE)
    small = np.zeros((2, 2), np.float32)
    assert L.ps_compute_descriptors(small.ctypes.data, 2, 2, 2, kp.ctypes.data, 1, 0, None, None, 0, C.byref(n)) == ps.PS_OK
    assert n.value == 0
    img = np.zeros((16, 16), np.float32)
    n.value = 7
    assert L.ps_compute_descriptors(img.ctypes.data, 16, 16, 16, kp.ctypes.data, 0, 0, None, None, 0, C.byref(n)) == ps.PS_OK
    assert n.value == 0
    assert L.ps_compute_descriptors(img.ctypes.data, 16, 16, 16, kp.ctypes.data, 1, 2, None, None, 0,
                                    C.byref(n)) == ps.PS_E_INVALID_ARG


def test_catalog_layouts():
    assert ps.CATALOG_HIT_DTYPE.itemsize == 120
    assert C.sizeof(ps._CatalogInfoC) == 32
    assert C.sizeof(ps._CatalogEntryC) == 32
    assert ps.CATALOG_HIT_DTYPE.fields["transform"][1] == 60
    assert ps.CATALOG_HIT_DTYPE.fields["homography"][1] == 84


@pytest.mark.parametrize("data,msg", [
    (b"PSIFT", "Unexpected EOF while reading binary catalog"),
    (b"NOTACAT!" + b"\0" * 40, "Invalid binary catalog magic header"),
    (b"PSIFTCAT" + bytes([2, 0, 0, 0]) + b"\0" * 25, "Unsupported binary catalog version 2 (expected 1)"),
    (b"PSIFTCAT" + bytes([1, 0, 0, 0]) + b"\0" * 12 + b"\x05" + b"\0" * 8, "Unknown descriptor kind 5 in binary catalog"),
    (b"PSIFTCAT" + bytes([1, 0, 0, 0]) + b"\0" * 13 + bytes([1] + [0] * 7) + bytes([3, 0, 0, 0]) + b"ab",
     "Truncated entry id in binary catalog"),
    (b"PSIFTCAT" + bytes([1, 0, 0, 0]) + b"\0" * 13 + bytes([1] + [0] * 7) + bytes([1, 0, 0, 0]) + b"\xc0" + b"\0" * 16,
     "Invalid UTF-8 entry id in binary catalog"),
    (b"PSIFTCAT" + bytes([1, 0, 0, 0]) + b"\0" * 13 + bytes([1] + [0] * 7) + bytes([1, 0, 0, 0]) + b"a" + b"\0" * 8
     + bytes([1] + [0] * 7) + b"\0" * 163, "Unexpected EOF while reading binary catalog"),
])
def test_catalog_format_errors_need_no_device(data, msg):
    """catalog.rs:447-540: a malformed PSIFTCAT file is rejected with the reference's message
    before any device work."""
    with pytest.raises(ValueError) as ei:
        ps.FeatureCatalog.from_binary_bytes(data)
    assert str(ei.value) == msg


def test_catalog_file_errors():
    with pytest.raises(OSError):
        ps.FeatureCatalog.load_binary("/nonexistent/dir/catalog.psift")
    L = ps.lib()
    for s in (ps.PS_E_IO, ps.PS_E_FORMAT):
        assert L.ps_status_string(s).decode() not in ("", "unknown status")


def test_catalog_needs_a_device():
    """A well-formed catalog still needs the GPU: no host copy is searched instead."""
    if has_gpu():
        pytest.skip("a GPU is present")
    empty = b"PSIFTCAT" + bytes([1, 0, 0, 0]) + b"\0" * 21
    with pytest.raises(ps.DeviceError) as ei:
        ps.FeatureCatalog.from_binary_bytes(empty)
    assert ei.value.status == ps.PS_E_NO_DEVICE
    with pytest.raises(ps.DeviceError):
        ps.FeatureCatalog(3, 3, 0.015)


def test_geometry_layouts():
    assert ps.FUNDAMENTAL_PROBLEM_DTYPE.itemsize == 16
    assert ps.FUNDAMENTAL_RESULT_DTYPE.itemsize == 88
    assert ps.GRIC_PROBLEM_DTYPE.itemsize == 128 and ps.GRIC_PROBLEM_DTYPE.fields["f"][1] == 56
    assert ps.GRIC_RESULT_DTYPE.itemsize == 24
    assert C.sizeof(ps._FundamentalOptionsC) == 32


def test_geometry_trivial_inputs_need_no_device():
    """geometry.rs:872-874 (< 7 matches: no model) and :1239-1241 (no matches: None)."""
    m = np.zeros(6, ps.MATCH_DTYPE)
    inl, F = ps.estimate_fundamental_ransac(m, 1.5)
    assert F is None and len(inl) == 0
    assert ps.select_model_gric(m[:0], np.eye(3), None) is None
    assert ps.select_model_gric(m, None, None) is None
    with pytest.raises(ValueError):
        ps.estimate_similarity(m, ps.RansacOptions(0.1))


def test_panorama_layouts_and_argument_errors():
    """stitch.rs:474-476: fewer than 2 images is the reference's error; bad options are refused
    before any device work."""
    assert C.sizeof(ps._PanoramaOptionsC) == 24
    assert ps.PANORAMA_PAIR_DTYPE.itemsize == 184 and ps.PAIR_EDGE_DTYPE.itemsize == 112
    assert ps.FEATURE_SET_DTYPE.itemsize == 8 and ps.MATCH_PROBLEM_DTYPE.itemsize == 16
    img = np.zeros((40, 40), np.uint8)
    with pytest.raises(ps.DeviceError) as ei:
        ps.panorama_pair_edges([img])
    assert ei.value.status == ps.PS_E_INVALID_ARG
    assert "at least 2 input images" in str(ei.value)
    with pytest.raises(ValueError):
        ps.panorama_pair_edges([img, img], ps.PanoramaOptions(octaves=0))
    with pytest.raises(ValueError):
        ps.panorama_pair_edges([img, img], ps.PanoramaOptions(ransac_threshold=float("nan")))
    if not has_gpu():
        with pytest.raises(ps.DeviceError) as ei:
            ps.panorama_pair_edges([img, img])
        assert ei.value.status == ps.PS_E_NO_DEVICE
