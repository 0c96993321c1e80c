"""Python host mirror of pure-sift 0.2.0's hot-path API over libps_sift.so.

Same names, argument meaning and error behaviour as the Rust API
(src/pipeline.rs:22-170):

    cfg = SiftConfig().octaves(4).scales(3).contrast_threshold(0.015)
    sift = Sift.new(cfg)                      # raises ValueError(msg) like Result::Err(String)
    feats = sift.detect_and_compute(img)      # img: (h, w) float32 in [0, 1] or uint8 luma
    matcher = Matcher.new(MatchOptions(ratio=0.8, mutual=False))
    matches = matcher.match_(feats_q, feats_t)

Features are returned as a `Features` record of numpy arrays (x, y, scale,
response, orientation, descriptors[n, 128]); `Features.to_list()` gives
per-feature `SiftFeature` objects like the reference's Vec<SiftFeature>.

Every compute call runs the HIP kernels; there is no CPU fallback. Without a
gfx950 device the calls raise DeviceError (PS_E_NO_DEVICE).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libps_sift.so")
REPO = os.path.dirname(HERE)
HEADER_PATH = os.path.join(REPO, "include", "ps_sift.h")

# pipeline.rs:10-19, :998, :1047-1054
BASE_SIGMA = 1.6
ASSUMED_INPUT_SIGMA = 0.5
DEFAULT_UPSAMPLE_FIRST_OCTAVE = True
DEFAULT_CONTRAST_THRESHOLD = 0.015
SIFT_DESCRIPTOR_LEN = 128
DEFAULT_MATCH_RATIO = 0.75
DEFAULT_RANSAC_REPROJ_THRESHOLD = 3.0
MAX_FEATURES_PER_IMAGE = 4096
BLUR_SLIDE = 65536  # PS_BLUR_SLIDE: flag bit of a sliding-window launch shape in the tuned table

PS_OK = 0
PS_E_INVALID_ARG = -1
PS_E_CONFIG = -2
PS_E_DEVICE = -3
PS_E_CAPACITY = -4
PS_E_NO_DEVICE = -5
PS_E_OUT_OF_MEMORY = -6
PS_E_INJECTED = -7
PS_E_IO = -8
PS_E_FORMAT = -9


class DeviceError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[{status}] {msg}")
        self.status = status


class DescriptorKind:
    Sift = 0
    RootSift = 1


class _SiftConfigC(C.Structure):
    _fields_ = [("octaves", C.c_uint32), ("scales", C.c_uint32), ("contrast_threshold", C.c_float),
                ("upsample_first_octave", C.c_uint8), ("descriptor_kind", C.c_uint8), ("_pad", C.c_uint8 * 2)]


class _MatchOptionsC(C.Structure):
    _fields_ = [("ratio", C.c_float), ("mutual", C.c_uint8), ("precision", C.c_uint8), ("_pad", C.c_uint8 * 2)]


class _RansacOptionsC(C.Structure):
    _fields_ = [("reproj_threshold", C.c_float), ("has_seed", C.c_uint8), ("_pad", C.c_uint8 * 3),
                ("seed", C.c_uint64)]


class _RansacParamsC(C.Structure):
    _fields_ = [("reproj_threshold", C.c_float), ("sampler", C.c_uint8), ("local_opt", C.c_uint8),
                ("sim_seed_model", C.c_uint8), ("adaptive", C.c_uint8), ("confidence", C.c_double),
                ("min_iters", C.c_uint32), ("max_iters", C.c_uint32)]


class _FundamentalOptionsC(C.Structure):
    _fields_ = [("sampson_threshold", C.c_double), ("max_iters", C.c_uint64), ("confidence", C.c_double),
                ("seed", C.c_uint64)]


FUNDAMENTAL_PROBLEM_DTYPE = np.dtype([("offset", "<u4"), ("n", "<u4"), ("seed", "<u8")])
FUNDAMENTAL_RESULT_DTYPE = np.dtype([("f", "<f8", (9,)), ("has_model", "<u4"), ("n_inliers", "<u4"),
                                     ("iters", "<u4"), ("_pad", "<u4")])
GRIC_PROBLEM_DTYPE = np.dtype([("offset", "<u4"), ("n", "<u4"), ("has_h", "u1"), ("has_f", "u1"), ("_pad", "u1", (6,)),
                               ("h", "<f4", (9,)), ("_pad2", "<u4"), ("f", "<f8", (9,))])
PAIR_PROBLEM_DTYPE = np.dtype([("offset", "<u4"), ("n", "<u4"), ("qfeat_offset", "<u4"), ("tfeat_offset", "<u4")])
PAIR_GEOMETRY_DTYPE = np.dtype([("raw_match_count", "<u4"), ("track_match_count", "<u4"), ("inlier_ratio", "<f4"),
                                ("mean_dx", "<f4"), ("mean_dy", "<f4"), ("median_dx", "<f4"), ("median_dy", "<f4"),
                                ("mean_scale_ratio", "<f4"), ("mean_rotation_deg", "<f4"), ("has_transform", "<u4"),
                                ("n_similarity", "<u4"), ("transform", "<f4", (6,)), ("has_homography", "<u4"),
                                ("n_homography", "<u4"), ("homography", "<f4", (9,)), ("has_fundamental", "<u4"),
                                ("n_fundamental", "<u4"), ("fundamental", "<f8", (9,)), ("has_gric", "<i4"),
                                ("gric_selected", "<i4"), ("homography_gric", "<f8"), ("fundamental_gric", "<f8"),
                                ("best_is_homography", "<u4"), ("n_best", "<u4")], align=True)
GRIC_RESULT_DTYPE = np.dtype([("has_comparison", "<i4"), ("selected", "<i4"), ("homography_gric", "<f8"),
                              ("fundamental_gric", "<f8")])


class _PanoramaOptionsC(C.Structure):
    _fields_ = [("octaves", C.c_uint32), ("scales", C.c_uint32), ("threshold", C.c_float), ("ratio", C.c_float),
                ("ransac_threshold", C.c_float), ("_pad", C.c_uint32)]


class _CatalogInfoC(C.Structure):
    _fields_ = [("n_entries", C.c_uint64), ("n_features", C.c_uint64), ("octaves", C.c_uint32),
                ("scales", C.c_uint32), ("contrast_threshold", C.c_float), ("descriptor_kind", C.c_uint8),
                ("_pad", C.c_uint8 * 3)]


class _CatalogEntryC(C.Structure):
    _fields_ = [("id", C.c_void_p), ("id_len", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32),
                ("feature_count", C.c_uint64)]


IMAGE_DESC_DTYPE = np.dtype([("offset", "<u8"), ("height", "<u4"), ("width", "<u4"), ("row_stride", "<u4"),
                             ("_pad", "<u4")])
BATCH_NO_GRAPH = 1
FEATURE_SET_DTYPE = np.dtype([("offset", "<u4"), ("count", "<u4")])
MATCH_PROBLEM_DTYPE = np.dtype([("query_set", "<u4"), ("target_set", "<u4"), ("out_offset", "<u4"), ("_pad", "<u4")])
PANORAMA_PAIR_DTYPE = np.dtype([("from", "<u4"), ("to", "<u4"), ("raw_matches", "<u4"), ("has_edge", "<u4"),
                                ("inliers", "<u8"), ("h_from_to", "<f8", (9,)), ("h_to_from", "<f8", (9,)),
                                ("mean_i_from", "<f8"), ("mean_i_to", "<f8")])
PAIR_EDGE_DTYPE = np.dtype([("from", "<u8"), ("to", "<u8"), ("h_from_to", "<f8", (9,)), ("inliers", "<u8"),
                            ("mean_i_from", "<f8"), ("mean_i_to", "<f8")])
assert PANORAMA_PAIR_DTYPE.itemsize == 184 and PAIR_EDGE_DTYPE.itemsize == 112

RANSAC_PROBLEM_DTYPE = np.dtype([("offset", "<u4"), ("n", "<u4"), ("sim_seed", "<u8"), ("hom_seed", "<u8")])
RANSAC_RESULT_DTYPE = np.dtype([("h", "<f4", (9,)), ("has_model", "<u4"), ("n_inliers", "<u4"), ("hyps", "<u4"),
                                ("sim", "<f4", (4,)), ("has_sim", "<u4"), ("n_sim", "<u4")])
SAMPLER_REFERENCE, SAMPLER_UNIFORM, SAMPLER_WEIGHTS = 0, 1, 2

FEAT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("scale", "<f4"), ("response", "<f4"),
                       ("orientation", "<f4")])
KP_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("scale", "<f4"), ("response", "<f4"),
                     ("octave", "<u4"), ("layer", "<u4")])
MATCH_DTYPE = np.dtype([("query_index", "<u8"), ("target_index", "<u8"), ("qx", "<f4"), ("qy", "<f4"),
                        ("tx", "<f4"), ("ty", "<f4"), ("distance", "<f4"), ("ratio", "<f4"),
                        ("dx", "<f4"), ("dy", "<f4")])

_lib = None


def lib() -> C.CDLL:
    """Loads the in-tree libps_sift.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C pure-sift_amd)")
        L = C.CDLL(LIB_PATH)
        vp, sz, i32, u32p = C.c_void_p, C.c_size_t, C.c_int32, C.POINTER(C.c_uint32)
        L.ps_init.argtypes = [C.c_int]
        L.ps_status_string.argtypes = [i32]
        L.ps_status_string.restype = C.c_char_p
        L.ps_last_error.restype = C.c_char_p
        L.ps_validate_sift_config.argtypes = [C.POINTER(_SiftConfigC), C.c_char_p, sz]
        L.ps_validate_match_options.argtypes = [C.POINTER(_MatchOptionsC), C.c_char_p, sz]
        L.ps_detect_and_compute.argtypes = [vp, sz, sz, sz, C.POINTER(_SiftConfigC), vp, vp, sz, C.POINTER(sz)]
        L.ps_detect_and_compute_u8.argtypes = [vp, sz, sz, sz, C.POINTER(_SiftConfigC), vp, vp, sz, C.POINTER(sz)]
        L.ps_detect_keypoints.argtypes = [vp, sz, sz, sz, C.POINTER(_SiftConfigC), vp, sz, C.POINTER(sz)]
        L.ps_compute_descriptors.argtypes = [vp, sz, sz, sz, vp, sz, C.c_uint8, vp, vp, sz, C.POINTER(sz)]
        L.ps_match.argtypes = [vp, vp, sz, vp, vp, sz, C.POINTER(_MatchOptionsC), vp, sz, C.POINTER(sz)]
        L.ps_build_pyramid.argtypes = [vp, sz, sz, sz, C.POINTER(_SiftConfigC), vp, sz, C.POINTER(sz),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.ps_nn2.argtypes = [vp, sz, vp, sz, vp, vp, vp]
        L.ps_engine_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.ps_engine_destroy.argtypes = [vp]
        L.ps_engine_stream.argtypes = [vp]
        L.ps_engine_stream.restype = vp
        L.ps_engine_sync.argtypes = [vp]
        L.ps_engine_check.argtypes = [vp]
        L.ps_engine_extract_u8.argtypes = [vp, vp, sz, sz, sz, C.POINTER(_SiftConfigC), vp, vp, sz, u32p]
        L.ps_engine_match.argtypes = [vp, vp, vp, u32p, sz, vp, vp, u32p, sz, C.POINTER(_MatchOptionsC), vp, u32p]
        L.ps_engine_match_static.argtypes = [vp, vp, vp, u32p, sz, vp, vp, u32p, sz, C.POINTER(_MatchOptionsC), vp,
                                             u32p]
        L.ps_debug_fail_next.argtypes = []
        L.ps_engine_profile_begin.argtypes = [vp, C.c_uint32]
        L.ps_engine_profile_end.argtypes = [vp, C.POINTER(C.c_float), u32p]
        L.ps_homography_estimate.argtypes = [vp, sz, C.POINTER(_RansacOptionsC), vp, C.POINTER(sz), vp,
                                             C.POINTER(C.c_int32)]
        L.ps_ransac_batch.argtypes = [vp, vp, vp, C.c_uint32, sz, C.POINTER(_RansacParamsC), vp, vp]
        L.ps_engine_ransac_batch.argtypes = [vp, vp, vp, vp, C.c_uint32, sz, C.POINTER(_RansacParamsC), vp, vp]
        L.ps_engine_mlp_features.argtypes = [vp, vp, vp, vp, vp, C.c_uint32, vp]
        L.ps_engine_mlp_forward.argtypes = [vp, vp, vp, sz, vp, vp]
        L.ps_engine_mlp_grad.argtypes = [vp, vp, vp, vp, sz, vp, vp]
        L.ps_engine_mlp_adam.argtypes = [vp, vp, vp, vp, vp, C.c_uint32, C.c_float]
        L.ps_mlp_forward.argtypes = [vp, vp, sz, vp]
        L.ps_mlp_loss_grad.argtypes = [vp, vp, vp, sz, vp, vp]
        L.ps_mlp_checkpoint_save.argtypes = [C.c_char_p, vp, vp, vp, C.c_uint32]
        L.ps_mlp_checkpoint_load.argtypes = [C.c_char_p, vp, vp, vp, u32p, u32p]
        L.ps_engine_extract_batch_u8.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(_SiftConfigC), vp, vp, sz, vp,
                                                 C.c_uint32]
        L.ps_engine_autotune_blur.argtypes = [vp, vp, C.c_uint32, C.POINTER(_SiftConfigC), u32p]
        L.ps_engine_autotune_batch.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(_SiftConfigC), u32p]
        L.ps_autotune_reset.argtypes = []
        L.ps_autotune_reset.restype = None
        L.ps_autotune_entries.argtypes = []
        L.ps_autotune_entries.restype = C.c_uint32
        L.ps_autotune_entry.argtypes = [C.c_uint32, C.POINTER(C.c_int32)]
        L.ps_autotune_entry.restype = i32
        L.ps_nn2_precision.argtypes = [vp, sz, vp, sz, C.c_int, vp, vp, vp, vp]
        L.ps_engine_nn2.argtypes = [vp, vp, u32p, sz, vp, u32p, sz, C.c_int, vp, vp, vp, vp]
        L.ps_quantize_u128.argtypes = [vp, sz, vp]
        L.ps_count_coarse_hamming_votes.argtypes = [vp, sz, vp, sz, C.POINTER(sz)]
        L.ps_hamming_topk.argtypes = [vp, sz, vp, sz, C.c_uint32, vp, vp]
        L.ps_rerank_l2.argtypes = [vp, sz, vp, sz, vp, C.c_uint32, vp, vp]
        L.ps_engine_quantize_u128.argtypes = [vp, vp, sz, vp]
        L.ps_engine_hamming_topk.argtypes = [vp, vp, sz, vp, sz, C.c_uint32, vp, vp]
        L.ps_engine_rerank_l2.argtypes = [vp, vp, sz, vp, sz, vp, C.c_uint32, vp, vp]
        L.ps_similarity_estimate.argtypes = [vp, sz, C.POINTER(_RansacOptionsC), vp, C.POINTER(sz), vp,
                                             C.POINTER(C.c_int32)]
        L.ps_fundamental_estimate.argtypes = [vp, sz, C.POINTER(_FundamentalOptionsC), vp, C.POINTER(sz), vp,
                                              C.POINTER(C.c_int32)]
        L.ps_fundamental_batch.argtypes = [vp, vp, C.c_uint32, sz, C.c_double, C.c_uint32, C.c_double, vp, vp]
        L.ps_engine_fundamental_batch.argtypes = [vp, vp, vp, C.c_uint32, sz, C.c_double, C.c_uint32, C.c_double,
                                                  vp, vp]
        L.ps_select_model_gric.argtypes = [vp, sz, vp, vp, C.c_double, vp]
        L.ps_engine_pair_geometry_batch.argtypes = [vp, vp, vp, vp, vp, C.c_uint32, sz, C.POINTER(_RansacOptionsC), vp,
                                                    vp]
        L.ps_pair_geometry_estimate.argtypes = [vp, sz, vp, sz, vp, sz, C.POINTER(_RansacOptionsC), vp, vp]
        L.ps_engine_match_sets.argtypes = [vp, vp, vp, sz, vp, C.c_uint32, vp, vp, sz, vp, C.c_uint32, C.c_uint32, vp,
                                           C.c_uint32, C.c_float, vp, vp]
        L.ps_engine_panorama_pairs.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(_PanoramaOptionsC), vp]
        L.ps_panorama_pair_edges.argtypes = [vp, vp, vp, vp, C.c_uint32, C.POINTER(_PanoramaOptionsC), vp, sz,
                                             C.POINTER(sz)]
        L.ps_engine_gric_batch.argtypes = [vp, vp, vp, C.c_uint32, C.c_double, vp]
        L.ps_catalog_create.argtypes = [C.c_uint32, C.c_uint32, C.c_float, C.c_uint8, C.POINTER(vp)]
        L.ps_catalog_load_binary.argtypes = [C.c_char_p, sz, C.POINTER(vp)]
        L.ps_catalog_load_file.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.ps_catalog_save_binary.argtypes = [vp, C.c_char_p]
        L.ps_catalog_destroy.argtypes = [vp]
        L.ps_catalog_get_info.argtypes = [vp, C.POINTER(_CatalogInfoC)]
        L.ps_catalog_entry.argtypes = [vp, sz, C.POINTER(_CatalogEntryC)]
        L.ps_catalog_entry_features.argtypes = [vp, sz, vp, vp, vp, sz]
        L.ps_catalog_upsert.argtypes = [vp, C.c_char_p, sz, C.c_uint32, C.c_uint32, vp, vp, sz]
        L.ps_catalog_remove.argtypes = [vp, C.c_char_p, sz, C.POINTER(C.c_int32)]
        L.ps_catalog_coarse_votes.argtypes = [vp, vp, sz, vp]
        L.ps_catalog_search.argtypes = [vp, vp, vp, sz, sz, sz, C.c_float, C.c_float, vp, sz, C.POINTER(sz)]
        L.ps_engine_catalog_search.argtypes = [vp, vp, vp, vp, sz, sz, sz, C.c_float, C.c_float, vp, sz,
                                               C.POINTER(sz)]
        for name in ("ps_engine_extract_batch_u8", "ps_engine_autotune_blur", "ps_engine_autotune_batch", "ps_engine_mlp_features", "ps_engine_mlp_forward", "ps_engine_mlp_grad", "ps_engine_mlp_adam",
                     "ps_mlp_forward", "ps_mlp_loss_grad", "ps_mlp_checkpoint_save", "ps_mlp_checkpoint_load", "ps_homography_estimate", "ps_ransac_batch", "ps_engine_ransac_batch", "ps_nn2_precision", "ps_engine_nn2", "ps_quantize_u128", "ps_count_coarse_hamming_votes", "ps_hamming_topk", "ps_rerank_l2",
                     "ps_engine_quantize_u128", "ps_engine_hamming_topk", "ps_engine_rerank_l2","ps_init", "ps_validate_sift_config", "ps_validate_match_options", "ps_detect_and_compute",
                     "ps_detect_and_compute_u8", "ps_detect_keypoints", "ps_match", "ps_build_pyramid", "ps_nn2",
                     "ps_engine_create", "ps_engine_destroy", "ps_engine_sync", "ps_engine_check",
                     "ps_engine_extract_u8", "ps_engine_match", "ps_engine_match_static", "ps_engine_profile_begin",
                     "ps_engine_profile_end", "ps_catalog_create", "ps_catalog_load_binary", "ps_catalog_load_file",
                     "ps_catalog_save_binary", "ps_catalog_destroy", "ps_catalog_get_info", "ps_catalog_entry",
                     "ps_catalog_entry_features", "ps_catalog_upsert", "ps_catalog_remove",
                     "ps_catalog_coarse_votes", "ps_catalog_search", "ps_engine_catalog_search",
                     "ps_similarity_estimate", "ps_fundamental_estimate", "ps_fundamental_batch",
                     "ps_engine_fundamental_batch", "ps_select_model_gric", "ps_engine_gric_batch",
                     "ps_engine_pair_geometry_batch", "ps_pair_geometry_estimate", "ps_engine_match_sets",
                     "ps_engine_panorama_pairs", "ps_panorama_pair_edges"):
            getattr(L, name).restype = i32
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != PS_OK:
        msg = lib().ps_last_error().decode() or lib().ps_status_string(status).decode()
        if status in (PS_E_CONFIG, PS_E_FORMAT):
            raise ValueError(msg)
        if status == PS_E_IO:
            raise OSError(msg)
        raise DeviceError(status, msg)


@dataclasses.dataclass
class SiftConfig:
    """pipeline.rs:22-104 (builder methods return a modified copy, like the Rust by-value builder)."""
    octaves_: int = 3
    scales_: int = 3
    contrast_threshold_: float = DEFAULT_CONTRAST_THRESHOLD
    upsample_first_octave_: bool = DEFAULT_UPSAMPLE_FIRST_OCTAVE
    descriptor_kind_: int = DescriptorKind.Sift

    def octaves(self, v: int) -> "SiftConfig":
        return dataclasses.replace(self, octaves_=int(v))

    def scales(self, v: int) -> "SiftConfig":
        return dataclasses.replace(self, scales_=int(v))

    def contrast_threshold(self, v: float) -> "SiftConfig":
        return dataclasses.replace(self, contrast_threshold_=float(v))

    def upsample_first_octave(self, v: bool) -> "SiftConfig":
        return dataclasses.replace(self, upsample_first_octave_=bool(v))

    def descriptor_kind(self, v: int) -> "SiftConfig":
        return dataclasses.replace(self, descriptor_kind_=int(v))

    def _c(self) -> _SiftConfigC:
        if not (0 <= self.octaves_ < 2 ** 32 and 0 <= self.scales_ < 2 ** 32):
            raise ValueError(f"Invalid `octaves`={self.octaves_}: must be between 1 and 8")
        return _SiftConfigC(self.octaves_, self.scales_, self.contrast_threshold_,
                            1 if self.upsample_first_octave_ else 0, self.descriptor_kind_)

    def validate(self) -> None:
        """Raises ValueError with the reference's message (pipeline.rs:80-103)."""
        c = self._c()
        buf = C.create_string_buffer(256)
        st = lib().ps_validate_sift_config(C.byref(c), buf, 256)
        if st == PS_E_CONFIG:
            raise ValueError(buf.value.decode())
        check(st)


@dataclasses.dataclass
class MatchOptions:
    """pipeline.rs:1720-1736"""
    ratio: float = DEFAULT_MATCH_RATIO
    mutual: bool = False
    precision: int = 0   # PRECISION_FP32 (the reference), PRECISION_BF16 or PRECISION_FP8

    def _c(self) -> _MatchOptionsC:
        return _MatchOptionsC(self.ratio, 1 if self.mutual else 0, int(self.precision))


@dataclasses.dataclass
class SiftFeature:
    """pipeline.rs:1058-1072"""
    x: float
    y: float
    scale: float
    response: float
    orientation: float
    descriptor: np.ndarray


@dataclasses.dataclass
class Features:
    """A Vec<SiftFeature> as struct-of-arrays: header records + [n, 128] descriptors."""
    header: np.ndarray        # FEAT_DTYPE
    descriptors: np.ndarray   # float32 [n, 128]

    def __len__(self) -> int:
        return len(self.header)

    def to_list(self) -> List[SiftFeature]:
        return [SiftFeature(float(h["x"]), float(h["y"]), float(h["scale"]), float(h["response"]),
                            float(h["orientation"]), d.copy()) for h, d in zip(self.header, self.descriptors)]

    @staticmethod
    def from_arrays(header, descriptors) -> "Features":
        return Features(np.ascontiguousarray(header, FEAT_DTYPE),
                        np.ascontiguousarray(descriptors, np.float32).reshape(-1, 128))


def _img_arg(img: np.ndarray):
    if img.ndim != 2:
        raise ValueError("image must be 2-D (h, w)")
    if img.dtype == np.uint8:
        a = img if img.strides[1] == 1 and img.strides[0] >= img.shape[1] else np.ascontiguousarray(img)
        return a, a.strides[0], True
    a = img.astype(np.float32, copy=False)
    if a.strides[1] != 4 or a.strides[0] % 4:
        a = np.ascontiguousarray(a)
    return a, a.strides[0] // 4, False


class Sift:
    """pipeline.rs:106-146"""

    def __init__(self, config: SiftConfig):
        config.validate()
        self._config = config
        self._c = config._c()

    @staticmethod
    def new(config: SiftConfig) -> "Sift":
        return Sift(config)

    def config(self) -> SiftConfig:
        return self._config

    def detect_keypoints(self, img: np.ndarray) -> np.ndarray:
        """Uncapped keypoints in scan order (KP_DTYPE records). img: float32 [0,1] luma."""
        a, stride, is_u8 = _img_arg(img)
        if is_u8:
            a = a.astype(np.float32) * np.float32(1.0 / 255.0)
            stride = a.strides[0] // 4
        n = C.c_size_t(0)
        cap = 1 << 16
        while True:
            out = np.zeros(cap, KP_DTYPE)
            st = lib().ps_detect_keypoints(a.ctypes.data, a.shape[0], a.shape[1], stride, C.byref(self._c),
                                           out.ctypes.data, cap, C.byref(n))
            if st == PS_E_CAPACITY and n.value > cap:
                cap = n.value
                continue
            check(st)
            return out[:n.value].copy()

    def detect_and_compute(self, img: np.ndarray) -> Features:
        """Sift::detect_and_compute: <= 4096 features with 128-D descriptors."""
        a, stride, is_u8 = _img_arg(img)
        hdr = np.zeros(MAX_FEATURES_PER_IMAGE, FEAT_DTYPE)
        desc = np.zeros((MAX_FEATURES_PER_IMAGE, 128), np.float32)
        n = C.c_size_t(0)
        fn = lib().ps_detect_and_compute_u8 if is_u8 else lib().ps_detect_and_compute
        check(fn(a.ctypes.data, a.shape[0], a.shape[1], stride, C.byref(self._c), hdr.ctypes.data, desc.ctypes.data,
                 MAX_FEATURES_PER_IMAGE, C.byref(n)))
        return Features(hdr[:n.value].copy(), desc[:n.value].copy())

    def compute_descriptors(self, img: np.ndarray, keypoints: np.ndarray) -> Features:
        """compute_sift_descriptors_with_kind (pipeline.rs:1581-1620) with this config's descriptor
        kind: orientations + descriptors for the given keypoints (KP_DTYPE records), uncapped; the
        pyramid parameters are inferred from the keypoints as the reference does. Paired with
        detect_keypoints it is the reference's CLI / HTTP extract path (cli.rs:174-212)."""
        a, stride, is_u8 = _img_arg(img)
        if is_u8:
            a = a.astype(np.float32) * np.float32(1.0 / 255.0)
            stride = a.strides[0] // 4
        kps = np.ascontiguousarray(keypoints, KP_DTYPE)
        cap = max(4 * len(kps), 1)
        hdr = np.zeros(cap, FEAT_DTYPE)
        desc = np.zeros((cap, 128), np.float32)
        n = C.c_size_t(0)
        check(lib().ps_compute_descriptors(a.ctypes.data, a.shape[0], a.shape[1], stride, kps.ctypes.data, len(kps),
                                           int(self._config.descriptor_kind_), hdr.ctypes.data, desc.ctypes.data,
                                           cap, C.byref(n)))
        return Features(hdr[:n.value].copy(), desc[:n.value].copy())

    def build_pyramid(self, img: np.ndarray):
        """Debug/parity: [(gauss[S+3,h,w], dog[S+2,h,w]), ...] of the device pyramid."""
        a, stride, is_u8 = _img_arg(img)
        if is_u8:
            a = a.astype(np.float32) * np.float32(1.0 / 255.0)
            stride = a.strides[0] // 4
        n = C.c_size_t(0)
        dims = (C.c_int32 * 16)()
        noct = C.c_int32(0)
        check(lib().ps_build_pyramid(a.ctypes.data, a.shape[0], a.shape[1], stride, C.byref(self._c), None, 0,
                                     C.byref(n), dims, C.byref(noct)))
        buf = np.zeros(max(n.value, 1), np.float32)
        check(lib().ps_build_pyramid(a.ctypes.data, a.shape[0], a.shape[1], stride, C.byref(self._c),
                                     buf.ctypes.data, buf.size, C.byref(n), dims, C.byref(noct)))
        S = self._config.scales_
        res, off = [], 0
        for o in range(noct.value):
            h, w = dims[2 * o], dims[2 * o + 1]
            g = buf[off:off + (S + 3) * h * w].reshape(S + 3, h, w)
            off += (S + 3) * h * w
            d = buf[off:off + (S + 2) * h * w].reshape(S + 2, h, w)
            off += (S + 2) * h * w
            res.append((g, d))
        return res


class Matcher:
    """pipeline.rs:148-170"""

    def __init__(self, options: MatchOptions):
        c = options._c()
        buf = C.create_string_buffer(160)
        st = lib().ps_validate_match_options(C.byref(c), buf, 160)
        if st == PS_E_CONFIG:
            raise ValueError(buf.value.decode())
        check(st)
        self._options = options
        self._c = c

    @staticmethod
    def new(options: MatchOptions) -> "Matcher":
        return Matcher(options)

    def match_(self, query: Features, target: Features) -> np.ndarray:
        """Matcher::match_: FeatureMatch records (MATCH_DTYPE) in query order."""
        qh = np.ascontiguousarray(query.header, FEAT_DTYPE)
        th = np.ascontiguousarray(target.header, FEAT_DTYPE)
        qd = np.ascontiguousarray(query.descriptors, np.float32)
        td = np.ascontiguousarray(target.descriptors, np.float32)
        out = np.zeros(max(len(qh), 1), MATCH_DTYPE)
        n = C.c_size_t(0)
        check(lib().ps_match(qh.ctypes.data, qd.ctypes.data, len(qh), th.ctypes.data, td.ctypes.data, len(th),
                             C.byref(self._c), out.ctypes.data, out.size, C.byref(n)))
        return out[:n.value].copy()


def nn2(q_desc: np.ndarray, t_desc: np.ndarray):
    """(best_idx, best_d2, second_d2) per query, bit-exact with the reference's f32 sums."""
    q = np.ascontiguousarray(q_desc, np.float32).reshape(-1, 128)
    t = np.ascontiguousarray(t_desc, np.float32).reshape(-1, 128)
    bi = np.zeros(len(q), np.uint32)
    b = np.zeros(len(q), np.float32)
    s = np.zeros(len(q), np.float32)
    check(lib().ps_nn2(q.ctypes.data, len(q), t.ctypes.data, len(t), bi.ctypes.data, b.ctypes.data, s.ctypes.data))
    return bi, b, s


PRECISION_FP32, PRECISION_BF16, PRECISION_FP8 = 0, 1, 2


def nn2_precision(q_desc: np.ndarray, t_desc: np.ndarray, precision: int, with_approx: bool = False):
    """nn2 with a selectable distance engine: (best_idx, best_d2, second_d2[, approx_best_d2]).
    BF16 / FP8 rank with MFMA distances and rescore the 8 best in exact f32."""
    q = np.ascontiguousarray(q_desc, np.float32).reshape(-1, 128)
    t = np.ascontiguousarray(t_desc, np.float32).reshape(-1, 128)
    bi = np.zeros(len(q), np.uint32)
    b = np.zeros(len(q), np.float32)
    s = np.zeros(len(q), np.float32)
    ap = np.zeros(len(q), np.float32) if with_approx else None
    check(lib().ps_nn2_precision(q.ctypes.data, len(q), t.ctypes.data, len(t), precision, bi.ctypes.data,
                                 b.ctypes.data, s.ctypes.data, ap.ctypes.data if with_approx else None))
    return (bi, b, s, ap) if with_approx else (bi, b, s)


@dataclasses.dataclass
class RansacOptions:
    """pipeline.rs:179-213"""
    reproj_threshold: float = 3.0
    seed: Optional[int] = None

    def _c(self) -> _RansacOptionsC:
        return _RansacOptionsC(self.reproj_threshold, 0 if self.seed is None else 1, (C.c_uint8 * 3)(),
                               0 if self.seed is None else int(self.seed))


def homography_estimate(matches: np.ndarray, options: RansacOptions = RansacOptions()):
    """Homography::estimate (pipeline.rs:225-235): (inlier indices, 3x3 H or None)."""
    m = np.ascontiguousarray(matches, MATCH_DTYPE)
    inl = np.zeros(max(len(m), 1), np.uint32)
    h = np.zeros(9, np.float32)
    n = C.c_size_t(0)
    has = C.c_int32(0)
    c = options._c()
    check(lib().ps_homography_estimate(m.ctypes.data, len(m), C.byref(c), inl.ctypes.data, C.byref(n), h.ctypes.data,
                                       C.byref(has)))
    return inl[:n.value].copy(), (h.reshape(3, 3).copy() if has.value else None)


def ransac_params(sampler=SAMPLER_REFERENCE, reproj_threshold=3.0, confidence=0.999, min_iters=64, max_iters=2000,
                  local_opt=True, sim_seed_model=True, adaptive=True) -> _RansacParamsC:
    return _RansacParamsC(reproj_threshold, sampler, int(local_opt), int(sim_seed_model), int(adaptive), confidence,
                          min_iters, max_iters)


def ransac_batch(matches: np.ndarray, problems: np.ndarray, params: _RansacParamsC, weights=None):
    """Batched RANSAC: (results [RANSAC_RESULT_DTYPE], inlier index array aligned with matches)."""
    m = np.ascontiguousarray(matches, MATCH_DTYPE)
    pr = np.ascontiguousarray(problems, RANSAC_PROBLEM_DTYPE)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    res = np.zeros(len(pr), RANSAC_RESULT_DTYPE)
    inl = np.zeros(max(len(m), 1), np.uint32)
    check(lib().ps_ransac_batch(m.ctypes.data, None if w is None else w.ctypes.data, pr.ctypes.data, len(pr), len(m),
                                C.byref(params), inl.ctypes.data, res.ctypes.data))
    return res, inl


MLP_PARAMS, MLP_FEATURES = 4801, 8
MLP_PAIR_DTYPE = np.dtype([("match_offset", "<u4"), ("n", "<u4"), ("qfeat_offset", "<u4"), ("tfeat_offset", "<u4"),
                           ("width", "<f4"), ("height", "<f4")])


def mlp_init(seed: int = 0) -> np.ndarray:
    """He-uniform initial parameters of the 8-64-64-1 match-scoring MLP (f32 masters)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_PARAMS, np.float32)
    p[0:512] = rng.uniform(-1, 1, 512) * np.sqrt(6.0 / 8)
    p[576:4672] = rng.uniform(-1, 1, 4096) * np.sqrt(6.0 / 64)
    p[4736:4800] = rng.uniform(-1, 1, 64) * np.sqrt(6.0 / 64)
    return p


def mlp_forward(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Logits of the match-scoring MLP (BF16 weights, FP32 accumulation) for n x 8 inputs."""
    p = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
    out = np.zeros(len(xx), np.float32)
    check(lib().ps_mlp_forward(p.ctypes.data, xx.ctypes.data, len(xx), out.ctypes.data))
    return out


def save_mlp_checkpoint(path: str, params: np.ndarray, m: np.ndarray = None, v: np.ndarray = None,
                        step: int = 0) -> None:
    """Writes the MLP masters (and, for resuming training, the Adam moments and step) to `path`
    atomically (ps_mlp_checkpoint_save)."""
    p = np.ascontiguousarray(params, np.float32).reshape(MLP_PARAMS)
    if (m is None) != (v is None):
        raise ValueError("give both Adam moments or neither")
    mm = None if m is None else np.ascontiguousarray(m, np.float32).reshape(MLP_PARAMS)
    vv = None if v is None else np.ascontiguousarray(v, np.float32).reshape(MLP_PARAMS)
    check(lib().ps_mlp_checkpoint_save(os.fsencode(path), p.ctypes.data, None if mm is None else mm.ctypes.data,
                                       None if vv is None else vv.ctypes.data, int(step)))


def load_mlp_checkpoint(path: str) -> dict:
    """{"params", "m", "v" (None without moments), "step"} from a checkpoint file."""
    p = np.zeros(MLP_PARAMS, np.float32)
    m = np.zeros(MLP_PARAMS, np.float32)
    v = np.zeros(MLP_PARAMS, np.float32)
    step, flags = C.c_uint32(0), C.c_uint32(0)
    check(lib().ps_mlp_checkpoint_load(os.fsencode(path), p.ctypes.data, m.ctypes.data, v.ctypes.data,
                                       C.byref(step), C.byref(flags)))
    has = bool(flags.value & 1)
    return {"params": p, "m": m if has else None, "v": v if has else None, "step": int(step.value)}


def mlp_loss_grad(params: np.ndarray, x: np.ndarray, y: np.ndarray):
    """(mean BCE-with-logits loss, gradient [4801]) computed by the GPU training kernels."""
    p = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
    yy = np.ascontiguousarray(y, np.float32)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = C.c_float(0)
    check(lib().ps_mlp_loss_grad(p.ctypes.data, xx.ctypes.data, yy.ctypes.data, len(xx), g.ctypes.data,
                                 C.byref(loss)))
    return loss.value, g


# ---------------------------------------------------------------- two-view geometry

@dataclasses.dataclass
class FundamentalRansacOptions:
    """geometry.rs:108-128"""
    sampson_threshold: float = 1.5
    max_iters: int = 2000
    confidence: float = 0.999
    seed: int = 0xF00DCAFE12345678

    def _c(self) -> _FundamentalOptionsC:
        return _FundamentalOptionsC(self.sampson_threshold, self.max_iters, self.confidence, self.seed)


def estimate_similarity(matches: np.ndarray, options: RansacOptions = RansacOptions()):
    """Homography::estimate_similarity (pipeline.rs:238-246): (inlier indices, SimilarityTransform2D
    as (a, b, tx, ty, scale, rotation_deg) or None)."""
    m = np.ascontiguousarray(matches, MATCH_DTYPE)
    inl = np.zeros(max(len(m), 1), np.uint32)
    model = np.zeros(6, np.float32)
    n = C.c_size_t(0)
    has = C.c_int32(0)
    o = options._c()
    check(lib().ps_similarity_estimate(m.ctypes.data, len(m), C.byref(o), inl.ctypes.data, C.byref(n),
                                       model.ctypes.data, C.byref(has)))
    return inl[:n.value].copy(), (tuple(float(v) for v in model) if has.value else None)


def estimate_fundamental_ransac_with_options(matches: np.ndarray, options: FundamentalRansacOptions = None):
    """geometry.rs:868-1016: (inlier indices, 3x3 f64 F or None)."""
    options = options or FundamentalRansacOptions()
    m = np.ascontiguousarray(matches, MATCH_DTYPE)
    inl = np.zeros(max(len(m), 1), np.uint32)
    f = np.zeros(9)
    n = C.c_size_t(0)
    has = C.c_int32(0)
    o = options._c()
    check(lib().ps_fundamental_estimate(m.ctypes.data, len(m), C.byref(o), inl.ctypes.data, C.byref(n), f.ctypes.data,
                                        C.byref(has)))
    return inl[:n.value].copy(), (f.reshape(3, 3).copy() if has.value else None)


def estimate_fundamental_ransac(matches: np.ndarray, sampson_threshold: float):
    """geometry.rs:857-866"""
    return estimate_fundamental_ransac_with_options(matches, FundamentalRansacOptions(sampson_threshold))


def fundamental_batch(matches: np.ndarray, problems: np.ndarray, sampson_threshold=1.5, max_iters=2000,
                      confidence=0.999):
    """Many fundamental RANSAC problems in one launch: (results [FUNDAMENTAL_RESULT_DTYPE], inliers)."""
    m = np.ascontiguousarray(matches, MATCH_DTYPE)
    pr = np.ascontiguousarray(problems, FUNDAMENTAL_PROBLEM_DTYPE)
    res = np.zeros(len(pr), FUNDAMENTAL_RESULT_DTYPE)
    inl = np.zeros(max(len(m), 1), np.uint32)
    check(lib().ps_fundamental_batch(m.ctypes.data, pr.ctypes.data, len(pr), len(m), sampson_threshold, max_iters,
                                     confidence, inl.ctypes.data, res.ctypes.data))
    return res, inl


def select_model_gric(matches: np.ndarray, homography=None, fundamental=None, sigma: float = 1.0):
    """geometry.rs:1233-1341: None or (selected: "Homography" / "Fundamental", h_gric, f_gric)."""
    m = np.ascontiguousarray(matches, MATCH_DTYPE)
    h = None if homography is None else np.ascontiguousarray(homography, np.float32).reshape(9)
    f = None if fundamental is None else np.ascontiguousarray(fundamental, np.float64).reshape(9)
    out = np.zeros(1, GRIC_RESULT_DTYPE)
    check(lib().ps_select_model_gric(m.ctypes.data, len(m), None if h is None else h.ctypes.data,
                                     None if f is None else f.ctypes.data, sigma, out.ctypes.data))
    r = out[0]
    if not r["has_comparison"]:
        return None
    return ("Fundamental" if r["selected"] else "Homography", float(r["homography_gric"]),
            float(r["fundamental_gric"]))


def pair_geometry(query: Features, target: Features, matches: np.ndarray,
                  options: RansacOptions = RansacOptions()):
    """The post-processing of run_match_from_bytes (cli.rs:290-318) on one pair's raw matches:
    (PAIR_GEOMETRY_DTYPE record, kept inlier indices)."""
    q = np.ascontiguousarray(query.header, FEAT_DTYPE)
    t = np.ascontiguousarray(target.header, FEAT_DTYPE)
    m = np.ascontiguousarray(matches, MATCH_DTYPE)
    out = np.zeros(1, PAIR_GEOMETRY_DTYPE)
    best = np.zeros(max(len(m), 1), np.uint32)
    o = options._c()
    check(lib().ps_pair_geometry_estimate(q.ctypes.data, len(q), t.ctypes.data, len(t), m.ctypes.data, len(m),
                                          C.byref(o), best.ctypes.data, out.ctypes.data))
    return out[0], best[:int(out[0]["n_best"])].copy()


@dataclasses.dataclass
class PanoramaOptions:
    """stitch.rs:45-86 PanoramaOptions: the fields the pairwise stage reads, with the
    reference's defaults (3 octaves, 3 scales, DEFAULT_CONTRAST_THRESHOLD, DEFAULT_MATCH_RATIO,
    DEFAULT_RANSAC_REPROJ_THRESHOLD)."""
    octaves: int = 3
    scales: int = 3
    threshold: float = 0.015
    ratio: float = 0.75
    ransac_threshold: float = 3.0

    def _c(self) -> _PanoramaOptionsC:
        return _PanoramaOptionsC(self.octaves, self.scales, self.threshold, self.ratio, self.ransac_threshold, 0)


def panorama_pair_edges(images, options: PanoramaOptions = PanoramaOptions()) -> np.ndarray:
    """stitch.rs:478-571 up to the pair edges: images are u8 luma planes (h x w); returns the
    PAIR_EDGE_DTYPE records in the reference's order (per pair i < j with an edge: i -> j,
    then j -> i)."""
    ims = [np.ascontiguousarray(im, np.uint8) for im in images]
    n = len(ims)
    ptrs = (C.c_void_p * max(n, 1))(*[im.ctypes.data for im in ims])
    hs = np.array([im.shape[0] for im in ims] or [0], np.uint32)
    ws = np.array([im.shape[1] for im in ims] or [0], np.uint32)
    cap = max(n * (n - 1), 1)
    out = np.zeros(cap, PAIR_EDGE_DTYPE)
    cnt = C.c_size_t(0)
    o = options._c()
    check(lib().ps_panorama_pair_edges(ptrs, hs.ctypes.data, ws.ctypes.data, None, n, C.byref(o), out.ctypes.data,
                                       cap, C.byref(cnt)))
    return out[:cnt.value].copy()


HAMMING_MAX_K = 1024


def _codes(a) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint64)
    return a.reshape(-1, 2)


def quantize_descriptor_u128(desc: np.ndarray) -> np.ndarray:
    """catalog.rs:37-45: (n, 2) u64 little-endian words, bit i = desc[i] > 0.065."""
    d = np.ascontiguousarray(desc, np.float32).reshape(-1, 128)
    out = np.zeros((len(d), 2), np.uint64)
    check(lib().ps_quantize_u128(d.ctypes.data, len(d), out.ctypes.data))
    return out


def count_coarse_hamming_votes(query_sigs: np.ndarray, entry_sigs: np.ndarray) -> int:
    """catalog.rs:707-728: queries whose nearest entry code is within 24 bits."""
    q, e = _codes(query_sigs), _codes(entry_sigs)
    n = C.c_size_t(0)
    check(lib().ps_count_coarse_hamming_votes(q.ctypes.data, len(q), e.ctypes.data, len(e), C.byref(n)))
    return n.value


def hamming_topk(catalog: np.ndarray, queries: np.ndarray, k: int = 100):
    """Exact top-k shortlist by (Hamming distance, catalog index): (idx u32 [nq,k], dist u8 [nq,k])."""
    c, q = _codes(catalog), _codes(queries)
    idx = np.zeros((len(q), k), np.uint32)
    dist = np.zeros((len(q), k), np.uint8)
    check(lib().ps_hamming_topk(c.ctypes.data, len(c), q.ctypes.data, len(q), k, idx.ctypes.data, dist.ctypes.data))
    return idx, dist


def rerank_l2(q_desc: np.ndarray, cat_desc: np.ndarray, shortlist: np.ndarray):
    """FP32 rerank of a shortlist by (sequential squared L2, index): (idx [nq,k], d2 [nq,k])."""
    q = np.ascontiguousarray(q_desc, np.float32).reshape(-1, 128)
    c = np.ascontiguousarray(cat_desc, np.float32).reshape(-1, 128)
    sl = np.ascontiguousarray(shortlist, np.uint32)
    nq, k = sl.shape
    idx = np.zeros((nq, k), np.uint32)
    d2 = np.zeros((nq, k), np.float32)
    check(lib().ps_rerank_l2(q.ctypes.data, nq, c.ctypes.data, len(c), sl.ctypes.data, k, idx.ctypes.data,
                             d2.ctypes.data))
    return idx, d2


def extract_sift_features(img, octaves, scales, threshold, upsample_first_octave=True,
                          kind=DescriptorKind.Sift) -> Features:
    """Free-function funnel extract_sift_features_with_kind (pipeline.rs:1633-1658)."""
    cfg = SiftConfig(octaves, scales, threshold, upsample_first_octave, kind)
    return Sift(cfg).detect_and_compute(img)


def match_descriptors_with_options(query: Features, target: Features, options: MatchOptions) -> np.ndarray:
    """Free-function funnel (pipeline.rs:1744-1841)."""
    return Matcher(options).match_(query, target)


class Engine:
    """Device-resident pipeline (one HIP stream + workspace); takes device pointers (ints)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().ps_engine_create(device, C.byref(h)))
        self.h = h

    def stream(self) -> int:
        return lib().ps_engine_stream(self.h) or 0

    def extract_u8(self, d_img: int, h: int, w: int, stride: int, cfg: SiftConfig, d_feat: int, d_desc: int,
                   cap: int, d_count: int) -> None:
        c = cfg._c()
        check(lib().ps_engine_extract_u8(self.h, d_img, h, w, stride, C.byref(c), d_feat, d_desc, cap,
                                         C.cast(d_count, C.POINTER(C.c_uint32))))

    def match(self, d_qf, d_qd, d_nq, nq_max, d_tf, d_td, d_nt, nt_max, opt: MatchOptions, d_out, d_out_count):
        c = opt._c()
        u = lambda p: C.cast(p, C.POINTER(C.c_uint32))
        check(lib().ps_engine_match(self.h, d_qf, d_qd, u(d_nq), nq_max, d_tf, d_td, u(d_nt), nt_max, C.byref(c),
                                    d_out, u(d_out_count)))

    def match_static(self, d_qf, d_qd, d_nq, nq_max, d_tf, d_td, d_nt, nt_max, opt: MatchOptions, d_out,
                     d_out_count):
        """ps_engine_match without device allocation (static scratch; nq_max, nt_max <= 4096)."""
        c = opt._c()
        u = lambda p: C.cast(p, C.POINTER(C.c_uint32))
        check(lib().ps_engine_match_static(self.h, d_qf, d_qd, u(d_nq), nq_max, d_tf, d_td, u(d_nt), nt_max,
                                           C.byref(c), d_out, u(d_out_count)))

    def sync(self) -> None:
        check(lib().ps_engine_sync(self.h))

    def extract_batch_u8(self, d_pixels: int, images: np.ndarray, cfg: SiftConfig, d_feat: int, d_desc: int,
                         cap: int, d_counts: int, flags: int = 0) -> None:
        """images: IMAGE_DESC_DTYPE array (host); outputs at i * cap per image."""
        im = np.ascontiguousarray(images, IMAGE_DESC_DTYPE)
        c = cfg._c()
        check(lib().ps_engine_extract_batch_u8(self.h, d_pixels, im.ctypes.data, len(im), C.byref(c), d_feat, d_desc,
                                               cap, d_counts, flags))

    def autotune_blur(self, sizes, cfg: SiftConfig) -> int:
        """Times the blur tile heights for the pyramid levels of (height, width) sizes."""
        hw = np.ascontiguousarray(np.asarray(sizes, np.uint32).reshape(-1, 2))
        c = cfg._c()
        n = C.c_uint32(0)
        check(lib().ps_engine_autotune_blur(self.h, hw.ctypes.data, len(hw), C.byref(c), C.byref(n)))
        return n.value

    def autotune_batch(self, d_pixels: int, images: np.ndarray, cfg: SiftConfig) -> int:
        """Times the grouped launch shapes of this batch (images as for extract_batch_u8)."""
        im = np.ascontiguousarray(images, IMAGE_DESC_DTYPE)
        c = cfg._c()
        n = C.c_uint32(0)
        check(lib().ps_engine_autotune_batch(self.h, d_pixels, im.ctypes.data, len(im), C.byref(c), C.byref(n)))
        return n.value

    def nn2(self, d_q: int, d_nq: int, nq_max: int, d_t: int, d_nt: int, nt_max: int, precision: int,
            d_best_idx: int, d_best: int, d_second: int, d_approx: int = 0) -> None:
        u = lambda p: C.cast(p, C.POINTER(C.c_uint32))
        check(lib().ps_engine_nn2(self.h, d_q, u(d_nq), nq_max, d_t, u(d_nt), nt_max, precision, d_best_idx,
                                  d_best, d_second, d_approx or None))

    def ransac_batch(self, d_matches: int, d_weights: int, d_problems: int, n_problems: int, total_matches: int,
                     params, d_inliers: int, d_results: int) -> None:
        check(lib().ps_engine_ransac_batch(self.h, d_matches, d_weights or None, d_problems, n_problems, total_matches,
                                           C.byref(params), d_inliers, d_results))

    def pair_geometry_batch(self, d_matches: int, d_qfeat: int, d_tfeat: int, d_pairs: int, n_pairs: int,
                            total_matches: int, options: RansacOptions, d_best: int, d_out: int) -> None:
        o = options._c()
        check(lib().ps_engine_pair_geometry_batch(self.h, d_matches, d_qfeat, d_tfeat, d_pairs, n_pairs, total_matches,
                                                  C.byref(o), d_best, d_out))

    def match_sets(self, d_qf: int, d_qd: int, q_rows: int, d_qsets: int, n_qsets: int, d_tf: int, d_td: int,
                   t_rows: int, d_tsets: int, n_tsets: int, set_cap: int, d_problems: int, n_problems: int,
                   ratio: float, d_out: int, d_out_counts: int) -> None:
        """Many set-against-set ratio matches in one launch chain (FEATURE_SET_DTYPE /
        MATCH_PROBLEM_DTYPE tables on the device)."""
        check(lib().ps_engine_match_sets(self.h, d_qf, d_qd, q_rows, d_qsets, n_qsets, d_tf, d_td, t_rows, d_tsets,
                                         n_tsets, set_cap, d_problems, n_problems, ratio, d_out, d_out_counts))

    def panorama_pairs(self, d_pixels: int, images: np.ndarray, options: PanoramaOptions, d_pairs: int) -> None:
        """The panorama's pairwise stage on device images (IMAGE_DESC_DTYPE host array):
        n (n - 1) / 2 PANORAMA_PAIR_DTYPE records at d_pairs."""
        im = np.ascontiguousarray(images, IMAGE_DESC_DTYPE)
        o = options._c()
        check(lib().ps_engine_panorama_pairs(self.h, d_pixels, im.ctypes.data, len(im), C.byref(o), d_pairs))

    def mlp_features(self, d_matches: int, d_qfeat: int, d_tfeat: int, d_pairs: int, n_pairs: int, d_x: int) -> None:
        check(lib().ps_engine_mlp_features(self.h, d_matches, d_qfeat, d_tfeat, d_pairs, n_pairs, d_x))

    def mlp_forward(self, d_params: int, d_x: int, n: int, d_logit: int, d_prob: int = 0) -> None:
        check(lib().ps_engine_mlp_forward(self.h, d_params, d_x, n, d_logit, d_prob or None))

    def mlp_grad(self, d_params: int, d_x: int, d_y: int, n: int, d_grad: int, d_loss: int = 0) -> None:
        check(lib().ps_engine_mlp_grad(self.h, d_params, d_x, d_y, n, d_grad, d_loss or None))

    def mlp_adam(self, d_params: int, d_m: int, d_v: int, d_grad: int, step: int, lr: float) -> None:
        check(lib().ps_engine_mlp_adam(self.h, d_params, d_m, d_v, d_grad, step, lr))

    def quantize_u128(self, d_desc: int, n: int, d_codes: int) -> None:
        check(lib().ps_engine_quantize_u128(self.h, d_desc, n, d_codes))

    def hamming_topk(self, d_catalog: int, n: int, d_queries: int, nq: int, k: int, d_idx: int, d_dist: int) -> None:
        check(lib().ps_engine_hamming_topk(self.h, d_catalog, n, d_queries, nq, k, d_idx, d_dist))

    def rerank_l2(self, d_qdesc: int, nq: int, d_cat_desc: int, n: int, d_shortlist: int, k: int, d_idx: int,
                  d_d2: int) -> None:
        check(lib().ps_engine_rerank_l2(self.h, d_qdesc, nq, d_cat_desc, n, d_shortlist, k, d_idx, d_d2))

    # "octave0" ends when octave 0's levels are written on the engine stream; "octaves_extrema"
    # covers the later octaves (on side streams, started during octave 0) and every octave's
    # extrema, up to the join -- their sum is the whole scale-space stage
    STAGES = ("octave0", "octaves_extrema", "order", "describe", "match")

    def profile_begin(self, max_calls: int) -> None:
        check(lib().ps_engine_profile_begin(self.h, max_calls))

    def profile_end(self):
        """({stage: total ms}, number of extract calls) since profile_begin (HIP events)."""
        ms = (C.c_float * 5)()
        calls = C.c_uint32(0)
        check(lib().ps_engine_profile_end(self.h, ms, C.byref(calls)))
        return dict(zip(self.STAGES, [float(v) for v in ms])), calls.value

    def check(self) -> None:
        check(lib().ps_engine_check(self.h))

    def close(self) -> None:
        if self.h:
            lib().ps_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- feature catalog (catalog.rs)

DEFAULT_CATALOG_SHORTLIST = 16   # catalog.rs:21
DEFAULT_CATALOG_TOP_K = 5        # catalog.rs:24

CATALOG_HIT_DTYPE = np.dtype([("entry", "<u4"), ("coarse_votes", "<u4"), ("raw_match_count", "<u8"),
                              ("match_count", "<u8"), ("inlier_ratio", "<f4"), ("mean_dx", "<f4"),
                              ("mean_dy", "<f4"), ("median_dx", "<f4"), ("median_dy", "<f4"),
                              ("mean_scale_ratio", "<f4"), ("mean_rotation_deg", "<f4"), ("has_transform", "<u4"),
                              ("has_homography", "<u4"), ("transform", "<f4", (6,)), ("homography", "<f4", (9,))],
                             align=True)


@dataclasses.dataclass
class TrackingSummary:
    """pipeline.rs:1137-1153"""
    raw_match_count: int
    match_count: int
    inlier_ratio: float
    mean_displacement: tuple
    median_displacement: tuple
    mean_scale_ratio: float
    mean_rotation_deg: float


@dataclasses.dataclass
class SimilarityTransform2D:
    """pipeline.rs:1096-1109"""
    a: float
    b: float
    tx: float
    ty: float
    scale: float
    rotation_deg: float


@dataclasses.dataclass
class CatalogEntrySummary:
    """catalog.rs:47-57"""
    id: str
    width: int
    height: int
    feature_count: int


@dataclasses.dataclass
class CatalogSearchHit:
    """catalog.rs:158-177 (homography as a row-major 3x3 array)"""
    id: str
    width: int
    height: int
    feature_count: int
    coarse_votes: int
    tracking: TrackingSummary
    transform: Optional[SimilarityTransform2D]
    homography: Optional[np.ndarray]


class FeatureCatalog:
    """catalog.rs:179-197 FeatureCatalog, resident in device memory (libps_sift's ps_catalog).

    Mutating calls (upsert_entry, remove) need exclusive access; searches may run concurrently."""

    def __init__(self, octaves: int = 3, scales: int = 3, threshold: float = DEFAULT_CONTRAST_THRESHOLD,
                 descriptor_kind: int = DescriptorKind.Sift, _handle=None):
        if _handle is None:
            _handle = C.c_void_p()
            check(lib().ps_catalog_create(octaves, scales, threshold, descriptor_kind, C.byref(_handle)))
        self.h = _handle

    @staticmethod
    def new(octaves: int, scales: int, threshold: float) -> "FeatureCatalog":
        return FeatureCatalog(octaves, scales, threshold)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _lib is not None:
            _lib.ps_catalog_destroy(h)
            self.h = None

    # -- persistence (catalog.rs:361-570)
    @staticmethod
    def from_binary_bytes(data: bytes) -> "FeatureCatalog":
        h = C.c_void_p()
        check(lib().ps_catalog_load_binary(data, len(data), C.byref(h)))
        return FeatureCatalog(_handle=h)

    @staticmethod
    def binary_bytes(entries, octaves: int = 3, scales: int = 3, threshold: float = DEFAULT_CONTRAST_THRESHOLD,
                     descriptor_kind: int = DescriptorKind.Sift) -> bytes:
        """PSIFTCAT bytes (catalog.rs:400-445) for entries (id, width, height, Features): u8
        descriptors (pipeline.rs:1024-1030) and signatures quantised on the device from the f32
        descriptors -- a built catalog as save_binary writes it."""
        rec_t = np.dtype([("f", "<f4", (5,)), ("sig", "<u8", (2,)), ("q", "u1", (128,))])
        assert rec_t.itemsize == 164
        parts = [b"PSIFTCAT", np.array([1, octaves, scales], "<u4").tobytes(), np.float32(threshold).tobytes(),
                 bytes([descriptor_kind]), np.uint64(len(entries)).tobytes()]
        for ident, w, h, feats in entries:
            b = ident.encode("utf-8")
            hdr = np.ascontiguousarray(feats.header, FEAT_DTYPE)
            d = np.ascontiguousarray(feats.descriptors, np.float32).reshape(-1, 128)
            rec = np.zeros(len(hdr), rec_t)
            rec["f"] = hdr.view("<f4").reshape(-1, 5)
            if len(hdr):
                rec["sig"] = quantize_descriptor_u128(d)
                x = (d * np.float32(512)).astype(np.float32).astype(np.float64)   # f32 product, then exact
                r = np.sign(x) * np.floor(np.abs(x) + 0.5)                         # half away from zero
                rec["q"] = np.where(np.isnan(r), 0, np.clip(r, 0, 255)).astype(np.uint8)
            parts += [np.uint32(len(b)).tobytes(), b, np.array([w, h], "<u4").tobytes(),
                      np.uint64(len(hdr)).tobytes(), rec.tobytes()]
        return b"".join(parts)

    @staticmethod
    def load_binary(path: str) -> "FeatureCatalog":
        h = C.c_void_p()
        check(lib().ps_catalog_load_file(os.fsencode(path), C.byref(h)))
        return FeatureCatalog(_handle=h)

    @staticmethod
    def load_json(path: str) -> "FeatureCatalog":
        """catalog.rs:386-395: serde JSON; signatures recomputed from the f32 descriptors."""
        import json
        with open(path, "r", encoding="utf-8") as f:
            raw = json.load(f)
        kind = raw.get("descriptor_kind", "Sift")
        cat = FeatureCatalog(int(raw["octaves"]), int(raw["scales"]), float(raw["threshold"]),
                             DescriptorKind.RootSift if kind == "RootSift" else DescriptorKind.Sift)
        for e in raw.get("entries", []):
            feats = e.get("features", [])
            hdr = np.zeros(len(feats), FEAT_DTYPE)
            desc = np.zeros((len(feats), 128), np.float32)
            for i, ft in enumerate(feats):
                hdr[i] = (ft["x"], ft["y"], ft["scale"], ft["response"], ft["orientation"])
                d = np.asarray(ft["descriptor"], np.float32)
                desc[i, :len(d)] = d[:128]
            cat.upsert_entry(e["id"], int(e["width"]), int(e["height"]), Features(hdr, desc))
        return cat

    @staticmethod
    def load_auto(path: str) -> "FeatureCatalog":
        """catalog.rs:555-570: PSIFTCAT by magic bytes, else JSON."""
        with open(path, "rb") as f:
            head = f.read(8)
        if head == b"PSIFTCAT":
            return FeatureCatalog.load_binary(path)
        return FeatureCatalog.load_json(path)

    def save_binary(self, path: str) -> None:
        check(lib().ps_catalog_save_binary(self.h, os.fsencode(path)))

    def save_json(self, path: str, pretty: bool = False) -> None:
        """catalog.rs:374-383 (written through a temporary file and a rename)."""
        import json
        info = self._info()
        out = {"octaves": info.octaves, "scales": info.scales, "threshold": float(np.float32(info.contrast_threshold)),
               "descriptor_kind": "RootSift" if info.descriptor_kind == DescriptorKind.RootSift else "Sift",
               "entries": []}
        for i in range(len(self)):
            s = self.summary(i)
            hdr, desc, _ = self.entry_features(i)
            out["entries"].append({"id": s.id, "width": s.width, "height": s.height, "features": [
                {"x": float(h["x"]), "y": float(h["y"]), "scale": float(h["scale"]),
                 "response": float(h["response"]), "orientation": float(h["orientation"]),
                 "descriptor": [float(v) for v in d]} for h, d in zip(hdr, desc)]})
        parent = os.path.dirname(path)
        if parent:
            os.makedirs(parent, exist_ok=True)
        tmp = os.path.join(parent, f"{os.path.basename(path) or 'catalog'}.tmp.{os.getpid()}")
        with open(tmp, "w", encoding="utf-8") as f:
            json.dump(out, f, indent=2 if pretty else None)
            f.write("\n")
        os.replace(tmp, path)

    def save_auto(self, path: str, pretty: bool = False) -> None:
        """catalog.rs:543-552: JSON when the extension is .json (any case), else binary."""
        if path.lower().endswith(".json"):
            self.save_json(path, pretty)
        else:
            self.save_binary(path)

    # -- contents
    def _info(self) -> _CatalogInfoC:
        info = _CatalogInfoC()
        check(lib().ps_catalog_get_info(self.h, C.byref(info)))
        return info

    def __len__(self) -> int:
        return int(self._info().n_entries)

    @property
    def n_features(self) -> int:
        return int(self._info().n_features)

    def summary(self, idx: int) -> CatalogEntrySummary:
        e = _CatalogEntryC()
        check(lib().ps_catalog_entry(self.h, idx, C.byref(e)))
        id_ = C.string_at(e.id, e.id_len).decode("utf-8") if e.id_len else ""
        return CatalogEntrySummary(id_, int(e.width), int(e.height), int(e.feature_count))

    def list_summaries(self) -> List[CatalogEntrySummary]:
        """catalog.rs:347-357"""
        return [self.summary(i) for i in range(len(self))]

    def entry_features(self, idx: int):
        """(FEAT_DTYPE records, [n, 128] f32 descriptors, [n, 2] u64 signatures) as stored."""
        n = self.summary(idx).feature_count
        hdr = np.zeros(n, FEAT_DTYPE)
        desc = np.zeros((n, 128), np.float32)
        sig = np.zeros((n, 2), np.uint64)
        check(lib().ps_catalog_entry_features(self.h, idx, hdr.ctypes.data, desc.ctypes.data, sig.ctypes.data, n))
        return hdr, desc, sig

    def upsert_entry(self, id: str, width: int, height: int, features: Features) -> None:
        """catalog.rs:309-328"""
        b = id.encode("utf-8")
        hdr = np.ascontiguousarray(features.header, FEAT_DTYPE)
        desc = np.ascontiguousarray(features.descriptors, np.float32).reshape(-1, 128)
        check(lib().ps_catalog_upsert(self.h, b, len(b), width, height, hdr.ctypes.data, desc.ctypes.data,
                                      len(hdr)))

    def remove(self, id: str) -> bool:
        """catalog.rs:331-339"""
        b = id.encode("utf-8")
        r = C.c_int32(0)
        check(lib().ps_catalog_remove(self.h, b, len(b), C.byref(r)))
        return bool(r.value)

    # -- search (catalog.rs:576-702)
    def coarse_votes(self, query_descriptors: np.ndarray) -> np.ndarray:
        """The vote of every entry (catalog.rs:595-633), as u32 [n_entries]."""
        q = np.ascontiguousarray(query_descriptors, np.float32).reshape(-1, 128)
        out = np.zeros(max(len(self), 1), np.uint32)
        check(lib().ps_catalog_coarse_votes(self.h, q.ctypes.data, len(q), out.ctypes.data))
        return out[:len(self)]

    def search_raw(self, query: Features, top_k: int = DEFAULT_CATALOG_TOP_K,
                   shortlist_k: int = DEFAULT_CATALOG_SHORTLIST, ratio_threshold: float = DEFAULT_MATCH_RATIO,
                   ransac_threshold: float = DEFAULT_RANSAC_REPROJ_THRESHOLD) -> np.ndarray:
        """search_features as CATALOG_HIT_DTYPE records (entry = catalog index)."""
        hdr = np.ascontiguousarray(query.header, FEAT_DTYPE)
        desc = np.ascontiguousarray(query.descriptors, np.float32).reshape(-1, 128)
        cap = max(min(top_k, len(self)), 1)
        out = np.zeros(cap, CATALOG_HIT_DTYPE)
        n = C.c_size_t(0)
        check(lib().ps_catalog_search(self.h, hdr.ctypes.data, desc.ctypes.data, len(hdr), top_k, shortlist_k,
                                      ratio_threshold, ransac_threshold, out.ctypes.data, cap, C.byref(n)))
        return out[:n.value].copy()

    def search_raw_device(self, engine: Engine, d_q: int, d_q_desc: int, nq: int, top_k: int = DEFAULT_CATALOG_TOP_K,
                          shortlist_k: int = DEFAULT_CATALOG_SHORTLIST, ratio_threshold: float = DEFAULT_MATCH_RATIO,
                          ransac_threshold: float = DEFAULT_RANSAC_REPROJ_THRESHOLD) -> np.ndarray:
        """search_raw with device-resident query features on an engine's stream."""
        cap = max(min(top_k, len(self)), 1)
        out = np.zeros(cap, CATALOG_HIT_DTYPE)
        n = C.c_size_t(0)
        check(lib().ps_engine_catalog_search(engine.h, self.h, d_q, d_q_desc, nq, top_k, shortlist_k,
                                             ratio_threshold, ransac_threshold, out.ctypes.data, cap, C.byref(n)))
        return out[:n.value].copy()

    def search_features(self, query: Features, top_k: int = DEFAULT_CATALOG_TOP_K,
                        shortlist_k: int = DEFAULT_CATALOG_SHORTLIST, ratio_threshold: float = DEFAULT_MATCH_RATIO,
                        ransac_threshold: float = DEFAULT_RANSAC_REPROJ_THRESHOLD) -> List[CatalogSearchHit]:
        """catalog.rs:576-702"""
        hits = []
        for r in self.search_raw(query, top_k, shortlist_k, ratio_threshold, ransac_threshold):
            s = self.summary(int(r["entry"]))
            tr = TrackingSummary(int(r["raw_match_count"]), int(r["match_count"]), float(r["inlier_ratio"]),
                                 (float(r["mean_dx"]), float(r["mean_dy"])),
                                 (float(r["median_dx"]), float(r["median_dy"])), float(r["mean_scale_ratio"]),
                                 float(r["mean_rotation_deg"]))
            tf = SimilarityTransform2D(*[float(v) for v in r["transform"]]) if r["has_transform"] else None
            hm = r["homography"].reshape(3, 3).copy() if r["has_homography"] else None
            hits.append(CatalogSearchHit(s.id, s.width, s.height, s.feature_count, int(r["coarse_votes"]), tr, tf,
                                         hm))
        return hits
