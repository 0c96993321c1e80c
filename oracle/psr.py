"""ctypes binding of the CPU oracle (libps_oracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libps_oracle.so")

DESC_LEN = 128
MAX_FEATURES = 4096


class Keypoint(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("scale", C.c_float), ("response", C.c_float),
                ("octave", C.c_uint32), ("layer", C.c_uint32)]


class Feature(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("scale", C.c_float), ("response", C.c_float),
                ("orientation", C.c_float)]


class Match(C.Structure):
    _fields_ = [("query_index", C.c_uint64), ("target_index", C.c_uint64),
                ("qx", C.c_float), ("qy", C.c_float), ("tx", C.c_float), ("ty", C.c_float),
                ("distance", C.c_float), ("ratio", C.c_float), ("dx", C.c_float), ("dy", C.c_float)]


class Pyramid(C.Structure):
    _fields_ = [("n_octaves", C.c_int), ("scales", C.c_int), ("h", C.c_int * 8), ("w", C.c_int * 8),
                ("scale_factor", C.c_float * 8), ("gauss", C.POINTER(C.c_float) * 8),
                ("dog", C.POINTER(C.c_float) * 8)]


class Similarity(C.Structure):
    _fields_ = [("a", C.c_float), ("b", C.c_float), ("tx", C.c_float), ("ty", C.c_float),
                ("scale", C.c_float), ("rotation_deg", C.c_float)]


class RansacParams(C.Structure):
    _fields_ = [("sampler", C.c_int), ("weights", C.POINTER(C.c_float)), ("confidence", C.c_double),
                ("min_iters", C.c_int), ("max_iters", C.c_int), ("local_opt", C.c_int),
                ("sim_seed_model", C.c_int), ("adaptive", C.c_int)]


KP_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("scale", "<f4"), ("response", "<f4"),
                     ("octave", "<u4"), ("layer", "<u4")])
FEAT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("scale", "<f4"), ("response", "<f4"),
                       ("orientation", "<f4")])
MATCH_DTYPE = np.dtype([("query_index", "<u8"), ("target_index", "<u8"), ("qx", "<f4"), ("qy", "<f4"),
                        ("tx", "<f4"), ("ty", "<f4"), ("distance", "<f4"), ("ratio", "<f4"),
                        ("dx", "<f4"), ("dy", "<f4")])

class CatalogC(C.Structure):
    _fields_ = [("octaves", C.c_uint32), ("scales", C.c_uint32), ("threshold", C.c_float), ("kind", C.c_uint32),
                ("n_entries", C.c_size_t), ("n_features", C.c_size_t), ("id_blob", C.c_void_p),
                ("id_off", C.c_void_p), ("width", C.c_void_p), ("height", C.c_void_p), ("off", C.c_void_p),
                ("feat", C.c_void_p), ("sig", C.c_void_p), ("desc", C.c_void_p), ("mih", C.c_void_p)]


TRACKING_DTYPE = np.dtype([("raw_match_count", "<u8"), ("match_count", "<u8"), ("inlier_ratio", "<f4"),
                           ("mean_dx", "<f4"), ("mean_dy", "<f4"), ("median_dx", "<f4"), ("median_dy", "<f4"),
                           ("mean_scale_ratio", "<f4"), ("mean_rotation_deg", "<f4")], align=True)
HIT_DTYPE = np.dtype([("entry", "<u4"), ("coarse_votes", "<u4"), ("raw_match_count", "<u8"),
                      ("match_count", "<u8"), ("inlier_ratio", "<f4"), ("mean_dx", "<f4"), ("mean_dy", "<f4"),
                      ("median_dx", "<f4"), ("median_dy", "<f4"), ("mean_scale_ratio", "<f4"),
                      ("mean_rotation_deg", "<f4"), ("has_transform", "<u4"), ("has_homography", "<u4"),
                      ("transform", "<f4", (6,)), ("homography", "<f4", (9,))], align=True)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        L.psr_gaussian_kernel_1d.argtypes = [C.c_float, fp, C.c_int]
        L.psr_gaussian_kernel_1d.restype = C.c_int
        L.psr_gaussian_blur.argtypes = [fp, C.c_int, C.c_int, C.c_float, fp]
        L.psr_downsample_2x.argtypes = [fp, C.c_int, C.c_int, fp]
        L.psr_upsample_2x.argtypes = [fp, C.c_int, C.c_int, fp]
        L.psr_luma_to_f32.argtypes = [C.c_void_p, C.c_size_t, fp]
        L.psr_build_pyramid.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Pyramid)]
        L.psr_free_pyramid.argtypes = [C.POINTER(Pyramid)]
        L.psr_extract_extrema.argtypes = [C.POINTER(Pyramid), C.c_float, C.POINTER(C.POINTER(Keypoint))]
        L.psr_extract_extrema.restype = C.c_size_t
        L.psr_detect_keypoints.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                           C.POINTER(C.POINTER(Keypoint))]
        L.psr_detect_keypoints.restype = C.c_size_t
        L.psr_gradient_planes.argtypes = [fp, C.c_int, C.c_int, fp, fp]
        L.psr_keypoint_orientations.argtypes = [fp, fp, C.c_int, C.c_int, C.POINTER(Keypoint), fp]
        L.psr_keypoint_orientations.restype = C.c_int
        L.psr_descriptor_128.argtypes = [fp, fp, C.c_int, C.c_int, C.POINTER(Keypoint), C.c_float, fp]
        L.psr_rootsift.argtypes = [fp, fp]
        L.psr_descriptors_from_pyramid.argtypes = [C.POINTER(Pyramid), C.POINTER(Keypoint), C.c_size_t, C.c_int,
                                                   C.POINTER(C.POINTER(Feature)), C.POINTER(fp)]
        L.psr_descriptors_from_pyramid.restype = C.c_size_t
        L.psr_compute_descriptors.argtypes = [fp, C.c_int, C.c_int, C.POINTER(Keypoint), C.c_size_t, C.c_int,
                                              C.POINTER(C.POINTER(Feature)), C.POINTER(C.POINTER(C.c_float))]
        L.psr_compute_descriptors.restype = C.c_size_t
        L.psr_extract_features.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                           C.POINTER(C.POINTER(Feature)), C.POINTER(fp)]
        L.psr_extract_features.restype = C.c_size_t
        L.psr_match_descriptors.argtypes = [C.POINTER(Feature), fp, C.c_size_t, C.POINTER(Feature), fp, C.c_size_t,
                                            C.c_float, C.c_int, C.POINTER(Match)]
        L.psr_match_descriptors.restype = C.c_size_t
        L.psr_nn2.argtypes = [fp, C.c_size_t, fp, C.c_size_t, C.POINTER(C.c_uint32), fp, fp]
        L.psr_quantize_u128.argtypes = [fp, C.c_size_t, C.POINTER(C.c_uint64)]
        L.psr_count_coarse_hamming_votes.argtypes = [C.POINTER(C.c_uint64), C.c_size_t, C.POINTER(C.c_uint64),
                                                     C.c_size_t]
        L.psr_count_coarse_hamming_votes.restype = C.c_size_t
        L.psr_hamming_topk.argtypes = [C.POINTER(C.c_uint64), C.c_size_t, C.POINTER(C.c_uint64), C.c_size_t,
                                       C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint8)]
        L.psr_rerank_l2.argtypes = [C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_float), C.c_size_t,
                                    C.POINTER(C.c_uint32), C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
        L.psr_mlp_forward.argtypes = [C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float), C.c_size_t,
                                      C.POINTER(C.c_float)]
        L.psr_mlp_loss_grad.argtypes = [C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                        C.c_size_t, C.POINTER(C.c_double)]
        L.psr_mlp_loss_grad.restype = C.c_double
        L.psr_splitmix64_next.argtypes = [C.POINTER(C.c_uint64)]
        L.psr_splitmix64_next.restype = C.c_uint64
        L.psr_filter_matches_ransac.argtypes = [C.POINTER(Match), C.c_size_t, C.c_float, C.c_uint64,
                                                C.POINTER(C.c_uint32), C.POINTER(Similarity), C.POINTER(C.c_int)]
        L.psr_filter_matches_ransac.restype = C.c_size_t
        L.psr_homography_ransac.argtypes = [C.POINTER(Match), C.c_size_t, C.c_float, C.c_uint64, C.c_uint64,
                                            C.POINTER(RansacParams), C.POINTER(C.c_uint32), fp, C.POINTER(C.c_int),
                                            C.POINTER(C.c_uint32)]
        L.psr_homography_ransac.restype = C.c_size_t
        L.psr_homography_ransac_ref.argtypes = [C.POINTER(Match), C.c_size_t, C.c_float, C.c_uint64, C.c_int,
                                                C.POINTER(C.c_uint32), fp, C.POINTER(C.c_int)]
        L.psr_homography_ransac_ref.restype = C.c_size_t
        L.psr_fit_homography_dlt.argtypes = [C.POINTER(Match), C.c_size_t, C.POINTER(C.c_double)]
        L.psr_fit_homography_dlt.restype = C.c_int
        L.psr_adaptive_ransac_iters.argtypes = [C.c_size_t, C.c_size_t, C.c_int, C.c_double, C.c_size_t, C.c_size_t]
        L.psr_adaptive_ransac_iters.restype = C.c_size_t
        L.psr_quantize_u8.argtypes = [fp, C.c_size_t, C.c_void_p]
        L.psr_dequantize_u8.argtypes = [C.c_void_p, C.c_size_t, fp]
        L.psr_catalog_parse.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(CatalogC), C.c_char_p, C.c_size_t]
        L.psr_catalog_parse.restype = C.c_int
        L.psr_catalog_free.argtypes = [C.POINTER(CatalogC)]
        L.psr_catalog_index.argtypes = [C.POINTER(CatalogC)]
        L.psr_catalog_index.restype = C.c_void_p
        L.psr_catalog_index_free.argtypes = [C.c_void_p]
        L.psr_catalog_write.argtypes = [C.POINTER(CatalogC), C.c_void_p, C.c_size_t]
        L.psr_catalog_write.restype = C.c_size_t
        L.psr_catalog_votes.argtypes = [C.POINTER(CatalogC), C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        L.psr_catalog_shortlist.argtypes = [C.POINTER(CatalogC), C.c_void_p, C.c_int, C.c_size_t, C.c_void_p,
                                            C.c_void_p]
        L.psr_catalog_shortlist.restype = C.c_size_t
        L.psr_summarize_tracking.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        L.psr_catalog_search.argtypes = [C.POINTER(CatalogC), C.c_void_p, fp, C.c_size_t, C.c_size_t, C.c_size_t,
                                         C.c_float, C.c_float, C.c_void_p]
        L.psr_catalog_search.restype = C.c_size_t
        dp = C.POINTER(C.c_double)
        L.psr_jacobi9.argtypes = [dp, dp, dp]
        L.psr_jacobi9.restype = C.c_int
        L.psr_sampson_sq.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.psr_sampson_sq.restype = C.c_double
        L.psr_fit_fundamental_8point.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, dp]
        L.psr_fit_fundamental_8point.restype = C.c_int
        L.psr_fit_fundamental_7point.argtypes = [C.c_void_p, dp]
        L.psr_fit_fundamental_7point.restype = C.c_int
        L.psr_adaptive_epipolar_iters.argtypes = [C.c_size_t, C.c_size_t, C.c_int, C.c_double, C.c_size_t,
                                                  C.c_size_t]
        L.psr_adaptive_epipolar_iters.restype = C.c_size_t
        L.psr_fundamental_ransac.argtypes = [C.c_void_p, C.c_size_t, C.c_double, C.c_size_t, C.c_double,
                                             C.c_uint64, C.c_void_p, dp, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]
        L.psr_fundamental_ransac.restype = C.c_size_t
        L.psr_invert3_f32.argtypes = [C.c_void_p, dp]
        L.psr_invert3_f32.restype = C.c_int
        L.psr_select_model_gric.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_double,
                                            C.POINTER(C.c_int), dp, dp]
        L.psr_select_model_gric.restype = C.c_int
        L.psr_is_sane_homography.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t]
        L.psr_is_sane_homography.restype = C.c_int
        L.psr_resolve_safe_alignment.argtypes = [C.c_int, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_size_t,
                                                 C.c_size_t, C.c_void_p]
        L.psr_resolve_safe_alignment.restype = C.c_int
        L.psr_sample_patch_intensity.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_float, C.c_float]
        L.psr_sample_patch_intensity.restype = C.c_double
        L.psr_set_threads.argtypes = [C.c_int]
        L.psr_get_threads.restype = C.c_int
        L.psr_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def set_threads(n: int) -> None:
    lib().psr_set_threads(int(n))


def get_threads() -> int:
    return lib().psr_get_threads()


def gaussian_kernel_1d(sigma: float) -> np.ndarray:
    out = np.zeros(1024, np.float32)
    n = lib().psr_gaussian_kernel_1d(C.c_float(sigma), _fp(out), 1024)
    assert n > 0
    return out[:n].copy()


def gaussian_blur(img: np.ndarray, sigma: float) -> np.ndarray:
    img = _f32(img)
    out = np.zeros_like(img)
    lib().psr_gaussian_blur(_fp(img), img.shape[0], img.shape[1], C.c_float(sigma), _fp(out))
    return out


def upsample_2x(img: np.ndarray) -> np.ndarray:
    img = _f32(img)
    out = np.zeros((img.shape[0] * 2, img.shape[1] * 2), np.float32)
    lib().psr_upsample_2x(_fp(img), img.shape[0], img.shape[1], _fp(out))
    return out


def downsample_2x(img: np.ndarray) -> np.ndarray:
    img = _f32(img)
    out = np.zeros((img.shape[0] // 2, img.shape[1] // 2), np.float32)
    if out.size:
        lib().psr_downsample_2x(_fp(img), img.shape[0], img.shape[1], _fp(out))
    return out


def luma_to_f32(img_u8: np.ndarray) -> np.ndarray:
    img_u8 = np.ascontiguousarray(img_u8, dtype=np.uint8)
    out = np.zeros(img_u8.shape, np.float32)
    lib().psr_luma_to_f32(img_u8.ctypes.data, img_u8.size, _fp(out))
    return out


def build_pyramid(img: np.ndarray, octaves: int, scales: int, upsample: bool):
    """Returns a list of octaves: dict(gauss=[S+3,h,w], dog=[S+2,h,w], scale_factor)."""
    img = _f32(img)
    p = Pyramid()
    lib().psr_build_pyramid(_fp(img), img.shape[0], img.shape[1], octaves, scales, int(upsample), C.byref(p))
    res = []
    for o in range(p.n_octaves):
        h, w = p.h[o], p.w[o]
        g = np.ctypeslib.as_array(p.gauss[o], shape=((scales + 3) * h * w,)).reshape(scales + 3, h, w).copy()
        d = np.ctypeslib.as_array(p.dog[o], shape=((scales + 2) * h * w,)).reshape(scales + 2, h, w).copy()
        res.append(dict(gauss=g, dog=d, scale_factor=p.scale_factor[o]))
    lib().psr_free_pyramid(C.byref(p))
    return res


def _take_kps(ptr, n):
    if n == 0:
        lib().psr_free(ptr)
        return np.zeros(0, KP_DTYPE)
    arr = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)), shape=(n * KP_DTYPE.itemsize,))
    out = arr.view(KP_DTYPE).copy()
    lib().psr_free(ptr)
    return out


def detect_keypoints(img, octaves=3, scales=3, threshold=0.015, upsample=True) -> np.ndarray:
    img = _f32(img)
    ptr = C.POINTER(Keypoint)()
    n = lib().psr_detect_keypoints(_fp(img), img.shape[0], img.shape[1], octaves, scales, C.c_float(threshold),
                                   int(upsample), C.byref(ptr))
    return _take_kps(ptr, n)


def gradient_planes(g: np.ndarray):
    g = _f32(g)
    mag = np.zeros_like(g)
    ang = np.zeros_like(g)
    lib().psr_gradient_planes(_fp(g), g.shape[0], g.shape[1], _fp(mag), _fp(ang))
    return mag, ang


def keypoint_orientations(mag, ang, kp) -> np.ndarray:
    mag, ang = _f32(mag), _f32(ang)
    k = Keypoint(*[kp[f] for f in KP_DTYPE.names])
    out = np.zeros(4, np.float32)
    n = lib().psr_keypoint_orientations(_fp(mag), _fp(ang), mag.shape[1], mag.shape[0], C.byref(k), _fp(out))
    return out[:n].copy()


def descriptor_128(mag, ang, kp, orientation: float) -> np.ndarray:
    mag, ang = _f32(mag), _f32(ang)
    k = Keypoint(*[kp[f] for f in KP_DTYPE.names])
    out = np.zeros(128, np.float32)
    lib().psr_descriptor_128(_fp(mag), _fp(ang), mag.shape[1], mag.shape[0], C.byref(k), C.c_float(orientation),
                             _fp(out))
    return out


def _take_feats(fptr, dptr, n):
    if n == 0:
        lib().psr_free(fptr)
        lib().psr_free(dptr)
        return np.zeros(0, FEAT_DTYPE), np.zeros((0, 128), np.float32)
    f = np.ctypeslib.as_array(C.cast(fptr, C.POINTER(C.c_uint8)), shape=(n * FEAT_DTYPE.itemsize,))
    feats = f.view(FEAT_DTYPE).copy()
    descs = np.ctypeslib.as_array(dptr, shape=(n * 128,)).reshape(n, 128).copy()
    lib().psr_free(fptr)
    lib().psr_free(dptr)
    return feats, descs


def extract_features(img, octaves=3, scales=3, threshold=0.015, upsample=True, kind=0):
    """extract_sift_features_with_kind: returns (features[FEAT_DTYPE], descriptors[n,128])."""
    img = _f32(img)
    fptr = C.POINTER(Feature)()
    dptr = C.POINTER(C.c_float)()
    n = lib().psr_extract_features(_fp(img), img.shape[0], img.shape[1], octaves, scales, C.c_float(threshold),
                                   int(upsample), int(kind), C.byref(fptr), C.byref(dptr))
    return _take_feats(fptr, dptr, n)


def descriptors_for_keypoints(img, kps, octaves, scales, upsample, kind=0):
    """compute_sift_descriptors_from_pyramid_with_kind on the pyramid of img (no truncation)."""
    img = _f32(img)
    p = Pyramid()
    lib().psr_build_pyramid(_fp(img), img.shape[0], img.shape[1], octaves, scales, int(upsample), C.byref(p))
    kps = np.ascontiguousarray(kps, KP_DTYPE)
    fptr = C.POINTER(Feature)()
    dptr = C.POINTER(C.c_float)()
    n = lib().psr_descriptors_from_pyramid(C.byref(p), kps.ctypes.data_as(C.POINTER(Keypoint)), len(kps), int(kind),
                                           C.byref(fptr), C.byref(dptr))
    lib().psr_free_pyramid(C.byref(p))
    return _take_feats(fptr, dptr, n)


def compute_descriptors(img, kps, kind=0):
    """compute_sift_descriptors_with_kind (pipeline.rs:1581-1620): pyramid parameters inferred
    from the keypoints, features for every keypoint in order."""
    img = _f32(img)
    kps = np.ascontiguousarray(kps, KP_DTYPE)
    fptr = C.POINTER(Feature)()
    dptr = C.POINTER(C.c_float)()
    n = lib().psr_compute_descriptors(_fp(img), img.shape[0], img.shape[1], kps.ctypes.data_as(C.POINTER(Keypoint)),
                                      len(kps), int(kind), C.byref(fptr), C.byref(dptr))
    return _take_feats(fptr, dptr, n)


def match(qf, qd, tf, td, ratio=0.75, mutual=False) -> np.ndarray:
    qf = np.ascontiguousarray(qf, FEAT_DTYPE)
    tf = np.ascontiguousarray(tf, FEAT_DTYPE)
    qd, td = _f32(qd), _f32(td)
    out = np.zeros(max(len(qf), 1), MATCH_DTYPE)
    n = lib().psr_match_descriptors(qf.ctypes.data_as(C.POINTER(Feature)), _fp(qd), len(qf),
                                    tf.ctypes.data_as(C.POINTER(Feature)), _fp(td), len(tf),
                                    C.c_float(ratio), int(mutual), out.ctypes.data_as(C.POINTER(Match)))
    return out[:n].copy()


def nn2(qd, td):
    qd, td = _f32(qd), _f32(td)
    nq = len(qd)
    bi = np.zeros(nq, np.uint32)
    b = np.zeros(nq, np.float32)
    s = np.zeros(nq, np.float32)
    lib().psr_nn2(_fp(qd), nq, _fp(td), len(td), bi.ctypes.data_as(C.POINTER(C.c_uint32)), _fp(b), _fp(s))
    return bi, b, s


def quantize_u128(desc) -> np.ndarray:
    desc = _f32(desc).reshape(-1, 128)
    out = np.zeros((len(desc), 2), np.uint64)
    lib().psr_quantize_u128(_fp(desc), len(desc), out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def count_coarse_hamming_votes(q, e) -> int:
    q = np.ascontiguousarray(q, np.uint64)
    e = np.ascontiguousarray(e, np.uint64)
    return lib().psr_count_coarse_hamming_votes(q.ctypes.data_as(C.POINTER(C.c_uint64)), len(q),
                                                e.ctypes.data_as(C.POINTER(C.c_uint64)), len(e))


def hamming_topk(catalog, queries, k=100):
    catalog = np.ascontiguousarray(catalog, np.uint64)
    queries = np.ascontiguousarray(queries, np.uint64)
    nq = len(queries)
    idx = np.zeros((nq, k), np.uint32)
    dist = np.zeros((nq, k), np.uint8)
    lib().psr_hamming_topk(catalog.ctypes.data_as(C.POINTER(C.c_uint64)), len(catalog),
                           queries.ctypes.data_as(C.POINTER(C.c_uint64)), nq, k,
                           idx.ctypes.data_as(C.POINTER(C.c_uint32)), dist.ctypes.data_as(C.POINTER(C.c_uint8)))
    return idx, dist


def rerank_l2(qd, cd, shortlist):
    """FP32 rerank of a shortlist (nq x k catalog indices) by (sequential squared L2, index)."""
    qd = _f32(qd).reshape(-1, 128)
    cd = _f32(cd).reshape(-1, 128)
    sl = np.ascontiguousarray(shortlist, np.uint32)
    nq, k = sl.shape
    idx = np.zeros((nq, k), np.uint32)
    d2 = np.zeros((nq, k), np.float32)
    lib().psr_rerank_l2(_fp(qd), nq, _fp(cd), len(cd), sl.ctypes.data_as(C.POINTER(C.c_uint32)), k,
                        idx.ctypes.data_as(C.POINTER(C.c_uint32)), _fp(d2))
    return idx, d2


MLP_PARAMS = 4801


def mlp_forward(params, x, bf16=True):
    """Logits of the 8-64-64-1 match-scoring MLP (f64 arithmetic, BF16-rounded weights)."""
    p = _f32(params)
    x = _f32(x).reshape(-1, 8)
    out = np.zeros(len(x), np.float32)
    lib().psr_mlp_forward(_fp(p), int(bf16), _fp(x), len(x), _fp(out))
    return out


def mlp_loss_grad(params, x, y, bf16=True):
    """(mean BCE-with-logits loss, gradient [4801] f64) of the match-scoring MLP."""
    p = _f32(params)
    x = _f32(x).reshape(-1, 8)
    y = _f32(y)
    g = np.zeros(MLP_PARAMS, np.float64)
    loss = lib().psr_mlp_loss_grad(_fp(p), int(bf16), _fp(x), _fp(y), len(x), g.ctypes.data_as(C.POINTER(C.c_double)))
    return loss, g


def mlp_features(m, qfeat, tfeat, width, height):
    """Match-scoring MLP inputs of one pair (float64 restatement of their definition):
    ratio (NaN -> 1, capped at 2), log(scale_t / scale_q) (scales floored at 1e-6), sin and
    cos of (orientation_t - orientation_q), distance, qx / width, qy / height, and the rank of
    the match by (ratio, index) within the pair divided by the match count."""
    m = np.asarray(m)
    n = len(m)
    r = np.where(np.isnan(m["ratio"]), 1.0, m["ratio"].astype(np.float64))
    order = np.lexsort((np.arange(n), r))
    rank = np.empty(n, np.float64)
    rank[order] = np.arange(n)
    qi = m["query_index"].astype(np.int64)
    ti = m["target_index"].astype(np.int64)
    sq = np.maximum(qfeat["scale"][qi].astype(np.float64), 1e-6)
    st = np.maximum(tfeat["scale"][ti].astype(np.float64), 1e-6)
    dth = tfeat["orientation"][ti].astype(np.float64) - qfeat["orientation"][qi].astype(np.float64)
    x = np.stack([np.minimum(r, 2.0), np.log(st / sq), np.sin(dth), np.cos(dth), m["distance"].astype(np.float64),
                  m["qx"] / width, m["qy"] / height, rank / max(n, 1)], 1)
    return x


def splitmix64(state: int, n: int):
    st = C.c_uint64(state)
    return [lib().psr_splitmix64_next(C.byref(st)) for _ in range(n)]


def make_matches(qpts, tpts, ratio=0.5, distance=0.0) -> np.ndarray:
    qpts = np.asarray(qpts, np.float32).reshape(-1, 2)
    tpts = np.asarray(tpts, np.float32).reshape(-1, 2)
    n = len(qpts)
    m = np.zeros(n, MATCH_DTYPE)
    m["query_index"] = np.arange(n)
    m["target_index"] = np.arange(n)
    m["qx"], m["qy"] = qpts[:, 0], qpts[:, 1]
    m["tx"], m["ty"] = tpts[:, 0], tpts[:, 1]
    m["distance"] = distance
    m["ratio"] = ratio
    m["dx"] = tpts[:, 0] - qpts[:, 0]
    m["dy"] = tpts[:, 1] - qpts[:, 1]
    return m


def filter_matches_ransac(m, thresh=3.0, seed=0xD1B54A32D192ED03):
    m = np.ascontiguousarray(m, MATCH_DTYPE)
    inl = np.zeros(max(len(m), 1), np.uint32)
    model = Similarity()
    has = C.c_int(0)
    n = lib().psr_filter_matches_ransac(m.ctypes.data_as(C.POINTER(Match)), len(m), C.c_float(thresh),
                                        C.c_uint64(seed), inl.ctypes.data_as(C.POINTER(C.c_uint32)),
                                        C.byref(model), C.byref(has))
    mod = dict(a=model.a, b=model.b, tx=model.tx, ty=model.ty, scale=model.scale,
               rotation_deg=model.rotation_deg) if has.value else None
    return inl[:n].copy(), mod


def homography_ransac_ref(m, thresh=3.0, seed=None):
    m = np.ascontiguousarray(m, MATCH_DTYPE)
    inl = np.zeros(max(len(m), 1), np.uint32)
    h = np.zeros(9, np.float32)
    has = C.c_int(0)
    n = lib().psr_homography_ransac_ref(m.ctypes.data_as(C.POINTER(Match)), len(m), C.c_float(thresh),
                                        C.c_uint64(0 if seed is None else seed), int(seed is None),
                                        inl.ctypes.data_as(C.POINTER(C.c_uint32)), _fp(h), C.byref(has))
    return inl[:n].copy(), (h.reshape(3, 3).copy() if has.value else None)


def homography_ransac(m, thresh=3.0, sim_seed=0, hom_seed=0, sampler=1, weights=None, confidence=0.99,
                      min_iters=1, max_iters=10000, local_opt=False, sim_seed_model=False, adaptive=True):
    """Generalised RANSAC: sampler 0 = reference 1/ratio CDF, 1 = uniform, 2 = weights (guided)."""
    m = np.ascontiguousarray(m, MATCH_DTYPE)
    w = None
    prm = RansacParams(sampler, None, confidence, min_iters, max_iters, int(local_opt), int(sim_seed_model),
                       int(adaptive))
    if sampler == 2:
        w = _f32(weights)
        prm.weights = _fp(w)
    inl = np.zeros(max(len(m), 1), np.uint32)
    h = np.zeros(9, np.float32)
    has = C.c_int(0)
    hyps = C.c_uint32(0)
    n = lib().psr_homography_ransac(m.ctypes.data_as(C.POINTER(Match)), len(m), C.c_float(thresh),
                                    C.c_uint64(sim_seed), C.c_uint64(hom_seed), C.byref(prm),
                                    inl.ctypes.data_as(C.POINTER(C.c_uint32)), _fp(h), C.byref(has), C.byref(hyps))
    return inl[:n].copy(), (h.reshape(3, 3).copy() if has.value else None), hyps.value


def fit_homography_dlt(m):
    m = np.ascontiguousarray(m, MATCH_DTYPE)
    h = np.zeros(9, np.float64)
    ok = lib().psr_fit_homography_dlt(m.ctypes.data_as(C.POINTER(Match)), len(m),
                                      h.ctypes.data_as(C.POINTER(C.c_double)))
    return h.reshape(3, 3) if ok else None


def adaptive_ransac_iters(inl, total, sample_size, conf, min_iters, max_cap) -> int:
    return lib().psr_adaptive_ransac_iters(inl, total, sample_size, conf, min_iters, max_cap)


# ---------------------------------------------------------------- feature catalog

def quantize_u8(desc) -> np.ndarray:
    desc = _f32(desc).reshape(-1, 128)
    out = np.zeros(desc.shape, np.uint8)
    lib().psr_quantize_u8(_fp(desc), len(desc), out.ctypes.data)
    return out


def dequantize_u8(q) -> np.ndarray:
    q = np.ascontiguousarray(q, np.uint8).reshape(-1, 128)
    out = np.zeros(q.shape, np.float32)
    lib().psr_dequantize_u8(q.ctypes.data, len(q), _fp(out))
    return out


class Catalog:
    """A catalog as flat arrays (the oracle's psr_catalog view): entry e owns the features
    off[e]:off[e+1] of feat (FEAT_DTYPE), sig (n x 2 u64) and desc (n x 128 f32)."""

    def __init__(self, ids, width, height, off, feat, sig, desc, octaves=3, scales=3, threshold=0.015, kind=0):
        self.ids = list(ids)
        self.width = np.ascontiguousarray(width, np.uint32)
        self.height = np.ascontiguousarray(height, np.uint32)
        self.off = np.ascontiguousarray(off, np.uint64)
        self.feat = np.ascontiguousarray(feat, FEAT_DTYPE)
        self.sig = np.ascontiguousarray(sig, np.uint64).reshape(-1, 2)
        self.desc = _f32(desc).reshape(-1, 128)
        self.octaves, self.scales, self.threshold, self.kind = octaves, scales, threshold, kind
        blobs = [i.encode("utf-8") if isinstance(i, str) else bytes(i) for i in self.ids]
        self._blob = np.frombuffer(b"".join(blobs) or b"\0", np.uint8).copy()
        self._id_off = np.zeros(len(blobs) + 1, np.uint64)
        self._id_off[1:] = np.cumsum([len(b) for b in blobs]) if blobs else []
        self._mih = None

    def build_index(self) -> None:
        """Builds the inverted index once (catalog.rs:280-297), as the reference keeps it."""
        if self._mih is None:
            cc = self.c()
            self._mih = lib().psr_catalog_index(C.byref(cc))

    def __del__(self):
        if getattr(self, "_mih", None) and _lib is not None:
            _lib.psr_catalog_index_free(self._mih)
            self._mih = None

    @staticmethod
    def from_entries(entries, octaves=3, scales=3, threshold=0.015, kind=0, sigs=None):
        """entries: list of (id, width, height, feat, desc); signatures quantised from the
        descriptors (catalog.rs:124-132) unless given."""
        counts = [len(e[3]) for e in entries]
        off = np.zeros(len(entries) + 1, np.uint64)
        off[1:] = np.cumsum(counts) if entries else []
        feat = np.concatenate([np.asarray(e[3], FEAT_DTYPE) for e in entries]) if entries else np.zeros(0, FEAT_DTYPE)
        desc = (np.concatenate([_f32(e[4]).reshape(-1, 128) for e in entries]) if entries
                else np.zeros((0, 128), np.float32))
        sig = quantize_u128(desc) if sigs is None else sigs
        return Catalog([e[0] for e in entries], [e[1] for e in entries], [e[2] for e in entries], off, feat,
                       sig, desc, octaves, scales, threshold, kind)

    def c(self) -> CatalogC:
        return CatalogC(self.octaves, self.scales, self.threshold, self.kind, len(self.ids), len(self.feat),
                        self._blob.ctypes.data, self._id_off.ctypes.data, self.width.ctypes.data,
                        self.height.ctypes.data, self.off.ctypes.data, self.feat.ctypes.data,
                        self.sig.ctypes.data, self.desc.ctypes.data, self._mih)


def catalog_parse(data: bytes) -> Catalog:
    """catalog.rs:447-540 load_binary; raises ValueError with the reference's message."""
    cc = CatalogC()
    err = C.create_string_buffer(256)
    if lib().psr_catalog_parse(data, len(data), C.byref(cc), err, 256) != 0:
        raise ValueError(err.value.decode())
    try:
        ne, nf = cc.n_entries, cc.n_features

        def arr(ptr, dtype, count):
            if count == 0:
                return np.zeros(0, dtype)
            return np.frombuffer((C.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr), dtype).copy()

        id_off = arr(cc.id_off, np.uint64, ne + 1)
        blob = arr(cc.id_blob, np.uint8, int(id_off[-1])).tobytes()
        ids = [blob[id_off[i]:id_off[i + 1]].decode("utf-8") for i in range(ne)]
        return Catalog(ids, arr(cc.width, np.uint32, ne), arr(cc.height, np.uint32, ne), arr(cc.off, np.uint64, ne + 1),
                       arr(cc.feat, FEAT_DTYPE, nf), arr(cc.sig, np.uint64, 2 * nf).reshape(-1, 2),
                       arr(cc.desc, np.float32, 128 * nf).reshape(-1, 128), cc.octaves, cc.scales, cc.threshold,
                       cc.kind)
    finally:
        lib().psr_catalog_free(C.byref(cc))


def catalog_write(cat: Catalog) -> bytes:
    """catalog.rs:400-445 save_binary (u8-quantised descriptors, stored signatures)."""
    cc = cat.c()
    need = lib().psr_catalog_write(C.byref(cc), None, 0)
    buf = np.zeros(need, np.uint8)
    lib().psr_catalog_write(C.byref(cc), buf.ctypes.data, need)
    return buf.tobytes()


def catalog_votes(cat: Catalog, q_codes, mih: bool) -> np.ndarray:
    q = np.ascontiguousarray(q_codes, np.uint64).reshape(-1, 2)
    votes = np.zeros(max(len(cat.ids), 1), np.uint32)
    cc = cat.c()
    lib().psr_catalog_votes(C.byref(cc), q.ctypes.data, len(q), int(mih), votes.ctypes.data)
    return votes[:len(cat.ids)]


def catalog_shortlist(cat: Catalog, votes, mih: bool, shortlist_k: int):
    v = np.ascontiguousarray(votes, np.uint32)
    ce = np.zeros(max(len(cat.ids), 1), np.uint32)
    cv = np.zeros(max(len(cat.ids), 1), np.uint32)
    cc = cat.c()
    n = lib().psr_catalog_shortlist(C.byref(cc), v.ctypes.data, int(mih), shortlist_k, ce.ctypes.data,
                                    cv.ctypes.data)
    return ce[:n].copy(), cv[:n].copy()


def summarize_tracking(qf, tf, m):
    qf = np.ascontiguousarray(qf, FEAT_DTYPE)
    tf = np.ascontiguousarray(tf, FEAT_DTYPE)
    m = np.ascontiguousarray(m, MATCH_DTYPE)
    out = np.zeros(1, TRACKING_DTYPE)
    lib().psr_summarize_tracking(qf.ctypes.data, tf.ctypes.data, m.ctypes.data, len(m), out.ctypes.data)
    return out[0]


def catalog_search(cat: Catalog, qf, qd, top_k=5, shortlist_k=16, ratio=0.75, ransac_threshold=3.0):
    """catalog.rs:576-702 search_features -> HIT_DTYPE records (entry = catalog index)."""
    qf = np.ascontiguousarray(qf, FEAT_DTYPE)
    qd = _f32(qd).reshape(-1, 128)
    out = np.zeros(max(min(top_k, len(cat.ids)), 1), HIT_DTYPE)
    cc = cat.c()
    n = lib().psr_catalog_search(C.byref(cc), qf.ctypes.data, _fp(qd), len(qf), top_k, shortlist_k,
                                 C.c_float(ratio), C.c_float(ransac_threshold), out.ctypes.data)
    return out[:n].copy()


# ---------------------------------------------------------------- two-view geometry (geometry.rs)

FUNDAMENTAL_DEFAULT_SEED = 0xF00DCAFE12345678   # geometry.rs:125


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def jacobi9(a):
    """geometry.rs:211-295: (eigenvalues ascending, eigenvectors as columns) or None."""
    a = np.ascontiguousarray(a, np.float64).reshape(9, 9)
    ev = np.zeros(9)
    vec = np.zeros((9, 9))
    ok = lib().psr_jacobi9(_dp(a), _dp(ev), _dp(vec))
    return (ev, vec) if ok else None


def sampson_sq(f, p1, p2) -> float:
    f = np.ascontiguousarray(f, np.float64).reshape(9)
    return lib().psr_sampson_sq(_dp(f), float(p1[0]), float(p1[1]), float(p2[0]), float(p2[1]))


def fit_fundamental_8point(m, idx=None):
    m = np.ascontiguousarray(m, MATCH_DTYPE)
    f = np.zeros(9)
    if idx is None:
        ok = lib().psr_fit_fundamental_8point(m.ctypes.data, None, len(m), _dp(f))
    else:
        ix = np.ascontiguousarray(idx, np.uint32)
        ok = lib().psr_fit_fundamental_8point(m.ctypes.data, ix.ctypes.data, len(ix), _dp(f))
    return f.reshape(3, 3) if ok else None


def fit_fundamental_7point(m):
    m = np.ascontiguousarray(m, MATCH_DTYPE)[:7]
    out = np.zeros((4, 9))
    k = lib().psr_fit_fundamental_7point(m.ctypes.data, _dp(out))
    return [out[i].reshape(3, 3).copy() for i in range(k)]


def adaptive_epipolar_iters(best, total, sample_size, conf, min_iters, max_iters) -> int:
    return lib().psr_adaptive_epipolar_iters(best, total, sample_size, conf, min_iters, max_iters)


def fundamental_ransac(m, sampson_threshold=1.5, max_iters=2000, confidence=0.999, seed=FUNDAMENTAL_DEFAULT_SEED):
    """geometry.rs:868-1016: (inlier indices, 3x3 F or None, iterations)."""
    m = np.ascontiguousarray(m, MATCH_DTYPE)
    inl = np.zeros(max(len(m), 1), np.uint32)
    f = np.zeros(9)
    has = C.c_int(0)
    it = C.c_size_t(0)
    n = lib().psr_fundamental_ransac(m.ctypes.data, len(m), sampson_threshold, max_iters, confidence,
                                     C.c_uint64(seed), inl.ctypes.data, _dp(f), C.byref(has), C.byref(it))
    return inl[:n].copy(), (f.reshape(3, 3).copy() if has.value else None), it.value


def invert3_f32(h):
    h = np.ascontiguousarray(h, np.float32).reshape(9)
    out = np.zeros(9)
    return out.reshape(3, 3) if lib().psr_invert3_f32(h.ctypes.data, _dp(out)) else None


def select_model_gric(m, h=None, f=None, sigma=1.0):
    """geometry.rs:1233-1341: None or (selected: 0 homography / 1 fundamental, h_gric, f_gric)."""
    m = np.ascontiguousarray(m, MATCH_DTYPE)
    hh = None if h is None else np.ascontiguousarray(h, np.float32).reshape(9)
    ff = None if f is None else np.ascontiguousarray(f, np.float64).reshape(9)
    sel = C.c_int(0)
    hg = C.c_double(0)
    fg = C.c_double(0)
    ok = lib().psr_select_model_gric(m.ctypes.data, len(m), None if hh is None else hh.ctypes.data,
                                     None if ff is None else ff.ctypes.data, sigma, C.byref(sel), C.byref(hg),
                                     C.byref(fg))
    return (sel.value, hg.value, fg.value) if ok else None


def pair_geometry(qf, tf, m, reproj_threshold=3.0, seed=None):
    """cli.rs:290-318 post-processing of one pair's raw matches (ransac enabled), restated
    with the oracle's functions: a dict of the tracking summary, similarity, homography,
    fundamental (8+ matches), GRIC and the kept inliers."""
    m = np.ascontiguousarray(m, MATCH_DTYPE)
    tr = summarize_tracking(qf, tf, m)
    sim_seed = 0xD1B54A32D192ED03 if seed is None else seed
    sim_inl, sim = filter_matches_ransac(m, reproj_threshold, sim_seed)
    hom_inl, H = homography_ransac_ref(m, reproj_threshold, seed)
    F, f_inl = None, None
    if len(m) >= 8:
        f_inl, F, _ = fundamental_ransac(m, min(float(np.float32(reproj_threshold)), 3.0))
    gric = select_model_gric(m, H, F, max(float(np.float32(reproj_threshold)), 1.0))
    best = hom_inl if len(hom_inl) > len(sim_inl) else sim_inl
    return dict(tracking=tr, sim_inliers=sim_inl, similarity=sim, hom_inliers=hom_inl, homography=H,
                fundamental=F, n_fundamental=None if F is None else len(f_inl), gric=gric, best=best,
                best_is_homography=len(hom_inl) > len(sim_inl))


# ---- panorama pairwise stage (stitch.rs:478-571) -------------------------------------------

def is_sane_homography(h, w: int, hgt: int) -> bool:
    """visualize.rs:303-355 is_sane_homography_for_dims."""
    h = np.ascontiguousarray(h, np.float32).reshape(9)
    return bool(lib().psr_is_sane_homography(h.ctypes.data, w, hgt))


def resolve_safe_alignment(h, hom_inliers: int, sim, w: int, hgt: int):
    """visualize.rs:361-374 resolve_safe_alignment_homography: h (3x3 f32 or None), sim
    (dict with a, b, tx, ty, or None); returns the 3x3 f32 alignment or None."""
    hh = np.zeros(9, np.float32) if h is None else np.ascontiguousarray(h, np.float32).reshape(9)
    ss = np.zeros(4, np.float32) if sim is None else np.array([sim["a"], sim["b"], sim["tx"], sim["ty"]], np.float32)
    out = np.zeros(9, np.float32)
    ok = lib().psr_resolve_safe_alignment(int(h is not None), hh.ctypes.data, hom_inliers, int(sim is not None),
                                          ss.ctypes.data, w, hgt, out.ctypes.data)
    return out.reshape(3, 3) if ok else None


def sample_patch_intensity(gray, x: float, y: float) -> float:
    """stitch.rs:203-226 over an f32 luma plane."""
    gray = np.ascontiguousarray(gray, np.float32)
    return lib().psr_sample_patch_intensity(gray.ctypes.data, gray.shape[0], gray.shape[1], C.c_float(x),
                                            C.c_float(y))


def panorama_pair(gray_i, gray_j, m, ransac_threshold=3.0):
    """stitch.rs:504-566 for one image pair's raw matches m (gray_*: f32 luma planes): None
    where the reference makes no edge, else a dict with h (3x3 f64 from the f32 alignment),
    h_inv, inliers, mean_i, mean_j."""
    m = np.ascontiguousarray(m, MATCH_DTYPE)
    if len(m) < 6:
        return None
    sim_inl, sim = filter_matches_ransac(m, ransac_threshold)
    hom_inl, H = homography_ransac_ref(m, ransac_threshold)
    best = hom_inl if len(hom_inl) >= len(sim_inl) else sim_inl
    if len(best) < 6:
        return None
    h_ij = resolve_safe_alignment(H, len(hom_inl), sim, gray_i.shape[1], gray_i.shape[0])
    if h_ij is None:
        return None
    h_ji = invert3_f32(h_ij)
    if h_ji is None:
        return None
    sum_i = 0.0
    sum_j = 0.0
    for k in best:   # f64 sums in inlier order
        sum_i += sample_patch_intensity(gray_i, float(m["qx"][k]), float(m["qy"][k]))
        sum_j += sample_patch_intensity(gray_j, float(m["tx"][k]), float(m["ty"][k]))
    inv_c = 1.0 / len(best)
    return dict(h=h_ij.astype(np.float64), h_inv=h_ji, inliers=len(best), mean_i=max(sum_i * inv_c, 1.0),
                mean_j=max(sum_j * inv_c, 1.0), best=best)


def panorama_pairs(images_u8, octaves=3, scales=3, threshold=0.015, ratio=0.75, ransac_threshold=3.0,
                   features=None):
    """stitch.rs:478-571 up to the pair edges: features per image (extract_sift_features:
    SIFT descriptors, upsampled first octave, cap 4096) unless given as [(feat, desc)], then
    every pair i < j in order; returns [(i, j, raw_match_count, panorama_pair(...) or None)]."""
    grays = [luma_to_f32(im) for im in images_u8]
    if features is None:
        features = [extract_features(g, octaves, scales, threshold, True, kind=0) for g in grays]
    out = []
    n = len(images_u8)
    for i in range(n):
        for j in range(i + 1, n):
            (qf, qd), (tf, td) = features[i], features[j]
            m = match(qf, qd, tf, td, ratio)
            out.append((i, j, len(m), panorama_pair(grays[i], grays[j], m, ransac_threshold)))
    return out
