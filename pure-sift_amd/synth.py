"""Seeded synthetic workloads for the SIFT extract + match path.

Every generator is a pure function of its seed (numpy PCG64), so the same
image/descriptor set is produced on every machine; tests pin the outputs by
SHA-256.

Workloads (names used by bench.py and the tests):
  blob_pair(640, 480, n_blobs=200)            u8 image + homography-warped partner
  blob_pair(1920, 1080, n_blobs=2000, band_noise=True)  1080p pair, band-limited noise
  unit_descriptors(50_000, 200_000)           ReLU(N(0,1)) unit-L2 128-D sets, 20% planted
  hamming_catalog(1_000_000, 1000)            128-bit codes + planted queries (0-20 flips)
  batch_sizes(256)                            (w, h) of the variable-size batch
"""
from __future__ import annotations

import hashlib
import math

import numpy as np


def _blur_gauss_f64(img: np.ndarray, sigma: float) -> np.ndarray:
    r = int(math.ceil(4 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-(x * x) / (2 * sigma * sigma))
    k /= k.sum()
    pad = np.pad(img, ((0, 0), (r, r)), mode="reflect")
    tmp = np.zeros_like(img)
    for i, w in enumerate(k):
        tmp += w * pad[:, i:i + img.shape[1]]
    pad = np.pad(tmp, ((r, r), (0, 0)), mode="reflect")
    out = np.zeros_like(img)
    for i, w in enumerate(k):
        out += w * pad[i:i + img.shape[0], :]
    return out


BAND_NOISE_STD = 12.0


def _gauss_energy(sigma: float) -> float:
    """Std gain of _blur_gauss_f64 on white noise: (sum of squared 1-D taps) for a 2-D separable blur."""
    r = int(math.ceil(4 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-(x * x) / (2 * sigma * sigma))
    k /= k.sum()
    return float((k * k).sum())


def blob_image(width: int, height: int, n_blobs: int, seed: int, band_noise: bool = False) -> np.ndarray:
    """Mid-gray 128 background + isotropic Gaussian blobs (sigma U[2,12], signed amplitude
    U[40,200]) + noise: i.i.d. N(0,3), or N(0,20) blurred with sigma 1.5 when band_noise.
    Rounded and clipped to u8."""
    rng = np.random.Generator(np.random.PCG64(seed))
    img = np.full((height, width), 128.0, np.float64)
    cx = rng.uniform(0, width, n_blobs)
    cy = rng.uniform(0, height, n_blobs)
    sg = rng.uniform(2.0, 12.0, n_blobs)
    amp = rng.uniform(40.0, 200.0, n_blobs) * np.where(rng.random(n_blobs) < 0.5, -1.0, 1.0)
    for x0, y0, s, a in zip(cx, cy, sg, amp):
        r = int(math.ceil(4 * s))
        xa, xb = max(0, int(x0) - r), min(width, int(x0) + r + 2)
        ya, yb = max(0, int(y0) - r), min(height, int(y0) + r + 2)
        if xa >= xb or ya >= yb:
            continue
        xs = np.arange(xa, xb, dtype=np.float64) - x0
        ys = np.arange(ya, yb, dtype=np.float64) - y0
        g = np.exp(-(ys[:, None] ** 2 + xs[None, :] ** 2) / (2 * s * s))
        img[ya:yb, xa:xb] += a * g
    if band_noise:
        # N(0,20) low-passed by a sigma-1.5 Gaussian, rescaled to std BAND_NOISE_STD (the filter
        # alone leaves std ~3.8, which yields ~2k keypoints at 1080p instead of the 8-15k band)
        noise = _blur_gauss_f64(rng.normal(0.0, 20.0, (height, width)), 1.5)
        img += noise * (BAND_NOISE_STD / (20.0 * _gauss_energy(1.5)))
    else:
        img += rng.normal(0.0, 3.0, (height, width))
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def random_homography(width: int, height: int, rng: np.random.Generator) -> np.ndarray:
    """H mapping image-1 pixel coords to image-2: rotation U[-30,30] deg and scale U[0.8,1.25]
    about the image centre, translation U[-40,40] px, perspective terms U[-1e-4,1e-4]."""
    ang = math.radians(rng.uniform(-30.0, 30.0))
    s = rng.uniform(0.8, 1.25)
    t = rng.uniform(-40.0, 40.0, 2)
    p = rng.uniform(-1e-4, 1e-4, 2)
    cx, cy = (width - 1) * 0.5, (height - 1) * 0.5
    A = np.array([[s * math.cos(ang), -s * math.sin(ang), 0.0],
                  [s * math.sin(ang), s * math.cos(ang), 0.0],
                  [p[0], p[1], 1.0]])
    T1 = np.array([[1.0, 0.0, -cx], [0.0, 1.0, -cy], [0.0, 0.0, 1.0]])
    T2 = np.array([[1.0, 0.0, cx + t[0]], [0.0, 1.0, cy + t[1]], [0.0, 0.0, 1.0]])
    H = T2 @ A @ T1
    return H / H[2, 2]


def warp_homography(img: np.ndarray, H: np.ndarray, fill: float = 128.0) -> np.ndarray:
    """Target(x') = bilinear source(H^-1 x'); out-of-bounds pixels are `fill`."""
    h, w = img.shape
    Hi = np.linalg.inv(H)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    den = Hi[2, 0] * xs + Hi[2, 1] * ys + Hi[2, 2]
    sx = (Hi[0, 0] * xs + Hi[0, 1] * ys + Hi[0, 2]) / den
    sy = (Hi[1, 0] * xs + Hi[1, 1] * ys + Hi[1, 2]) / den
    inside = (sx >= 0) & (sx <= w - 1) & (sy >= 0) & (sy <= h - 1)
    sxc = np.clip(sx, 0, w - 1)
    syc = np.clip(sy, 0, h - 1)
    x0 = np.floor(sxc).astype(np.int64)
    y0 = np.floor(syc).astype(np.int64)
    x1 = np.minimum(x0 + 1, w - 1)
    y1 = np.minimum(y0 + 1, h - 1)
    fx = sxc - x0
    fy = syc - y0
    src = img.astype(np.float64)
    val = (1 - fy) * ((1 - fx) * src[y0, x0] + fx * src[y0, x1]) + fy * ((1 - fx) * src[y1, x0] + fx * src[y1, x1])
    out = np.where(inside, val, fill)
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def blob_pair(width: int, height: int, n_blobs: int, seed: int = 0, band_noise: bool = False):
    """(image, warped partner, H) for the extract + match workloads."""
    img = blob_image(width, height, n_blobs, seed, band_noise)
    rng = np.random.Generator(np.random.PCG64(seed + 0x5EED))
    H = random_homography(width, height, rng)
    return img, warp_homography(img, H), H


def panorama_tiles(width: int, height: int, n_tiles: int, n_blobs: int, seed: int = 0, tile_frac: float = 0.42,
                   exposure=None):
    """Overlapping horizontal crops of one wide blob image, as the reference's 4-tile panorama
    test builds them (tests/panorama_tests.rs:29-47): tile width = width * tile_frac, step =
    (width - tile width) / (n_tiles - 1), each tile's pixels scaled by its exposure factor
    (rounded, clamped to u8). Returns (tiles, step)."""
    base = blob_image(width, height, n_blobs, seed, band_noise=True)
    tw = width * int(round(tile_frac * 100)) // 100
    step = (width - tw) // max(n_tiles - 1, 1)
    tiles = []
    for i in range(n_tiles):
        crop = base[:, i * step:i * step + tw].astype(np.float32)
        f = 1.0 if exposure is None else float(exposure[i])
        tiles.append(np.clip(np.round(crop * np.float32(f)), 0, 255).astype(np.uint8))
    return tiles, step


def unit_descriptors(n_query: int, n_db: int, seed: int = 0, planted_frac: float = 0.2, noise: float = 0.05):
    """ReLU(N(0,1)) L2-normalised 128-D database; queries are random, except a planted
    fraction that are database rows + N(0, noise), renormalised. Returns (q, db, planted_idx)
    with planted_idx = -1 for unplanted queries."""
    rng = np.random.Generator(np.random.PCG64(seed))
    db = np.maximum(rng.standard_normal((n_db, 128), dtype=np.float32), 0)
    db /= np.maximum(np.linalg.norm(db, axis=1, keepdims=True), 1e-12)
    q = np.maximum(rng.standard_normal((n_query, 128), dtype=np.float32), 0)
    planted = np.full(n_query, -1, np.int64)
    npl = int(round(planted_frac * n_query))
    sel = rng.choice(n_query, npl, replace=False)
    src = rng.integers(0, n_db, npl)
    q[sel] = db[src] + rng.normal(0, noise, (npl, 128)).astype(np.float32)
    planted[sel] = src
    q /= np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    return q.astype(np.float32), db.astype(np.float32), planted


def hamming_catalog(n_codes: int, n_queries: int, seed: int = 0, max_flips: int = 20):
    """1M random 128-bit codes (two little-endian u64 words each, every bit Bernoulli(0.5))
    and queries planted as catalog codes with 0..max_flips random bits flipped."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cat = rng.integers(0, 2 ** 63, (n_codes, 2), dtype=np.int64).astype(np.uint64)
    cat ^= (rng.integers(0, 2, (n_codes, 2), dtype=np.int64).astype(np.uint64) << np.uint64(63))
    src = rng.integers(0, n_codes, n_queries)
    q = cat[src].copy()
    nflip = rng.integers(0, max_flips + 1, n_queries)
    for i in range(n_queries):
        bits = rng.choice(128, nflip[i], replace=False)
        for b in bits:
            q[i, b // 64] ^= np.uint64(1) << np.uint64(b % 64)
    return cat, q, src, nflip


def batch_sizes(n: int, seed: int = 0):
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.integers(320, 2049, n)
    h = rng.integers(240, 1537, n)
    return list(zip(w.tolist(), h.tolist()))


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
