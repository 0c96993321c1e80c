"""Seeded two-view scenes for the geometry tests: 3-D points seen by two pinhole cameras
(the shape of the reference's tests/geometry_tests.rs scenes), as FeatureMatch records."""
from __future__ import annotations

import math

import numpy as np

import psr


def rotation(yaw_deg: float, pitch_deg: float, roll_deg: float) -> np.ndarray:
    """R = Rz(yaw) Ry(pitch) Rx(roll)."""
    z, y, x = (math.radians(v) for v in (yaw_deg, pitch_deg, roll_deg))
    rz = np.array([[math.cos(z), -math.sin(z), 0], [math.sin(z), math.cos(z), 0], [0, 0, 1]])
    ry = np.array([[math.cos(y), 0, math.sin(y)], [0, 1, 0], [-math.sin(y), 0, math.cos(y)]])
    rx = np.array([[1, 0, 0], [0, math.cos(x), -math.sin(x)], [0, math.sin(x), math.cos(x)]])
    return rz @ ry @ rx


def intrinsics(f=720.0, cx=400.0, cy=300.0) -> np.ndarray:
    return np.array([[f, 0, cx], [0, f, cy], [0, 0, 1.0]])


def project(K, R, t, P):
    Xc = P @ R.T + t
    return np.stack([K[0, 0] * Xc[:, 0] / Xc[:, 2] + K[0, 2], K[1, 1] * Xc[:, 1] / Xc[:, 2] + K[1, 2]], 1)


def true_fundamental(K, R, t) -> np.ndarray:
    tx = np.array([[0, -t[2], t[1]], [t[2], 0, -t[0]], [-t[1], t[0], 0]])
    Ki = np.linalg.inv(K)
    F = Ki.T @ tx @ R @ Ki
    nrm = np.linalg.norm(F)
    return F / nrm if nrm > 0 else F


def two_view(n_inliers, n_outliers, R, t, K=None, seed=0, planar_depth=None, noise=0.0):
    """Matches: n_inliers projections of random 3-D points (depth 4..9, or one plane), then
    n_outliers random point pairs in an 800 x 600 frame (ratio 0.45 vs 0.78)."""
    K = intrinsics() if K is None else K
    rng = np.random.default_rng(seed)
    P = np.stack([(rng.random(n_inliers) - 0.5) * 4.0, (rng.random(n_inliers) - 0.5) * 3.5,
                  (4.0 + rng.random(n_inliers) * 5.0) if planar_depth is None else np.full(n_inliers, planar_depth)], 1)
    p1 = project(K, np.eye(3), np.zeros(3), P)
    p2 = project(K, R, np.asarray(t, float), P) + rng.normal(0, noise, (n_inliers, 2))
    o1 = rng.random((n_outliers, 2)) * [800, 600]
    o2 = rng.random((n_outliers, 2)) * [800, 600]
    q = np.concatenate([p1, o1]).astype(np.float32)
    tt = np.concatenate([p2, o2]).astype(np.float32)
    m = psr.make_matches(q, tt)
    m["ratio"] = np.concatenate([np.full(n_inliers, 0.45), np.full(n_outliers, 0.78)]).astype(np.float32)
    m["distance"] = np.concatenate([np.full(n_inliers, 0.1), np.full(n_outliers, 0.4)]).astype(np.float32)
    return m, true_fundamental(K, R, np.asarray(t, float))


def same_up_to_scale(a, b) -> float:
    a = np.asarray(a, float).ravel() / np.linalg.norm(a)
    b = np.asarray(b, float).ravel() / np.linalg.norm(b)
    return min(np.abs(a - b).max(), np.abs(a + b).max())
