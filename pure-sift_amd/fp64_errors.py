"""Float64 restatements of the hot path, used to state the error of each f32 result against
an FP64 evaluation of the same algorithm (pure-sift 0.2.0, src/pipeline.rs):

* scale space (pipeline.rs:275-659): Gaussian taps, the 2x bilinear upsample, the separable
  clamped blur and the stride-2 downsample, all in float64 with float64 sigmas;
* keypoints (pipeline.rs:697-770, :881-893): the 3-D Taylor step solved in float64, once on
  the f32 DoG samples (the error of the f32 solve alone) and once on the float64 pyramid's
  DoG samples (the error of the whole f32 chain), at the f32 result's integer extremum;
* orientations and descriptors (pipeline.rs:1168-1480): gradient planes, the 36-bin
  orientation histogram with its peak rule and the 4x4x8 descriptor in float64 on the
  float64 Gaussian level;
* matches (pipeline.rs:1744-1841): nearest / second-nearest distances of the same f32
  descriptors in float64;
* homographies (pipeline.rs:2276-2630): the normalised DLT of the final inliers by a float64
  SVD, compared with the f32 model by reprojection.

These are checkers of the f32 results: the product path never calls them.
"""
import math

import numpy as np
from scipy.ndimage import correlate1d

BASE_SIGMA = 1.6
ASSUMED_INPUT_SIGMA = 0.5
TWO_PI = 2.0 * math.pi


def gaussian_taps64(sigma: float) -> np.ndarray:
    """pipeline.rs:275-296 in float64: radius max(ceil(3 sigma), 1), weights normalised to 1."""
    r = max(int(math.ceil(3.0 * sigma)), 1)
    x = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-(x * x) / (2.0 * sigma * sigma)) / (math.sqrt(2.0 * math.pi) * sigma)
    return w / w.sum()


def blur64(img: np.ndarray, sigma: float) -> np.ndarray:
    """Separable blur with clamped borders (pipeline.rs:314-425): horizontal, then vertical."""
    w = gaussian_taps64(sigma)
    tmp = correlate1d(img, w, axis=1, mode="nearest")
    return correlate1d(tmp, w, axis=0, mode="nearest")


def upsample64(img: np.ndarray) -> np.ndarray:
    """upsample_2x_bilinear (pipeline.rs:455-483): 0.25 * the 4 clamped neighbours."""
    h, w = img.shape
    yy, xx = np.mgrid[0:2 * h, 0:2 * w]
    y0, x0 = yy >> 1, xx >> 1
    y1 = np.minimum(y0 + (yy & 1), h - 1)
    x1 = np.minimum(x0 + (xx & 1), w - 1)
    return 0.25 * (img[y0, x0] + img[y0, x1] + img[y1, x0] + img[y1, x1])


def pyramid64(img: np.ndarray, octaves: int, scales: int, upsample: bool):
    """build_scale_space_pyramid_with_options (pipeline.rs:550-638) in float64: a list of
    (gaussians [S+3, h, w] float64, scale_factor) per octave."""
    img = np.asarray(img, np.float64)
    if upsample:
        base = blur64(upsample64(img), math.sqrt(max(BASE_SIGMA ** 2 - (2 * ASSUMED_INPUT_SIGMA) ** 2, 1e-4)))
        mult0 = 0.5
    else:
        base = blur64(img, math.sqrt(max(BASE_SIGMA ** 2 - ASSUMED_INPUT_SIGMA ** 2, 1e-4)))
        mult0 = 1.0
    k = 2.0 ** (1.0 / scales)
    out = []
    for o in range(octaves):
        h, w = base.shape
        if h < 3 or w < 3:
            break
        g = [base]
        prev = BASE_SIGMA
        for s in range(1, scales + 3):
            tgt = BASE_SIGMA * k ** s
            g.append(blur64(g[-1], math.sqrt(max(tgt * tgt - prev * prev, 1e-4))))
            prev = tgt
        out.append((np.stack(g), mult0 * (1 << o)))
        base = g[scales][0::2, 0::2][: h // 2, : w // 2]
    return out


def _refine64(cube):
    """One 3-D Taylor step (pipeline.rs:697-770) in float64 on the 3x3x3 DoG neighbourhood
    cube[slice, row, column] of an extremum: (dx, dy, ds, interp, used_3d)."""
    p, c, n = cube
    cc = c[1, 1]
    gx = 0.5 * (c[1, 2] - c[1, 0])
    gy = 0.5 * (c[2, 1] - c[0, 1])
    gs = 0.5 * (n[1, 1] - p[1, 1])
    dxx = c[1, 2] - 2 * cc + c[1, 0]
    dyy = c[2, 1] - 2 * cc + c[0, 1]
    dss = n[1, 1] - 2 * cc + p[1, 1]
    dxy = 0.25 * ((c[2, 2] - c[2, 0]) - (c[0, 2] - c[0, 0]))
    dxs = 0.25 * ((n[1, 2] - n[1, 0]) - (p[1, 2] - p[1, 0]))
    dys = 0.25 * ((n[2, 1] - n[0, 1]) - (p[2, 1] - p[0, 1]))
    Hm = np.array([[dxx, dxy, dxs], [dxy, dyy, dys], [dxs, dys, dss]])
    g = np.array([gx, gy, gs])
    if abs(np.linalg.det(Hm)) > 1e-9:
        o = -np.linalg.solve(Hm, g)
        if np.all(np.abs(o) <= 0.5):
            return o[0], o[1], o[2], cc + 0.5 * float(g @ o), True
    c22 = dxx * dyy - dxy * dxy
    if abs(c22) > 1e-9:
        ox = min(max(-(dyy * gx - dxy * gy) / c22, -0.5), 0.5)
        oy = min(max(-(-dxy * gx + dxx * gy) / c22, -0.5), 0.5)
        os_ = min(max(-gs / dss, -0.5), 0.5) if abs(dss) > 1e-6 else 0.0
        return ox, oy, os_, cc + 0.5 * (gx * ox + gy * oy + gs * os_), False
    return 0.0, 0.0, 0.0, cc, False


def _cube(dog, s, y, x):
    """float64 3x3x3 neighbourhood of DoG slices s-1..s+1 around (y, x)."""
    return np.asarray(dog[s - 1:s + 2, y - 1:y + 2, x - 1:x + 2], np.float64)


def _is_extremum(cube):
    v = cube[1, 1, 1]
    others = np.delete(cube.ravel(), 13)
    return bool(np.all(v > others)) if v > 0 else bool(np.all(v < others))


def locate_extrema(kps, dog32, mults, scales):
    """The integer extremum (octave, slice, y, x) of each f32 keypoint: the strict 3x3x3
    extremum of the f32 DoG within half a pixel of the keypoint's octave-local position and
    one slice of its layer whose f64 Taylor step lands nearest to the keypoint."""
    found = []
    for kp in kps:
        o = int(kp["octave"])
        d = dog32[o]
        m = mults[o]
        xf, yf = float(kp["x"]) / m, float(kp["y"]) / m
        best = None
        for s in range(max(1, int(kp["layer"]) - 1), min(scales, int(kp["layer"]) + 1) + 1):
            for yi in {int(math.floor(yf)), int(math.ceil(yf))}:
                for xi in {int(math.floor(xf)), int(math.ceil(xf))}:
                    if not (1 <= yi < d.shape[1] - 1 and 1 <= xi < d.shape[2] - 1):
                        continue
                    if abs(yi - yf) > 0.5 + 1e-4 or abs(xi - xf) > 0.5 + 1e-4:
                        continue
                    cube = _cube(d, s, yi, xi)
                    if not _is_extremum(cube):
                        continue
                    dx, dy, ds, _, _ = _refine64(cube)
                    err = abs((xi + dx) * m - float(kp["x"])) + abs((yi + dy) * m - float(kp["y"]))
                    if best is None or err < best[0]:
                        best = (err, o, s, yi, xi)
        found.append(best[1:] if best else None)
    return found


def _scale64(o_mult, s, ds, scales):
    k = 2.0 ** (1.0 / scales)
    return max(BASE_SIGMA * o_mult * k ** s * k ** ds, 0.25 * BASE_SIGMA)


def keypoint_errors(kps, pyr32, pyr64, scales):
    """Errors of f32 keypoints (x, y, scale) against the float64 Taylor step.
    pyr32: per octave (gauss [S+3, h, w] f32, dog [S+2, h, w] f32) as the pyramid dump gives
    them; pyr64: pyramid64(...) of the same image. Positions in input-image pixels, scale in
    octaves (|log2(scale / scale64)|). 'solve': f64 solve on the f32 DoG samples; 'total': on
    the float64 pyramid."""
    dog32 = [np.asarray(d) for _, d in pyr32]
    mults = [m for _, m in pyr64]
    locs = locate_extrema(kps, dog32, mults, scales)
    res = {"solve": {"dxy": [], "dscale": []}, "total": {"dxy": [], "dscale": []}, "unlocated": 0}
    for kp, loc in zip(kps, locs):
        if loc is None:
            res["unlocated"] += 1
            continue
        o, s, yi, xi = loc
        m = mults[o]
        g64 = pyr64[o][0]
        for name, cube in (("solve", _cube(dog32[o], s, yi, xi)),
                           ("total", g64[s:s + 3, yi - 1:yi + 2, xi - 1:xi + 2] -
                            g64[s - 1:s + 2, yi - 1:yi + 2, xi - 1:xi + 2])):
            dx, dy, ds, _, _ = _refine64(cube)
            x64, y64 = (xi + dx) * m, (yi + dy) * m
            sc = _scale64(m, s, ds, scales)
            res[name]["dxy"].append(max(abs(x64 - float(kp["x"])), abs(y64 - float(kp["y"]))))
            res[name]["dscale"].append(abs(math.log2(float(kp["scale"]) / sc)))
    out = {"n": len(kps), "unlocated": res["unlocated"]}
    for name in ("solve", "total"):
        a, b = np.array(res[name]["dxy"]), np.array(res[name]["dscale"])
        out[name] = {"pos_px_max": float(a.max()) if len(a) else 0.0,
                     "pos_px_p99": float(np.percentile(a, 99)) if len(a) else 0.0,
                     "scale_oct_max": float(b.max()) if len(b) else 0.0,
                     "scale_oct_p99": float(np.percentile(b, 99)) if len(b) else 0.0,
                     # keypoints outside the 1e-3 px / 1e-4 octave tolerance (the float64 step
                     # can take the other branch of the |offset| <= 0.5 test near its edge)
                     "over_1e-3_px": int((a > 1e-3).sum()), "over_1e-4_oct": int((b > 1e-4).sum())}
    return out


def gradient_planes64(g: np.ndarray):
    """pipeline.rs:1168-1203: central differences, zero border, angle in [0, 2 pi)."""
    gx = np.zeros_like(g)
    gy = np.zeros_like(g)
    gx[1:-1, 1:-1] = g[1:-1, 2:] - g[1:-1, :-2]
    gy[1:-1, 1:-1] = g[2:, 1:-1] - g[:-2, 1:-1]
    return np.sqrt(gx * gx + gy * gy), np.mod(np.arctan2(gy, gx), TWO_PI)


def orientations64(mag, ang, x, y, scale):
    """compute_keypoint_orientations (pipeline.rs:1233-1347) in float64 (octave-local kp)."""
    h, w = mag.shape
    so = max(1.5 * scale, 1.0)
    R = max(int(math.ceil(3 * so)), 2)
    cx, cy = int(round(x)), int(round(y))
    y0, y1 = max(cy - R, 1), max(min(cy + R, h - 2), 1)
    x0, x1 = max(cx - R, 1), max(min(cx + R, w - 2), 1)
    yy, xx = np.mgrid[y0:y1 + 1, x0:x1 + 1]
    r2 = (xx - cx) ** 2 + (yy - cy) ** 2
    mm = mag[y0:y1 + 1, x0:x1 + 1]
    ok = (r2 <= R * R) & (mm > 1e-7)
    bins = np.floor(ang[y0:y1 + 1, x0:x1 + 1] * 36 / TWO_PI).astype(np.int64) % 36
    hist = np.bincount(bins[ok], weights=(mm * np.exp(-r2 / (2 * so * so)))[ok], minlength=36)
    sm = (np.roll(hist, 1) + hist + np.roll(hist, -1)) / 3.0
    best = int(np.argmax(sm))
    if sm[best] <= 1e-7:
        return [0.0]
    peaks = [(b, sm[b]) for b in range(36)
             if sm[b] >= 0.8 * sm[best] and sm[b] > sm[(b - 1) % 36] and sm[b] >= sm[(b + 1) % 36]]
    if not peaks:
        peaks = [(best, sm[best])]
    peaks = sorted(peaks, key=lambda t: (-t[1], t[0]))[:4]
    out = []
    for b, v in peaks:
        left, right = sm[(b - 1) % 36], sm[(b + 1) % 36]
        den = left - 2 * v + right
        off = min(max(0.5 * (left - right) / den if abs(den) > 1e-6 else 0.0, -0.5), 0.5)
        out.append(((b + 0.5 + off) * TWO_PI / 36) % TWO_PI)
    return out


def descriptor64(mag, ang, x, y, scale, theta):
    """compute_128d_descriptor (pipeline.rs:1352-1480) in float64 (octave-local kp)."""
    h, w = mag.shape
    c, s = math.cos(theta), math.sin(theta)
    hw = max(3 * scale, 1.5)
    R = max(int(math.ceil(hw * math.sqrt(2) * 5 * 0.5)), 4)
    cx, cy = int(round(x)), int(round(y))
    y0, y1 = max(cy - R, 1), max(min(cy + R, h - 2), 1)
    x0, x1 = max(cx - R, 1), max(min(cx + R, w - 2), 1)
    yy, xx = np.mgrid[y0:y1 + 1, x0:x1 + 1].astype(np.float64)
    dx, dy = xx - x, yy - y
    rr = (dx * c + dy * s) / hw
    cc = (-dx * s + dy * c) / hw
    rb, cb = rr + 1.5, cc + 1.5
    mm = mag[y0:y1 + 1, x0:x1 + 1]
    ok = (rb > -1) & (rb < 4) & (cb > -1) & (cb < 4) & (mm > 1e-7)
    rb, cb, rr, cc, mm = rb[ok], cb[ok], rr[ok], cc[ok], mm[ok]
    ob = np.mod(ang[y0:y1 + 1, x0:x1 + 1][ok] - theta, TWO_PI) * 8 / TWO_PI
    wm = mm * np.exp(-(rr * rr + cc * cc) / 8)
    r0, c0, o0 = np.floor(rb).astype(np.int64), np.floor(cb).astype(np.int64), np.floor(ob).astype(np.int64)
    fr, fc, fo = rb - r0, cb - c0, ob - o0
    desc = np.zeros(128)
    for ir in (0, 1):
        wr = fr if ir else 1 - fr
        for ic in (0, 1):
            wc = fc if ic else 1 - fc
            for io in (0, 1):
                wo = fo if io else 1 - fo
                r, cl = r0 + ir, c0 + ic
                v = (r >= 0) & (r < 4) & (cl >= 0) & (cl < 4)
                idx = ((r * 4 + cl) * 8 + (o0 + io) % 8)[v]
                desc += np.bincount(idx, weights=(wm * wr * wc * wo)[v], minlength=128)
    n1 = np.linalg.norm(desc)
    if n1 > 1e-7:
        desc = np.minimum(desc / n1, 0.2)
        n2 = np.linalg.norm(desc)
        if n2 > 1e-7:
            desc = desc / n2
    return desc


def descriptor_errors(feats, descs, kps, pyr64, n_octaves, sample=64, seed=0):
    """Descriptor L2 and orientation error of f32 features against the float64 pipeline on
    the float64 Gaussian level (octave min(octave, O-1), layer min(layer, 5)); a sample of
    features, each paired with the float64 orientation nearest to its own. Features carry the
    image-space keypoint (x, y, scale, response); its octave and layer come from the keypoint
    list `kps` it was computed from, and the octave coordinates as pipeline.rs:1519-1528."""
    key = lambda r: tuple(np.float32(r[n]).view(np.uint32).item() for n in ("x", "y", "scale", "response"))
    kmap = {key(k): (int(k["octave"]), int(k["layer"])) for k in kps}
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(len(feats), min(sample, len(feats)), replace=False))
    planes = {}
    l2, l2o, dori = [], [], []
    for i in idx:
        f = feats[i]
        o, layer = kmap[key(f)]
        o, layer = min(o, n_octaves - 1), min(layer, 5)
        g, m = pyr64[o]
        if (o, layer) not in planes:
            planes[(o, layer)] = gradient_planes64(g[layer])
        mag, ang = planes[(o, layer)]
        x, y = float(f["x"]) / m, float(f["y"]) / m
        sc = max(float(f["scale"]) / m, 0.8)
        oris = orientations64(mag, ang, x, y, sc)
        t = float(f["orientation"])
        diff = [min(abs(t - a), TWO_PI - abs(t - a)) for a in oris]
        j = int(np.argmin(diff))
        dori.append(diff[j])
        # at the feature's own orientation: the descriptor arithmetic alone
        l2.append(float(np.linalg.norm(descriptor64(mag, ang, x, y, sc, t) - descs[i])))
        # at the float64 orientation: the whole chain (orientation bin assignments included)
        l2o.append(float(np.linalg.norm(descriptor64(mag, ang, x, y, sc, oris[j]) - descs[i])))
    l2, l2o, dori = np.array(l2), np.array(l2o), np.array(dori)
    return {"n": int(len(idx)),
            "l2_max": float(l2.max()), "l2_p99": float(np.percentile(l2, 99)),
            "l2_with_f64_orientation_max": float(l2o.max()),
            "l2_with_f64_orientation_p99": float(np.percentile(l2o, 99)),
            "l2_with_f64_orientation_median": float(np.median(l2o)),
            "orientation_rad_max": float(dori.max()), "orientation_rad_p99": float(np.percentile(dori, 99)),
            "orientation_rad_median": float(np.median(dori)),
            # features whose float64 peak set differs (a bin assignment or the 80% peak rule flips)
            "orientation_over_1e-3_rad": int((dori > 1e-3).sum())}


def match_errors(qd, td, matches):
    """Distance and ratio of f32 match records against float64 nearest / second-nearest of the
    same descriptors; and how often the float64 nearest index differs."""
    q = np.asarray(qd, np.float64)[matches["query_index"].astype(np.int64)]
    t = np.asarray(td, np.float64)
    d2 = (q * q).sum(1)[:, None] + (t * t).sum(1)[None, :] - 2.0 * q @ t.T
    # exact float64 distances of the two best candidates (the GEMM form only ranks them)
    part = np.argpartition(d2, 2, axis=1)[:, :3]
    ex = ((q[:, None, :] - t[part]) ** 2).sum(-1)
    order = np.argsort(ex, axis=1, kind="stable")
    best = np.take_along_axis(ex, order, 1)
    bi = np.take_along_axis(part, order, 1)[:, 0]
    dist = np.sqrt(best[:, 0])
    ratio = np.sqrt(best[:, 0]) / np.sqrt(best[:, 1])
    derr = np.abs(matches["distance"].astype(np.float64) - dist)
    rerr = np.abs(matches["ratio"].astype(np.float64) - ratio)
    return {"n": int(len(matches)), "distance_abs_max": float(derr.max(initial=0.0)),
            "ratio_abs_max": float(rerr.max(initial=0.0)),
            "nearest_index_differs": float(np.mean(bi != matches["target_index"].astype(np.int64)))
            if len(matches) else 0.0}


def homography_dlt64(q, t):
    """Normalised DLT (Hartley) by float64 SVD: H with t ~ H q, H[2, 2] = 1."""
    def norm(p):
        c = p.mean(0)
        d = np.sqrt(((p - c) ** 2).sum(1)).mean()
        s = math.sqrt(2) / d if d > 0 else 1.0
        return np.array([[s, 0, -s * c[0]], [0, s, -s * c[1]], [0, 0, 1.0]])
    Tq, Tt = norm(q), norm(t)
    qh = (Tq @ np.c_[q, np.ones(len(q))].T).T
    th = (Tt @ np.c_[t, np.ones(len(t))].T).T
    A = []
    for (x, y, _), (u, v, _) in zip(qh, th):
        A.append([-x, -y, -1, 0, 0, 0, u * x, u * y, u])
        A.append([0, 0, 0, -x, -y, -1, v * x, v * y, v])
    _, _, vt = np.linalg.svd(np.array(A))
    Hn = vt[-1].reshape(3, 3)
    H = np.linalg.inv(Tt) @ Hn @ Tq
    return H / H[2, 2]


def homography_errors(matches, inliers, H32):
    """Reprojection difference (px) over the inlier query points between the f32 model and
    the float64 least-squares DLT of the same inliers. The reference's model is its last
    local-optimisation refit (pipeline.rs:2700-2740), fitted on the inlier set before the final
    rescoring, so where that set and the final one differ this also measures that difference."""
    m = matches[np.asarray(inliers, np.int64)]
    q = np.c_[m["qx"], m["qy"]].astype(np.float64)
    t = np.c_[m["tx"], m["ty"]].astype(np.float64)
    H64 = homography_dlt64(q, t)

    def proj(H):
        p = (np.asarray(H, np.float64) @ np.c_[q, np.ones(len(q))].T).T
        return p[:, :2] / p[:, 2:3]
    return {"n_inliers": int(len(m)), "reproj_px_max": float(np.abs(proj(H32) - proj(H64)).max())}
