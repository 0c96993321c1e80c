// Fundamental-matrix LO-RANSAC (geometry.rs:857-1016) with the normalised 8-point and the
// 7-point solvers (geometry.rs:145-805), and the GRIC model comparison (geometry.rs:1233-1341),
// batched: one workgroup (one wave) per problem.
//
// Exactness: every f64 expression is evaluated in the reference's order with separate
// roundings, sums run sequentially in index order, hypot is the glibc algorithm the reference
// binds (a correction step after sqrt(x^2 + y^2), no fused multiply-add), and the SplitMix64
// stream is consumed draw for draw, so models and inliers equal the sequential algorithm's.
// The adaptive bound uses the device f64 log (a last-ulp difference from glibc moves ceil()
// only within ~1e-15 of an integer); the 7-point path (exactly 7 matches) uses device cbrt /
// acos / cos before its Newton polish.
//
// Parallel structure: the sample stream never depends on the evaluation, so thread 0 groups
// draws into the next 64 iterations, each thread fits and scores one hypothesis (its 9x9
// Jacobi matrices live in LDS, element-major, so every lane's dynamic pivot indexing is
// bank-conflict free), and thread 0 replays the 64 in order with the adaptive bound.
#include <hip/hip_runtime.h>

#include "ps_internal.h"

#pragma clang fp contract(off)

namespace ps {

namespace {

constexpr int kFT = 64;   // threads (and hypotheses per chunk) per problem
constexpr uint32_t kFStage = 256;   // points per staging chunk
constexpr uint64_t kGammaF = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t fmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// k-th output (k = 0, 1, ...) of the SplitMix64 generator started at `seed`
__device__ __forceinline__ uint64_t fsplitmix_at(uint64_t seed, uint64_t k) { return fmix64(seed + (k + 1) * kGammaF); }

__device__ uint32_t fcdf_pick(const double *cdf, uint32_t n, double u) {   // pipeline.rs:2031-2036
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    return lo < n - 1 ? lo : n - 1;
}

// glibc's hypot for finite arguments: h = sqrt(x^2 + y^2) corrected by one step that recovers
// the rounding error of the squares (Borges' corrected algorithm, no fma); arguments whose
// ratio exceeds 2^54 return the larger plus the smaller.
__device__ double glibc_hypot(double x, double y) {
    double ax = fabs(x), ay = fabs(y);
    if (ax < ay) {
        const double t = ax;
        ax = ay;
        ay = t;
    }
    if (!isfinite(ax) || !isfinite(ay)) return isinf(ax) || isinf(ay) ? INFINITY : ax + ay;
    if (ay == 0.0) return ax;
    if (ay <= ax * 5.551115123125783e-17) return ax + ay;   // 2^-54
    double h = sqrt(ax * ax + ay * ay);
    double t1, t2;
    if (h <= 2.0 * ay) {
        const double delta = h - ay;
        t1 = ax * (2.0 * delta - ax);
        t2 = (delta - 2.0 * (ax - ay)) * delta;
    } else {
        const double delta = h - ax;
        t1 = 2.0 * delta * (ax - 2.0 * ay);
        t2 = (4.0 * delta - ay) * ay + delta * delta;
    }
    h -= (t1 + t2) / (2.0 * h);
    return h;
}

// geometry.rs:617-635
__device__ __forceinline__ double sampson_sq(const double *f, double x1, double y1, double x2, double y2) {
    const double fx0 = f[0] * x1 + f[1] * y1 + f[2];
    const double fx1 = f[3] * x1 + f[4] * y1 + f[5];
    const double fx2 = f[6] * x1 + f[7] * y1 + f[8];
    const double ft0 = f[0] * x2 + f[3] * y2 + f[6];
    const double ft1 = f[1] * x2 + f[4] * y2 + f[7];
    const double e = x2 * fx0 + y2 * fx1 + fx2;
    const double den = fx0 * fx0 + fx1 * fx1 + ft0 * ft0 + ft1 * ft1;
    if (den < 1e-18 || !isfinite(den)) return INFINITY;
    return (e * e) / den;
}

__device__ __forceinline__ double msampson(const double *f, const ps_feature_match &p) {
    return sampson_sq(f, (double)p.qx, (double)p.qy, (double)p.tx, (double)p.ty);
}

__device__ void mul3(const double a[9], const double b[9], double o[9]) {   // geometry.rs:145-153
    double t[9];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) t[r * 3 + c] = a[r * 3] * b[c] + a[r * 3 + 1] * b[3 + c] + a[r * 3 + 2] * b[6 + c];
#pragma unroll
    for (int i = 0; i < 9; i++) o[i] = t[i];
}

__device__ void transpose3(const double m[9], double o[9]) {
    double t[9];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) t[r * 3 + c] = m[c * 3 + r];
#pragma unroll
    for (int i = 0; i < 9; i++) o[i] = t[i];
}

__device__ __forceinline__ double det3(const double m[9]) {   // geometry.rs:165-169
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}

__device__ bool normalize_frobenius(double m[9]) {   // geometry.rs:172-184
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 9; i++) s += m[i] * m[i];
    if (!isfinite(s) || s < 1e-24) return false;
    const double inv = 1.0 / sqrt(s);
#pragma unroll
    for (int i = 0; i < 9; i++) m[i] *= inv;
    return true;
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3]) {
    const double t0 = a[1] * b[2] - a[2] * b[1], t1 = a[2] * b[0] - a[0] * b[2], t2 = a[0] * b[1] - a[1] * b[0];
    o[0] = t0;
    o[1] = t1;
    o[2] = t2;
}

__device__ bool normalize_vec3(const double v[3], double o[3]) {   // geometry.rs:201-207
    const double n = sqrt(dot3(v, v));
    if (!isfinite(n) || n < 1e-12) return false;
    const double t0 = v[0] / n, t1 = v[1] / n, t2 = v[2] / n;
    o[0] = t0;
    o[1] = t1;
    o[2] = t2;
    return true;
}

// One rotation of the 3x3 classical Jacobi with pivot (P, Q) and the remaining row R.
template <int P, int Q, int R>
__device__ __forceinline__ void jrot3(double (&a)[3][3], double (&v)[3][3]) {
    const double app = a[P][P], aqq = a[Q][Q], apq = a[P][Q];
    const double tau = (aqq - app) / (2.0 * apq);
    const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
    const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
    a[P][P] = app - t * apq;
    a[Q][Q] = aqq + t * apq;
    a[P][Q] = 0.0;
    a[Q][P] = 0.0;
    const double arp = a[R][P], arq = a[R][Q];
    const double np = c * arp - s * arq, nq = s * arp + c * arq;
    a[R][P] = np;
    a[P][R] = np;
    a[R][Q] = nq;
    a[Q][R] = nq;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const double vrp = v[r][P], vrq = v[r][Q];
        v[r][P] = c * vrp - s * vrq;
        v[r][Q] = s * vrp + c * vrq;
    }
}

// geometry.rs:297-373: classical Jacobi, 1e-14, up to 100 rotations; eigenpairs sorted by
// descending eigenvalue (stable)
__device__ bool jacobi3(double (&a)[3][3], double ev[3], double (&vec)[3][3]) {
    double v[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    for (int it = 0; it < 100; it++) {
        int piv = 0;   // 0: (0,1), 1: (0,2), 2: (1,2)
        double mo = 0.0;
        if (fabs(a[0][1]) > mo) {
            mo = fabs(a[0][1]);
            piv = 0;
        }
        if (fabs(a[0][2]) > mo) {
            mo = fabs(a[0][2]);
            piv = 1;
        }
        if (fabs(a[1][2]) > mo) {
            mo = fabs(a[1][2]);
            piv = 2;
        }
        if (!isfinite(mo)) return false;
        if (mo < 1e-14) break;
        if (piv == 0) jrot3<0, 1, 2>(a, v);
        else if (piv == 1) jrot3<0, 2, 1>(a, v);
        else jrot3<1, 2, 0>(a, v);
    }
    int ord[3] = {0, 1, 2};
    const double d[3] = {a[0][0], a[1][1], a[2][2]};
    // insertion sort, descending: shift while the new key is greater than the left one
    for (int i = 1; i < 3; i++) {
        const int k = ord[i];
        int j = i - 1;
        while (j >= 0 && d[k] > d[ord[j]]) {
            ord[j + 1] = ord[j];
            j--;
        }
        ord[j + 1] = k;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int o = ord[c];
        ev[c] = o == 0 ? d[0] : (o == 1 ? d[1] : d[2]);
#pragma unroll
        for (int r = 0; r < 3; r++) vec[r][c] = o == 0 ? v[r][0] : (o == 1 ? v[r][1] : v[r][2]);
    }
    return true;
}

// geometry.rs:463-531
__device__ bool svd3(const double m[9], double u[9], double s[3], double vt[9]) {
    double mt[9], mtm9[9];
    transpose3(m, mt);
    mul3(mt, m, mtm9);
    double mtm[3][3], vm[3][3], ev[3];
#pragma unroll
    for (int i = 0; i < 9; i++) mtm[i / 3][i % 3] = mtm9[i];
    if (!jacobi3(mtm, ev, vm)) return false;
#pragma unroll
    for (int i = 0; i < 3; i++) s[i] = sqrt(ev[i] > 0.0 ? ev[i] : 0.0);
    auto mulmv = [&](const double vec[3], double out[3]) {
        const double o0 = m[0] * vec[0] + m[1] * vec[1] + m[2] * vec[2];
        const double o1 = m[3] * vec[0] + m[4] * vec[1] + m[5] * vec[2];
        const double o2 = m[6] * vec[0] + m[7] * vec[1] + m[8] * vec[2];
        out[0] = o0;
        out[1] = o1;
        out[2] = o2;
    };
    const double v0[3] = {vm[0][0], vm[1][0], vm[2][0]}, v1[3] = {vm[0][1], vm[1][1], vm[2][1]},
                 v2[3] = {vm[0][2], vm[1][2], vm[2][2]};
    double u0[3] = {1.0, 0.0, 0.0}, u1[3], u2[3], tmp[3];
    if (s[0] > 1e-12) {
        mulmv(v0, tmp);
        if (!normalize_vec3(tmp, u0)) {
            u0[0] = 1.0;
            u0[1] = 0.0;
            u0[2] = 0.0;
        }
    }
    bool have_u1 = false;
    if (s[1] > 1e-12) {
        double raw[3];
        mulmv(v1, raw);
        const double proj = dot3(raw, u0);
        const double orth[3] = {raw[0] - proj * u0[0], raw[1] - proj * u0[1], raw[2] - proj * u0[2]};
        have_u1 = normalize_vec3(orth, u1);
    }
    if (!have_u1) {
        double basis[3] = {0.0, 1.0, 0.0};
        if (fabs(u0[0]) < 0.9) {
            basis[0] = 1.0;
            basis[1] = 0.0;
        }
        double cr[3];
        cross3(u0, basis, cr);
        if (!normalize_vec3(cr, u1)) {
            u1[0] = 0.0;
            u1[1] = 1.0;
            u1[2] = 0.0;
        }
    }
    double c01[3];
    cross3(u0, u1, c01);
    if (!normalize_vec3(c01, u2)) return false;
    if (s[2] > 1e-12) {
        double mv2[3];
        mulmv(v2, mv2);
        if (dot3(u2, mv2) < 0.0) {
            u2[0] = -u2[0];
            u2[1] = -u2[1];
            u2[2] = -u2[2];
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        u[r * 3] = u0[r];
        u[r * 3 + 1] = u1[r];
        u[r * 3 + 2] = u2[r];
    }
    double vmf[9];
#pragma unroll
    for (int i = 0; i < 9; i++) vmf[i] = vm[i / 3][i % 3];
    transpose3(vmf, vt);
    return true;
}

// geometry.rs:533-544
__device__ bool rank2_project(double f[9]) {
    double u[9], s[3], vt[9];
    if (!svd3(f, u, s, vt)) return false;
    const double diag[9] = {s[0], 0.0, 0.0, 0.0, s[1], 0.0, 0.0, 0.0, 0.0};
    double dv[9];
    mul3(diag, vt, dv);
    mul3(u, dv, f);
    return normalize_frobenius(f);
}

// Per-thread 9x9 storage in LDS, element-major: element e of lane l at [e][l].
struct JacobiLds {
    double a[81][kFT];
    double v[81][kFT];
};

// geometry.rs:211-295 on lane `l`'s matrix (already in J.a): classical Jacobi (largest
// off-diagonal pivot, first in row-major order on ties), up to 500 rotations. On success
// the eigenvalue order (ascending, stable) is in ord[].
__device__ bool jacobi9_lane(JacobiLds &J, int l, int (&ord)[9]) {
#pragma unroll
    for (int i = 0; i < 81; i++) J.v[i][l] = (i / 9 == i % 9) ? 1.0 : 0.0;
    for (int it = 0; it < 500; it++) {
        int p = 0, q = 1;
        double mo = 0.0;
#pragma unroll
        for (int i = 0; i < 9; i++)
#pragma unroll
            for (int j = i + 1; j < 9; j++) {
                const double av = fabs(J.a[i * 9 + j][l]);
                if (av > mo) {
                    mo = av;
                    p = i;
                    q = j;
                }
            }
        if (!isfinite(mo)) return false;
        if (mo < 1e-13) break;
        const double app = J.a[p * 10][l], aqq = J.a[q * 10][l], apq = J.a[p * 9 + q][l];
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        J.a[p * 10][l] = app - t * apq;
        J.a[q * 10][l] = aqq + t * apq;
        J.a[p * 9 + q][l] = 0.0;
        J.a[q * 9 + p][l] = 0.0;
#pragma unroll
        for (int r = 0; r < 9; r++) {
            if (r != p && r != q) {
                const double arp = J.a[r * 9 + p][l], arq = J.a[r * 9 + q][l];
                const double np = c * arp - s * arq, nq = s * arp + c * arq;
                J.a[r * 9 + p][l] = np;
                J.a[p * 9 + r][l] = np;
                J.a[r * 9 + q][l] = nq;
                J.a[q * 9 + r][l] = nq;
            }
            const double vrp = J.v[r * 9 + p][l], vrq = J.v[r * 9 + q][l];
            J.v[r * 9 + p][l] = c * vrp - s * vrq;
            J.v[r * 9 + q][l] = s * vrp + c * vrq;
        }
    }
#pragma unroll
    for (int i = 0; i < 9; i++) ord[i] = i;
    // insertion sort, ascending: shift while the new key is less than the left one
    for (int i = 1; i < 9; i++) {
        const int k = ord[i];
        const double dk = J.a[k * 10][l];
        int j = i - 1;
        while (j >= 0 && dk < J.a[ord[j] * 10][l]) {
            ord[j + 1] = ord[j];
            j--;
        }
        ord[j + 1] = k;
    }
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 81; i++) ok = ok && isfinite(J.v[i][l]);
    return ok;
}

// Hartley normalisation of n points given by get(i) -> (x, y) (geometry.rs:547-583); writes
// (cx, cy, scale) and returns false like the reference's None.
template <typename Get>
__device__ bool normalize_params(uint32_t n, Get get, double &cx, double &cy, double &sc) {
    if (n == 0) return false;
    const double inv_n = 1.0 / (double)n;
    cx = 0.0;
    cy = 0.0;
#pragma unroll 8
    for (uint32_t i = 0; i < n; i++) {
        double x, y;
        get(i, x, y);
        if (!isfinite(x) || !isfinite(y)) return false;
        cx += x;
        cy += y;
    }
    cx *= inv_n;
    cy *= inv_n;
    double md = 0.0;
#pragma unroll 8
    for (uint32_t i = 0; i < n; i++) {
        double x, y;
        get(i, x, y);
        md += glibc_hypot(x - cx, y - cy);
    }
    md *= inv_n;
    if (!isfinite(md) || md < 1e-9) return false;
    sc = 1.41421356237309504880 / md;
    return true;
}

struct Norm {
    double qcx, qcy, qs, tcx, tcy, ts;
};

// row r of the 8-point system for the k-th point (geometry.rs:586-604)
__device__ __forceinline__ void fund_row(const Norm &N, const ps_feature_match &p, double r[9]) {
    const double x1 = ((double)p.qx - N.qcx) * N.qs, y1 = ((double)p.qy - N.qcy) * N.qs;
    const double x2 = ((double)p.tx - N.tcx) * N.ts, y2 = ((double)p.ty - N.tcy) * N.ts;
    r[0] = x2 * x1;
    r[1] = x2 * y1;
    r[2] = x2;
    r[3] = y2 * x1;
    r[4] = y2 * y1;
    r[5] = y2;
    r[6] = x1;
    r[7] = y1;
    r[8] = 1.0;
}

// From the eigenvector column `col` of lane l's Jacobi result to the denormalised,
// rank-2, Frobenius-normalised F (geometry.rs:645-663).
__device__ bool finish_8point(JacobiLds &J, int l, int col, const Norm &N, double f[9]) {
#pragma unroll
    for (int i = 0; i < 9; i++) f[i] = J.v[i * 9 + col][l];
    if (!rank2_project(f)) return false;
    const double tq[9] = {N.qs, 0.0, -N.qs * N.qcx, 0.0, N.qs, -N.qs * N.qcy, 0.0, 0.0, 1.0};
    const double tt[9] = {N.ts, 0.0, -N.ts * N.tcx, 0.0, N.ts, -N.ts * N.tcy, 0.0, 0.0, 1.0};
    double ttt[9], fq[9];
    transpose3(tt, ttt);
    mul3(f, tq, fq);
    mul3(ttt, fq, f);
    return normalize_frobenius(f);
}

// 8-point fit over the 8 matches m[idx[0..8)] by one lane (geometry.rs:638-664)
__device__ bool fit8_lane(const ps_feature_match *m, const uint32_t idx[8], JacobiLds &J, int l, double f[9]) {
    Norm N;
    if (!normalize_params(8, [&](uint32_t i, double &x, double &y) { x = (double)m[idx[i]].qx; y = (double)m[idx[i]].qy; },
                          N.qcx, N.qcy, N.qs))
        return false;
    if (!normalize_params(8, [&](uint32_t i, double &x, double &y) { x = (double)m[idx[i]].tx; y = (double)m[idx[i]].ty; },
                          N.tcx, N.tcy, N.ts))
        return false;
    double ata[45];
#pragma unroll
    for (int k = 0; k < 45; k++) ata[k] = 0.0;
    for (int k = 0; k < 8; k++) {
        double r[9];
        fund_row(N, m[idx[k]], r);
        int e = 0;
#pragma unroll
        for (int i = 0; i < 9; i++)
#pragma unroll
            for (int j = i; j < 9; j++) ata[e++] += r[i] * r[j];
    }
    int e = 0;
#pragma unroll
    for (int i = 0; i < 9; i++)
#pragma unroll
        for (int j = i; j < 9; j++) {
            J.a[i * 9 + j][l] = ata[e];
            J.a[j * 9 + i][l] = ata[e];
            e++;
        }
    int ord[9];
    if (!jacobi9_lane(J, l, ord)) return false;
    return finish_8point(J, l, ord[0], N, f);
}

// (count, truncated error) of F over all n matches, sequentially (geometry.rs:808-828)
__device__ uint32_t eval_f_seq(const ps_feature_match *m, uint32_t n, const double *f, double tsq, double *err) {
    uint32_t c = 0;
    double e = 0.0;
#pragma unroll 8
    for (uint32_t i = 0; i < n; i++) {
        const double d2 = msampson(f, m[i]);
        if (d2 <= tsq) {
            c++;
            e += d2;
        } else {
            e += tsq;
        }
    }
    *err = e;
    return c;
}

// geometry.rs:831-852 (powi by binary exponentiation, as compiler-rt)
__device__ uint32_t adaptive_epi(uint32_t best, uint32_t total, int s, double conf, uint32_t min_it, uint32_t cap) {
    if (total <= (uint32_t)s || best == 0) return cap;
    const double w = (double)best / (double)total;
    if (w >= 0.9999) return min_it;
    double r = 1.0, a = w;
    int b = s;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    if (r <= 1e-12) return cap;
    const double d = 1.0 - r;
    const double den = log(d > 1e-15 ? d : 1e-15);
    const double cc = conf < 0.5 ? 0.5 : (conf > 0.999999 ? 0.999999 : conf);
    const double num = log(1.0 - cc);
    const double it = ceil(num / den);
    double k = it > 0.0 ? it : 0.0;   // f64 -> usize saturates, NaN -> 0
    if (k < (double)min_it) k = (double)min_it;
    if (k > (double)cap) k = (double)cap;
    return (uint32_t)k;
}

// geometry.rs:667-724 (device cbrt / acos / cos)
__device__ int solve_cubic(double c3, double c2, double c1, double c0, double roots[3]) {
    int nr = 0;
    if (fabs(c3) < 1e-12) {
        if (fabs(c2) < 1e-12) {
            if (fabs(c1) < 1e-12) return 0;
            roots[0] = -c0 / c1;
            return 1;
        }
        const double disc = c1 * c1 - 4.0 * c2 * c0;
        if (disc < 0.0) return 0;
        const double sq = sqrt(disc);
        roots[0] = (-c1 + sq) / (2.0 * c2);
        roots[1] = (-c1 - sq) / (2.0 * c2);
        return 2;
    }
    const double a = c2 / c3, b = c1 / c3, c = c0 / c3;
    const double a3 = a / 3.0;
    const double p = b - a * a3;
    const double q = (2.0 * a * a * a) / 27.0 - (a * b) / 3.0 + c;
    const double p3 = p / 3.0;
    const double disc = (q * 0.5) * (q * 0.5) + p3 * (p3 * p3);
    if (disc > 1e-14) {
        const double sq = sqrt(disc);
        const double u = cbrt(-0.5 * q + sq), v = cbrt(-0.5 * q - sq);
        roots[nr++] = u + v - a3;
    } else if (disc >= -1e-14) {
        const double u = cbrt(-0.5 * q);
        roots[nr++] = 2.0 * u - a3;
        roots[nr++] = -u - a3;
    } else {
        const double r = sqrt(-p / 3.0);
        double cp = (-0.5 * q) / (r * r * r);
        if (cp < -1.0) cp = -1.0;
        if (cp > 1.0) cp = 1.0;
        const double phi = acos(cp), two_r = 2.0 * r;
        const double pi = 3.14159265358979323846;
        roots[nr++] = two_r * cos(phi / 3.0) - a3;
        roots[nr++] = two_r * cos((phi + 2.0 * pi) / 3.0) - a3;
        roots[nr++] = two_r * cos((phi + 4.0 * pi) / 3.0) - a3;
    }
    for (int k = 0; k < nr; k++)
        for (int it = 0; it < 3; it++) {
            const double t = roots[k];
            const double fv = ((c3 * t + c2) * t + c1) * t + c0;
            const double df = (3.0 * c3 * t + 2.0 * c2) * t + c1;
            if (fabs(df) > 1e-14) roots[k] = t - fv / df;
        }
    return nr;
}

// geometry.rs:727-805 on m[0..7) by lane l; returns the candidate count (<= 4)
__device__ int fit7_lane(const ps_feature_match *m, JacobiLds &J, int l, double out[4][9]) {
    Norm N;
    if (!normalize_params(7, [&](uint32_t i, double &x, double &y) { x = (double)m[i].qx; y = (double)m[i].qy; }, N.qcx,
                          N.qcy, N.qs))
        return 0;
    if (!normalize_params(7, [&](uint32_t i, double &x, double &y) { x = (double)m[i].tx; y = (double)m[i].ty; }, N.tcx,
                          N.tcy, N.ts))
        return 0;
    double ata[45];
    for (int k = 0; k < 45; k++) ata[k] = 0.0;
    for (int k = 0; k < 7; k++) {
        double r[9];
        fund_row(N, m[k], r);
        int e = 0;
        for (int i = 0; i < 9; i++)
            for (int j = i; j < 9; j++) ata[e++] += r[i] * r[j];
    }
    int e = 0;
    for (int i = 0; i < 9; i++)
        for (int j = i; j < 9; j++) {
            J.a[i * 9 + j][l] = ata[e];
            J.a[j * 9 + i][l] = ata[e];
            e++;
        }
    int ord[9];
    if (!jacobi9_lane(J, l, ord)) return 0;
    double f1[9], f2[9], b[9];
    for (int i = 0; i < 9; i++) {
        f1[i] = J.v[i * 9 + ord[0]][l];
        f2[i] = J.v[i * 9 + ord[1]][l];
    }
    for (int i = 0; i < 9; i++) b[i] = f1[i] - f2[i];
    auto eval_mat = [&](double t, double o[9]) {
        for (int i = 0; i < 9; i++) o[i] = f2[i] + t * b[i];
    };
    double e1[9], em1[9], e2[9];
    eval_mat(1.0, e1);
    eval_mat(-1.0, em1);
    eval_mat(2.0, e2);
    const double p0 = det3(f2), p1 = det3(e1), pm1 = det3(em1), p2 = det3(e2);
    const double c0 = p0, c2 = 0.5 * (p1 + pm1) - c0, s1 = 0.5 * (p1 - pm1), s2 = p2 - c0 - 4.0 * c2;
    const double c3 = (s2 - 2.0 * s1) / 6.0, c1 = s1 - c3;
    double roots[3];
    const int nr = solve_cubic(c3, c2, c1, c0, roots);
    const double tq[9] = {N.qs, 0.0, -N.qs * N.qcx, 0.0, N.qs, -N.qs * N.qcy, 0.0, 0.0, 1.0};
    const double tt[9] = {N.ts, 0.0, -N.ts * N.tcx, 0.0, N.ts, -N.ts * N.tcy, 0.0, 0.0, 1.0};
    double ttt[9];
    transpose3(tt, ttt);
    int nout = 0;
    for (int k = 0; k < nr; k++) {
        if (!isfinite(roots[k])) continue;
        double fc[9], fq[9], f[9];
        eval_mat(roots[k], fc);
        mul3(fc, tq, fq);
        mul3(ttt, fq, f);
        if (normalize_frobenius(f)) {
            for (int i = 0; i < 9; i++) out[nout][i] = f[i];
            nout++;
        }
    }
    if (fabs(c3) < 1e-12 && fabs(det3(b)) < 1e-10) {
        double bq[9], f[9];
        mul3(b, tq, bq);
        mul3(ttt, bq, f);
        if (normalize_frobenius(f)) {
            for (int i = 0; i < 9; i++) out[nout][i] = f[i];
            nout++;
        }
    }
    return nout;
}

struct FundShared {
    JacobiLds J;
    double hf[kFT][9];           // hypothesis models
    double herr[kFT];
    uint32_t hcnt[kFT], hok[kFT];
    uint32_t quad[kFT][8], qvalid[kFT];
    uint32_t draws[1024 + 32];   // sampler draws plus one sample's worth of lookahead
    uint8_t len[1024];           // draws consumed by the sample starting at each position
    uint32_t start[kFT];         // buffer position of each sample of the batch
    uint32_t n_quads, n_draws, draw_pos, wave_cnt;
    double best[9], be;
    uint32_t bc, have, ok, ni;
    double norm[6];
    // chunk staging for the order-dependent sums: all lanes load / compute per-point terms,
    // lane 0 adds them in point order
    double stage[4][kFStage];
    uint32_t bad;
};

// entry j of the 8-point row (geometry.rs:586-604) of a normalised correspondence, selected
// without a dynamically indexed array
__device__ __forceinline__ double fund_row_at(int j, double x1, double y1, double x2, double y2) {
    const double v[9] = {x2 * x1, x2 * y1, x2, y2 * x1, y2 * y1, y2, x1, y1, 1.0};
    double r = v[0];
#pragma unroll
    for (int i = 1; i < 9; i++) r = j == i ? v[i] : r;
    return r;
}

// ordered compaction (single wave): indices i in [0, n) with pred(i), in increasing order
template <typename Pred>
__device__ uint32_t compact_wave(uint32_t n, uint32_t *out, Pred pred) {
    const int lane = threadIdx.x;
    uint32_t base = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += kFT) {
        const uint32_t i = c0 + lane;
        const bool p = i < n && pred(i);
        const uint64_t mask = __ballot(p);
        if (p) out[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = i;
        base += (uint32_t)__popcll(mask);
    }
    return base;
}

// 8-point fit over m[idx[0..ni)] by the whole wave (geometry.rs:638-664): normalisation by
// lane 0, the 45 AtA entries each summed in point order by one lane, the Jacobi and the
// rank-2 step on lane 0's LDS matrices. Result in S.best-like storage `fout` (shared).
__device__ bool fit8_block(const ps_feature_match *m, const uint32_t *idx, uint32_t ni, FundShared &S, double *fout) {
    const int lane = threadIdx.x;
    // Hartley normalisation of both point sets (geometry.rs:547-583): the points and their
    // glibc-hypot distances staged in chunks, lane 0 adding them in point order; a non-finite
    // point, or a degenerate mean distance, of either set fails the fit
    if (lane == 0) S.bad = ni == 0 ? 1u : 0u;
    __syncthreads();
    double cq[2] = {0.0, 0.0}, ct[2] = {0.0, 0.0};   // lane 0's running sums
    for (uint32_t c0 = 0; c0 < ni; c0 += kFStage) {
        const uint32_t cn = min(kFStage, ni - c0);
        for (uint32_t k = lane; k < cn; k += kFT) {
            const ps_feature_match &p = m[idx[c0 + k]];
            const double qx = p.qx, qy = p.qy, tx = p.tx, ty = p.ty;
            if (!isfinite(qx) || !isfinite(qy) || !isfinite(tx) || !isfinite(ty)) S.bad = 1u;
            S.stage[0][k] = qx;
            S.stage[1][k] = qy;
            S.stage[2][k] = tx;
            S.stage[3][k] = ty;
        }
        __syncthreads();
        if (lane == 0 && !S.bad) {
#pragma unroll 8
            for (uint32_t k = 0; k < cn; k++) {
                cq[0] += S.stage[0][k];
                cq[1] += S.stage[1][k];
                ct[0] += S.stage[2][k];
                ct[1] += S.stage[3][k];
            }
        }
        __syncthreads();
    }
    if (S.bad) {
        __syncthreads();
        return false;
    }
    const double inv_n = 1.0 / (double)ni;
    if (lane == 0) {
        S.norm[0] = cq[0] * inv_n;
        S.norm[1] = cq[1] * inv_n;
        S.norm[3] = ct[0] * inv_n;
        S.norm[4] = ct[1] * inv_n;
    }
    __syncthreads();
    {
        const double qcx = S.norm[0], qcy = S.norm[1], tcx = S.norm[3], tcy = S.norm[4];
        double mdq = 0.0, mdt = 0.0;
        for (uint32_t c0 = 0; c0 < ni; c0 += kFStage) {
            const uint32_t cn = min(kFStage, ni - c0);
            for (uint32_t k = lane; k < cn; k += kFT) {
                const ps_feature_match &p = m[idx[c0 + k]];
                S.stage[0][k] = glibc_hypot((double)p.qx - qcx, (double)p.qy - qcy);
                S.stage[1][k] = glibc_hypot((double)p.tx - tcx, (double)p.ty - tcy);
            }
            __syncthreads();
            if (lane == 0) {
#pragma unroll 8
                for (uint32_t k = 0; k < cn; k++) {
                    mdq += S.stage[0][k];
                    mdt += S.stage[1][k];
                }
            }
            __syncthreads();
        }
        if (lane == 0) {
            mdq *= inv_n;
            mdt *= inv_n;
            const bool ok = isfinite(mdq) && !(mdq < 1e-9) && isfinite(mdt) && !(mdt < 1e-9);
            S.norm[2] = 1.41421356237309504880 / mdq;
            S.norm[5] = 1.41421356237309504880 / mdt;
            S.ok = ok;
        }
        __syncthreads();
    }
    if (!S.ok) {
        __syncthreads();
        return false;
    }
    const Norm N{S.norm[0], S.norm[1], S.norm[2], S.norm[3], S.norm[4], S.norm[5]};
    {
        int a = 0, r = lane < 45 ? lane : 0;
        while (r >= 9 - a) {
            r -= 9 - a;
            a++;
        }
        const int b = a + r;
        double acc = 0.0;
        for (uint32_t c0 = 0; c0 < ni; c0 += kFStage) {
            const uint32_t cn = min(kFStage, ni - c0);
            for (uint32_t k = lane; k < cn; k += kFT) {
                const ps_feature_match &p = m[idx[c0 + k]];
                S.stage[0][k] = ((double)p.qx - N.qcx) * N.qs;
                S.stage[1][k] = ((double)p.qy - N.qcy) * N.qs;
                S.stage[2][k] = ((double)p.tx - N.tcx) * N.ts;
                S.stage[3][k] = ((double)p.ty - N.tcy) * N.ts;
            }
            __syncthreads();
            if (lane < 45) {
#pragma unroll 4
                for (uint32_t k = 0; k < cn; k++) {
                    const double x1 = S.stage[0][k], y1 = S.stage[1][k], x2 = S.stage[2][k], y2 = S.stage[3][k];
                    acc += fund_row_at(a, x1, y1, x2, y2) * fund_row_at(b, x1, y1, x2, y2);
                }
            }
            __syncthreads();
        }
        if (lane < 45) {
            S.J.a[a * 9 + b][0] = acc;
            S.J.a[b * 9 + a][0] = acc;
        }
    }
    __syncthreads();
    if (lane == 0) {
        int ord[9];
        double f[9];
        bool ok = jacobi9_lane(S.J, 0, ord) && finish_8point(S.J, 0, ord[0], N, f);
        if (ok)
            for (int i = 0; i < 9; i++) fout[i] = f[i];
        S.ok = ok;
    }
    __syncthreads();
    const bool ok = S.ok;
    __syncthreads();
    return ok;
}

// count / truncated error of the model in fsh over all n matches: per-match values computed
// in parallel into the staging buffer, chunk by chunk, lane 0 adding them in index order
__device__ uint32_t eval_f_block(const ps_feature_match *m, uint32_t n, const double *fsh, double tsq, FundShared &S,
                                 double *err) {
    const int lane = threadIdx.x;
    double f[9];
    for (int i = 0; i < 9; i++) f[i] = fsh[i];
    __shared__ double s_e;
    __shared__ uint32_t s_c;
    uint32_t c = 0;   // lane 0
    double e = 0.0;
    for (uint32_t c0 = 0; c0 < n; c0 += kFStage) {
        const uint32_t cn = min(kFStage, n - c0);
        for (uint32_t k = lane; k < cn; k += kFT) S.stage[0][k] = msampson(f, m[c0 + k]);
        __syncthreads();
        if (lane == 0) {
#pragma unroll 8
            for (uint32_t k = 0; k < cn; k++) {
                const double d2 = S.stage[0][k];
                if (d2 <= tsq) {
                    c++;
                    e += d2;
                } else {
                    e += tsq;
                }
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        s_c = c;
        s_e = e;
    }
    __syncthreads();
    *err = s_e;
    const uint32_t cc = s_c;
    __syncthreads();
    return cc;
}

// The sample that starts at draws[0]: distinct match indices in draw order until K of them
// (valid) or A attempts (invalid); returns the number of draws it consumes. The sampler's
// sequential grouping is the chain of such samples, each starting where the previous ended.
template <int K, int A>
__device__ __forceinline__ uint32_t scan_sample(const uint32_t *draws, uint32_t (&q)[K], bool &valid) {
    uint32_t ch = 0, at = 0;
#pragma unroll
    for (int j = 0; j < K; j++) q[j] = 0;
    while (true) {
        const uint32_t c = draws[at++];
        bool dup = false;
#pragma unroll
        for (int j = 0; j < K; j++) dup |= (uint32_t)j < ch && q[j] == c;
        if (!dup) {
#pragma unroll
            for (int j = 0; j < K; j++)
                if ((uint32_t)j == ch) q[j] = c;
            ch++;
        }
        if (ch == (uint32_t)K || at == (uint32_t)A) break;
    }
    valid = ch == (uint32_t)K;
    return at;
}

__global__ __launch_bounds__(kFT) void fundamental_kernel(const ps_feature_match *matches,
                                                          const ps_fundamental_problem *probs, double thresh_in,
                                                          uint32_t max_iters, double confidence, double *work_d,
                                                          uint32_t *work_u, uint32_t *inliers,
                                                          ps_fundamental_result *outs) {
    __shared__ FundShared S;
    const int lane = threadIdx.x;
    const ps_fundamental_problem P = probs[blockIdx.x];
    const uint32_t n = P.n;
    const ps_feature_match *m = matches + P.offset;
    double *cdf = work_d + (size_t)P.offset * 2;
    uint32_t *ibuf = work_u + P.offset;
    uint32_t *out_inl = inliers + P.offset;
    ps_fundamental_result &O = outs[blockIdx.x];
    if (lane == 0) {
        O.has_model = 0;
        O.n_inliers = 0;
        O.iters = 0;
        for (int i = 0; i < 9; i++) O.f[i] = 0.0;
    }
    if (n < 7) return;
    double th = thresh_in > 0.25 ? thresh_in : 0.25;   // f64::max(x, 0.25), NaN -> 0.25
    if (thresh_in != thresh_in) th = 0.25;
    const double tsq = th * th;

    if (n == 7) {   // geometry.rs:880-906
        if (lane == 0) {
            double cand[4][9];
            const int nc = fit7_lane(m, S.J, 0, cand);
            uint32_t bc = 0;
            double be = INFINITY;
            int have = -1;
            for (int k = 0; k < nc; k++) {
                double e;
                const uint32_t c = eval_f_seq(m, n, cand[k], tsq, &e);
                if (c > bc || (c == bc && e < be)) {
                    bc = c;
                    be = e;
                    have = k;
                }
            }
            S.have = have >= 0;
            if (have >= 0)
                for (int i = 0; i < 9; i++) S.best[i] = cand[have][i];
        }
        __syncthreads();
        if (!S.have) return;
        const uint32_t ni = compact_wave(n, out_inl, [&](uint32_t i) { return msampson(S.best, m[i]) <= tsq; });
        if (lane == 0) {
            O.has_model = 1;
            O.n_inliers = ni;
            for (int i = 0; i < 9; i++) O.f[i] = S.best[i];
        }
        return;
    }

    // reference weights 1 / clamp(ratio, 0.05, 1), prefix sums in f64 in order (pipeline.rs:2006-2028)
    if (lane == 0) {
        double acc = 0.0;
        for (uint32_t i = 0; i < n; i++) {
            const float rr = m[i].ratio;
            double r = 1.0;
            if (isfinite(rr)) {
                r = (double)rr;
                if (r < 0.05) r = 0.05;
                if (r > 1.0) r = 1.0;
            }
            acc += 1.0 / r;
            cdf[i] = acc;
        }
        if (acc > 0.0) {
            const double inv = 1.0 / acc;
            for (uint32_t i = 0; i < n; i++) cdf[i] *= inv;
        }
        cdf[n - 1] = 1.0;
        S.have = 0;
        S.bc = 0;
        S.be = INFINITY;
    }
    __syncthreads();
    const uint32_t cap = max_iters > 64 ? max_iters : 64, min_it = 64 < cap ? 64 : cap;
    uint32_t max_it = cap, it = 0;
    // the SplitMix64 stream: the buffer holds draws [base, base + 1024 + 32) and the next sample
    // starts at stream position pos
    uint64_t base = 0, pos = 0;
    bool filled = false;
    while (it < max_it) {
        // 1. the next kFT samples (geometry.rs:929-940: 8 distinct matches within 32 draws), the
        // sequential grouping done in parallel: every buffer position gets the length of a sample
        // starting there, lane 0 follows the chain of sample starts, each lane rebuilds one sample
        uint32_t nq = 0;
        while (nq < (uint32_t)kFT) {
            if (!filled || pos >= base + 1024) {
                base = pos;
                filled = true;
                for (uint32_t k = lane; k < 1024 + 32; k += kFT) {
                    const uint64_t z = fsplitmix_at(P.seed, base + k);
                    S.draws[k] = fcdf_pick(cdf, n, (double)(z >> 11) * (1.0 / 9007199254740992.0));
                }
                __syncthreads();
                for (uint32_t k = lane; k < 1024; k += kFT) {
                    uint32_t qq[8];
                    bool v;
                    S.len[k] = (uint8_t)scan_sample<8, 32>(S.draws + k, qq, v);
                }
                __syncthreads();
            }
            if (lane == 0) {
                uint64_t p = pos;
                uint32_t k = nq;
                while (p < base + 1024 && k < (uint32_t)kFT) {
                    S.start[k++] = (uint32_t)(p - base);
                    p += S.len[p - base];
                }
                S.n_quads = k;
                S.draw_pos = (uint32_t)(p - base);
            }
            __syncthreads();
            const uint32_t nq2 = S.n_quads;
            if ((uint32_t)lane >= nq && (uint32_t)lane < nq2) {
                uint32_t qq[8];
                bool v;
                scan_sample<8, 32>(S.draws + S.start[lane], qq, v);
                S.qvalid[lane] = v;
#pragma unroll
                for (int j = 0; j < 8; j++) S.quad[lane][j] = qq[j];
            }
            pos = base + S.draw_pos;
            nq = nq2;
            __syncthreads();
        }
        // 2. fit and score one hypothesis per lane
        {
            S.hok[lane] = 0;
            // the adaptive bound only shrinks as the best improves: past it nothing is replayed
            if (S.qvalid[lane] && it + (uint32_t)lane < max_it) {
                uint32_t idx[8];
                for (int j = 0; j < 8; j++) idx[j] = S.quad[lane][j];
                double f[9];
                if (fit8_lane(m, idx, S.J, lane, f)) {
                    double e;
                    S.hcnt[lane] = eval_f_seq(m, n, f, tsq, &e);
                    S.herr[lane] = e;
                    for (int i = 0; i < 9; i++) S.hf[lane][i] = f[i];
                    S.hok[lane] = 1;
                }
            }
        }
        __syncthreads();
        // 3. replay in order (geometry.rs:927-972)
        if (lane == 0) {
            uint32_t bc = S.bc, mx = max_it, k = 0, itl = it;
            double be = S.be;
            for (; k < (uint32_t)kFT && itl < mx; k++) {
                itl++;
                if (!S.hok[k]) continue;
                const uint32_t c = S.hcnt[k];
                const double e = S.herr[k];
                if (c > bc || (c == bc && e < be)) {
                    bc = c;
                    be = e;
                    for (int i = 0; i < 9; i++) S.best[i] = S.hf[k][i];
                    S.have = 1;
                    mx = adaptive_epi(bc, n, 8, confidence, min_it, cap);
                }
            }
            S.bc = bc;
            S.be = be;
            S.n_quads = mx;   // broadcast the bound
            S.n_draws = itl;  // and the iteration count
        }
        __syncthreads();
        max_it = S.n_quads;
        it = S.n_draws;
        __syncthreads();
    }
    if (lane == 0) O.iters = it;
    if (!S.have) return;
    // LO-RANSAC refit on the consensus set, twice (geometry.rs:979-1004)
    for (int lo = 0; lo < 2; lo++) {
        const uint32_t ni = compact_wave(n, ibuf, [&](uint32_t i) { return msampson(S.best, m[i]) <= tsq; });
        if (ni < 8) break;
        __shared__ double s_ref[9];
        if (!fit8_block(m, ibuf, ni, S, s_ref)) break;
        double e;
        const uint32_t c = eval_f_block(m, n, s_ref, tsq, S, &e);
        bool take = c > S.bc || (c == S.bc && e < S.be);
        __syncthreads();
        if (!take) break;
        if (lane == 0) {
            S.bc = c;
            S.be = e;
            for (int i = 0; i < 9; i++) S.best[i] = s_ref[i];
        }
        __syncthreads();
    }
    const uint32_t ni = compact_wave(n, out_inl, [&](uint32_t i) { return msampson(S.best, m[i]) <= tsq; });
    if (lane == 0) {
        O.has_model = 1;
        O.n_inliers = ni;
        for (int i = 0; i < 9; i++) O.f[i] = S.best[i];
    }
}


__device__ __forceinline__ double fmin_rust(double a, double b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); }

// geometry.rs:1233-1341, one wave per problem: per-match terms in parallel into LDS chunks,
// lane 0 adds them in match order.
__global__ __launch_bounds__(kFT) void gric_kernel(const ps_feature_match *matches, const ps_gric_problem *probs,
                                                   double sigma, ps_gric_result *outs) {
    __shared__ double s_h[kFT], s_f[kFT];
    const int lane = threadIdx.x;
    const ps_gric_problem P = probs[blockIdx.x];
    const ps_feature_match *m = matches + P.offset;
    const uint32_t n = P.n;
    ps_gric_result &O = outs[blockIdx.x];
    double sg = sigma > 0.25 ? sigma : 0.25;
    if (sigma != sigma) sg = 0.25;
    const double s2 = sg * sg, r = 4.0, l1 = 2.0, l2 = 2.0;
    const double caph = l1 * (r - 2.0), capf = l1 * (r - 3.0);
    const bool hh = P.has_h, hf = P.has_f;
    double hd[9], inv[9], f[9];
    bool has_inv = false;
    for (int i = 0; i < 9; i++) {
        hd[i] = (double)P.h[i];
        f[i] = P.f[i];
    }
    if (hh) has_inv = invert3_f32(P.h, inv);
    double dh = 0.0, df = 0.0;
    for (uint32_t c0 = 0; c0 < n; c0 += kFT) {
        const uint32_t i = c0 + lane;
        if (i < n) {
            const double qx = m[i].qx, qy = m[i].qy, tx = m[i].tx, ty = m[i].ty;
            if (hh) {
                const double fw = hd[6] * qx + hd[7] * qy + hd[8];
                double fwd, bwd;
                if (fabs(fw) > 1e-12) {
                    const double px = (hd[0] * qx + hd[1] * qy + hd[2]) / fw, py = (hd[3] * qx + hd[4] * qy + hd[5]) / fw;
                    fwd = (px - tx) * (px - tx) + (py - ty) * (py - ty);
                } else {
                    fwd = INFINITY;
                }
                if (has_inv) {
                    const double bw = inv[6] * tx + inv[7] * ty + inv[8];
                    if (fabs(bw) > 1e-12) {
                        const double px = (inv[0] * tx + inv[1] * ty + inv[2]) / bw;
                        const double py = (inv[3] * tx + inv[4] * ty + inv[5]) / bw;
                        bwd = (px - qx) * (px - qx) + (py - qy) * (py - qy);
                    } else {
                        bwd = INFINITY;
                    }
                } else {
                    bwd = fwd;
                }
                s_h[lane] = fmin_rust(0.5 * (fwd + bwd) / s2, caph);
            }
            if (hf) s_f[lane] = fmin_rust(sampson_sq(f, qx, qy, tx, ty) / s2, capf);
        }
        __syncthreads();
        if (lane == 0) {
            const uint32_t cnt = n - c0 < (uint32_t)kFT ? n - c0 : (uint32_t)kFT;
            for (uint32_t j = 0; j < cnt; j++) {
                if (hh) dh += s_h[j];
                if (hf) df += s_f[j];
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        const double hg = hh ? dh + l1 * 2.0 * (double)n + l2 * 8.0 : INFINITY;
        const double fg = hf ? df + l1 * 3.0 * (double)n + l2 * 7.0 : INFINITY;
        const bool some = n > 0 && (isfinite(hg) || isfinite(fg));
        O.has_comparison = some ? 1 : 0;
        O.selected = fg < hg ? 1 : 0;
        O.homography_gric = hg;
        O.fundamental_gric = fg;
    }
}

}  // namespace

size_t fundamental_work_bytes(size_t total_matches) {
    return total_matches * 2 * sizeof(double) + total_matches * sizeof(uint32_t) + 512;
}

hipError_t launch_fundamental(const ps_feature_match *matches, const ps_fundamental_problem *probs, uint32_t n_probs,
                              double sampson_threshold, uint32_t max_iters, double confidence, size_t total_matches,
                              void *work, uint32_t *inliers, ps_fundamental_result *outs, hipStream_t st) {
    if (n_probs == 0) return hipSuccess;
    double *wd = static_cast<double *>(work);
    uint32_t *wu = reinterpret_cast<uint32_t *>(wd + total_matches * 2);
    fundamental_kernel<<<n_probs, kFT, 0, st>>>(matches, probs, sampson_threshold, max_iters, confidence, wd, wu,
                                                inliers, outs);
    return hipGetLastError();
}

hipError_t launch_gric(const ps_feature_match *matches, const ps_gric_problem *probs, uint32_t n_probs, double sigma,
                       ps_gric_result *outs, hipStream_t st) {
    if (n_probs == 0) return hipSuccess;
    gric_kernel<<<n_probs, kFT, 0, st>>>(matches, probs, sigma, outs);
    return hipGetLastError();
}

}  // namespace ps
