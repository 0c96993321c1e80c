/*
 * ps_oracle_geometry.c -- CPU oracle (test infrastructure) for the two-view geometry the
 * per-pair match post-processing runs (src/cli.rs:290-318): fundamental-matrix LO-RANSAC
 * with the normalised 8-point and 7-point solvers (src/geometry.rs:145-1016) and the GRIC
 * model comparison (src/geometry.rs:1233-1341, with src/visualize.rs:202-232
 * invert_matrix_3x3). Never linked into the product library.
 *
 * Numerics follow the Rust f64 code: no contraction, operations in source order, glibc libm
 * for hypot / ln / cbrt / acos / cos (the libm Rust binds on Linux), stable sorts, sequential
 * sums. How each part is pinned: tests/test_oracle_geometry.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ps_oracle.h"

#define PI_D 3.14159265358979323846

/* ------------------------------------------------------------------ small linear algebra */

static void mul3(const double a[3][3], const double b[3][3], double out[3][3]) {   /* geometry.rs:145-153 */
    double o[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) o[r][c] = a[r][0] * b[0][c] + a[r][1] * b[1][c] + a[r][2] * b[2][c];
    memcpy(out, o, sizeof o);
}

static void transpose3(const double m[3][3], double out[3][3]) {
    double o[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) o[r][c] = m[c][r];
    memcpy(out, o, sizeof o);
}

static double det3(const double m[3][3]) {   /* geometry.rs:165-169 */
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

static int normalize_frobenius(double m[3][3]) {   /* geometry.rs:172-184 */
    double s = 0.0;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) s += m[r][c] * m[r][c];
    if (!isfinite(s) || s < 1e-24) return 0;
    double inv = 1.0 / sqrt(s);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) m[r][c] *= inv;
    return 1;
}

static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void cross3(const double a[3], const double b[3], double o[3]) {
    double t[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    memcpy(o, t, sizeof t);
}

static int normalize_vec3(const double v[3], double o[3]) {   /* geometry.rs:201-207 */
    double n = sqrt(dot3(v, v));
    if (!isfinite(n) || n < 1e-12) return 0;
    double t[3] = {v[0] / n, v[1] / n, v[2] / n};
    memcpy(o, t, sizeof t);
    return 1;
}

/* stable insertion sort of index array by key (asc when dir > 0, desc when dir < 0);
 * partial_cmp with unordered -> Equal */
static void stable_order(const double *key, int *order, int n, int dir) {
    for (int i = 0; i < n; i++) order[i] = i;
    for (int i = 1; i < n; i++) {
        int k = order[i], j = i - 1;
        while (j >= 0) {
            double a = key[order[j]], b = key[k];
            int greater = dir > 0 ? (a > b) : (b > a);
            if (!greater) break;
            order[j + 1] = order[j];
            j--;
        }
        order[j + 1] = k;
    }
}

/* geometry.rs:211-295: classical Jacobi (largest off-diagonal pivot, first in row-major order
 * on ties), up to 500 rotations, eigenpairs sorted by ascending eigenvalue */
int psr_jacobi9(const double ain[81], double eigvals[9], double eigvecs[81]) {
    double a[9][9], v[9][9];
    memcpy(a, ain, sizeof a);
    for (int i = 0; i < 9; i++)
        for (int j = 0; j < 9; j++) v[i][j] = i == j ? 1.0 : 0.0;
    for (int it = 0; it < 500; it++) {
        int p = 0, q = 1;
        double mo = 0.0;
        for (int i = 0; i < 9; i++)
            for (int j = i + 1; j < 9; j++) {
                double av = fabs(a[i][j]);
                if (av > mo) { mo = av; p = i; q = j; }
            }
        if (!isfinite(mo)) return 0;
        if (mo < 1e-13) break;
        double app = a[p][p], aqq = a[q][q], apq = a[p][q];
        double tau = (aqq - app) / (2.0 * apq);
        double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        a[p][p] = app - t * apq;
        a[q][q] = aqq + t * apq;
        a[p][q] = 0.0;
        a[q][p] = 0.0;
        for (int r = 0; r < 9; r++)
            if (r != p && r != q) {
                double arp = a[r][p], arq = a[r][q];
                double np = c * arp - s * arq, nq = s * arp + c * arq;
                a[r][p] = np; a[p][r] = np;
                a[r][q] = nq; a[q][r] = nq;
            }
        for (int r = 0; r < 9; r++) {
            double vrp = v[r][p], vrq = v[r][q];
            v[r][p] = c * vrp - s * vrq;
            v[r][q] = s * vrp + c * vrq;
        }
    }
    double diag[9];
    int order[9];
    for (int i = 0; i < 9; i++) diag[i] = a[i][i];
    stable_order(diag, order, 9, 1);
    for (int d = 0; d < 9; d++) {
        eigvals[d] = a[order[d]][order[d]];
        for (int r = 0; r < 9; r++) {
            double val = v[r][order[d]];
            if (!isfinite(val)) return 0;
            eigvecs[r * 9 + d] = val;
        }
    }
    return 1;
}

/* geometry.rs:297-373 (100 rotations, 1e-14, sorted descending) */
static int jacobi3(double a[3][3], double eigvals[3], double eigvecs[3][3]) {
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int it = 0; it < 100; it++) {
        int p = 0, q = 1;
        double mo = 0.0;
        for (int i = 0; i < 3; i++)
            for (int j = i + 1; j < 3; j++)
                if (fabs(a[i][j]) > mo) { mo = fabs(a[i][j]); p = i; q = j; }
        if (!isfinite(mo)) return 0;
        if (mo < 1e-14) break;
        double app = a[p][p], aqq = a[q][q], apq = a[p][q];
        double tau = (aqq - app) / (2.0 * apq);
        double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        a[p][p] = app - t * apq;
        a[q][q] = aqq + t * apq;
        a[p][q] = 0.0;
        a[q][p] = 0.0;
        for (int r = 0; r < 3; r++)
            if (r != p && r != q) {
                double arp = a[r][p], arq = a[r][q];
                double np = c * arp - s * arq, nq = s * arp + c * arq;
                a[r][p] = np; a[p][r] = np;
                a[r][q] = nq; a[q][r] = nq;
            }
        for (int r = 0; r < 3; r++) {
            double vrp = v[r][p], vrq = v[r][q];
            v[r][p] = c * vrp - s * vrq;
            v[r][q] = s * vrp + c * vrq;
        }
    }
    double diag[3] = {a[0][0], a[1][1], a[2][2]};
    int order[3];
    stable_order(diag, order, 3, -1);
    for (int d = 0; d < 3; d++) {
        eigvals[d] = a[order[d]][order[d]];
        for (int r = 0; r < 3; r++) eigvecs[r][d] = v[r][order[d]];
    }
    return 1;
}

/* geometry.rs:463-531 */
static int svd3(const double m[3][3], double u[3][3], double s[3], double vt[3][3]) {
    double mt[3][3], mtm[3][3], ev[3], vm[3][3];
    transpose3(m, mt);
    mul3(mt, m, mtm);
    if (!jacobi3(mtm, ev, vm)) return 0;
    for (int i = 0; i < 3; i++) s[i] = sqrt(ev[i] > 0.0 ? ev[i] : 0.0);   /* max(0.0) then sqrt */
    double v0[3] = {vm[0][0], vm[1][0], vm[2][0]}, v1[3] = {vm[0][1], vm[1][1], vm[2][1]},
           v2[3] = {vm[0][2], vm[1][2], vm[2][2]};
#define MULMV(vec, out)                                                             \
    do {                                                                            \
        out[0] = m[0][0] * vec[0] + m[0][1] * vec[1] + m[0][2] * vec[2];            \
        out[1] = m[1][0] * vec[0] + m[1][1] * vec[1] + m[1][2] * vec[2];            \
        out[2] = m[2][0] * vec[0] + m[2][1] * vec[1] + m[2][2] * vec[2];            \
    } while (0)
    double u0[3] = {1.0, 0.0, 0.0}, u1[3], u2[3], tmp[3];
    if (s[0] > 1e-12) {
        MULMV(v0, tmp);
        if (!normalize_vec3(tmp, u0)) { u0[0] = 1.0; u0[1] = 0.0; u0[2] = 0.0; }
    }
    int have_u1 = 0;
    if (s[1] > 1e-12) {
        double raw[3];
        MULMV(v1, raw);
        double proj = dot3(raw, u0);
        double orth[3] = {raw[0] - proj * u0[0], raw[1] - proj * u0[1], raw[2] - proj * u0[2]};
        have_u1 = normalize_vec3(orth, u1);
    }
    if (!have_u1) {
        double basis[3] = {0.0, 1.0, 0.0};
        if (fabs(u0[0]) < 0.9) { basis[0] = 1.0; basis[1] = 0.0; }
        double cr[3];
        cross3(u0, basis, cr);
        if (!normalize_vec3(cr, u1)) { u1[0] = 0.0; u1[1] = 1.0; u1[2] = 0.0; }
    }
    double c01[3];
    cross3(u0, u1, c01);
    if (!normalize_vec3(c01, u2)) return 0;
    if (s[2] > 1e-12) {
        double mv2[3];
        MULMV(v2, mv2);
        if (dot3(u2, mv2) < 0.0) { u2[0] = -u2[0]; u2[1] = -u2[1]; u2[2] = -u2[2]; }
    }
#undef MULMV
    for (int r = 0; r < 3; r++) { u[r][0] = u0[r]; u[r][1] = u1[r]; u[r][2] = u2[r]; }
    transpose3(vm, vt);
    return 1;
}

/* geometry.rs:533-544 */
static int rank2_project(double f[3][3]) {
    double u[3][3], s[3], vt[3][3];
    if (!svd3(f, u, s, vt)) return 0;
    double diag[3][3] = {{s[0], 0.0, 0.0}, {0.0, s[1], 0.0}, {0.0, 0.0, 0.0}}, dv[3][3];
    mul3(diag, vt, dv);
    mul3(u, dv, f);
    return normalize_frobenius(f);
}

/* geometry.rs:547-583 (Hartley normalisation; mean distance by glibc hypot) */
static int normalize_points(const double *x, const double *y, size_t n, double *nx, double *ny, double t[3][3]) {
    if (n == 0) return 0;
    double inv_n = 1.0 / (double)n, cx = 0.0, cy = 0.0;
    for (size_t i = 0; i < n; i++) {
        if (!isfinite(x[i]) || !isfinite(y[i])) return 0;
        cx += x[i];
        cy += y[i];
    }
    cx *= inv_n;
    cy *= inv_n;
    double md = 0.0;
    for (size_t i = 0; i < n; i++) md += hypot(x[i] - cx, y[i] - cy);
    md *= inv_n;
    if (!isfinite(md) || md < 1e-9) return 0;
    double sc = 1.41421356237309504880 / md;
    for (size_t i = 0; i < n; i++) {
        nx[i] = (x[i] - cx) * sc;
        ny[i] = (y[i] - cy) * sc;
    }
    double tt[3][3] = {{sc, 0.0, -sc * cx}, {0.0, sc, -sc * cy}, {0.0, 0.0, 1.0}};
    memcpy(t, tt, sizeof tt);
    return 1;
}

/* geometry.rs:586-604 */
static void fundamental_ata(const double *x1, const double *y1, const double *x2, const double *y2, size_t n,
                            double ata[81]) {
    memset(ata, 0, 81 * sizeof(double));
    for (size_t k = 0; k < n; k++) {
        double r[9] = {x2[k] * x1[k], x2[k] * y1[k], x2[k], y2[k] * x1[k], y2[k] * y1[k], y2[k], x1[k], y1[k], 1.0};
        for (int i = 0; i < 9; i++)
            for (int j = i; j < 9; j++) ata[i * 9 + j] += r[i] * r[j];
    }
    for (int i = 0; i < 9; i++)
        for (int j = 0; j < i; j++) ata[i * 9 + j] = ata[j * 9 + i];
}

/* geometry.rs:617-635 */
double psr_sampson_sq(const double f[9], double x1, double y1, double x2, double y2) {
    double fx0 = f[0] * x1 + f[1] * y1 + f[2];
    double fx1 = f[3] * x1 + f[4] * y1 + f[5];
    double fx2 = f[6] * x1 + f[7] * y1 + f[8];
    double ft0 = f[0] * x2 + f[3] * y2 + f[6];
    double ft1 = f[1] * x2 + f[4] * y2 + f[7];
    double e = x2 * fx0 + y2 * fx1 + fx2;
    double den = fx0 * fx0 + fx1 * fx1 + ft0 * ft0 + ft1 * ft1;
    if (den < 1e-18 || !isfinite(den)) return INFINITY;
    return (e * e) / den;
}

/* normalised points of m[idx[0..n)] (or of m[0..n) when idx is NULL) */
static int normalized_pair_points(const psr_match *m, const uint32_t *idx, size_t n, double *buf, double tq[3][3],
                                  double tt[3][3]) {
    double *qx = buf, *qy = buf + n, *tx = buf + 2 * n, *ty = buf + 3 * n;
    for (size_t i = 0; i < n; i++) {
        const psr_match *p = &m[idx ? idx[i] : i];
        qx[i] = (double)p->qx; qy[i] = (double)p->qy; tx[i] = (double)p->tx; ty[i] = (double)p->ty;
    }
    if (!normalize_points(qx, qy, n, qx, qy, tq)) return 0;
    if (!normalize_points(tx, ty, n, tx, ty, tt)) return 0;
    return 1;
}

/* geometry.rs:638-664 over m[idx[0..n)] (n >= 8) */
int psr_fit_fundamental_8point(const psr_match *m, const uint32_t *idx, size_t n, double fout[9]) {
    if (n < 8) return 0;
    double *buf = (double *)malloc(sizeof(double) * 4 * n);
    double tq[3][3], tt[3][3], ata[81], ev[9], vec[81];
    int ok = normalized_pair_points(m, idx, n, buf, tq, tt);
    if (ok) {
        fundamental_ata(buf, buf + n, buf + 2 * n, buf + 3 * n, n, ata);
        ok = psr_jacobi9(ata, ev, vec);
    }
    free(buf);
    if (!ok) return 0;
    double f[3][3] = {{vec[0], vec[9], vec[18]}, {vec[27], vec[36], vec[45]}, {vec[54], vec[63], vec[72]}};
    if (!rank2_project(f)) return 0;
    double ttt[3][3], fq[3][3], out[3][3];
    transpose3(tt, ttt);
    mul3(f, tq, fq);
    mul3(ttt, fq, out);
    if (!normalize_frobenius(out)) return 0;
    memcpy(fout, out, sizeof out);
    return 1;
}

/* geometry.rs:667-724 */
static int solve_cubic(double c3, double c2, double c1, double c0, double roots[3]) {
    int nr = 0;
    if (fabs(c3) < 1e-12) {
        if (fabs(c2) < 1e-12) {
            if (fabs(c1) < 1e-12) return 0;
            roots[0] = -c0 / c1;
            return 1;
        }
        double disc = c1 * c1 - 4.0 * c2 * c0;
        if (disc < 0.0) return 0;
        double sq = sqrt(disc);
        roots[0] = (-c1 + sq) / (2.0 * c2);
        roots[1] = (-c1 - sq) / (2.0 * c2);
        return 2;   /* returned before the Newton polish, as there */
    }
    double a = c2 / c3, b = c1 / c3, c = c0 / c3;
    double a3 = a / 3.0;
    double p = b - a * a3;
    double q = (2.0 * a * a * a) / 27.0 - (a * b) / 3.0 + c;
    double p3 = p / 3.0;
    double disc = (q * 0.5) * (q * 0.5) + (p3 * p3) * p3;   /* powi(3): (x * x) * x */
    if (disc > 1e-14) {
        double sq = sqrt(disc);
        double u = cbrt(-0.5 * q + sq), v = cbrt(-0.5 * q - sq);
        roots[nr++] = u + v - a3;
    } else if (disc >= -1e-14) {
        double u = cbrt(-0.5 * q);
        roots[nr++] = 2.0 * u - a3;
        roots[nr++] = -u - a3;
    } else {
        double r = sqrt(-p / 3.0);
        double cp = (-0.5 * q) / (r * r * r);
        if (cp < -1.0) cp = -1.0;
        if (cp > 1.0) cp = 1.0;
        double phi = acos(cp), two_r = 2.0 * r;
        roots[nr++] = two_r * cos(phi / 3.0) - a3;
        roots[nr++] = two_r * cos((phi + 2.0 * PI_D) / 3.0) - a3;
        roots[nr++] = two_r * cos((phi + 4.0 * PI_D) / 3.0) - a3;
    }
    for (int k = 0; k < nr; k++)
        for (int it = 0; it < 3; it++) {
            double t = roots[k];
            double f = ((c3 * t + c2) * t + c1) * t + c0;
            double df = (3.0 * c3 * t + 2.0 * c2) * t + c1;
            if (fabs(df) > 1e-14) roots[k] = t - f / df;
        }
    return nr;
}

/* geometry.rs:727-805 over m[0..7); returns the candidate count (<= 4) */
int psr_fit_fundamental_7point(const psr_match *m, double fout[4][9]) {
    double buf[28], tq[3][3], tt[3][3], ata[81], ev[9], vec[81];
    if (!normalized_pair_points(m, NULL, 7, buf, tq, tt)) return 0;
    fundamental_ata(buf, buf + 7, buf + 14, buf + 21, 7, ata);
    if (!psr_jacobi9(ata, ev, vec)) return 0;
    double f1[3][3], f2[3][3], b[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            f1[r][c] = vec[(r * 3 + c) * 9 + 0];
            f2[r][c] = vec[(r * 3 + c) * 9 + 1];
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) b[r][c] = f1[r][c] - f2[r][c];
#define EVAL_MAT(t, out)                                                      \
    do {                                                                      \
        for (int r_ = 0; r_ < 3; r_++)                                        \
            for (int c_ = 0; c_ < 3; c_++) out[r_][c_] = f2[r_][c_] + (t) * b[r_][c_]; \
    } while (0)
    double e1[3][3], em1[3][3], e2[3][3];
    EVAL_MAT(1.0, e1);
    EVAL_MAT(-1.0, em1);
    EVAL_MAT(2.0, e2);
    double p0 = det3(f2), p1 = det3(e1), pm1 = det3(em1), p2 = det3(e2);
    double c0 = p0, c2 = 0.5 * (p1 + pm1) - c0, s1 = 0.5 * (p1 - pm1), s2 = p2 - c0 - 4.0 * c2;
    double c3 = (s2 - 2.0 * s1) / 6.0, c1 = s1 - c3;
    double roots[3];
    int nr = solve_cubic(c3, c2, c1, c0, roots), nout = 0;
    double ttt[3][3];
    transpose3(tt, ttt);
    for (int k = 0; k < nr; k++) {
        if (!isfinite(roots[k])) continue;
        double fc[3][3], fq[3][3], f[3][3];
        EVAL_MAT(roots[k], fc);
        mul3(fc, tq, fq);
        mul3(ttt, fq, f);
        if (normalize_frobenius(f)) memcpy(fout[nout++], f, sizeof f);
    }
#undef EVAL_MAT
    if (fabs(c3) < 1e-12 && fabs(det3(b)) < 1e-10) {
        double bq[3][3], f[3][3];
        mul3(b, tq, bq);
        mul3(ttt, bq, f);
        if (normalize_frobenius(f)) memcpy(fout[nout++], f, sizeof f);
    }
    return nout;
}

/* geometry.rs:808-828: (count, truncated error) */
static size_t eval_fundamental(const psr_match *m, size_t n, const double f[9], double tsq, double *err) {
    size_t count = 0;
    double e = 0.0;
    for (size_t i = 0; i < n; i++) {
        double d2 = psr_sampson_sq(f, (double)m[i].qx, (double)m[i].qy, (double)m[i].tx, (double)m[i].ty);
        if (d2 <= tsq) { count++; e += d2; }
        else e += tsq;
    }
    *err = e;
    return count;
}

/* compiler-rt __powidf2 (binary exponentiation), as Rust's f64::powi */
static double powi_d(double a, int b) {
    int recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

/* geometry.rs:831-852 */
size_t psr_adaptive_epipolar_iters(size_t best, size_t total, int sample_size, double confidence, size_t min_iters,
                                   size_t max_iters) {
    if (total <= (size_t)sample_size || best == 0) return max_iters;
    double w = (double)best / (double)total;
    if (w >= 0.9999) return min_iters;
    double pg = powi_d(w, sample_size);
    if (pg <= 1e-12) return max_iters;
    double d = 1.0 - pg;
    double denom = log(d > 1e-15 ? d : 1e-15);
    double cc = confidence < 0.5 ? 0.5 : (confidence > 0.999999 ? 0.999999 : confidence);
    double num = log(1.0 - cc);
    double it = ceil(num / denom);
    size_t k = it >= 1.8e19 ? (size_t)-1 : (it > 0 ? (size_t)it : 0);   /* f64 as usize saturates */
    if (k < min_iters) k = min_iters;
    if (k > max_iters) k = max_iters;
    return k;
}

static uint64_t smx_next(uint64_t *s) {
    *s += 0x9E3779B97F4A7C15ull;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* pipeline.rs:2006-2036 */
static void ratio_cdf(const psr_match *m, size_t n, double *cdf) {
    double acc = 0.0;
    for (size_t i = 0; i < n; i++) {
        double r = 1.0;
        if (isfinite(m[i].ratio)) {
            r = (double)m[i].ratio;
            if (r < 0.05) r = 0.05;
            if (r > 1.0) r = 1.0;
        }
        acc += 1.0 / r;
        cdf[i] = acc;
    }
    if (acc > 0.0) {
        double inv = 1.0 / acc;
        for (size_t i = 0; i < n; i++) cdf[i] *= inv;
    }
    if (n) cdf[n - 1] = 1.0;
}

static size_t weighted_pick(const double *cdf, size_t n, uint64_t *state) {
    double u = (double)(smx_next(state) >> 11) * (1.0 / 9007199254740992.0);
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    return lo < n - 1 ? lo : n - 1;
}

static size_t collect_sampson_inliers(const psr_match *m, size_t n, const double f[9], double tsq, uint32_t *out) {
    size_t k = 0;
    for (size_t i = 0; i < n; i++)
        if (psr_sampson_sq(f, (double)m[i].qx, (double)m[i].qy, (double)m[i].tx, (double)m[i].ty) <= tsq)
            out[k++] = (uint32_t)i;
    return k;
}

/* geometry.rs:868-1016 estimate_fundamental_ransac_with_options. inl has capacity n; returns
 * the inlier count; *iters = iterations run (0 on the 7-match path). */
size_t psr_fundamental_ransac(const psr_match *m, size_t n, double sampson_threshold, size_t max_iters,
                              double confidence, uint64_t seed, uint32_t *inl, double fout[9], int *has_model,
                              size_t *iters) {
    *has_model = 0;
    if (iters) *iters = 0;
    if (n < 7) return 0;
    double th = sampson_threshold > 0.25 ? sampson_threshold : 0.25;
    if (sampson_threshold != sampson_threshold) th = 0.25;   /* f64::max(NaN, 0.25) = 0.25 */
    const double tsq = th * th;
    if (n == 7) {
        double cand[4][9], best[9], be = INFINITY;
        size_t bc = 0;
        int nc = psr_fit_fundamental_7point(m, cand), have = 0;
        for (int k = 0; k < nc; k++) {
            double e;
            size_t c = eval_fundamental(m, n, cand[k], tsq, &e);
            if (c > bc || (c == bc && e < be)) { bc = c; be = e; memcpy(best, cand[k], sizeof best); have = 1; }
        }
        if (!have) return 0;
        memcpy(fout, best, sizeof best);
        *has_model = 1;
        return collect_sampson_inliers(m, n, fout, tsq, inl);
    }
    double *cdf = (double *)malloc(sizeof(double) * n);
    ratio_cdf(m, n, cdf);
    uint64_t rng = seed;
    const size_t cap = max_iters > 64 ? max_iters : 64, min_it = 64 < cap ? 64 : cap;
    size_t max_it = cap, bc = 0, it = 0;
    double be = INFINITY, best[9];
    int have = 0;
    while (it < max_it) {
        it++;
        uint32_t idx[8];
        int chosen = 0, att = 0;
        while (chosen < 8 && att < 32) {
            att++;
            uint32_t c = (uint32_t)weighted_pick(cdf, n, &rng);
            int dup = 0;
            for (int j = 0; j < chosen; j++) dup |= idx[j] == c;
            if (!dup) idx[chosen++] = c;
        }
        if (chosen < 8) continue;
        double f[9];
        if (psr_fit_fundamental_8point(m, idx, 8, f)) {
            double e;
            size_t c = eval_fundamental(m, n, f, tsq, &e);
            if (c > bc || (c == bc && e < be)) {
                bc = c;
                be = e;
                memcpy(best, f, sizeof best);
                have = 1;
                max_it = psr_adaptive_epipolar_iters(bc, n, 8, confidence, min_it, cap);
            }
        }
    }
    free(cdf);
    if (iters) *iters = it;
    if (!have) return 0;
    for (int lo = 0; lo < 2; lo++) {   /* LO refit on the consensus set */
        size_t ni = collect_sampson_inliers(m, n, best, tsq, inl);
        if (ni < 8) break;
        double r[9];
        if (!psr_fit_fundamental_8point(m, inl, ni, r)) break;
        double e;
        size_t c = eval_fundamental(m, n, r, tsq, &e);
        if (c > bc || (c == bc && e < be)) { bc = c; be = e; memcpy(best, r, sizeof best); }
        else break;
    }
    memcpy(fout, best, sizeof best);
    *has_model = 1;
    return collect_sampson_inliers(m, n, fout, tsq, inl);
}

/* visualize.rs:202-232 (f32 matrix in, f64 inverse out) */
int psr_invert3_f32(const float h[9], double out[9]) {
    double a[3][3];
    for (int i = 0; i < 9; i++) a[i / 3][i % 3] = (double)h[i];
    double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    double c01 = -(a[1][0] * a[2][2] - a[1][2] * a[2][0]);
    double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    double c10 = -(a[0][1] * a[2][2] - a[0][2] * a[2][1]);
    double c11 = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    double c12 = -(a[0][0] * a[2][1] - a[0][1] * a[2][0]);
    double c20 = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    double c21 = -(a[0][0] * a[1][2] - a[0][2] * a[1][0]);
    double c22 = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    if (fabs(det) < 1e-12 || !isfinite(det)) return 0;
    double id = 1.0 / det;
    double o[9] = {c00 * id, c10 * id, c20 * id, c01 * id, c11 * id, c21 * id, c02 * id, c12 * id, c22 * id};
    memcpy(out, o, sizeof o);
    return 1;
}

static double fmin_rust(double a, double b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); }

/* geometry.rs:1233-1341. h: row-major f32 homography or NULL; f: row-major fundamental or
 * NULL. Returns 0 (None) or 1 with *selected = 0 (Homography) / 1 (Fundamental). */
int psr_select_model_gric(const psr_match *m, size_t n, const float *h, const double *f, double sigma,
                          int *selected, double *h_gric, double *f_gric) {
    if (n == 0) return 0;
    double sg = sigma > 0.25 ? sigma : 0.25;
    if (sigma != sigma) sg = 0.25;
    const double s2 = sg * sg, r = 4.0, l1 = 2.0, l2 = 2.0;
    double hg = INFINITY, fg = INFINITY;
    if (h) {
        double hd[9], inv[9];
        for (int i = 0; i < 9; i++) hd[i] = (double)h[i];
        int has_inv = psr_invert3_f32(h, inv);
        const double dh = 2.0, kh = 8.0, caph = l1 * (r - dh);
        double data = 0.0;
        for (size_t i = 0; i < n; i++) {
            double qx = m[i].qx, qy = m[i].qy, tx = m[i].tx, ty = m[i].ty;
            double fw = hd[6] * qx + hd[7] * qy + hd[8], fwd, bwd;
            if (fabs(fw) > 1e-12) {
                double px = (hd[0] * qx + hd[1] * qy + hd[2]) / fw, py = (hd[3] * qx + hd[4] * qy + hd[5]) / fw;
                fwd = (px - tx) * (px - tx) + (py - ty) * (py - ty);
            } else {
                fwd = INFINITY;
            }
            if (has_inv) {
                double bw = inv[6] * tx + inv[7] * ty + inv[8];
                if (fabs(bw) > 1e-12) {
                    double px = (inv[0] * tx + inv[1] * ty + inv[2]) / bw, py = (inv[3] * tx + inv[4] * ty + inv[5]) / bw;
                    bwd = (px - qx) * (px - qx) + (py - qy) * (py - qy);
                } else {
                    bwd = INFINITY;
                }
            } else {
                bwd = fwd;
            }
            double sym = 0.5 * (fwd + bwd);
            data += fmin_rust(sym / s2, caph);
        }
        hg = data + l1 * dh * (double)n + l2 * kh;
    }
    if (f) {
        const double df = 3.0, kf = 7.0, capf = l1 * (r - df);
        double data = 0.0;
        for (size_t i = 0; i < n; i++) {
            double d2 = psr_sampson_sq(f, (double)m[i].qx, (double)m[i].qy, (double)m[i].tx, (double)m[i].ty);
            data += fmin_rust(d2 / s2, capf);
        }
        fg = data + l1 * df * (double)n + l2 * kf;
    }
    if (!isfinite(hg) && !isfinite(fg)) return 0;
    *selected = fg < hg ? 1 : 0;
    *h_gric = hg;
    *f_gric = fg;
    return 1;
}
