/*
 * ps_oracle.c -- CPU oracle (test infrastructure) for the SIFT hot path.
 *
 * A plain-C restatement of pure-sift 0.2.0 (Rust) src/pipeline.rs and the
 * Hamming primitive of src/catalog.rs. Each function cites the reference
 * lines it restates. Build: oracle/Makefile (gcc -O2 -ffp-contract=off
 * -fno-fast-math -fopenmp). Never linked into the product library.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *  - gaussian_kernel_1d: closed form in float64, sum == 1 +- 1e-5, symmetry,
 *    len(sigma=4) == 25 (reference tests/pipeline_tests.rs:191-215, 1067-1068).
 *  - gaussian_blur: DC invariance on a constant image (pipeline_tests.rs:217-225)
 *    and a direct float64 2-D convolution with clamped borders.
 *  - upsample/downsample/DoG/luma: closed forms, bit-exact in numpy.
 *  - pyramid: octave dimensions (pipeline_tests.rs:167-189) and level-by-level
 *    equality with a numpy float32 restatement.
 *  - extrema: brute-force 26-neighbour predicate in numpy; refine vs a float64
 *    numpy.linalg.solve of the 3x3 Taylor system; sub-pixel blob test
 *    < 0.12 px (pipeline_tests.rs:941-987); scale band (pipeline_tests.rs:125-146);
 *    upsampled-vs-native position within 0.5 px (pipeline_tests.rs:1138-1187).
 *  - orientation: range [0, 2pi), >=80% splitting on the cross image
 *    (pipeline_tests.rs:989-1033), a pure-Python restatement on small cases.
 *  - descriptor: unit norm +-1e-4 and range (pipeline_tests.rs:254-279),
 *    rotation of a synthetic patch, pure-Python restatement on small cases.
 *  - matching: float64 brute force (indices), the mutual-NN known answer
 *    (pipeline_tests.rs:1221-1280).
 *  - Hamming: numpy bitwise_count brute force, quantisation closed form.
 *  - RANSAC: reference known answers (pipeline_tests.rs:343-524,
 *    tests/homography_properties.rs:70-237), SplitMix64 published vector.
 * The exact libm values (expf/atan2f/powf/sinf/cosf) are glibc's, as in the
 * Rust binary on Linux; the reference binary itself cannot be built here
 * (no Rust toolchain), so whole-pipeline outputs are "parity unpinned"
 * against the Rust binary beyond the invariants above.
 */
#include "ps_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define PI_F 3.14159265358979323846f
#define TWO_PI_F (2.0f * PI_F)
#define BASE_SIGMA 1.6f           /* pipeline.rs:10 */
#define ASSUMED_INPUT_SIGMA 0.5f  /* pipeline.rs:13 */
#define KMAX 512

static int g_threads = 0;

void psr_set_threads(int n) {
    g_threads = n;
    if (n > 0) omp_set_num_threads(n);
    else omp_set_num_threads(omp_get_num_procs());
}
int psr_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_num_procs(); }
void psr_free(void *p) { free(p); }

/* Rust f32::rem_euclid */
static inline float rem_euclid_f(float x, float m) {
    float r = fmodf(x, m);
    return r < 0.0f ? r + fabsf(m) : r;
}
/* Rust f32::clamp */
static inline float clampf_rs(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
/* Rust `f as usize` / `as isize`: saturating, NaN -> 0 */
static inline long f2isize(float f) {
    if (f != f) return 0;
    if (f >= 9.2e18f) return 9000000000000000000L;
    if (f <= -9.2e18f) return -9000000000000000000L;
    return (long)f;
}
static inline size_t f2usize(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 1.8e19f) return (size_t)-1;
    return (size_t)f;
}
/* Rust f32::powi (compiler-rt __powisf2: binary exponentiation) */
static float powi_rs(float a, int b) {
    int recip = b < 0;
    float r = 1.0f;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0f / r : r;
}

/* ---------------------------------------------------------------- kernels */

/* pipeline.rs:275-296 */
int psr_gaussian_kernel_1d(float sigma, float *out, int cap) {
    float cs = fmaxf(sigma, 1e-3f);
    long radius = f2isize(ceilf(3.0f * cs));
    if (radius < 1) radius = 1;
    float two_sigma_sq = 2.0f * cs * cs;
    float norm_factor = 1.0f / (sqrtf(2.0f * PI_F) * cs);
    int n = (int)(2 * radius + 1);
    if (n > cap) return -n;
    for (int i = 0; i < n; i++) {
        float xf = (float)(i - radius);
        out[i] = norm_factor * expf(-((xf * xf) / two_sigma_sq));
    }
    float sum = 0.0f;
    for (int i = 0; i < n; i++) sum += out[i];
    if (sum > 0.0f) {
        float inv = 1.0f / sum;
        for (int i = 0; i < n; i++) out[i] *= inv;
    }
    return n;
}

/* out[x] = sum_k in[clamp(x+k-r)] * w[k], summed in k order (pipeline.rs:331-377).
 * The loop is written k-outer/x-inner so it vectorises across x; every output
 * still accumulates its taps in the reference's order, so values are identical. */
static void horiz_row(const float *in, float *o, int w, const float *k, int n, int r) {
    for (int x = 0; x < w; x++) o[x] = 0.0f;
    for (int kk = 0; kk < n; kk++) {
        int off = kk - r;
        float wk = k[kk];
        int x0 = -off; if (x0 < 0) x0 = 0; if (x0 > w) x0 = w;
        int x1 = w - off; if (x1 > w) x1 = w; if (x1 < x0) x1 = x0;
        float lv = in[0], rv = in[w - 1];
        for (int x = 0; x < x0; x++) o[x] += lv * wk;
        for (int x = x0; x < x1; x++) o[x] += in[x + off] * wk;
        for (int x = x1; x < w; x++) o[x] += rv * wk;
    }
}

/* pipeline.rs:314-425 */
void psr_gaussian_blur(const float *in, int h, int w, float sigma, float *out) {
    if (h == 0 || w == 0) return;
    float k[KMAX];
    int n = psr_gaussian_kernel_1d(sigma, k, KMAX);
    if (n < 0) abort();
    int r = n / 2;
    float *tmp = (float *)malloc(sizeof(float) * (size_t)h * w);
    int par = (long)h * w >= 16384;
#pragma omp parallel for if (par) schedule(static)
    for (int y = 0; y < h; y++) horiz_row(in + (size_t)y * w, tmp + (size_t)y * w, w, k, n, r);
#pragma omp parallel for if (par) schedule(static)
    for (int y = 0; y < h; y++) {
        float *o = out + (size_t)y * w;
        for (int x = 0; x < w; x++) o[x] = 0.0f;
        for (int kk = 0; kk < n; kk++) {
            long sy = (long)y + kk - r;
            if (sy < 0) sy = 0;
            if (sy > h - 1) sy = h - 1;
            const float *row = tmp + (size_t)sy * w;
            float wk = k[kk];
            for (int x = 0; x < w; x++) o[x] += row[x] * wk;
        }
    }
    free(tmp);
}

/* pipeline.rs:429-449 */
void psr_downsample_2x(const float *in, int h, int w, float *out) {
    int nh = h / 2, nw = w / 2;
    if (nh == 0 || nw == 0) return;
    for (int y = 0; y < nh; y++)
        for (int x = 0; x < nw; x++) out[(size_t)y * nw + x] = in[(size_t)(2 * y) * w + 2 * x];
}

/* pipeline.rs:455-483 */
void psr_upsample_2x(const float *in, int h, int w, float *out) {
    int nh = h * 2, nw = w * 2;
    for (int uy = 0; uy < nh; uy++) {
        int y0 = uy / 2;
        int y1 = y0 + (uy & 1); if (y1 > h - 1) y1 = h - 1;
        const float *r0 = in + (size_t)y0 * w, *r1 = in + (size_t)y1 * w;
        float *o = out + (size_t)uy * nw;
        for (int ux = 0; ux < nw; ux++) {
            int x0 = ux / 2;
            int x1 = x0 + (ux & 1); if (x1 > w - 1) x1 = w - 1;
            o[ux] = 0.25f * (r0[x0] + r0[x1] + r1[x0] + r1[x1]);
        }
    }
}

/* pipeline.rs:957-966 */
void psr_luma_to_f32(const uint8_t *in, size_t n, float *out) {
    const float inv_255 = 1.0f / 255.0f;
    for (size_t i = 0; i < n; i++) out[i] = (float)in[i] * inv_255;
}

/* ---------------------------------------------------------------- pyramid */

/* pipeline.rs:550-638 */
int psr_build_pyramid(const float *img, int h, int w, int octaves, int scales, int upsample,
                      psr_pyramid *p) {
    memset(p, 0, sizeof(*p));
    p->scales = scales;
    if (h == 0 || w == 0 || octaves == 0 || scales == 0) return 0;
    if (octaves > PSR_MAX_OCTAVES) octaves = PSR_MAX_OCTAVES;
    int nslices = scales + 2;
    float k = powf(2.0f, 1.0f / (float)scales);
    float *base;
    int bh, bw;
    float base_mult;
    if (upsample) {
        bh = 2 * h; bw = 2 * w;
        float *up = (float *)malloc(sizeof(float) * (size_t)bh * bw);
        psr_upsample_2x(img, h, w, up);
        float sigma_in = 2.0f * ASSUMED_INPUT_SIGMA;
        float init_sigma = sqrtf(fmaxf(BASE_SIGMA * BASE_SIGMA - sigma_in * sigma_in, 1e-4f));
        base = (float *)malloc(sizeof(float) * (size_t)bh * bw);
        psr_gaussian_blur(up, bh, bw, init_sigma, base);
        free(up);
        base_mult = 0.5f;
    } else {
        bh = h; bw = w;
        float sigma_in = ASSUMED_INPUT_SIGMA;
        float init_sigma = sqrtf(fmaxf(BASE_SIGMA * BASE_SIGMA - sigma_in * sigma_in, 1e-4f));
        base = (float *)malloc(sizeof(float) * (size_t)bh * bw);
        psr_gaussian_blur(img, bh, bw, init_sigma, base);
        base_mult = 1.0f;
    }
    for (int o = 0; o < octaves; o++) {
        if (bh < 3 || bw < 3) break;
        size_t P = (size_t)bh * bw;
        float *G = (float *)malloc(sizeof(float) * P * (nslices + 1));
        float *D = (float *)malloc(sizeof(float) * P * nslices);
        memcpy(G, base, sizeof(float) * P);
        free(base);
        float prev_sigma = BASE_SIGMA;
        for (int s = 1; s <= nslices; s++) {
            float target = BASE_SIGMA * powi_rs(k, s);
            float delta = sqrtf(fmaxf(target * target - prev_sigma * prev_sigma, 1e-4f));
            psr_gaussian_blur(G + (s - 1) * P, bh, bw, delta, G + s * P);
            const float *nx = G + s * P, *pv = G + (s - 1) * P;
            float *d = D + (s - 1) * P;
            for (size_t i = 0; i < P; i++) d[i] = nx[i] - pv[i];
            prev_sigma = target;
        }
        p->h[o] = bh; p->w[o] = bw;
        p->gauss[o] = G; p->dog[o] = D;
        p->scale_factor[o] = base_mult * (float)(1u << o);
        p->n_octaves = o + 1;
        if (o + 1 < octaves) {
            int nh = bh / 2, nw = bw / 2;
            if (nh == 0 || nw == 0) { base = NULL; bh = 0; bw = 0; }
            else {
                base = (float *)malloc(sizeof(float) * (size_t)nh * nw);
                psr_downsample_2x(G + (size_t)scales * P, bh, bw, base);
                bh = nh; bw = nw;
            }
        } else {
            base = NULL; /* octave_base = prev_gaussian is never used again */
            break;
        }
    }
    free(base);
    return p->n_octaves;
}

void psr_free_pyramid(psr_pyramid *p) {
    for (int o = 0; o < PSR_MAX_OCTAVES; o++) {
        free(p->gauss[o]); free(p->dog[o]);
        p->gauss[o] = p->dog[o] = NULL;
    }
    p->n_octaves = 0;
}

/* ---------------------------------------------------------------- extrema */

/* pipeline.rs:664-687 */
static int passes_hessian_edge_test(const float *c, int w, int x, int y, float center) {
    const float *rp = c + (size_t)(y - 1) * w, *rc = c + (size_t)y * w, *rn = c + (size_t)(y + 1) * w;
    float dxx = rc[x + 1] - 2.0f * center + rc[x - 1];
    float dyy = rn[x] - 2.0f * center + rp[x];
    float dxy = 0.25f * ((rn[x + 1] - rn[x - 1]) - (rp[x + 1] - rp[x - 1]));
    float trace = dxx + dyy;
    float det = dxx * dyy - dxy * dxy;
    const float THRESHOLD_RATIO = ((10.0f + 1.0f) * (10.0f + 1.0f)) / 10.0f;
    return det > 1e-9f && (trace * trace) < THRESHOLD_RATIO * det;
}

/* pipeline.rs:697-770 */
static void refine_extremum_3d(const float *ps, const float *cs, const float *ns, int w, int x, int y,
                               float *odx, float *ody, float *ods, float *ointerp) {
    const float *c_prev = cs + (size_t)(y - 1) * w, *c_curr = cs + (size_t)y * w, *c_next = cs + (size_t)(y + 1) * w;
    const float *p_prev = ps + (size_t)(y - 1) * w, *p_curr = ps + (size_t)y * w, *p_next = ps + (size_t)(y + 1) * w;
    const float *n_prev = ns + (size_t)(y - 1) * w, *n_curr = ns + (size_t)y * w, *n_next = ns + (size_t)(y + 1) * w;
    float center = c_curr[x];
    float gx = 0.5f * (c_curr[x + 1] - c_curr[x - 1]);
    float gy = 0.5f * (c_next[x] - c_prev[x]);
    float gs = 0.5f * (n_curr[x] - p_curr[x]);
    float dxx = c_curr[x + 1] - 2.0f * center + c_curr[x - 1];
    float dyy = c_next[x] - 2.0f * center + c_prev[x];
    float dss = n_curr[x] - 2.0f * center + p_curr[x];
    float dxy = 0.25f * ((c_next[x + 1] - c_next[x - 1]) - (c_prev[x + 1] - c_prev[x - 1]));
    float dxs = 0.25f * ((n_curr[x + 1] - n_curr[x - 1]) - (p_curr[x + 1] - p_curr[x - 1]));
    float dys = 0.25f * ((n_next[x] - n_prev[x]) - (p_next[x] - p_prev[x]));
    float c00 = dyy * dss - dys * dys;
    float c01 = dys * dxs - dxy * dss;
    float c02 = dxy * dys - dyy * dxs;
    float c11 = dxx * dss - dxs * dxs;
    float c12 = dxy * dxs - dxx * dys;
    float c22 = dxx * dyy - dxy * dxy;
    float det3 = dxx * c00 + dxy * c01 + dxs * c02;
    if (fabsf(det3) > 1e-9f) {
        float inv_det = -1.0f / det3;
        float ox = (c00 * gx + c01 * gy + c02 * gs) * inv_det;
        float oy = (c01 * gx + c11 * gy + c12 * gs) * inv_det;
        float os = (c02 * gx + c12 * gy + c22 * gs) * inv_det;
        if (fabsf(ox) <= 0.5f && fabsf(oy) <= 0.5f && fabsf(os) <= 0.5f) {
            *odx = ox; *ody = oy; *ods = os;
            *ointerp = center + 0.5f * (gx * ox + gy * oy + gs * os);
            return;
        }
    }
    if (fabsf(c22) > 1e-9f) {
        float inv_det2 = -1.0f / c22;
        float ox = clampf_rs((dyy * gx - dxy * gy) * inv_det2, -0.5f, 0.5f);
        float oy = clampf_rs((-dxy * gx + dxx * gy) * inv_det2, -0.5f, 0.5f);
        float os = fabsf(dss) > 1e-6f ? clampf_rs(-gs / dss, -0.5f, 0.5f) : 0.0f;
        *odx = ox; *ody = oy; *ods = os;
        *ointerp = center + 0.5f * (gx * ox + gy * oy + gs * os);
        return;
    }
    *odx = 0.0f; *ody = 0.0f; *ods = 0.0f; *ointerp = center;
}

typedef struct { psr_keypoint *v; size_t n, cap; } kpvec;
static void kpvec_push(kpvec *a, psr_keypoint k) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 256;
        a->v = (psr_keypoint *)realloc(a->v, a->cap * sizeof(psr_keypoint));
    }
    a->v[a->n++] = k;
}

/* pipeline.rs:829-899 (one slice) */
static void scan_slice(int octave_idx, float *const *sl, int nsl, int h, int w, float mult, float thr,
                       int s, kpvec *out) {
    int S = nsl - 2; if (S < 1) S = 1;
    float k = powf(2.0f, 1.0f / (float)S);
    const float *ps = sl[s - 1], *cs = sl[s], *ns = sl[s + 1];
    float effective_scale = BASE_SIGMA * mult * powi_rs(k, s);
    for (int y = 1; y < h - 1; y++) {
        const float *ra = cs + (size_t)(y - 1) * w, *rc = cs + (size_t)y * w, *rb = cs + (size_t)(y + 1) * w;
        for (int x = 1; x < w - 1; x++) {
            float val = rc[x];
            if (fabsf(val) <= thr) continue;
            int ext;
            if (val > 0.0f) {
                ext = val > rc[x - 1] && val > rc[x + 1] && val > ra[x - 1] && val > ra[x] &&
                      val > ra[x + 1] && val > rb[x - 1] && val > rb[x] && val > rb[x + 1];
                if (ext) {
                    for (int dy = -1; dy <= 1 && ext; dy++)
                        for (int dx = -1; dx <= 1; dx++)
                            if (!(val > ps[(size_t)(y + dy) * w + x + dx])) { ext = 0; break; }
                    for (int dy = -1; dy <= 1 && ext; dy++)
                        for (int dx = -1; dx <= 1; dx++)
                            if (!(val > ns[(size_t)(y + dy) * w + x + dx])) { ext = 0; break; }
                }
            } else {
                ext = val < rc[x - 1] && val < rc[x + 1] && val < ra[x - 1] && val < ra[x] &&
                      val < ra[x + 1] && val < rb[x - 1] && val < rb[x] && val < rb[x + 1];
                if (ext) {
                    for (int dy = -1; dy <= 1 && ext; dy++)
                        for (int dx = -1; dx <= 1; dx++)
                            if (!(val < ps[(size_t)(y + dy) * w + x + dx])) { ext = 0; break; }
                    for (int dy = -1; dy <= 1 && ext; dy++)
                        for (int dx = -1; dx <= 1; dx++)
                            if (!(val < ns[(size_t)(y + dy) * w + x + dx])) { ext = 0; break; }
                }
            }
            if (ext && passes_hessian_edge_test(cs, w, x, y, val)) {
                float dx, dy, ds, interp;
                refine_extremum_3d(ps, cs, ns, w, x, y, &dx, &dy, &ds, &interp);
                if (fabsf(interp) > thr) {
                    float refined_scale = effective_scale * powf(k, ds);
                    float lf = clampf_rs(roundf((float)s + ds), 1.0f, (float)S);
                    psr_keypoint kp;
                    kp.x = ((float)x + dx) * mult;
                    kp.y = ((float)y + dy) * mult;
                    kp.scale = fmaxf(refined_scale, 0.25f * BASE_SIGMA);
                    kp.response = interp;
                    kp.octave = (uint32_t)octave_idx;
                    kp.layer = (uint32_t)f2usize(lf);
                    kpvec_push(out, kp);
                }
            }
        }
    }
}

/* pipeline.rs:810-916 */
static void extract_octave(int octave_idx, float *const *sl, int nsl, int h, int w, float mult,
                           float thr, kpvec *out) {
    if (nsl < 3) return;
    if (h < 3 || w < 3) return;
    int ns = nsl - 2;
    kpvec per[16];
    memset(per, 0, sizeof(per));
    int par = (long)h * w >= 16384;
#pragma omp parallel for if (par) schedule(static, 1)
    for (int s = 1; s <= ns; s++) scan_slice(octave_idx, sl, nsl, h, w, mult, thr, s, &per[s]);
    for (int s = 1; s <= ns; s++) {
        for (size_t i = 0; i < per[s].n; i++) kpvec_push(out, per[s].v[i]);
        free(per[s].v);
    }
}

size_t psr_extract_extrema(const psr_pyramid *p, float thr, psr_keypoint **out) {
    kpvec v = {0};
    for (int o = 0; o < p->n_octaves; o++) {
        size_t P = (size_t)p->h[o] * p->w[o];
        int nsl = p->scales + 2;
        float *sl[16];
        for (int s = 0; s < nsl; s++) sl[s] = p->dog[o] + s * P;
        extract_octave(o, sl, nsl, p->h[o], p->w[o], p->scale_factor[o], thr, &v);
    }
    *out = v.v;
    return v.n;
}

size_t psr_detect_keypoints(const float *img, int h, int w, int octaves, int scales, float thr,
                            int upsample, psr_keypoint **out) {
    psr_pyramid p;
    psr_build_pyramid(img, h, w, octaves, scales, upsample, &p);
    size_t n = psr_extract_extrema(&p, thr, out);
    psr_free_pyramid(&p);
    return n;
}

/* ---------------------------------------------------------------- descriptors */

/* pipeline.rs:1168-1203 */
void psr_gradient_planes(const float *g, int h, int w, float *mag, float *ang) {
    memset(mag, 0, sizeof(float) * (size_t)h * w);
    memset(ang, 0, sizeof(float) * (size_t)h * w);
    if (h < 3 || w < 3) return;
    for (int y = 1; y < h - 1; y++) {
        const float *ra = g + (size_t)(y - 1) * w, *rc = g + (size_t)y * w, *rb = g + (size_t)(y + 1) * w;
        for (int x = 1; x < w - 1; x++) {
            float dx = rc[x + 1] - rc[x - 1];
            float dy = rb[x] - ra[x];
            mag[(size_t)y * w + x] = sqrtf(dx * dx + dy * dy);
            ang[(size_t)y * w + x] = rem_euclid_f(atan2f(dy, dx), TWO_PI_F);
        }
    }
}

/* pipeline.rs:1233-1347 */
int psr_keypoint_orientations(const float *mag, const float *ang, int width, int height,
                              const psr_keypoint *kp, float out[4]) {
    enum { NB = 36 };
    const float two_pi = TWO_PI_F;
    float sigma_ori = fmaxf(1.5f * kp->scale, 1.0f);
    float two_sigma_sq = 2.0f * sigma_ori * sigma_ori;
    long radius = f2isize(ceilf(3.0f * sigma_ori));
    if (radius < 2) radius = 2;
    long cx = f2isize(roundf(kp->x)), cy = f2isize(roundf(kp->y));
    float hist[NB];
    memset(hist, 0, sizeof(hist));
    float bin_scale = (float)NB / two_pi;
    long y_min = cy - radius; if (y_min < 1) y_min = 1;
    long y_max = cy + radius; if (y_max > (long)height - 2) y_max = (long)height - 2; if (y_max < 1) y_max = 1;
    long x_min = cx - radius; if (x_min < 1) x_min = 1;
    long x_max = cx + radius; if (x_max > (long)width - 2) x_max = (long)width - 2; if (x_max < 1) x_max = 1;
    float r_sq_max = (float)(radius * radius);
    for (long y = y_min; y <= y_max; y++) {
        float dy = (float)(y - cy);
        size_t ro = (size_t)y * width;
        for (long x = x_min; x <= x_max; x++) {
            float dx = (float)(x - cx);
            float r_sq = dx * dx + dy * dy;
            if (r_sq > r_sq_max) continue;
            float m = mag[ro + x];
            if (m <= 1e-7f) continue;
            float weight = expf(-r_sq / two_sigma_sq);
            float a = ang[ro + x];
            size_t bin = f2usize(floorf(a * bin_scale)) % NB;
            hist[bin] += m * weight;
        }
    }
    float sm[NB];
    for (int b = 0; b < NB; b++) {
        float prev = hist[(b + NB - 1) % NB], curr = hist[b], next = hist[(b + 1) % NB];
        sm[b] = (prev + curr + next) * (float)(1.0 / 3.0);
    }
    int best_bin = 0;
    float best_val = sm[0];
    for (int b = 1; b < NB; b++)
        if (sm[b] > best_val) { best_val = sm[b]; best_bin = b; }
    if (best_val <= 1e-7f) { out[0] = 0.0f; return 1; }
    float split = 0.80f * best_val;
    int pb[NB]; float pv[NB]; int np = 0;
    for (int b = 0; b < NB; b++) {
        float v = sm[b];
        if (v < split) continue;
        float l = sm[(b + NB - 1) % NB], r = sm[(b + 1) % NB];
        if (v > l && v >= r) { pb[np] = b; pv[np] = v; np++; }
    }
    if (np == 0) { pb[0] = best_bin; pv[0] = best_val; np = 1; }
    else {
        /* stable insertion sort: value desc, then bin asc (pipeline.rs:1323-1327) */
        for (int i = 1; i < np; i++) {
            int b = pb[i]; float v = pv[i]; int j = i - 1;
            while (j >= 0 && (pv[j] < v || (!(pv[j] > v) && !(pv[j] < v) && pb[j] > b))) {
                pb[j + 1] = pb[j]; pv[j + 1] = pv[j]; j--;
            }
            pb[j + 1] = b; pv[j + 1] = v;
        }
        if (np > 4) np = 4;
    }
    for (int i = 0; i < np; i++) {
        int bin = pb[i];
        float val = pv[i];
        float l = sm[(bin + NB - 1) % NB], r = sm[(bin + 1) % NB];
        float denom = l - 2.0f * val + r;
        float off = fabsf(denom) > 1e-6f ? 0.5f * (l - r) / denom : 0.0f;
        off = clampf_rs(off, -0.5f, 0.5f);
        out[i] = rem_euclid_f(((float)bin + 0.5f + off) * (two_pi / (float)NB), two_pi);
    }
    return np;
}

/* pipeline.rs:1352-1480 */
void psr_descriptor_128(const float *mag, const float *ang, int width, int height,
                        const psr_keypoint *kp, float orientation, float desc[PSR_DESC_LEN]) {
    const float two_pi = TWO_PI_F;
    memset(desc, 0, sizeof(float) * PSR_DESC_LEN);
    if (width < 3 || height < 3) return;
    float cos_t = cosf(orientation), sin_t = sinf(orientation);
    float hist_width = fmaxf(3.0f * kp->scale, 1.5f);
    float inv_hist_width = 1.0f / hist_width;
    float exp_scale = -1.0f / (2.0f * 2.0f * 2.0f);
    long max_radius = f2isize(ceilf(hist_width * 1.41421356237309504880f * (4.0f + 1.0f) * 0.5f));
    if (max_radius < 4) max_radius = 4;
    long cx = f2isize(roundf(kp->x)), cy = f2isize(roundf(kp->y));
    long y_min = cy - max_radius; if (y_min < 1) y_min = 1;
    long y_max = cy + max_radius; if (y_max > (long)height - 2) y_max = (long)height - 2; if (y_max < 1) y_max = 1;
    long x_min = cx - max_radius; if (x_min < 1) x_min = 1;
    long x_max = cx + max_radius; if (x_max > (long)width - 2) x_max = (long)width - 2; if (x_max < 1) x_max = 1;
    float bin_per_rad = 8.0f / two_pi;
    for (long y = y_min; y <= y_max; y++) {
        float dy = (float)y - kp->y;
        size_t ro = (size_t)y * width;
        for (long x = x_min; x <= x_max; x++) {
            float dx = (float)x - kp->x;
            float r_rot = (dx * cos_t + dy * sin_t) * inv_hist_width;
            float c_rot = (-dx * sin_t + dy * cos_t) * inv_hist_width;
            float r_bin = r_rot + 4.0f * 0.5f - 0.5f;
            float c_bin = c_rot + 4.0f * 0.5f - 0.5f;
            if (r_bin <= -1.0f || r_bin >= 4.0f || c_bin <= -1.0f || c_bin >= 4.0f) continue;
            float m = mag[ro + x];
            if (m <= 1e-7f) continue;
            float w_gauss = expf((r_rot * r_rot + c_rot * c_rot) * exp_scale);
            float wm = m * w_gauss;
            float rel = rem_euclid_f(ang[ro + x] - orientation, two_pi);
            float o_bin = rel * bin_per_rad;
            long r0 = f2isize(floorf(r_bin)), c0 = f2isize(floorf(c_bin)), o0 = f2isize(floorf(o_bin));
            float dr = r_bin - (float)r0, dc = c_bin - (float)c0, dof = o_bin - (float)o0;
            for (int ir = 0; ir <= 1; ir++) {
                long rb = r0 + ir;
                if (rb < 0 || rb >= 4) continue;
                float wr = ir == 0 ? 1.0f - dr : dr;
                for (int ic = 0; ic <= 1; ic++) {
                    long cb = c0 + ic;
                    if (cb < 0 || cb >= 4) continue;
                    float wc = ic == 0 ? 1.0f - dc : dc;
                    for (int io = 0; io <= 1; io++) {
                        long ob = (o0 + io) % 8; if (ob < 0) ob += 8;
                        float wo = io == 0 ? 1.0f - dof : dof;
                        size_t idx = (size_t)((rb * 4 + cb) * 8 + ob);
                        desc[idx] += wm * wr * wc * wo;
                    }
                }
            }
        }
    }
    float n1 = 0.0f;
    for (int i = 0; i < PSR_DESC_LEN; i++) n1 += desc[i] * desc[i];
    n1 = sqrtf(n1);
    if (n1 > 1e-7f) {
        float inv1 = 1.0f / n1;
        for (int i = 0; i < PSR_DESC_LEN; i++) desc[i] = fminf(desc[i] * inv1, 0.2f);
        float n2 = 0.0f;
        for (int i = 0; i < PSR_DESC_LEN; i++) n2 += desc[i] * desc[i];
        n2 = sqrtf(n2);
        if (n2 > 1e-7f) {
            float inv2 = 1.0f / n2;
            for (int i = 0; i < PSR_DESC_LEN; i++) desc[i] *= inv2;
        }
    }
}

/* pipeline.rs:1012-1019 */
void psr_rootsift(const float *in, float *out) {
    float l1 = 0.0f;
    for (int i = 0; i < PSR_DESC_LEN; i++) l1 += fmaxf(in[i], 0.0f);
    if (l1 <= 1e-7f) { memmove(out, in, sizeof(float) * PSR_DESC_LEN); return; }
    float inv = 1.0f / l1;
    for (int i = 0; i < PSR_DESC_LEN; i++) out[i] = sqrtf(fmaxf(in[i], 0.0f) * inv);
}

/* pipeline.rs:1206-1227 */
static void resolve_octave_and_layer(const psr_keypoint *kp, const psr_pyramid *p, int *oct, int *layer) {
    if (p->n_octaves == 0) { *oct = 0; *layer = 0; return; }
    int o;
    if (kp->layer > 0) o = (int)(kp->octave < (uint32_t)(p->n_octaves - 1) ? kp->octave : (uint32_t)(p->n_octaves - 1));
    else {
        float ratio = fmaxf(kp->scale / BASE_SIGMA, 1.0f);
        size_t v = f2usize(floorf(log2f(ratio)));
        o = (int)(v < (size_t)(p->n_octaves - 1) ? v : (size_t)(p->n_octaves - 1));
    }
    int g_len = p->scales + 3;
    int ss = g_len - 3; if (ss < 1) ss = 1;
    int l;
    if (kp->layer > 0) l = (int)(kp->layer < (uint32_t)(g_len - 1) ? kp->layer : (uint32_t)(g_len - 1));
    else {
        float oct_mult = (float)(1u << o);
        float local_sigma = fmaxf(kp->scale / (BASE_SIGMA * oct_mult), 1.0f);
        size_t v = f2usize(roundf(log2f(local_sigma) * (float)ss));
        size_t hi = (size_t)(g_len - 2 > 1 ? g_len - 2 : 1);
        if (v < 1) v = 1;
        if (v > hi) v = hi;
        l = (int)v;
    }
    *oct = o; *layer = l;
}

/* pipeline.rs:1482-1577 */
size_t psr_descriptors_from_pyramid(const psr_pyramid *p, const psr_keypoint *kps, size_t n, int kind,
                                    psr_feature **feats_out, float **descs_out) {
    *feats_out = NULL; *descs_out = NULL;
    if (p->n_octaves == 0 || n == 0) return 0;
    int nlev = 0, lev_o[128], lev_l[128];
    int *kp_level = (int *)malloc(sizeof(int) * n);
    for (size_t i = 0; i < n; i++) {
        int o, l, j;
        resolve_octave_and_layer(&kps[i], p, &o, &l);
        for (j = 0; j < nlev; j++) if (lev_o[j] == o && lev_l[j] == l) break;
        if (j == nlev) { lev_o[nlev] = o; lev_l[nlev] = l; nlev++; }
        kp_level[i] = j;
    }
    float *mag[128], *angp[128];
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < nlev; j++) {
        int o = lev_o[j];
        size_t P = (size_t)p->h[o] * p->w[o];
        mag[j] = (float *)malloc(sizeof(float) * P);
        angp[j] = (float *)malloc(sizeof(float) * P);
        psr_gradient_planes(p->gauss[o] + (size_t)lev_l[j] * P, p->h[o], p->w[o], mag[j], angp[j]);
    }
    float *ori = (float *)malloc(sizeof(float) * 4 * n);
    int *nori = (int *)malloc(sizeof(int) * n);
    float *dtmp = (float *)malloc(sizeof(float) * PSR_DESC_LEN * 4 * n);
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t i = 0; i < n; i++) {
        int j = kp_level[i];
        int o = lev_o[j];
        float oct_mult = p->scale_factor[o];
        float inv = 1.0f / oct_mult;
        psr_keypoint ok;
        ok.x = kps[i].x * inv;
        ok.y = kps[i].y * inv;
        ok.scale = fmaxf(kps[i].scale * inv, 0.5f * BASE_SIGMA);
        ok.response = kps[i].response;
        ok.octave = (uint32_t)o;
        ok.layer = (uint32_t)lev_l[j];
        int no = psr_keypoint_orientations(mag[j], angp[j], p->w[o], p->h[o], &ok, ori + 4 * i);
        nori[i] = no;
        for (int q = 0; q < no; q++) {
            float *d = dtmp + (size_t)(4 * i + q) * PSR_DESC_LEN;
            psr_descriptor_128(mag[j], angp[j], p->w[o], p->h[o], &ok, ori[4 * i + q], d);
            if (kind == 1) psr_rootsift(d, d);
        }
    }
    size_t total = 0;
    for (size_t i = 0; i < n; i++) total += (size_t)nori[i];
    psr_feature *f = (psr_feature *)malloc(sizeof(psr_feature) * (total ? total : 1));
    float *ds = (float *)malloc(sizeof(float) * PSR_DESC_LEN * (total ? total : 1));
    size_t c = 0;
    for (size_t i = 0; i < n; i++)
        for (int q = 0; q < nori[i]; q++) {
            f[c].x = kps[i].x; f[c].y = kps[i].y; f[c].scale = kps[i].scale;
            f[c].response = kps[i].response; f[c].orientation = ori[4 * i + q];
            memcpy(ds + c * PSR_DESC_LEN, dtmp + (size_t)(4 * i + q) * PSR_DESC_LEN, sizeof(float) * PSR_DESC_LEN);
            c++;
        }
    for (int j = 0; j < nlev; j++) { free(mag[j]); free(angp[j]); }
    free(ori); free(nori); free(dtmp); free(kp_level);
    *feats_out = f; *descs_out = ds;
    return total;
}

/* pipeline.rs:1581-1620 compute_sift_descriptors_with_kind: pyramid parameters inferred from
 * the keypoints, then the features of every keypoint (no truncation) */
size_t psr_compute_descriptors(const float *img, int h, int w, const psr_keypoint *kps, size_t n, int kind,
                               psr_feature **feats_out, float **descs_out) {
    *feats_out = NULL; *descs_out = NULL;
    if (h < 3 || w < 3 || n == 0) return 0;
    uint32_t max_layer = 0;
    for (size_t i = 0; i < n; i++) if (kps[i].layer > max_layer) max_layer = kps[i].layer;
    size_t scales = max_layer > 3 ? max_layer : 3;
    float k = powf(2.0f, 1.0f / (float)scales);
    int upsample = 0;
    for (size_t i = 0; i < n; i++)
        if (kps[i].layer > 0) {   /* the first keypoint with a layer decides */
            float expected = BASE_SIGMA * (float)((size_t)1 << kps[i].octave) * powi_rs(k, (int)kps[i].layer);
            upsample = (kps[i].scale / expected) < 0.707f;
            break;
        }
    size_t max_oct = 0;
    for (size_t i = 0; i < n; i++) {
        size_t o = kps[i].layer > 0 ? (size_t)kps[i].octave : f2usize(floorf(log2f(fmaxf(kps[i].scale / BASE_SIGMA, 1.0f))));
        if (o > max_oct) max_oct = o;
    }
    psr_pyramid p;
    psr_build_pyramid(img, h, w, (int)(max_oct + 1), (int)scales, upsample, &p);
    size_t nf = psr_descriptors_from_pyramid(&p, kps, n, kind, feats_out, descs_out);
    psr_free_pyramid(&p);
    return nf;
}

/* stable merge sort of keypoints by |response| descending (pipeline.rs:1643-1651) */
static int resp_before(const psr_keypoint *a, const psr_keypoint *b) {
    /* true if a must come strictly before b */
    return fabsf(a->response) > fabsf(b->response);
}
static void stable_sort_kp(psr_keypoint *v, size_t n) {
    if (n < 2) return;
    psr_keypoint *tmp = (psr_keypoint *)malloc(sizeof(psr_keypoint) * n);
    for (size_t width = 1; width < n; width *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * width) {
            size_t mid = lo + width < n ? lo + width : n;
            size_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            size_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                if (resp_before(&v[j], &v[i])) tmp[k++] = v[j++];
                else tmp[k++] = v[i++];
            }
            while (i < mid) tmp[k++] = v[i++];
            while (j < hi) tmp[k++] = v[j++];
        }
        memcpy(v, tmp, sizeof(psr_keypoint) * n);
    }
    free(tmp);
}

/* pipeline.rs:1633-1658 */
size_t psr_extract_features(const float *img, int h, int w, int octaves, int scales, float thr,
                            int upsample, int kind, psr_feature **feats, float **descs) {
    psr_pyramid p;
    psr_build_pyramid(img, h, w, octaves, scales, upsample, &p);
    psr_keypoint *kps = NULL;
    size_t n = psr_extract_extrema(&p, thr, &kps);
    if (n > PSR_MAX_FEATURES) {
        stable_sort_kp(kps, n);
        n = PSR_MAX_FEATURES;
    }
    size_t nf = psr_descriptors_from_pyramid(&p, kps, n, kind, feats, descs);
    if (nf > PSR_MAX_FEATURES) nf = PSR_MAX_FEATURES;
    free(kps);
    psr_free_pyramid(&p);
    return nf;
}

/* ---------------------------------------------------------------- matching */

/* pipeline.rs:1700-1717 */
static inline float dist_sq_early_exit(const float *a, const float *b, float cutoff) {
    float sum = 0.0f;
    for (int c = 0; c < 4; c++) {
        for (int i = 32 * c; i < 32 * c + 32; i++) {
            float d = a[i] - b[i];
            sum += d * d;
        }
        if (sum >= cutoff) return sum;
    }
    return sum;
}

void psr_nn2(const float *qd, size_t nq, const float *td, size_t nt,
             uint32_t *best_idx, float *best_d2, float *second_d2) {
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t q = 0; q < nq; q++) {
        float best = INFINITY, second = INFINITY;
        size_t bi = 0;
        const float *a = qd + q * PSR_DESC_LEN;
        for (size_t t = 0; t < nt; t++) {
            float d = dist_sq_early_exit(a, td + t * PSR_DESC_LEN, second);
            if (d < best) { second = best; best = d; bi = t; }
            else if (d < second) second = d;
        }
        best_idx[q] = (uint32_t)bi; best_d2[q] = best; second_d2[q] = second;
    }
}

/* pipeline.rs:1744-1841 */
size_t psr_match_descriptors(const psr_feature *q, const float *qd, size_t nq, const psr_feature *t,
                             const float *td, size_t nt, float ratio_thr, int mutual, psr_match *out) {
    if (nq == 0 || nt < 2) return 0;
    uint32_t *bi = (uint32_t *)malloc(sizeof(uint32_t) * nq);
    float *b = (float *)malloc(sizeof(float) * nq), *s = (float *)malloc(sizeof(float) * nq);
    psr_nn2(qd, nq, td, nt, bi, b, s);
    size_t c = 0;
    for (size_t i = 0; i < nq; i++) {
        if (s[i] > 1e-12f && isfinite(b[i])) {
            float d1 = sqrtf(b[i]), d2 = sqrtf(s[i]);
            float r = d1 / d2;
            if (r < ratio_thr) {
                psr_match m;
                m.query_index = i; m.target_index = bi[i];
                m.qx = q[i].x; m.qy = q[i].y; m.tx = t[bi[i]].x; m.ty = t[bi[i]].y;
                m.distance = d1; m.ratio = r;
                m.dx = t[bi[i]].x - q[i].x; m.dy = t[bi[i]].y - q[i].y;
                out[c++] = m;
            }
        }
    }
    free(bi); free(b); free(s);
    if (!mutual || c == 0) return c;
    char *keep = (char *)malloc(c);
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t k = 0; k < c; k++) {
        const float *tdesc = td + out[k].target_index * PSR_DESC_LEN;
        float best = INFINITY;
        size_t bq = (size_t)-1;
        for (size_t i = 0; i < nq; i++) {
            float d = dist_sq_early_exit(tdesc, qd + i * PSR_DESC_LEN, best);
            if (d < best) { best = d; bq = i; }
        }
        keep[k] = bq == out[k].query_index;
    }
    size_t m = 0;
    for (size_t k = 0; k < c; k++) if (keep[k]) out[m++] = out[k];
    free(keep);
    return m;
}

/* ---------------------------------------------------------------- Hamming */

/* catalog.rs:37-45: bit i set iff desc[i] > 0.065 */
void psr_quantize_u128(const float *desc, size_t n, uint64_t *out) {
    for (size_t j = 0; j < n; j++) {
        uint64_t lo = 0, hi = 0;
        const float *d = desc + j * PSR_DESC_LEN;
        for (int i = 0; i < 128; i++)
            if (d[i] > 0.065f) { if (i < 64) lo |= 1ull << i; else hi |= 1ull << (i - 64); }
        out[2 * j] = lo; out[2 * j + 1] = hi;
    }
}

static inline unsigned ham128(const uint64_t *a, const uint64_t *b) {
    return (unsigned)(__builtin_popcountll(a[0] ^ b[0]) + __builtin_popcountll(a[1] ^ b[1]));
}

/* catalog.rs:707-728 */
size_t psr_count_coarse_hamming_votes(const uint64_t *q, size_t nq, const uint64_t *e, size_t ne) {
    if (ne == 0) return 0;
    size_t votes = 0;
    for (size_t i = 0; i < nq; i++) {
        unsigned mind = 0xffffffffu;
        for (size_t j = 0; j < ne; j++) {
            unsigned d = ham128(q + 2 * i, e + 2 * j);
            if (d < mind) { mind = d; if (mind <= 8) break; }
        }
        if (mind <= 24) votes++;
    }
    return votes;
}

/* top-k by (distance asc, index asc): counting select over the 129 distance values */
void psr_hamming_topk(const uint64_t *cat, size_t n, const uint64_t *qs, size_t nq, int k,
                      uint32_t *out_idx, uint8_t *out_dist) {
#pragma omp parallel
    {
        uint8_t *d = (uint8_t *)malloc(n ? n : 1);
#pragma omp for schedule(dynamic, 1)
        for (size_t qi = 0; qi < nq; qi++) {
            size_t hist[129];
            memset(hist, 0, sizeof(hist));
            for (size_t j = 0; j < n; j++) { d[j] = (uint8_t)ham128(qs + 2 * qi, cat + 2 * j); hist[d[j]]++; }
            size_t kk = (size_t)k < n ? (size_t)k : n;
            /* threshold distance T: count(d < T) < kk <= count(d <= T) */
            size_t below = 0; int T = 0;
            for (T = 0; T <= 128; T++) { if (below + hist[T] >= kk) break; below += hist[T]; }
            size_t take_eq = kk - below;
            /* emit in (distance, index) order */
            size_t start[129]; size_t acc = 0;
            for (int v = 0; v < T; v++) { start[v] = acc; acc += hist[v]; }
            start[T] = acc;
            for (size_t j = 0; j < n; j++) {
                int v = d[j];
                if (v < T) { size_t p = start[v]++; out_idx[qi * k + p] = (uint32_t)j; out_dist[qi * k + p] = (uint8_t)v; }
                else if (v == T && take_eq) { size_t p = start[T]++; out_idx[qi * k + p] = (uint32_t)j; out_dist[qi * k + p] = (uint8_t)v; take_eq--; }
            }
            for (size_t p = kk; p < (size_t)k; p++) { out_idx[qi * k + p] = 0xffffffffu; out_dist[qi * k + p] = 0xff; }
        }
        free(d);
    }
}

/* FP32 rerank of a Hamming shortlist: sequential squared L2 (the matcher's sum,
 * pipeline.rs:1700-1717 with no cutoff), ordered by (distance asc, index asc); slots with
 * index >= n go last as (UINT32_MAX, +inf). */
typedef struct { float d; uint32_t idx; } psr_rr;
static int rr_cmp(const void *a, const void *b) {
    const psr_rr *x = (const psr_rr *)a, *y = (const psr_rr *)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return x->idx < y->idx ? -1 : x->idx > y->idx;
}

void psr_rerank_l2(const float *qd, size_t nq, const float *cd, size_t n, const uint32_t *sl, int k,
                   uint32_t *out_idx, float *out_d2) {
#pragma omp parallel
    {
        psr_rr *r = (psr_rr *)malloc(sizeof(psr_rr) * (size_t)(k > 0 ? k : 1));
#pragma omp for schedule(dynamic, 4)
        for (size_t q = 0; q < nq; q++) {
            int m = 0;
            for (int c = 0; c < k; c++) {
                uint32_t idx = sl[q * (size_t)k + c];
                if (idx >= n) continue;
                r[m].d = dist_sq_early_exit(qd + q * PSR_DESC_LEN, cd + (size_t)idx * PSR_DESC_LEN, INFINITY);
                r[m].idx = idx;
                m++;
            }
            qsort(r, (size_t)m, sizeof(psr_rr), rr_cmp);
            for (int c = 0; c < k; c++) {
                out_idx[q * (size_t)k + c] = c < m ? r[c].idx : 0xffffffffu;
                out_d2[q * (size_t)k + c] = c < m ? r[c].d : INFINITY;
            }
        }
        free(r);
    }
}

/* ---------------------------------------------------------------- RANSAC */

/* pipeline.rs:1862-1868 */
uint64_t psr_splitmix64_next(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* pipeline.rs:1873-1897 */
static size_t eval_similarity(const psr_match *m, size_t n, float a, float b, float tx, float ty,
                              float thresh_sq, float *err) {
    size_t count = 0;
    float es = 0.0f;
    for (size_t i = 0; i < n; i++) {
        float px = a * m[i].qx - b * m[i].qy + tx;
        float py = b * m[i].qx + a * m[i].qy + ty;
        float dx = px - m[i].tx, dy = py - m[i].ty;
        float e = dx * dx + dy * dy;
        if (e <= thresh_sq) { count++; es += e; }
    }
    *err = es;
    return count;
}

/* pipeline.rs:1900-1974 */
static int fit_procrustes(const psr_match *m, const uint32_t *idx, size_t n, psr_similarity *out) {
    if (n == 0) return 0;
    if (n == 1) {
        out->a = 1.0f; out->b = 0.0f; out->tx = m[idx[0]].dx; out->ty = m[idx[0]].dy;
        out->scale = 1.0f; out->rotation_deg = 0.0f;
        return 1;
    }
    float inv_n = 1.0f / (float)n;
    float mqx = 0, mqy = 0, mtx = 0, mty = 0;
    for (size_t i = 0; i < n; i++) {
        const psr_match *p = &m[idx[i]];
        mqx += p->qx; mqy += p->qy; mtx += p->tx; mty += p->ty;
    }
    mqx *= inv_n; mqy *= inv_n; mtx *= inv_n; mty *= inv_n;
    float sqq = 0, na = 0, nb = 0;
    for (size_t i = 0; i < n; i++) {
        const psr_match *p = &m[idx[i]];
        float qx = p->qx - mqx, qy = p->qy - mqy, tx = p->tx - mtx, ty = p->ty - mty;
        sqq += qx * qx + qy * qy;
        na += qx * tx + qy * ty;
        nb += qx * ty - qy * tx;
    }
    float a = 1.0f, b = 0.0f;
    if (sqq > 1e-5f) {
        float af = na / sqq, bf = nb / sqq;
        float s2 = af * af + bf * bf;
        if (s2 >= 0.04f && s2 <= 25.0f) { a = af; b = bf; }
    }
    out->a = a; out->b = b;
    out->tx = mtx - (a * mqx - b * mqy);
    out->ty = mty - (b * mqx + a * mqy);
    out->scale = sqrtf(a * a + b * b);
    out->rotation_deg = atan2f(b, a) * 57.2957795130823208767981548141051703f;   /* f32::to_degrees */
    return 1;
}

/* pipeline.rs:1979-2002 */
size_t psr_adaptive_ransac_iters(size_t inl, size_t total, int sample_size, double conf,
                                 size_t min_iters, size_t max_cap) {
    if (total == 0 || inl == 0) return max_cap;
    double w = (double)inl / (double)total;
    if (w < 1e-3) w = 1e-3;
    if (w > 1.0 - 1e-6) w = 1.0 - 1e-6;
    double p = 1.0;
    { /* f64::powi, binary exponentiation */
        double a = w; int b = sample_size; p = 1.0;
        for (;;) { if (b & 1) p *= a; b /= 2; if (b == 0) break; a *= a; }
    }
    if (p < 1e-12) p = 1e-12;
    if (p > 1.0 - 1e-9) p = 1.0 - 1e-9;
    double num = log(1.0 - conf), den = log(1.0 - p);
    if (fabs(den) < 1e-12) return min_iters;
    double est = ceil(num / den);
    if (!isfinite(est)) return max_cap;
    size_t e = est <= 0 ? 0 : (est >= 1.8e19 ? (size_t)-1 : (size_t)est);
    if (e < min_iters) e = min_iters;
    if (e > max_cap) e = max_cap;
    return e;
}

/* weighted CDF (pipeline.rs:2006-2028 for the reference 1/ratio weights) */
static double *build_cdf(const psr_match *m, size_t n, const float *wts) {
    double *cdf = (double *)malloc(sizeof(double) * (n ? n : 1));
    double acc = 0.0;
    for (size_t i = 0; i < n; i++) {
        double wv;
        if (wts) wv = (double)wts[i];
        else {
            double r = isfinite(m[i].ratio) ? (double)m[i].ratio : 1.0;
            if (isfinite(m[i].ratio)) { if (r < 0.05) r = 0.05; if (r > 1.0) r = 1.0; }
            wv = 1.0 / r;
        }
        acc += wv;
        cdf[i] = acc;
    }
    if (acc > 0.0) { double inv = 1.0 / acc; for (size_t i = 0; i < n; i++) cdf[i] *= inv; }
    if (n) cdf[n - 1] = 1.0;
    return cdf;
}

/* pipeline.rs:2031-2036 */
static size_t sample_cdf(const double *cdf, size_t n, uint64_t *st) {
    double u = (double)(psr_splitmix64_next(st) >> 11) * (1.0 / (double)(1ull << 53));
    size_t lo = 0, hi = n; /* partition_point(c < u) */
    while (lo < hi) { size_t mid = lo + (hi - lo) / 2; if (cdf[mid] < u) lo = mid + 1; else hi = mid; }
    size_t mx = n ? n - 1 : 0;
    return lo < mx ? lo : mx;
}

static size_t sample_uniform(size_t n, uint64_t *st) {
    /* build-defined uniform sampler: same u in [0,1) as the reference, index floor(u*n) */
    double u = (double)(psr_splitmix64_next(st) >> 11) * (1.0 / (double)(1ull << 53));
    size_t i = (size_t)(u * (double)n);
    return i < n ? i : n - 1;
}

/* stable sort of indices by ratio ascending (pipeline.rs:2072-2078) */
static void sort_idx_by_ratio(const psr_match *m, uint32_t *idx, size_t n) {
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (size_t width = 1; width < n; width *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * width) {
            size_t mid = lo + width < n ? lo + width : n, hi = lo + 2 * width < n ? lo + 2 * width : n;
            size_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                if (m[idx[j]].ratio < m[idx[i]].ratio) tmp[k++] = idx[j++];
                else tmp[k++] = idx[i++];
            }
            while (i < mid) tmp[k++] = idx[i++];
            while (j < hi) tmp[k++] = idx[j++];
        }
        memcpy(idx, tmp, sizeof(uint32_t) * n);
    }
    free(tmp);
}

/* pipeline.rs:2055-2191 */
size_t psr_filter_matches_ransac(const psr_match *m, size_t n, float reproj, uint64_t seed,
                                 uint32_t *inl, psr_similarity *model, int *has_model) {
    *has_model = 0;
    if (n == 0) return 0;
    if (n == 1) { inl[0] = 0; *has_model = fit_procrustes(m, inl, 1, model); return 1; }
    float thresh = fmaxf(reproj, 0.5f), tsq = thresh * thresh;
    uint32_t *si = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (size_t i = 0; i < n; i++) si[i] = (uint32_t)i;
    sort_idx_by_ratio(m, si, n);
    float ba = 1.0f, bb = 0.0f, btx = m[si[0]].dx, bty = m[si[0]].dy, berr;
    size_t bc = eval_similarity(m, n, ba, bb, btx, bty, tsq, &berr);
    for (size_t t = 0; t < n && t < 32; t++) {
        float tx = m[si[t]].dx, ty = m[si[t]].dy, e;
        size_t c = eval_similarity(m, n, 1.0f, 0.0f, tx, ty, tsq, &e);
        if (c > bc || (c == bc && e < berr)) { bc = c; berr = e; ba = 1.0f; bb = 0.0f; btx = tx; bty = ty; }
    }
    double *cdf = build_cdf(m, n, NULL);
    uint64_t st = seed;
    size_t max_it = psr_adaptive_ransac_iters(bc, n, 2, 0.999, 64, 1000), it = 0;
    while (it < max_it) {
        it++;
        size_t i1 = sample_cdf(cdf, n, &st), i2 = sample_cdf(cdf, n, &st);
        if (i2 == i1) i2 = (i1 + 1) % n;
        const psr_match *m1 = &m[i1], *m2 = &m[i2];
        float dxq = m2->qx - m1->qx, dyq = m2->qy - m1->qy, dq = dxq * dxq + dyq * dyq;
        if (dq < 16.0f) continue;
        float dxt = m2->tx - m1->tx, dyt = m2->ty - m1->ty;
        float a = (dxq * dxt + dyq * dyt) / dq, b = (dxq * dyt - dyq * dxt) / dq;
        float s2 = a * a + b * b;
        if (!(s2 >= 0.04f && s2 <= 25.0f)) continue;
        float tx = m1->tx - (a * m1->qx - b * m1->qy), ty = m1->ty - (b * m1->qx + a * m1->qy), e;
        size_t c = eval_similarity(m, n, a, b, tx, ty, tsq, &e);
        if (c > bc || (c == bc && e < berr)) {
            bc = c; berr = e; ba = a; bb = b; btx = tx; bty = ty;
            max_it = psr_adaptive_ransac_iters(bc, n, 2, 0.999, 64, 1000);
        }
    }
    size_t ni = 0;
    for (size_t i = 0; i < n; i++) {
        float dx = (ba * m[i].qx - bb * m[i].qy + btx) - m[i].tx, dy = (bb * m[i].qx + ba * m[i].qy + bty) - m[i].ty;
        if (dx * dx + dy * dy <= tsq) inl[ni++] = (uint32_t)i;
    }
    psr_similarity ref;
    int has_ref = fit_procrustes(m, inl, ni, &ref);
    if (has_ref) {
        uint32_t *ri = (uint32_t *)malloc(sizeof(uint32_t) * n);
        size_t nr = 0;
        for (size_t i = 0; i < n; i++) {
            float dx = (ref.a * m[i].qx - ref.b * m[i].qy + ref.tx) - m[i].tx;
            float dy = (ref.b * m[i].qx + ref.a * m[i].qy + ref.ty) - m[i].ty;
            if (dx * dx + dy * dy <= tsq) ri[nr++] = (uint32_t)i;
        }
        if (nr >= ni) { memcpy(inl, ri, sizeof(uint32_t) * nr); ni = nr; }
        free(ri);
    }
    psr_similarity fin;
    if (fit_procrustes(m, inl, ni, &fin)) { *model = fin; *has_model = 1; }
    else if (has_ref) { *model = ref; *has_model = 1; }
    free(si); free(cdf);
    return ni;
}

/* pipeline.rs:2276-2353 */
static int smallest_eigvec9(double a[9][9], double h[9]) {
    double v[9][9];
    memset(v, 0, sizeof(v));
    for (int i = 0; i < 9; i++) v[i][i] = 1.0;
    for (int sweep = 0; sweep < 20; sweep++) {
        double mo = 0.0;
        for (int p = 0; p < 9; p++) for (int q = p + 1; q < 9; q++) { double av = fabs(a[p][q]); if (av > mo) mo = av; }
        if (mo < 1e-13) break;
        for (int p = 0; p < 8; p++)
            for (int q = p + 1; q < 9; q++) {
                double apq = a[p][q];
                if (fabs(apq) < 1e-15) continue;
                double app = a[p][p], aqq = a[q][q];
                double tau = (aqq - app) / (2.0 * apq);
                double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
                double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                a[p][p] = app - t * apq;
                a[q][q] = aqq + t * apq;
                a[p][q] = 0.0; a[q][p] = 0.0;
                for (int r = 0; r < 9; r++) {
                    if (r != p && r != q) {
                        double arp = a[r][p], arq = a[r][q];
                        a[r][p] = c * arp - s * arq; a[p][r] = a[r][p];
                        a[r][q] = s * arp + c * arq; a[q][r] = a[r][q];
                    }
                    double vrp = v[r][p], vrq = v[r][q];
                    v[r][p] = c * vrp - s * vrq;
                    v[r][q] = s * vrp + c * vrq;
                }
            }
    }
    int mi = 0; double me = a[0][0];
    for (int i = 1; i < 9; i++) if (a[i][i] < me) { me = a[i][i]; mi = i; }
    for (int r = 0; r < 9; r++) { h[r] = v[r][mi]; if (!isfinite(h[r])) return 0; }
    return 1;
}

/* pipeline.rs:2357-2404 */
static int degenerate_tri(const double px[4], const double py[4]) {
    static const int T[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    for (int t = 0; t < 4; t++) {
        int i = T[t][0], j = T[t][1], k = T[t][2];
        if (!isfinite(px[i]) || !isfinite(py[i]) || !isfinite(px[j]) || !isfinite(py[j]) ||
            !isfinite(px[k]) || !isfinite(py[k]))
            return 1;
        double dx1 = px[j] - px[i], dy1 = py[j] - py[i], dx2 = px[k] - px[i], dy2 = py[k] - py[i];
        double dx3 = px[k] - px[j], dy3 = py[k] - py[j];
        double cross = fabs(dx1 * dy2 - dy1 * dx2);
        double e1 = dx1 * dx1 + dy1 * dy1, e2 = dx2 * dx2 + dy2 * dy2, e3 = dx3 * dx3 + dy3 * dy3;
        double me = fmax(fmax(e1, e2), e3);
        if (cross < 4.0 || cross < 1e-3 * me) return 1;
    }
    return 0;
}
static int degenerate_quad(const psr_match *m, const size_t idx[4]) {
    double qx[4], qy[4], tx[4], ty[4];
    for (int i = 0; i < 4; i++) {
        qx[i] = m[idx[i]].qx; qy[i] = m[idx[i]].qy; tx[i] = m[idx[i]].tx; ty[i] = m[idx[i]].ty;
    }
    return degenerate_tri(qx, qy) || degenerate_tri(tx, ty);
}

/* pipeline.rs:2408-2528 over m[idx[0..n]] */
static int fit_dlt(const psr_match *m, const size_t *idx, size_t n, double H[3][3]) {
    if (n < 4) return 0;
    double inv_n = 1.0 / (double)n, cxq = 0, cyq = 0, cxt = 0, cyt = 0;
    for (size_t i = 0; i < n; i++) {
        const psr_match *p = &m[idx ? idx[i] : i];
        double qx = p->qx, qy = p->qy, tx = p->tx, ty = p->ty;
        if (!isfinite(qx) || !isfinite(qy) || !isfinite(tx) || !isfinite(ty)) return 0;
        cxq += qx; cyq += qy; cxt += tx; cyt += ty;
    }
    cxq *= inv_n; cyq *= inv_n; cxt *= inv_n; cyt *= inv_n;
    double mdq = 0, mdt = 0;
    for (size_t i = 0; i < n; i++) {
        const psr_match *p = &m[idx ? idx[i] : i];
        double dxq = (double)p->qx - cxq, dyq = (double)p->qy - cyq, dxt = (double)p->tx - cxt, dyt = (double)p->ty - cyt;
        mdq += sqrt(dxq * dxq + dyq * dyq);
        mdt += sqrt(dxt * dxt + dyt * dyt);
    }
    mdq *= inv_n; mdt *= inv_n;
    if (!isfinite(mdq) || !isfinite(mdt) || mdq < 1e-6 || mdt < 1e-6) return 0;
    const double SQ2 = 1.41421356237309504880;
    double sq = SQ2 / mdq, stt = SQ2 / mdt;
    double ata[9][9];
    memset(ata, 0, sizeof(ata));
    for (size_t i = 0; i < n; i++) {
        const psr_match *p = &m[idx ? idx[i] : i];
        double x = ((double)p->qx - cxq) * sq, y = ((double)p->qy - cyq) * sq;
        double xp = ((double)p->tx - cxt) * stt, yp = ((double)p->ty - cyt) * stt;
        double r1[9] = {-x, -y, -1.0, 0.0, 0.0, 0.0, x * xp, y * xp, xp};
        double r2[9] = {0.0, 0.0, 0.0, -x, -y, -1.0, x * yp, y * yp, yp};
        for (int a = 0; a < 9; a++) for (int b = a; b < 9; b++) ata[a][b] += r1[a] * r1[b] + r2[a] * r2[b];
    }
    for (int a = 0; a < 9; a++) for (int b = 0; b < a; b++) ata[a][b] = ata[b][a];
    double hv[9];
    if (!smallest_eigvec9(ata, hv)) return 0;
    double hn[3][3] = {{hv[0], hv[1], hv[2]}, {hv[3], hv[4], hv[5]}, {hv[6], hv[7], hv[8]}};
    double tq[3][3] = {{sq, 0.0, -sq * cxq}, {0.0, sq, -sq * cyq}, {0.0, 0.0, 1.0}};
    double ist = 1.0 / stt;
    double tti[3][3] = {{ist, 0.0, cxt}, {0.0, ist, cyt}, {0.0, 0.0, 1.0}};
    double tmp[3][3];
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) tmp[r][c] = hn[r][0] * tq[0][c] + hn[r][1] * tq[1][c] + hn[r][2] * tq[2][c];
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) H[r][c] = tti[r][0] * tmp[0][c] + tti[r][1] * tmp[1][c] + tti[r][2] * tmp[2][c];
    if (fabs(H[2][2]) > 1e-12) {
        double inv = 1.0 / H[2][2];
        for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) H[r][c] *= inv;
    }
    return 1;
}

int psr_fit_homography_dlt(const psr_match *m, size_t n, double h_out[9]) {
    double H[3][3];
    if (!fit_dlt(m, NULL, n, H)) return 0;
    for (int i = 0; i < 9; i++) h_out[i] = H[i / 3][i % 3];
    return 1;
}

/* pipeline.rs:2550-2576; optionally returns the sampling mass of the inliers */
static size_t eval_h(const psr_match *m, size_t n, double H[3][3], double tsq, double *err,
                     const double *wts, double *mass) {
    size_t c = 0;
    double es = 0.0, ms = 0.0;
    for (size_t i = 0; i < n; i++) {
        double qx = m[i].qx, qy = m[i].qy;
        double w = H[2][0] * qx + H[2][1] * qy + H[2][2];
        if (fabs(w) < 1e-9) continue;
        double iw = 1.0 / w;
        double px = (H[0][0] * qx + H[0][1] * qy + H[0][2]) * iw;
        double py = (H[1][0] * qx + H[1][1] * qy + H[1][2]) * iw;
        double dx = px - (double)m[i].tx, dy = py - (double)m[i].ty;
        double e = dx * dx + dy * dy;
        if (e <= tsq) { c++; es += e; if (wts) ms += wts[i]; }
    }
    *err = es;
    if (mass) *mass = ms;
    return c;
}

static size_t collect_h(const psr_match *m, size_t n, double H[3][3], double tsq, size_t *out) {
    size_t c = 0;
    for (size_t i = 0; i < n; i++) {
        double qx = m[i].qx, qy = m[i].qy;
        double w = H[2][0] * qx + H[2][1] * qy + H[2][2];
        if (fabs(w) < 1e-9) continue;
        double iw = 1.0 / w;
        double px = (H[0][0] * qx + H[0][1] * qy + H[0][2]) * iw;
        double py = (H[1][0] * qx + H[1][1] * qy + H[1][2]) * iw;
        double dx = px - (double)m[i].tx, dy = py - (double)m[i].ty;
        if (dx * dx + dy * dy <= tsq) out[c++] = i;
    }
    return c;
}

/* pipeline.rs:2605-2630 */
static void local_opt(const psr_match *m, size_t n, double H[3][3], size_t *bc, double *be, double tsq,
                      size_t *scratch) {
    for (int it = 0; it < 2; it++) {
        size_t ni = collect_h(m, n, H, tsq, scratch);
        if (ni < 4) break;
        double R[3][3];
        if (!fit_dlt(m, scratch, ni, R)) break;
        double re;
        size_t rc = eval_h(m, n, R, tsq, &re, NULL, NULL);
        if (rc > *bc || (rc == *bc && re < *be)) { memcpy(H, R, sizeof(R)); *bc = rc; *be = re; }
        else break;
    }
}

static void h_to_f32(double H[3][3], float out[9]) {
    for (int i = 0; i < 9; i++) out[i] = (float)H[i / 3][i % 3];
}

/* pipeline.rs:2663-2755, generalised over the sampler and termination rule */
size_t psr_homography_ransac(const psr_match *m, size_t n, float reproj, uint64_t sim_seed,
                             uint64_t hom_seed, const psr_ransac_params *prm, uint32_t *inl,
                             float h_out[9], int *has_model, uint32_t *hyps_used) {
    *has_model = 0;
    if (hyps_used) *hyps_used = 0;
    if (n == 0) return 0;
    double thresh = (double)fmaxf(reproj, 0.5f), tsq = thresh * thresh;
    double H[3][3];
    int have = 0;
    size_t bc = 0;
    double be = INFINITY;
    size_t n_sim = 0;
    uint32_t *sim_inl = (uint32_t *)malloc(sizeof(uint32_t) * n);
    double *wd = NULL;
    double wsum = 0.0;
    int sampler = prm->sampler;
    if (sampler == 2) {
        /* caller weights with no usable mass (a weight negative or not finite, or a total
           that is not finite and positive) fall back to uniform sampling: the reference's
           sampler always has positive mass (pipeline.rs:2006-2028) */
        int ok = 1;
        wd = (double *)malloc(sizeof(double) * n);
        for (size_t i = 0; i < n; i++) {
            wd[i] = (double)prm->weights[i];
            if (!(wd[i] >= 0.0) || !isfinite(wd[i])) ok = 0;
            wsum += wd[i];
        }
        if (!ok || !isfinite(wsum) || !(wsum > 0.0)) { sampler = 1; wsum = 0.0; }
    }
    if (prm->sim_seed_model) {
        psr_similarity s; int hs;
        n_sim = psr_filter_matches_ransac(m, n, reproj, sim_seed, sim_inl, &s, &hs);
        if (hs) {
            double Hs[3][3] = {{(double)s.a, -(double)s.b, (double)s.tx}, {(double)s.b, (double)s.a, (double)s.ty}, {0.0, 0.0, 1.0}};
            memcpy(H, Hs, sizeof(H));
            have = 1;
            bc = eval_h(m, n, H, tsq, &be, NULL, NULL);
        }
        if (n < 4) {
            memcpy(inl, sim_inl, sizeof(uint32_t) * n_sim);
            if (have) { h_to_f32(H, h_out); *has_model = 1; }
            free(sim_inl); free(wd);
            return n_sim;
        }
    } else if (n < 4) {
        free(sim_inl); free(wd);
        return 0;
    }
    double *cdf = sampler == 1 ? NULL : build_cdf(m, n, sampler == 2 ? prm->weights : NULL);
    uint64_t st = hom_seed;
    size_t min_it = (size_t)prm->min_iters, cap = (size_t)prm->max_iters;
    /* termination: reference uses inlier_count/total; the uniform and guided variants use
       the inlier mass under their own sampling distribution (equal for uniform) */
    double bmass = 0.0;
    size_t max_it;
    if (!prm->adaptive) max_it = cap;
    else if (sampler == 2) {
        if (have) { double e; eval_h(m, n, H, tsq, &e, wd, &bmass); }
        max_it = have && wsum > 0 ? psr_adaptive_ransac_iters((size_t)(bmass / wsum * 1e9), 1000000000, 4, prm->confidence, min_it, cap) : cap;
    } else max_it = psr_adaptive_ransac_iters(bc, n, 4, prm->confidence, min_it, cap);
    size_t it = 0;
    size_t *scratch = (size_t *)malloc(sizeof(size_t) * n);
    while (it < max_it) {
        it++;
        size_t si[4] = {0, 0, 0, 0};
        size_t chosen = 0, att = 0;
        while (chosen < 4 && att < 16) {
            att++;
            size_t c = sampler == 1 ? sample_uniform(n, &st) : sample_cdf(cdf, n, &st);
            int dup = 0;
            for (size_t j = 0; j < chosen; j++) if (si[j] == c) dup = 1;
            if (!dup) si[chosen++] = c;
        }
        if (chosen < 4) continue;
        if (degenerate_quad(m, si)) continue;
        double Hc[3][3];
        if (!fit_dlt(m, si, 4, Hc)) continue;
        double e, ms = 0.0;
        size_t c = eval_h(m, n, Hc, tsq, &e, wd, &ms);
        if (c > bc || (c == bc && e < be)) {
            if (prm->local_opt) local_opt(m, n, Hc, &c, &e, tsq, scratch);
            bc = c; be = e;
            memcpy(H, Hc, sizeof(H));
            have = 1;
            if (prm->adaptive) {
                if (sampler == 2) {
                    double e2; eval_h(m, n, H, tsq, &e2, wd, &bmass);
                    max_it = wsum > 0 ? psr_adaptive_ransac_iters((size_t)(bmass / wsum * 1e9), 1000000000, 4, prm->confidence, min_it, cap) : cap;
                } else max_it = psr_adaptive_ransac_iters(bc, n, 4, prm->confidence, min_it, cap);
            }
        }
    }
    if (hyps_used) *hyps_used = (uint32_t)it;
    size_t ni;
    if (!have) {
        memcpy(inl, sim_inl, sizeof(uint32_t) * n_sim);
        ni = n_sim;
    } else {
        if (prm->local_opt) local_opt(m, n, H, &bc, &be, tsq, scratch);
        ni = collect_h(m, n, H, tsq, scratch);
        for (size_t i = 0; i < ni; i++) inl[i] = (uint32_t)scratch[i];
        h_to_f32(H, h_out);
        *has_model = 1;
    }
    free(scratch); free(cdf); free(sim_inl); free(wd);
    return ni;
}

size_t psr_homography_ransac_ref(const psr_match *m, size_t n, float thresh, uint64_t seed,
                                 int seed_is_default, uint32_t *inl, float h_out[9], int *has_model) {
    psr_ransac_params prm = {0, NULL, 0.999, 64, 2000, 1, 1, 1};
    uint64_t sim_seed, hom_seed;
    if (seed_is_default) { sim_seed = 0xD1B54A32D192ED03ull; hom_seed = 0x517CC1B727220A95ull; } /* pipeline.rs:173-175 */
    else { sim_seed = seed; hom_seed = seed ^ 0x9E3779B97F4A7C15ull; }                         /* pipeline.rs:2650-2661 */
    return psr_homography_ransac(m, n, thresh, sim_seed, hom_seed, &prm, inl, h_out, has_model, NULL);
}

/* ---------------------------------------------------------------- match-scoring MLP
 * 8 -> 64 (ReLU) -> 64 (ReLU) -> 1 logit; parameters (4801 f32) laid out as
 * W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3. Weights are used as their BF16
 * (round-to-nearest-even) values when bf16 != 0; arithmetic in f64. Loss: mean binary
 * cross-entropy with logits. The gradient is with respect to the (rounded) weights. */
static double bf16_round(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    b = (b + 0x7fffu + ((b >> 16) & 1u)) & 0xffff0000u;
    float r;
    memcpy(&r, &b, 4);
    return (double)r;
}

void psr_mlp_forward(const float *params, int bf16, const float *x, size_t n, float *logit) {
    double P[PSR_MLP_PARAMS];
    for (int i = 0; i < PSR_MLP_PARAMS; i++) P[i] = bf16 ? bf16_round(params[i]) : (double)params[i];
    const double *W1 = P, *b1 = P + 512, *W2 = P + 576, *b2 = P + 4672, *w3 = P + 4736, *b3 = P + 4800;
    for (size_t s = 0; s < n; s++) {
        double h1[64], z = b3[0];
        for (int i = 0; i < 64; i++) {
            double a = b1[i];
            for (int k = 0; k < 8; k++) a += W1[i * 8 + k] * (double)x[s * 8 + k];
            h1[i] = a > 0 ? a : 0;
        }
        for (int i = 0; i < 64; i++) {
            double a = b2[i];
            for (int j = 0; j < 64; j++) a += W2[i * 64 + j] * h1[j];
            z += w3[i] * (a > 0 ? a : 0);
        }
        logit[s] = (float)z;
    }
}

double psr_mlp_loss_grad(const float *params, int bf16, const float *x, const float *y, size_t n, double *grad) {
    double P[PSR_MLP_PARAMS];
    for (int i = 0; i < PSR_MLP_PARAMS; i++) { P[i] = bf16 ? bf16_round(params[i]) : (double)params[i]; grad[i] = 0.0; }
    const double *W1 = P, *b1 = P + 512, *W2 = P + 576, *b2 = P + 4672, *w3 = P + 4736, *b3 = P + 4800;
    double *gW1 = grad, *gb1 = grad + 512, *gW2 = grad + 576, *gb2 = grad + 4672, *gw3 = grad + 4736, *gb3 = grad + 4800;
    double loss = 0.0, inv = n ? 1.0 / (double)n : 0.0;
    for (size_t s = 0; s < n; s++) {
        double z1[64], h1[64], z2[64], h2[64], z = b3[0];
        for (int i = 0; i < 64; i++) {
            double a = b1[i];
            for (int k = 0; k < 8; k++) a += W1[i * 8 + k] * (double)x[s * 8 + k];
            z1[i] = a; h1[i] = a > 0 ? a : 0;
        }
        for (int i = 0; i < 64; i++) {
            double a = b2[i];
            for (int j = 0; j < 64; j++) a += W2[i * 64 + j] * h1[j];
            z2[i] = a; h2[i] = a > 0 ? a : 0;
            z += w3[i] * h2[i];
        }
        double t = (double)y[s];
        loss += (z > 0 ? z : 0) - z * t + log1p(exp(-fabs(z)));
        double dz = (1.0 / (1.0 + exp(-z)) - t) * inv;
        gb3[0] += dz;
        double dz2[64], dh1[64];
        for (int i = 0; i < 64; i++) {
            gw3[i] += dz * h2[i];
            dz2[i] = z2[i] > 0 ? dz * w3[i] : 0.0;
            gb2[i] += dz2[i];
        }
        for (int j = 0; j < 64; j++) dh1[j] = 0.0;
        for (int i = 0; i < 64; i++)
            for (int j = 0; j < 64; j++) { gW2[i * 64 + j] += dz2[i] * h1[j]; dh1[j] += W2[i * 64 + j] * dz2[i]; }
        for (int i = 0; i < 64; i++) {
            double d1 = z1[i] > 0 ? dh1[i] : 0.0;
            gb1[i] += d1;
            for (int k = 0; k < 8; k++) gW1[i * 8 + k] += d1 * (double)x[s * 8 + k];
        }
    }
    return loss * inv;
}
