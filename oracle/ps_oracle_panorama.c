/*
 * CPU restatement of the panorama pairwise stage's per-pair helpers (test infrastructure
 * only: the checker for the GPU path, never shipped or called by it):
 *   visualize.rs:290-299  resolve_alignment_homography (similarity as a homography)
 *   visualize.rs:303-355  is_sane_homography_for_dims
 *   visualize.rs:361-374  resolve_safe_alignment_homography
 *   stitch.rs:203-226     sample_patch_intensity
 * The f64 expressions keep the reference's order; built with -ffp-contract=off.
 */
#include <math.h>
#include <stddef.h>

#include "ps_oracle.h"

/* visualize.rs:303-355 */
int psr_is_sane_homography(const float m[9], size_t src_w, size_t src_h) {
    if (src_w < 2 || src_h < 2) return 0;
    const double w = (double)(src_w - 1), ht = (double)(src_h - 1);
    const double cx[4] = {0.0, w, w, 0.0}, cy[4] = {0.0, 0.0, ht, ht};
    const double wc = (double)m[6] * (w * 0.5) + (double)m[7] * (ht * 0.5) + (double)m[8];
    if (!isfinite(wc) || fabs(wc) < 1e-6) return 0;
    double px[4], py[4];
    for (int idx = 0; idx < 4; idx++) {
        const double wi = (double)m[6] * cx[idx] + (double)m[7] * cy[idx] + (double)m[8];
        if (!isfinite(wi) || wi * wc <= 0.0 || fabs(wi) < 0.1 * fabs(wc)) return 0;
        px[idx] = ((double)m[0] * cx[idx] + (double)m[1] * cy[idx] + (double)m[2]) / wi;
        py[idx] = ((double)m[3] * cx[idx] + (double)m[4] * cy[idx] + (double)m[5]) / wi;
        if (!isfinite(px[idx]) || !isfinite(py[idx])) return 0;
    }
    int sign_pos = 0, sign_neg = 0;
    double shoelace = 0.0;
    for (int i = 0; i < 4; i++) {
        const int b = (i + 1) % 4, c = (i + 2) % 4;
        shoelace += px[i] * py[b] - px[b] * py[i];
        const double cross = (px[b] - px[i]) * (py[c] - py[b]) - (py[b] - py[i]) * (px[c] - px[b]);
        if (cross > 1e-3)
            sign_pos++;
        else if (cross < -1e-3)
            sign_neg++;
        else
            return 0;
    }
    if (!(sign_pos == 4 || sign_neg == 4)) return 0;
    const double quad_area = 0.5 * fabs(shoelace), src_area = w * ht;
    return quad_area >= 0.04 * src_area && quad_area <= 25.0 * src_area;
}

/* visualize.rs:361-374 with :290-299; sim = (a, b, tx, ty). Returns 1 with out set, or 0. */
int psr_resolve_safe_alignment(int has_h, const float h[9], size_t hom_inliers, int has_sim, const float sim[4],
                               size_t src_w, size_t src_h, float out[9]) {
    if (has_h && hom_inliers >= 6 && psr_is_sane_homography(h, src_w, src_h)) {
        for (int i = 0; i < 9; i++) out[i] = h[i];
        return 1;
    }
    if (!has_sim) return 0;
    const float t[9] = {sim[0], -sim[1], sim[2], sim[1], sim[0], sim[3], 0.0f, 0.0f, 1.0f};
    for (int i = 0; i < 9; i++) out[i] = t[i];
    return 1;
}

/* stitch.rs:203-226 over an f32 luma plane (h x w, row-major) */
double psr_sample_patch_intensity(const float *gray, size_t h, size_t w, float x, float y) {
    if (h == 0 || w == 0 || !isfinite(x) || !isfinite(y)) return 0.0;
    /* (v.round() as isize).clamp(0, n - 1): Rust's round takes halves away from zero, as roundf */
    const float rx = roundf(x), ry = roundf(y);
    const size_t cx = rx <= 0.0f ? 0 : (rx >= (float)(w - 1) ? w - 1 : (size_t)rx);
    const size_t cy = ry <= 0.0f ? 0 : (ry >= (float)(h - 1) ? h - 1 : (size_t)ry);
    const size_t y0 = cy > 0 ? cy - 1 : 0, y1 = cy + 1 < h - 1 ? cy + 1 : h - 1;
    const size_t x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < w - 1 ? cx + 1 : w - 1;
    double sum = 0.0;
    size_t cnt = 0;
    for (size_t yy = y0; yy <= y1; yy++)
        for (size_t xx = x0; xx <= x1; xx++) {
            sum += (double)gray[yy * w + xx];
            cnt++;
        }
    return cnt == 0 ? 0.0 : (sum / (double)cnt) * 255.0;
}
