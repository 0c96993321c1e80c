// Device side of the panorama's pairwise stage (stitch.rs:497-571): the feature-set and RANSAC
// problem lists of all image pairs, and one pair edge per pair (the safe alignment homography,
// its inverse and the mean patch intensities of the kept inliers).
#include <hip/hip_runtime.h>

#include "ps_internal.h"

#pragma clang fp contract(off)

namespace ps {

namespace {

// image i's features: rows [i * cap, i * cap + count_i) of the batch arrays
__global__ void pano_sets_kernel(const uint32_t *counts, uint32_t n, uint32_t cap, ps_feature_set *sets) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sets[i] = ps_feature_set{i * cap, min(counts[i], cap)};
}

// stitch.rs:504-512: pairs with fewer than 6 raw matches make no edge, so their RANSAC
// problems are empty
__global__ void pano_ransac_problems_kernel(const ps_match_problem *probs, const uint32_t *mcount, uint32_t np,
                                            uint64_t sim_seed, uint64_t hom_seed, ps_ransac_problem *rp) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const uint32_t n = mcount[k];
    rp[k] = ps_ransac_problem{probs[k].out_offset, n >= 6 ? n : 0u, sim_seed, hom_seed};
}

// stitch.rs:203-226: mean of the up-to-3x3 pixel patch around the rounded point (clamped into
// the image), the pixels read as luma_to_array2 does (u8 * (1/255) in f32), summed in f64 in
// row order, times 255
__device__ double patch_intensity(const uint8_t *img, uint32_t h, uint32_t w, uint32_t stride, float x, float y) {
    if (h == 0 || w == 0 || !isfinite(x) || !isfinite(y)) return 0.0;
    const float rx = roundf(x), ry = roundf(y);   // f32::round: halves away from zero
    const uint32_t cx = rx <= 0.0f ? 0u : (rx >= (float)(w - 1) ? w - 1 : (uint32_t)rx);
    const uint32_t cy = ry <= 0.0f ? 0u : (ry >= (float)(h - 1) ? h - 1 : (uint32_t)ry);
    const uint32_t y0 = cy > 0 ? cy - 1 : 0u, y1 = min(cy + 1, h - 1);
    const uint32_t x0 = cx > 0 ? cx - 1 : 0u, x1 = min(cx + 1, w - 1);
    const float inv255 = 1.0f / 255.0f;
    double sum = 0.0;
    uint32_t cnt = 0;
    for (uint32_t yy = y0; yy <= y1; yy++)
        for (uint32_t xx = x0; xx <= x1; xx++) {
            sum += (double)((float)img[(size_t)yy * stride + xx] * inv255);
            cnt++;
        }
    return (sum / (double)cnt) * 255.0;
}

// visualize.rs:303-355: the source rectangle's corners stay on one side of the horizon, map
// to a strictly convex quadrilateral of consistent orientation, with an area between 0.04x
// and 25x the source's
__device__ bool sane_homography(const float *m, uint32_t src_w, uint32_t src_h) {
    if (src_w < 2 || src_h < 2) return false;
    const double W = (double)(src_w - 1), H = (double)(src_h - 1);
    const double cxs[4] = {0.0, W, W, 0.0}, cys[4] = {0.0, 0.0, H, H};
    const double wc = (double)m[6] * (W * 0.5) + (double)m[7] * (H * 0.5) + (double)m[8];
    if (!isfinite(wc) || fabs(wc) < 1e-6) return false;
    double px[4], py[4];
    for (int c = 0; c < 4; c++) {
        const double wi = (double)m[6] * cxs[c] + (double)m[7] * cys[c] + (double)m[8];
        if (!isfinite(wi) || wi * wc <= 0.0 || fabs(wi) < 0.1 * fabs(wc)) return false;
        px[c] = ((double)m[0] * cxs[c] + (double)m[1] * cys[c] + (double)m[2]) / wi;
        py[c] = ((double)m[3] * cxs[c] + (double)m[4] * cys[c] + (double)m[5]) / wi;
        if (!isfinite(px[c]) || !isfinite(py[c])) return false;
    }
    int pos = 0, neg = 0;
    double shoelace = 0.0;
    for (int i = 0; i < 4; i++) {
        const int b = (i + 1) & 3, c = (i + 2) & 3;
        shoelace += px[i] * py[b] - px[b] * py[i];
        const double cross = (px[b] - px[i]) * (py[c] - py[b]) - (py[b] - py[i]) * (px[c] - px[b]);
        if (cross > 1e-3) pos++;
        else if (cross < -1e-3) neg++;
        else return false;
    }
    if (!(pos == 4 || neg == 4)) return false;
    const double quad = 0.5 * fabs(shoelace), src = W * H;
    return quad >= 0.04 * src && quad <= 25.0 * src;
}

// stitch.rs:504-566 for one pair per workgroup. Thread 0 decides the edge (kept inliers: the
// homography's when at least as many as the similarity's; alignment by
// resolve_safe_alignment_homography, visualize.rs:361-374; its f64 inverse); the patch
// intensities of the kept inliers are sampled in parallel, 256 at a time through LDS, and
// thread 0 adds them in inlier order.
__global__ __launch_bounds__(256) void pano_edge_kernel(const uint8_t *pixels, const ps_image_desc *imgs,
                                                        const ps_match_problem *probs, const uint32_t *mcount,
                                                        const ps_feature_match *matches, const ps_ransac_result *rr,
                                                        const uint32_t *hom_inl, const void *work, size_t total,
                                                        ps_panorama_pair *out) {
    __shared__ double s_a[256], s_b[256];
    __shared__ int s_ok;
    const uint32_t k = blockIdx.x, tid = threadIdx.x;
    const ps_match_problem P = probs[k];
    const uint32_t i = P.query_set, j = P.target_set, n = mcount[k];
    const ps_ransac_result &R = rr[k];
    const uint32_t n_sim = R.n_sim, n_hom = R.n_inliers;
    const bool hom_best = n_hom >= n_sim;
    const uint32_t nb = hom_best ? n_hom : n_sim;
    ps_panorama_pair &O = out[k];
    if (tid == 0) {
        O.from = i;
        O.to = j;
        O.raw_matches = n;
        O.has_edge = 0;
        O.inliers = 0;
        O.mean_i_from = O.mean_i_to = 0.0;
        for (int c = 0; c < 9; c++) O.h_from_to[c] = O.h_to_from[c] = 0.0;
        bool ok = n >= 6 && nb >= 6;
        float h[9];
        if (ok) {
            if (R.has_model && n_hom >= 6 && sane_homography(R.h, imgs[i].width, imgs[i].height)) {
                for (int c = 0; c < 9; c++) h[c] = R.h[c];
            } else if (R.has_sim) {   // visualize.rs:290-299
                const float a = R.sim[0], b = R.sim[1];
                const float t[9] = {a, -b, R.sim[2], b, a, R.sim[3], 0.0f, 0.0f, 1.0f};
                for (int c = 0; c < 9; c++) h[c] = t[c];
            } else {
                ok = false;
            }
        }
        double inv[9];
        if (ok) ok = invert3_f32(h, inv);
        if (ok) {
            for (int c = 0; c < 9; c++) {
                O.h_from_to[c] = (double)h[c];
                O.h_to_from[c] = inv[c];
            }
            O.inliers = nb;
        }
        s_ok = ok ? 1 : 0;
    }
    __syncthreads();
    if (!s_ok) return;
    const uint32_t *src = hom_best ? hom_inl + P.out_offset : ransac_sim_inliers(work, total, P.out_offset, n);
    const ps_feature_match *m = matches + P.out_offset;
    const ps_image_desc A = imgs[i], B = imgs[j];
    double sum_i = 0.0, sum_j = 0.0;
    for (uint32_t c0 = 0; c0 < nb; c0 += 256) {
        const uint32_t c = c0 + tid;
        if (c < nb) {
            const ps_feature_match &M = m[src[c]];
            s_a[tid] = patch_intensity(pixels + A.offset, A.height, A.width, A.row_stride, M.qx, M.qy);
            s_b[tid] = patch_intensity(pixels + B.offset, B.height, B.width, B.row_stride, M.tx, M.ty);
        }
        __syncthreads();
        if (tid == 0) {
            const uint32_t cnt = nb - c0 < 256u ? nb - c0 : 256u;
            for (uint32_t t = 0; t < cnt; t++) {
                sum_i += s_a[t];
                sum_j += s_b[t];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double inv_c = 1.0 / (double)nb;
        const double mi = sum_i * inv_c, mj = sum_j * inv_c;
        O.mean_i_from = mi > 1.0 ? mi : 1.0;   // f64::max(1.0); a NaN mean gives 1.0, as there
        O.mean_i_to = mj > 1.0 ? mj : 1.0;
        O.has_edge = 1;
    }
}

}  // namespace

hipError_t launch_pano_sets(const uint32_t *counts, uint32_t n, uint32_t cap, ps_feature_set *sets, hipStream_t st) {
    if (n == 0) return hipSuccess;
    pano_sets_kernel<<<(n + 255) / 256, 256, 0, st>>>(counts, n, cap, sets);
    return hipGetLastError();
}

hipError_t launch_pano_ransac_problems(const ps_match_problem *probs, const uint32_t *mcount, uint32_t np,
                                       uint64_t sim_seed, uint64_t hom_seed, ps_ransac_problem *rp, hipStream_t st) {
    if (np == 0) return hipSuccess;
    pano_ransac_problems_kernel<<<(np + 255) / 256, 256, 0, st>>>(probs, mcount, np, sim_seed, hom_seed, rp);
    return hipGetLastError();
}

hipError_t launch_pano_edges(const uint8_t *pixels, const ps_image_desc *imgs, const ps_match_problem *probs,
                             const uint32_t *mcount, uint32_t np, const ps_feature_match *matches,
                             const ps_ransac_result *rr, const uint32_t *hom_inl, const void *work, size_t total,
                             ps_panorama_pair *out, hipStream_t st) {
    if (np == 0) return hipSuccess;
    pano_edge_kernel<<<np, 256, 0, st>>>(pixels, imgs, probs, mcount, matches, rr, hom_inl, work, total, out);
    return hipGetLastError();
}

}  // namespace ps
