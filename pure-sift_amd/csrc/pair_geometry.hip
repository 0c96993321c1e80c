// Glue kernels of the per-pair match post-processing (cli.rs:290-318): problem lists for the
// batched RANSAC / fundamental / tracking launches, GRIC problems from their models, and the
// per-pair result record with the kept inlier list. One thread per pair (setup, GRIC
// problems), one workgroup per pair (finish: the inlier copy).
#include <hip/hip_runtime.h>

#include "ps_internal.h"

#pragma clang fp contract(off)

namespace ps {

namespace {

constexpr uint64_t kTrackSimSeed = 0xD1B54A32D192ED03ull;   // pipeline.rs:173 (summarize_tracking's own run)

__global__ void pair_setup_kernel(const ps_pair_problem *pairs, uint32_t n_pairs, uint64_t sim_seed, uint64_t hom_seed,
                                  uint64_t fund_seed, ps_ransac_problem *rp, ps_ransac_problem *rp3,
                                  ps_fundamental_problem *fp, TrackJob *jobs) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_pairs) return;
    const ps_pair_problem P = pairs[k];
    rp[k] = ps_ransac_problem{P.offset, P.n, sim_seed, hom_seed};
    rp3[k] = ps_ransac_problem{P.offset, P.n, kTrackSimSeed, 0};
    // cli.rs:298: the fundamental matrix only for 8 or more raw matches
    fp[k] = ps_fundamental_problem{P.offset, P.n >= 8 ? P.n : 0u, fund_seed};
    jobs[k] = TrackJob{P.offset, P.n, P.tfeat_offset, P.qfeat_offset};
}

__global__ void pair_gric_problems_kernel(const ps_pair_problem *pairs, uint32_t n_pairs, const ps_ransac_result *rr,
                                          const ps_fundamental_result *fr, ps_gric_problem *gp) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_pairs) return;
    ps_gric_problem g;
    g.offset = pairs[k].offset;
    g.n = pairs[k].n;
    g.has_h = rr[k].has_model ? 1 : 0;
    g.has_f = fr[k].has_model ? 1 : 0;
    for (int i = 0; i < 6; i++) g._pad[i] = 0;
    g._pad2 = 0;
    for (int i = 0; i < 9; i++) {
        g.h[i] = rr[k].h[i];
        g.f[i] = fr[k].f[i];
    }
    gp[k] = g;
}

__global__ __launch_bounds__(256) void pair_finish_kernel(const ps_pair_problem *pairs, const ps_ransac_result *rr,
                                                          const ps_ransac_result *rr3, const void *work_t, size_t total,
                                                          const uint32_t *hom_inl, const ps_fundamental_result *fr,
                                                          const ps_gric_result *gr, const TrackOut *tr,
                                                          uint32_t *best_inl, ps_pair_geometry *out) {
    const uint32_t k = blockIdx.x;
    const ps_pair_problem P = pairs[k];
    const ps_ransac_result &R = rr[k];
    const uint32_t n_sim = R.n_sim, n_hom = R.n_inliers;
    // cli.rs:313-317: the homography's inliers when they outnumber the similarity's
    const bool hom_best = n_hom > n_sim;
    const uint32_t nb = hom_best ? n_hom : n_sim;
    const uint32_t *src = hom_best ? hom_inl + P.offset : ransac_sim_inliers(work_t, total, P.offset, P.n);
    for (uint32_t i = threadIdx.x; i < nb; i += 256) best_inl[P.offset + i] = src[i];
    if (threadIdx.x != 0) return;
    ps_pair_geometry o;
    const TrackOut &T = tr[k];
    o.raw_match_count = P.n;
    if (P.n == 0) {   // pipeline.rs:2215-2225
        o.track_match_count = 0;
        o.inlier_ratio = 0.0f;
        o.mean_dx = o.mean_dy = o.median_dx = o.median_dy = 0.0f;
        o.mean_scale_ratio = 1.0f;
        o.mean_rotation_deg = 0.0f;
    } else {
        const uint32_t na = rr3[k].n_sim > 0 ? rr3[k].n_sim : P.n;
        o.track_match_count = na;
        o.inlier_ratio = (float)na / (float)P.n;
        o.mean_dx = T.mean_dx;
        o.mean_dy = T.mean_dy;
        o.median_dx = T.median_dx;
        o.median_dy = T.median_dy;
        o.mean_scale_ratio = T.mean_scale_ratio;
        o.mean_rotation_deg = T.mean_rotation_deg;
    }
    o.has_transform = R.has_sim ? 1 : 0;
    o.n_similarity = n_sim;
    const float a = R.sim[0], b = R.sim[1];
    o.transform[0] = a;
    o.transform[1] = b;
    o.transform[2] = R.sim[2];
    o.transform[3] = R.sim[3];
    o.transform[4] = sqrtf(a * a + b * b);                                  // pipeline.rs:1962
    o.transform[5] = atan2f(b, a) * 57.2957795130823208767981548141051703f;   // pipeline.rs:1963
    if (!R.has_sim)
        for (int i = 0; i < 6; i++) o.transform[i] = 0.0f;
    o.has_homography = R.has_model ? 1 : 0;
    o.n_homography = n_hom;
    for (int i = 0; i < 9; i++) o.homography[i] = R.has_model ? R.h[i] : 0.0f;
    o.has_fundamental = fr[k].has_model ? 1 : 0;
    o.n_fundamental = fr[k].has_model ? fr[k].n_inliers : 0u;
    for (int i = 0; i < 9; i++) o.fundamental[i] = fr[k].f[i];
    o.has_gric = gr[k].has_comparison;
    o.gric_selected = gr[k].selected;
    o.homography_gric = gr[k].homography_gric;
    o.fundamental_gric = gr[k].fundamental_gric;
    o.best_is_homography = hom_best ? 1 : 0;
    o.n_best = nb;
    out[k] = o;
}

}  // namespace

hipError_t launch_pair_setup(const ps_pair_problem *pairs, uint32_t n_pairs, uint64_t sim_seed, uint64_t hom_seed,
                             uint64_t fund_seed, ps_ransac_problem *rp, ps_ransac_problem *rp3,
                             ps_fundamental_problem *fp, TrackJob *jobs, hipStream_t st) {
    if (n_pairs == 0) return hipSuccess;
    pair_setup_kernel<<<(n_pairs + 255) / 256, 256, 0, st>>>(pairs, n_pairs, sim_seed, hom_seed, fund_seed, rp, rp3, fp,
                                                             jobs);
    return hipGetLastError();
}

hipError_t launch_pair_gric_problems(const ps_pair_problem *pairs, uint32_t n_pairs, const ps_ransac_result *rr,
                                     const ps_fundamental_result *fr, ps_gric_problem *gp, hipStream_t st) {
    if (n_pairs == 0) return hipSuccess;
    pair_gric_problems_kernel<<<(n_pairs + 255) / 256, 256, 0, st>>>(pairs, n_pairs, rr, fr, gp);
    return hipGetLastError();
}

hipError_t launch_pair_finish(const ps_pair_problem *pairs, uint32_t n_pairs, const ps_ransac_result *rr,
                              const ps_ransac_result *rr3, const void *work_t, size_t total, const uint32_t *hom_inl,
                              const ps_fundamental_result *fr, const ps_gric_result *gr, const TrackOut *tr,
                              uint32_t *best_inl, ps_pair_geometry *out, hipStream_t st) {
    if (n_pairs == 0) return hipSuccess;
    pair_finish_kernel<<<n_pairs, 256, 0, st>>>(pairs, rr, rr3, work_t, total, hom_inl, fr, gr, tr, best_inl, out);
    return hipGetLastError();
}

}  // namespace ps
