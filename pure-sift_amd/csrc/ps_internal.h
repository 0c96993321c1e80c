// Internal declarations shared by the HIP translation units of libps_sift.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/ps_sift.h"

namespace ps {

constexpr int kMaxOct = 8;
constexpr int kMaxScales = 10;
constexpr int kMaxLevels = kMaxScales + 3;    // Gaussians per octave
constexpr int kMaxTaps = 33;                  // 2*16+1
constexpr uint32_t kCandCap = 1u << 20;       // extremum candidates per image before ordering
constexpr uint32_t kErrCandOverflow = 1u;
constexpr uint32_t kErrRadius = 2u;

// Keypoint plus its scan-order key: (octave << 56) | (slice << 48) | (y << 24) | x.
struct Cand {
    float x, y, scale, response;
    uint32_t octave, layer;
    uint64_t key;
};
static_assert(sizeof(Cand) == 32, "Cand layout");

struct Taps {
    float w[kMaxTaps];
};

// Geometry of one image's pyramid inside the engine's pyramid buffer. Octave o holds
// (S+3) Gaussian planes followed by (S+2) DoG planes of h[o] x w[o] floats.
struct PyrPlan {
    int n_oct = 0, S = 0, upsample = 0, in_h = 0, in_w = 0;
    int h[kMaxOct] = {}, w[kMaxOct] = {};
    size_t off[kMaxOct] = {};                 // float offset of octave o
    float scale_factor[kMaxOct] = {};
    int init_r = 0;
    Taps init_taps{};
    int lvl_r[kMaxLevels] = {};               // radius of the blur producing Gaussian s (s >= 1)
    Taps lvl_taps[kMaxLevels]{};
    float eff_scale[kMaxOct][kMaxLevels] = {};// BASE_SIGMA * mult * k^s (pipeline.rs:839)
    float k = 0.f;                            // 2^(1/S) (glibc powf, pipeline.rs:561/827)
    double log2k = 0.0;
    size_t total_floats = 0;
    size_t gauss(int o, int s) const { return off[o] + (size_t)s * h[o] * w[o]; }
    size_t dog(int o, int s) const { return off[o] + (size_t)(S + 3 + s) * h[o] * w[o]; }
};

// Read-only view of the pyramid for the descriptor kernels (passed by value).
struct PyrView {
    const float *base;
    size_t off[kMaxOct];
    int h[kMaxOct], w[kMaxOct];
    float scale_factor[kMaxOct];
    int n_oct, S;
};

enum SrcMode { SRC_F32 = 0, SRC_U8 = 1, SRC_F32_UP = 2, SRC_U8_UP = 3, SRC_DOWN2 = 4 };

struct BlurArgs {
    const void *src;
    size_t src_stride;   // elements of the source type
    int src_h, src_w;    // source dims (pre-upsample / pre-downsample)
    int h, w;            // output dims
    float *dst_g;        // Gaussian out (nullable)
    float *dst_d;        // DoG out = G - src (nullable)
    float *dst_src;      // the source value itself (nullable; materialises G0 for SRC_DOWN2)
    // words the launch's first workgroup zeroes before anything else (the extraction's counters,
    // cleared by its first blur instead of a launch of their own); nullable
    uint32_t *clear = nullptr;
    int n_clear = 0;
};

// host helpers (engine.hip)
bool make_plan(int h, int w, const ps_sift_config &cfg, PyrPlan &p);

// launches (pyramid.hip)
// Blur launch shape: 16, 32 or 64 selects the tile kernel with that output tile height;
// kBlurSlide | rows selects the sliding-window kernel with `rows` output rows per workgroup
// (a positive multiple of kBlurSlideStep).
constexpr int kBlurSlide = 1 << 16;
constexpr int kBlurSlideStep = 16;
hipError_t launch_blur(int radius, int src_mode, const BlurArgs &a, const Taps &t, int shape, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st);
// Extrema of one octave from Gaussian levels 0..S+1; level S+2 is not read: where a slice-S
// candidate needs it, it is formed from level S+1 with top_taps (radius top_r).
hipError_t launch_extrema(const float *gauss_octave, int h, int w, int S, float thr, float mult,
                          const float *eff_scale, double log2k, int octave, Cand *out, uint32_t *count,
                          uint32_t cap, uint32_t *err, const Taps &top_taps, int top_r, hipStream_t st);
// Orders candidates: scan order if n <= limit, else (|response| desc, scan order); writes the
// first min(n, limit) into out and that count into out_count.
size_t rank_scratch_bytes(uint32_t cap_in);
// packed_keys: octave-0 rows and columns < 4096, fewer than 16 slices and 8 octaves, so a
// candidate's scan key packs into 31 bits (enables the direct ordering of up to 32768 candidates).
hipError_t launch_rank(const Cand *cand, const uint32_t *count, uint32_t cap_in, uint32_t limit,
                       Cand *out, uint32_t *out_count, void *scratch, bool packed_keys, hipStream_t st);

// describe.hip
hipError_t launch_orientations(const PyrView &pv, const Cand *kps, const uint32_t *nkp, uint32_t max_kp,
                               float *ori, uint32_t *nori, hipStream_t st);
hipError_t launch_feature_scan(const uint32_t *nkp, const uint32_t *nori, uint32_t max_kp, uint32_t cap,
                               uint32_t *feat_kp, uint32_t *feat_j, uint32_t *nfeat, hipStream_t st);
hipError_t launch_descriptors(const PyrView &pv, const Cand *kps, const float *ori, const uint32_t *feat_kp,
                              const uint32_t *feat_j, const uint32_t *nfeat, uint32_t cap, int kind,
                              ps_feature *out_f, float *out_d, hipStream_t st);
hipError_t launch_cand_to_keypoints(const Cand *c, const uint32_t *n, uint32_t cap, ps_keypoint *out,
                                    hipStream_t st);

// ---- grouped persistent launches (batched extraction): one stage of many images per launch.
// Units (tiles / strips) of a launch are numbered job by job: job j owns units
// [unit0_j, unit0_{j+1}); jobs are sorted by unit0 and job 0 starts at unit 0.
struct BlurJob {
    BlurArgs a;
    uint32_t units_x;   // units per row: ceil(w / 64)
    uint32_t unit0;
};
// units of one level under a launch shape (see launch_blur)
uint32_t group_units(int h, int w, int shape);
// center: every job writes a DoG or source plane (dst_d / dst_src)
hipError_t launch_blur_group(int radius, int src_mode, const BlurJob *jobs, uint32_t n_jobs, uint32_t n_units,
                             const Taps &t, int shape, bool center, hipStream_t st);
struct ExtremaJob {
    const float *gauss;   // Gaussian level 0 of the octave (levels 0..S+1 follow, h x w each)
    Cand *out;
    uint32_t *count, *err;
    int h, w;
    uint32_t units_x, unit0;
};
uint32_t extrema_group_units(int h, int w, int rows);
// one octave index of many images (same config: same thresholds, scales and taps); rows 16 or 32
hipError_t launch_extrema_group(const ExtremaJob *jobs, uint32_t n_jobs, uint32_t n_units, int S, int rows, float thr,
                                float mult, const float *eff_scale, double log2k, int octave, uint32_t cap,
                                const Taps &top_taps, int top_r, hipStream_t st);
struct RankJob {
    const Cand *cand;
    const uint32_t *count;
    Cand *out;
    uint32_t *out_count;
    void *scratch;        // rank_scratch_bytes(cap_in)
};
hipError_t launch_rank_group(const RankJob *jobs, uint32_t n_jobs, uint32_t cap_in, uint32_t limit, hipStream_t st);
struct DescJob {
    PyrView pv;
    const Cand *kps;          // ordered keypoints
    const uint32_t *nkp;      // their count
    float *ori;               // 4 per keypoint
    uint32_t *nori;
    uint32_t *feat_kp, *feat_j;
    uint32_t *nfeat;          // the image's feature count (written)
    ps_feature *out_f;
    float *out_d;
};
// A pointer read from a job table is generic to the compiler, so its accesses would be flat
// loads and stores with 64-bit per-lane addresses. The round trip through the global address
// space lets address-space inference make them global accesses (scalar base + lane offset), as
// for kernel-argument pointers.
template <typename T>
__device__ __forceinline__ T *global_ptr(T *p) {
    return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}
__device__ __forceinline__ BlurArgs global_args(BlurArgs a) {
    a.src = global_ptr(a.src);
    a.dst_g = global_ptr(a.dst_g);
    a.dst_d = global_ptr(a.dst_d);
    a.dst_src = global_ptr(a.dst_src);
    return a;
}
// threadIdx.x through an opaque move: in a persistent loop it keeps the compiler from hoisting
// every thread-index-derived address of the loop body out of the loop, where they would all
// stay live (more than doubling the registers of a tile body) for the whole walk
__device__ __forceinline__ int opaque_tid() {
    int v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "v"((int)threadIdx.x));
    return v;
}
// off[i] = sum over j < i of ceil(min(counts[j * stride], cap) / div), off[n] the total
hipError_t launch_group_offsets(const uint32_t *counts, uint32_t stride, uint32_t cap, uint32_t div, uint32_t n,
                                uint32_t *off, hipStream_t st);
// item g of a grouped walk belongs to image find_off(off, n, g): the last i with off[i] <= g
// (off non-decreasing, off[0] = 0)
__device__ __forceinline__ uint32_t find_off(const uint32_t *off, uint32_t n, uint32_t g) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}
// orientations, feature flattening and descriptors of many images: kp_off / feat_off get the
// n_jobs + 1 exclusive offsets of min(kp_counts[i * kp_stride], max_kp) and of the features
hipError_t launch_describe_group(const DescJob *jobs, uint32_t n_jobs, uint32_t max_kp, uint32_t cap, int kind,
                                 uint32_t *kp_off, uint32_t *feat_off, const uint32_t *kp_counts,
                                 uint32_t kp_stride, const uint32_t *feat_counts, hipStream_t st);

// match.hip
// nq_max <= 262144 (1024 blocks of 256 queries)
size_t match_finish_scratch_bytes(size_t nq_max);
hipError_t launch_match_finish(const ps_feature *qf, const ps_feature *tf, const uint32_t *nq, size_t nq_max,
                               const uint32_t *nt, const uint32_t *best_idx, const float *best,
                               const float *second, const uint32_t *rev_idx, float ratio, int mutual,
                               ps_feature_match *out, uint32_t *out_count, void *scratch, hipStream_t st);

// match_mfma.hip: exact NN/2NN through an FP32 MFMA distance GEMM + exact rescoring
int nn_mfma_splits(size_t nq_max, size_t nt_max);
size_t nn_mfma_scratch_bytes(size_t nq_max, size_t nt_max);
// gen: a per-call number, distinct from the previous call's on the same scratch
hipError_t launch_nn_mfma(const float *qd, const uint32_t *nq, size_t nq_max, const float *td, const uint32_t *nt,
                          size_t nt_max, void *scratch, uint32_t *best_idx, float *best, float *second, uint32_t gen,
                          hipStream_t st);
// Grouped form over feature sets (one chain for many problems); shared = the query and target
// arrays and set tables are the same, so one prep pass serves both sides.
struct MatchSetsArgs {
    const ps_feature *qf, *tf;
    const float *qd, *td;
    size_t q_rows, t_rows;
    const ps_feature_set *qsets, *tsets;
    uint32_t n_qsets, n_tsets, cap;
    const ps_match_problem *probs;
    uint32_t n_probs;
    float ratio;
    bool shared;
};
size_t match_sets_scratch_bytes(const MatchSetsArgs &a);
hipError_t launch_match_sets(const MatchSetsArgs &a, void *scratch, uint32_t gen, ps_feature_match *out,
                             uint32_t *out_counts, hipStream_t st);

// Squared L2 distance of two 128-D f32 rows summed sequentially in element order
// with separate roundings (pipeline.rs:1764-1786 without the early exit, which never
// changes the result because partial sums are monotone).
__device__ __forceinline__ float seq_dist128(const float *a, const float *b) {
    float sum = 0.0f;
    const float4 *pa = reinterpret_cast<const float4 *>(a), *pb = reinterpret_cast<const float4 *>(b);
    for (int i = 0; i < 32; i++) {
        float4 x = pa[i], y = pb[i];
        float e;
        e = x.x - y.x; sum += e * e;
        e = x.y - y.y; sum += e * e;
        e = x.z - y.z; sum += e * e;
        e = x.w - y.w; sum += e * e;
    }
    return sum;
}

// precision.hip: BF16 / FP8 MFMA distances with an exact f32 rescore of the top 8
size_t nn_precision_scratch_bytes(int prec, size_t nq_max, size_t nt_max);
hipError_t launch_nn_precision(int prec, const float *qd, const uint32_t *nq, size_t nq_max, const float *td,
                               const uint32_t *nt, size_t nt_max, void *scratch, uint32_t *best_idx, float *best,
                               float *second, float *approx_best, hipStream_t st);

// ransac.hip: batched similarity-seeded LO-RANSAC, one workgroup per problem
size_t ransac_work_bytes(size_t total_matches);
// sim_only: run the similarity stage alone (sim_seed_model must be set); after either form,
// ransac_sim_inliers() points at a problem's similarity inliers (result.n_sim of them) in work.
hipError_t launch_ransac(const ps_feature_match *matches, const float *weights, const ps_ransac_problem *probs,
                         uint32_t n_probs, const ps_ransac_params &p, size_t total_matches, void *work,
                         uint32_t *inliers, ps_ransac_result *outs, hipStream_t st, bool sim_only = false);
__host__ __device__ inline const uint32_t *ransac_sim_inliers(const void *work, size_t total_matches,
                                                              uint32_t offset, uint32_t n) {
    // workspace: [2 * total f64 (CDF, errors)][per problem at 2 * offset u32: n scratch, n similarity inliers]
    const uint32_t *wu = reinterpret_cast<const uint32_t *>(static_cast<const double *>(work) + total_matches * 2);
    return wu + (size_t)offset * 2 + n;
}

// visualize.rs:202-232 (the f32 homography promoted to f64)
__device__ inline bool invert3_f32(const float *h, double o[9]) {
    double a[9];
    for (int i = 0; i < 9; i++) a[i] = (double)h[i];
    const double c00 = a[4] * a[8] - a[5] * a[7];
    const double c01 = -(a[3] * a[8] - a[5] * a[6]);
    const double c02 = a[3] * a[7] - a[4] * a[6];
    const double c10 = -(a[1] * a[8] - a[2] * a[7]);
    const double c11 = a[0] * a[8] - a[2] * a[6];
    const double c12 = -(a[0] * a[7] - a[1] * a[6]);
    const double c20 = a[1] * a[5] - a[2] * a[4];
    const double c21 = -(a[0] * a[5] - a[2] * a[3]);
    const double c22 = a[0] * a[4] - a[1] * a[3];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (fabs(det) < 1e-12 || !isfinite(det)) return false;
    const double id = 1.0 / det;
    const double t[9] = {c00 * id, c10 * id, c20 * id, c01 * id, c11 * id, c21 * id, c02 * id, c12 * id, c22 * id};
    for (int i = 0; i < 9; i++) o[i] = t[i];
    return true;
}

// fundamental.hip: fundamental-matrix LO-RANSAC and GRIC, one wave per problem
size_t fundamental_work_bytes(size_t total_matches);
hipError_t launch_fundamental(const ps_feature_match *matches, const ps_fundamental_problem *probs, uint32_t n_probs,
                              double sampson_threshold, uint32_t max_iters, double confidence, size_t total_matches,
                              void *work, uint32_t *inliers, ps_fundamental_result *outs, hipStream_t st);
hipError_t launch_gric(const ps_feature_match *matches, const ps_gric_problem *probs, uint32_t n_probs, double sigma,
                       ps_gric_result *outs, hipStream_t st);

// catalog.hip: GPU-resident feature catalog
struct TrackJob {
    uint32_t moff, n, tfoff, qfoff;   // matches at moff (n of them), target / query features at tfoff / qfoff
};
struct TrackOut {
    float mean_dx, mean_dy, median_dx, median_dy, mean_scale_ratio, mean_rotation_deg;
    uint32_t active, pad;
};
hipError_t launch_dequant_u8(const uint8_t *q, uint32_t n, float *out, hipStream_t st);
// votes[e] for entries [0, n_entries): off has n_entries + 1 u32 feature offsets
hipError_t launch_catalog_votes(const uint64_t *sig, const uint32_t *off, uint32_t n_entries, const uint64_t *q,
                                uint32_t nq, bool mih, uint32_t *votes, hipStream_t st);
// sim_res / sim_work: results and workspace of a RANSAC launch over the same matches at 3 px
hipError_t launch_catalog_problems(const uint32_t *cnt, const ps_feature_set *tsets, uint32_t ns, uint32_t nq,
                                   uint64_t sim_seed, uint64_t hom_seed, ps_ransac_problem *rp, TrackJob *jobs,
                                   hipStream_t st);
hipError_t launch_tracking(const ps_feature_match *matches, const TrackJob *jobs, uint32_t n_jobs,
                           const ps_ransac_result *sim_res, const void *sim_work, size_t sim_total,
                           const ps_feature *qf, const ps_feature *tf, TrackOut *out, hipStream_t st);

// pair_geometry.hip: the per-pair post-processing of a raw match set
hipError_t launch_pair_setup(const ps_pair_problem *pairs, uint32_t n_pairs, uint64_t sim_seed, uint64_t hom_seed,
                             uint64_t fund_seed, ps_ransac_problem *rp, ps_ransac_problem *rp3,
                             ps_fundamental_problem *fp, TrackJob *jobs, hipStream_t st);
hipError_t launch_pair_gric_problems(const ps_pair_problem *pairs, uint32_t n_pairs, const ps_ransac_result *rr,
                                     const ps_fundamental_result *fr, ps_gric_problem *gp, hipStream_t st);
hipError_t launch_pair_finish(const ps_pair_problem *pairs, uint32_t n_pairs, const ps_ransac_result *rr,
                              const ps_ransac_result *rr3, const void *work_t, size_t total, const uint32_t *hom_inl,
                              const ps_fundamental_result *fr, const ps_gric_result *gr, const TrackOut *tr,
                              uint32_t *best_inl, ps_pair_geometry *out, hipStream_t st);

// panorama.hip: the panorama's pairwise stage (feature sets, RANSAC problems, pair edges)
hipError_t launch_pano_sets(const uint32_t *counts, uint32_t n, uint32_t cap, ps_feature_set *sets, hipStream_t st);
hipError_t launch_pano_ransac_problems(const ps_match_problem *probs, const uint32_t *mcount, uint32_t np,
                                       uint64_t sim_seed, uint64_t hom_seed, ps_ransac_problem *rp, hipStream_t st);
hipError_t launch_pano_edges(const uint8_t *pixels, const ps_image_desc *imgs, const ps_match_problem *probs,
                             const uint32_t *mcount, uint32_t np, const ps_feature_match *matches,
                             const ps_ransac_result *rr, const uint32_t *hom_inl, const void *work, size_t total,
                             ps_panorama_pair *out, hipStream_t st);


// mlp.hip: match-scoring MLP (features, forward, gradient, Adam)
size_t mlp_grad_scratch_bytes();
hipError_t launch_mlp_forward(const float *params, const float *x, uint32_t n, float *logit, float *prob,
                              hipStream_t st);
hipError_t launch_mlp_grad(const float *params, const float *x, const float *y, uint32_t n, void *scratch,
                           float *grad, float *loss, hipStream_t st);
hipError_t launch_mlp_adam(float *params, float *m1, float *m2, const float *grad, uint32_t step, float lr,
                           hipStream_t st);
hipError_t launch_mlp_features(const ps_feature_match *m, const ps_feature *qf, const ps_feature *tf,
                               const ps_mlp_pair *pairs, uint32_t n_pairs, float *x, hipStream_t st);

// hamming.hip: 128-bit codes (catalog.rs:37-45, :707-728) and the top-k shortlist
constexpr uint32_t kHamMaxK = 1024;       // largest shortlist length
constexpr uint32_t kHamCandCap = 4096;    // candidate slots per query in the fast path
constexpr uint32_t kHamQueryChunk = 8192; // queries per pass (bounds the candidate buffer)
hipError_t launch_quantize_u128(const float *desc, uint32_t n, uint64_t *codes, hipStream_t st);
hipError_t launch_hamming_votes(const uint64_t *q, uint32_t nq, const uint64_t *e, uint32_t ne, uint32_t *scratch,
                                hipStream_t st);
size_t hamming_topk_scratch_bytes(uint32_t nq_chunk);
hipError_t launch_hamming_topk(const uint64_t *cat, uint32_t n, const uint64_t *qs, uint32_t nq, uint32_t k,
                               void *scratch, uint32_t *out_idx, uint8_t *out_dist, hipStream_t st);
hipError_t launch_rerank_l2(const float *qd, uint32_t nq, const float *cd, uint32_t n, const uint32_t *shortlist,
                            uint32_t k, uint32_t *out_idx, float *out_d2, hipStream_t st);

}  // namespace ps
