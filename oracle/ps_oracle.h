/*
 * ps_oracle.h -- CPU oracle for the SIFT extract + match hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of pure-sift 0.2.0's
 * CPU pipeline (src/pipeline.rs, src/catalog.rs) used as the parity checker
 * for the HIP implementation and, timed, as the CPU baseline. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product library (libps_sift.so) never links or calls it.
 *
 * Numerics follow the reference's f32 code exactly: no FMA contraction
 * (built with -ffp-contract=off), left-to-right evaluation as written in the
 * Rust source, glibc libm for exp/atan2/pow/sin/cos (the platform libm Rust
 * binds on Linux), binary-exponentiation powi (compiler-rt __powisf2 order),
 * stable sorts, sequential f32 sums.
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * comment of ps_oracle.c.
 */
#ifndef PS_ORACLE_H
#define PS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSR_DESC_LEN 128
#define PSR_MAX_FEATURES 4096
#define PSR_MAX_OCTAVES 8

typedef struct { float x, y, scale, response; uint32_t octave, layer; } psr_keypoint;
typedef struct { float x, y, scale, response, orientation; } psr_feature;
typedef struct {
    uint64_t query_index, target_index;
    float qx, qy, tx, ty, distance, ratio, dx, dy;
} psr_match;

typedef struct {
    int n_octaves;          /* octaves actually built (stops when h<3 || w<3) */
    int scales;
    int h[PSR_MAX_OCTAVES], w[PSR_MAX_OCTAVES];
    float scale_factor[PSR_MAX_OCTAVES];
    float *gauss[PSR_MAX_OCTAVES]; /* (scales+3) planes of h*w, level-major */
    float *dog[PSR_MAX_OCTAVES];   /* (scales+2) planes of h*w */
} psr_pyramid;

/* threading (mirrors the reference's Rayon sites); n<=0 -> all cores */
void psr_set_threads(int n);
int psr_get_threads(void);

/* pipeline.rs:275-296 */
int psr_gaussian_kernel_1d(float sigma, float *out, int cap);
/* pipeline.rs:314-425 */
void psr_gaussian_blur(const float *in, int h, int w, float sigma, float *out);
/* pipeline.rs:429-449 (out is (h/2)x(w/2)) */
void psr_downsample_2x(const float *in, int h, int w, float *out);
/* pipeline.rs:455-483 (out is 2h x 2w) */
void psr_upsample_2x(const float *in, int h, int w, float *out);
/* pipeline.rs:957-966 */
void psr_luma_to_f32(const uint8_t *in, size_t n, float *out);
/* pipeline.rs:550-638 */
int psr_build_pyramid(const float *img, int h, int w, int octaves, int scales,
                      int upsample, psr_pyramid *out);
void psr_free_pyramid(psr_pyramid *p);
/* pipeline.rs:535-547 / 810-916; *out is malloc'd, caller frees with psr_free */
size_t psr_extract_extrema(const psr_pyramid *p, float threshold, psr_keypoint **out);
/* pipeline.rs:969-978 */
size_t psr_detect_keypoints(const float *img, int h, int w, int octaves, int scales,
                            float threshold, int upsample, psr_keypoint **out);
/* pipeline.rs:1233-1347, on caller-provided gradient planes; returns count (<=4) */
int psr_keypoint_orientations(const float *mag, const float *ang, int w, int h,
                              const psr_keypoint *octave_local_kp, float out[4]);
/* pipeline.rs:1168-1203 */
void psr_gradient_planes(const float *gauss, int h, int w, float *mag, float *ang);
/* pipeline.rs:1352-1480 */
void psr_descriptor_128(const float *mag, const float *ang, int w, int h,
                        const psr_keypoint *octave_local_kp, float orientation,
                        float out[PSR_DESC_LEN]);
/* pipeline.rs:1012-1019 */
void psr_rootsift(const float *in, float *out);
/* pipeline.rs:1482-1577: features for an ordered keypoint list (no 4096 truncation) */
size_t psr_descriptors_from_pyramid(const psr_pyramid *p, const psr_keypoint *kps, size_t n,
                                    int kind, psr_feature **feats, float **descs);
/* pipeline.rs:1581-1620 (compute_sift_descriptors_with_kind): pyramid parameters inferred from
 * the keypoints, features for every keypoint in order (no truncation) */
size_t psr_compute_descriptors(const float *img, int h, int w, const psr_keypoint *kps, size_t n, int kind,
                               psr_feature **feats, float **descs);
/* pipeline.rs:1633-1658 (extract_sift_features_with_kind); cap + truncate to 4096 */
size_t psr_extract_features(const float *img, int h, int w, int octaves, int scales,
                            float threshold, int upsample, int kind,
                            psr_feature **feats, float **descs);
/* pipeline.rs:1744-1841; out has capacity nq; returns count */
size_t psr_match_descriptors(const psr_feature *q, const float *qd, size_t nq,
                             const psr_feature *t, const float *td, size_t nt,
                             float ratio, int mutual, psr_match *out);
/* per-query nearest / second-nearest squared distance (pipeline.rs:1764-1786) */
void psr_nn2(const float *qd, size_t nq, const float *td, size_t nt,
             uint32_t *best_idx, float *best_d2, float *second_d2);
/* catalog.rs:37-45 (two little-endian u64 words per descriptor) */
void psr_quantize_u128(const float *desc, size_t n, uint64_t *out);
/* catalog.rs:707-728 */
size_t psr_count_coarse_hamming_votes(const uint64_t *q, size_t nq, const uint64_t *e, size_t ne);
/* top-k Hamming shortlist, ordered by (distance asc, index asc) */
void psr_hamming_topk(const uint64_t *catalog, size_t n, const uint64_t *queries, size_t nq,
                      int k, uint32_t *out_idx, uint8_t *out_dist);
/* FP32 rerank of a shortlist by (sequential squared L2, index) */
void psr_rerank_l2(const float *qd, size_t nq, const float *cd, size_t n, const uint32_t *shortlist, int k,
                   uint32_t *out_idx, float *out_d2);

/* RANSAC (pipeline.rs:1860-2755) */
uint64_t psr_splitmix64_next(uint64_t *state);
typedef struct { float a, b, tx, ty, scale, rotation_deg; } psr_similarity;
/* returns inlier count; inlier_idx (cap n) gets indices into matches; *has_model set */
size_t psr_filter_matches_ransac(const psr_match *m, size_t n, float thresh, uint64_t seed,
                                 uint32_t *inlier_idx, psr_similarity *model, int *has_model);
/* sampler: 0 = reference 1/ratio CDF, 1 = uniform, 2 = caller weights (guided) */
typedef struct {
    int sampler;
    const float *weights;   /* sampler 2: per-match non-negative weights */
    double confidence;      /* adaptive-termination confidence (reference: 0.999) */
    int min_iters, max_iters;
    int local_opt;          /* 1 = LO-RANSAC as the reference */
    int sim_seed_model;     /* 1 = seed with the similarity RANSAC model as the reference */
    int adaptive;           /* 0 = run exactly max_iters hypotheses */
} psr_ransac_params;
size_t psr_homography_ransac(const psr_match *m, size_t n, float thresh, uint64_t sim_seed,
                             uint64_t hom_seed, const psr_ransac_params *prm,
                             uint32_t *inlier_idx, float h_out[9], int *has_model,
                             uint32_t *hyps_used);
/* reference entry point estimate_homography_ransac_with_seed semantics */
size_t psr_homography_ransac_ref(const psr_match *m, size_t n, float thresh, uint64_t seed,
                                 int seed_is_default, uint32_t *inlier_idx, float h_out[9],
                                 int *has_model);
int psr_fit_homography_dlt(const psr_match *m, size_t n, double h_out[9]);
size_t psr_adaptive_ransac_iters(size_t inl, size_t total, int sample_size, double conf,
                                 size_t min_iters, size_t max_cap);

/* match-scoring MLP 8-64-64-1 (4801 parameters), f64 arithmetic on (optionally BF16-rounded)
 * weights; mean BCE-with-logits loss and its gradient */
#define PSR_MLP_PARAMS 4801
void psr_mlp_forward(const float *params, int bf16, const float *x, size_t n, float *logit);
double psr_mlp_loss_grad(const float *params, int bf16, const float *x, const float *y, size_t n, double *grad);
void psr_free(void *p);

/* ---- feature catalog (ps_oracle_catalog.c; catalog.rs) ---- */
/* pipeline.rs:1024-1030 / 1034-1044, n rows of 128 */
void psr_quantize_u8(const float *desc, size_t n, uint8_t *out);
void psr_dequantize_u8(const uint8_t *q, size_t n, float *out);
/* A catalog as flat arrays: entry e has id id_blob[id_off[e] .. id_off[e+1]) and the
 * features off[e] .. off[e+1] (records, 2-word signatures, 128-float descriptors). */
typedef struct {
    uint32_t octaves, scales;
    float threshold;
    uint32_t kind;
    size_t n_entries, n_features;
    const char *id_blob;
    const uint64_t *id_off;
    const uint32_t *width, *height;
    const uint64_t *off;
    const psr_feature *feat;
    const uint64_t *sig;
    const float *desc;
    const void *mih;    /* optional inverted index from psr_catalog_index (else built per call) */
} psr_catalog;
/* catalog.rs:280-297 the inverted index of a catalog (free with psr_catalog_index_free) */
void *psr_catalog_index(const psr_catalog *c);
void psr_catalog_index_free(void *mih);
/* catalog.rs:447-540 over an in-memory file; 0 or -1 with the message in err */
int psr_catalog_parse(const uint8_t *buf, size_t len, psr_catalog *out, char *err, size_t cap);
void psr_catalog_free(psr_catalog *c);
/* catalog.rs:400-445; returns the size (writes only when cap suffices) */
size_t psr_catalog_write(const psr_catalog *c, uint8_t *buf, size_t cap);
/* catalog.rs:595-633 (mih: inverted-index path) */
void psr_catalog_votes(const psr_catalog *c, const uint64_t *q, size_t nq, int mih, uint32_t *votes);
size_t psr_catalog_shortlist(const psr_catalog *c, const uint32_t *votes, int mih, size_t shortlist_k,
                             uint32_t *cand_entry, uint32_t *cand_votes);
/* pipeline.rs:1137-1153 TrackingSummary / :2210-2272 summarize_tracking */
typedef struct {
    uint64_t raw_match_count, match_count;
    float inlier_ratio, mean_dx, mean_dy, median_dx, median_dy, mean_scale_ratio, mean_rotation_deg;
} psr_tracking;
void psr_summarize_tracking(const psr_feature *q, const psr_feature *t, const psr_match *m, size_t n,
                            psr_tracking *out);
/* catalog.rs:158-177 CatalogSearchHit, the layout of the product's ps_catalog_hit */
typedef struct {
    uint32_t entry, coarse_votes;
    uint64_t raw_match_count, match_count;
    float inlier_ratio, mean_dx, mean_dy, median_dx, median_dy, mean_scale_ratio, mean_rotation_deg;
    uint32_t has_transform, has_homography;
    float transform[6];   /* a, b, tx, ty, scale, rotation_deg */
    float homography[9];
} psr_catalog_hit;
/* catalog.rs:576-702; out capacity min(top_k, n_entries); returns the hit count */
size_t psr_catalog_search(const psr_catalog *c, const psr_feature *qf, const float *qd, size_t nq,
                          size_t top_k, size_t shortlist_k, float ratio, float ransac_thresh,
                          psr_catalog_hit *out);

/* ---- two-view geometry (ps_oracle_geometry.c; geometry.rs) ---- */
int psr_jacobi9(const double a[81], double eigvals[9], double eigvecs[81]);
double psr_sampson_sq(const double f[9], double x1, double y1, double x2, double y2);
int psr_fit_fundamental_8point(const psr_match *m, const uint32_t *idx, size_t n, double f[9]);
int psr_fit_fundamental_7point(const psr_match *m, double f[4][9]);
size_t psr_adaptive_epipolar_iters(size_t best, size_t total, int sample_size, double confidence, size_t min_iters,
                                   size_t max_iters);
size_t psr_fundamental_ransac(const psr_match *m, size_t n, double sampson_threshold, size_t max_iters,
                              double confidence, uint64_t seed, uint32_t *inl, double f[9], int *has_model,
                              size_t *iters);
int psr_invert3_f32(const float h[9], double out[9]);

/* ---- panorama pairwise stage helpers (ps_oracle_panorama.c) ---- */
int psr_is_sane_homography(const float m[9], size_t src_w, size_t src_h);
int psr_resolve_safe_alignment(int has_h, const float h[9], size_t hom_inliers, int has_sim, const float sim[4],
                               size_t src_w, size_t src_h, float out[9]);
double psr_sample_patch_intensity(const float *gray, size_t h, size_t w, float x, float y);
int psr_select_model_gric(const psr_match *m, size_t n, const float *h, const double *f, double sigma,
                          int *selected, double *h_gric, double *f_gric);

#ifdef __cplusplus
}
#endif
#endif
