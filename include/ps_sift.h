/*
 * ps_sift.h -- C ABI of the MI355X SIFT extract + match engine (libps_sift.so).
 *
 * Drop-in boundary for pure-sift 0.2.0's hot path. Every entry point below
 * names the reference interface it replaces (file:line in pure-sift 0.2.0).
 * Plain pointers and sizes only; all functions return ps_status (0 = OK,
 * negative = error; ps_status_string / ps_last_error explain). The library
 * never allocates memory that crosses the boundary: callers own every buffer.
 *
 * Thread safety: every function may be called concurrently from any thread.
 * Host-pointer entry points borrow an engine (device stream + workspace) from
 * an internal pool for the duration of the call. Engine entry points run
 * asynchronously on that engine's stream and take device pointers.
 *
 * There is no CPU fallback: without a usable gfx950 device every compute call
 * returns PS_E_NO_DEVICE.
 */
#ifndef PS_SIFT_H
#define PS_SIFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t ps_status;
#define PS_OK 0
#define PS_E_INVALID_ARG (-1)   /* null pointer / bad size */
#define PS_E_CONFIG (-2)        /* SiftConfig / MatchOptions validation failure */
#define PS_E_DEVICE (-3)        /* HIP runtime error */
#define PS_E_CAPACITY (-4)      /* output buffer too small; *out_count holds the needed size */
#define PS_E_NO_DEVICE (-5)     /* no usable GPU */
#define PS_E_OUT_OF_MEMORY (-6) /* device allocation failed */
#define PS_E_INJECTED (-7)      /* failure forced by ps_debug_fail_next (tests only) */
#define PS_E_IO (-8)            /* catalog file could not be read or written */
#define PS_E_FORMAT (-9)        /* malformed catalog bytes (message as catalog.rs:447-540) */

#define PS_DESCRIPTOR_LEN 128      /* pipeline.rs:998 SIFT_DESCRIPTOR_LEN */
#define PS_MAX_FEATURES 4096       /* pipeline.rs:1054 MAX_FEATURES_PER_IMAGE */

#define PS_DESCRIPTOR_SIFT 0       /* pipeline.rs:1002-1008 DescriptorKind::Sift */
#define PS_DESCRIPTOR_ROOTSIFT 1   /* DescriptorKind::RootSift */

/* pipeline.rs:22-34 SiftConfig (same fields, same order) */
typedef struct {
    uint32_t octaves;
    uint32_t scales;
    float contrast_threshold;
    uint8_t upsample_first_octave;
    uint8_t descriptor_kind;
    uint8_t _pad[2];
} ps_sift_config;

/* pipeline.rs:1720-1727 MatchOptions. `precision` (in the reference struct's padding,
 * 0 for reference callers) selects the distance engine: PS_PRECISION_FP32 is the reference's
 * exact f32 search; BF16 / FP8 rank all targets with low-precision MFMA distances and
 * rescore the 8 best with the exact f32 distance before the ratio test. */
#define PS_PRECISION_FP32 0
#define PS_PRECISION_BF16 1
#define PS_PRECISION_FP8 2   /* E4M3 with a per-row power-of-two scale */
typedef struct {
    float ratio;
    uint8_t mutual;
    uint8_t precision;
    uint8_t _pad[2];
} ps_match_options;

/* pipeline.rs:251-267 Keypoint */
typedef struct {
    float x, y, scale, response;
    uint32_t octave, layer;
} ps_keypoint;

/* pipeline.rs:1058-1072 SiftFeature without the descriptor Vec; descriptors travel
 * as a separate row-major [n][128] float array (feature i <-> row i). */
typedef struct {
    float x, y, scale, response, orientation;
} ps_feature;

/* pipeline.rs:1075-1091 FeatureMatch */
typedef struct {
    uint64_t query_index, target_index;
    float qx, qy;   /* query_point */
    float tx, ty;   /* target_point */
    float distance, ratio;
    float dx, dy;   /* displacement */
} ps_feature_match;

/* ------------------------------------------------------------------ setup */

/* Selects and initialises a device (idempotent, thread-safe). device < 0 -> current/0. */
ps_status ps_init(int device);
const char *ps_status_string(ps_status s);
/* Message of the last error raised on the calling thread ("" if none). */
const char *ps_last_error(void);
/* pipeline.rs:80-103 SiftConfig::validate; msg (optional) gets the reference's message. */
ps_status ps_validate_sift_config(const ps_sift_config *cfg, char *msg, size_t msg_cap);
/* pipeline.rs:156-164 Matcher::new validation. */
ps_status ps_validate_match_options(const ps_match_options *opt, char *msg, size_t msg_cap);

/* --------------------------------------------- host-pointer entry points */

/* pipeline.rs:136-145 Sift::detect_and_compute -> extract_sift_features_with_kind
 * (pipeline.rs:1633-1658). img: h x w f32 luma in [0,1], row stride in elements.
 * Writes at most `cap` features (<= 4096 are ever produced; cap >= 4096 always suffices). */
ps_status ps_detect_and_compute(const float *img, size_t h, size_t w, size_t row_stride,
                                const ps_sift_config *cfg, ps_feature *out_feat,
                                float *out_desc, size_t cap, size_t *out_count);

/* Same as ps_detect_and_compute on u8 luma; the (px * (1/255)) conversion of
 * luma_to_array2 (pipeline.rs:957-966) runs fused on the device, identical values. */
ps_status ps_detect_and_compute_u8(const uint8_t *img, size_t h, size_t w, size_t row_stride,
                                   const ps_sift_config *cfg, ps_feature *out_feat,
                                   float *out_desc, size_t cap, size_t *out_count);

/* pipeline.rs:125-133 Sift::detect_keypoints (uncapped, scan order). If more than `cap`
 * keypoints exist, returns PS_E_CAPACITY and sets *out_count to the needed count. */
ps_status ps_detect_keypoints(const float *img, size_t h, size_t w, size_t row_stride,
                              const ps_sift_config *cfg, ps_keypoint *out, size_t cap,
                              size_t *out_count);

/* pipeline.rs:1581-1620 compute_sift_descriptors_with_kind (the uncapped CLI / HTTP extract path
 * of cli.rs:174-212 pairs it with ps_detect_keypoints): orientations and 128-D descriptors for
 * the given keypoints. The pyramid is rebuilt with the parameters the reference infers from the
 * keypoints -- scales = max(max layer, 3); the first octave upsampled when the first keypoint
 * with layer > 0 has scale / (1.6 * 2^octave * k^layer) < 0.707; max octave + 1 octaves, a
 * layer-0 keypoint counting as octave floor(log2(max(scale / 1.6, 1)))) -- and each keypoint's
 * level is resolve_octave_and_layer (pipeline.rs:1206-1227). Features follow keypoint order, each
 * keypoint's orientations in peak order; kind is PS_DESCRIPTOR_SIFT or PS_DESCRIPTOR_ROOTSIFT.
 * Inferred pyramids beyond 8 octaves or 10 scales return PS_E_CONFIG. If more than `cap`
 * features result, returns PS_E_CAPACITY and sets *out_count to the needed count. */
ps_status ps_compute_descriptors(const float *img, size_t h, size_t w, size_t row_stride,
                                 const ps_keypoint *kps, size_t n, uint8_t kind, ps_feature *out_feat,
                                 float *out_desc, size_t cap, size_t *out_count);

/* pipeline.rs:167-169 Matcher::match_ -> match_descriptors_with_options (pipeline.rs:1744-1841).
 * Descriptors are [n][128] f32. out needs capacity >= nq. */
ps_status ps_match(const ps_feature *q, const float *q_desc, size_t nq,
                   const ps_feature *t, const float *t_desc, size_t nt,
                   const ps_match_options *opt, ps_feature_match *out, size_t cap, size_t *out_count);

/* Debug/parity: the full Gaussian + DoG pyramid of build_scale_space_pyramid_with_options
 * (pipeline.rs:550-638). Layout: for each octave o, (scales+3) Gaussian planes then
 * (scales+2) DoG planes of dims[2o] x dims[2o+1] floats, octaves back to back.
 * Pass out=NULL to query *out_len (floats) and dims (cap 2*8 ints) / *n_octaves. */
ps_status ps_build_pyramid(const float *img, size_t h, size_t w, size_t row_stride,
                           const ps_sift_config *cfg, float *out, size_t out_cap, size_t *out_len,
                           int32_t *dims, int32_t *n_octaves);

/* Nearest / second-nearest squared distances and nearest index per query
 * (pipeline.rs:1764-1786), bit-exact with the reference's sequential f32 sum. */
ps_status ps_nn2(const float *q_desc, size_t nq, const float *t_desc, size_t nt,
                 uint32_t *best_idx, float *best_d2, float *second_d2);

/* ---------------------------------------------------- engine entry points */
/* An engine owns one HIP stream and a workspace. Engine calls enqueue work on
 * that stream and return immediately; device pointers only. */
typedef struct ps_engine ps_engine;

ps_status ps_engine_create(int device, ps_engine **out);
ps_status ps_engine_destroy(ps_engine *e);
/* hipStream_t of the engine, as an opaque pointer. */
void *ps_engine_stream(ps_engine *e);
ps_status ps_engine_sync(ps_engine *e);

/* Device-resident extract: d_img u8 h x w (row stride in bytes). Writes up to cap
 * features + descriptors and *d_count (device u32). */
ps_status ps_engine_extract_u8(ps_engine *e, const uint8_t *d_img, size_t h, size_t w,
                               size_t row_stride, const ps_sift_config *cfg,
                               ps_feature *d_feat, float *d_desc, size_t cap, uint32_t *d_count);

/* Device-resident match. Counts are device u32s (as written by ps_engine_extract_u8);
 * nq_max / nt_max bound them (array capacities). Writes d_out[0..*d_out_count). */
ps_status ps_engine_match(ps_engine *e, const ps_feature *d_qf, const float *d_qd,
                          const uint32_t *d_nq, size_t nq_max, const ps_feature *d_tf,
                          const float *d_td, const uint32_t *d_nt, size_t nt_max,
                          const ps_match_options *opt, ps_feature_match *d_out, uint32_t *d_out_count);

/* No counterpart in the reference: ps_engine_match for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, results and validation as ps_engine_match with the exact FP32
 * path (opt->precision must be PS_PRECISION_FP32); every scratch buffer comes from a static
 * 17 MiB device array that each GPU holds until the process ends, so the call allocates
 * nothing (the array is set up by ps_init / ps_engine_create, else by the first call on a
 * device). Limits: nq_max <= 4096 and nt_max <= 4096 (PS_E_INVALID_ARG beyond). All calls on
 * one GPU share the array: calls that overlap on one GPU (several engines or threads) race
 * and must be serialised by the caller. */
ps_status ps_engine_match_static(ps_engine *e, const ps_feature *d_qf, const float *d_qd,
                                 const uint32_t *d_nq, size_t nq_max, const ps_feature *d_tf,
                                 const float *d_td, const uint32_t *d_nt, size_t nt_max,
                                 const ps_match_options *opt, ps_feature_match *d_out,
                                 uint32_t *d_out_count);

/* Many ratio-match problems in one launch chain. A feature set is a row range of a feature /
 * descriptor array (rows [offset, offset + count); count is device data, clamped to set_cap);
 * problem k matches query set problems[k].query_set against target set
 * problems[k].target_set exactly as match_descriptors does (pipeline.rs:1845-1858: Lowe's
 * ratio, no mutual check) and writes its accepted matches in query order to
 * d_out[out_offset ..] (room for the query set's count) and their number to
 * d_out_counts[k]; the records' indices are relative to the two sets. Query and target
 * arrays (and set tables) may be the same. No counterpart in the reference, whose panorama
 * calls match_descriptors once per image pair (stitch.rs:500-503). */
typedef struct {
    uint32_t offset, count;
} ps_feature_set;
typedef struct {
    uint32_t query_set, target_set, out_offset, _pad;
} ps_match_problem;
ps_status ps_engine_match_sets(ps_engine *e, const ps_feature *d_qf, const float *d_qd, size_t q_rows,
                               const ps_feature_set *d_qsets, uint32_t n_qsets, const ps_feature *d_tf,
                               const float *d_td, size_t t_rows, const ps_feature_set *d_tsets, uint32_t n_tsets,
                               uint32_t set_cap, const ps_match_problem *d_problems, uint32_t n_problems,
                               float ratio, ps_feature_match *d_out, uint32_t *d_out_counts);

/* Device-side error word of the engine (0 = none; set when an internal capacity
 * bound was exceeded during an async call). Synchronises the engine. */
ps_status ps_engine_check(ps_engine *e);

/* Stage timing with HIP events recorded on the engine's stream (profiling hook; off by
 * default). Between begin and end, every engine extract/match call records its stage
 * boundaries; end synchronises and returns per-stage totals in milliseconds:
 * [0] octave 0's Gaussian + DoG levels (engine stream), [1] the later octaves (side streams
 * forked during octave 0) and every octave's extrema, up to the join -- [0] + [1] is the
 * whole scale space --, [2] order + cap, [3] orientation + descriptors, [4] match; *calls
 * gets the number of extract calls recorded. */
#define PS_NUM_STAGES 5
ps_status ps_engine_profile_begin(ps_engine *e, uint32_t max_calls);
ps_status ps_engine_profile_end(ps_engine *e, float stage_ms[PS_NUM_STAGES], uint32_t *calls);

/* Nearest / second-nearest search of ps_nn2 with a selectable distance engine
 * (PS_PRECISION_*). approx_best_d2 (may be null) receives the low-precision squared
 * distance of the returned best target, for error reporting. */
ps_status ps_nn2_precision(const float *q_desc, size_t nq, const float *t_desc, size_t nt, int precision,
                           uint32_t *best_idx, float *best_d2, float *second_d2, float *approx_best_d2);

/* Device-resident form on an engine's stream; counts are device u32s bounded by the
 * capacities nq_max / nt_max. d_approx may be null. */
ps_status ps_engine_nn2(ps_engine *e, const float *d_q_desc, const uint32_t *d_nq, size_t nq_max,
                        const float *d_t_desc, const uint32_t *d_nt, size_t nt_max, int precision,
                        uint32_t *d_best_idx, float *d_best_d2, float *d_second_d2, float *d_approx);

/* ---- batched extraction ----------------------------------------------------------------
 * Many uneven-size u8 images in one call: image i is d_pixels[offset ..] (height x width,
 * row_stride bytes); its features / descriptors go to d_feat[i * cap ..] / d_desc[i * cap * 128 ..]
 * and its count to d_counts[i]. Same results as one ps_engine_extract_u8 per image. Every
 * stage runs once for the whole batch as a grouped persistent launch (one pyramid level, one
 * octave's extrema, the ordering, orientations and descriptors of all images per launch); the
 * sequence is captured into a HIP graph on first use and replayed while images, buffers, config
 * and tuned launch shapes are unchanged (PS_BATCH_NO_GRAPH launches eagerly instead). The
 * workspace grows with the batch: about 56 MB per image for extremum candidates and their
 * ordering, plus 20 bytes per pyramid pixel. No counterpart in the reference, whose callers loop
 * over images (cli.rs:2093-2113, stitch.rs:493-499). */
typedef struct {
    uint64_t offset;
    uint32_t height, width, row_stride, _pad;
} ps_image_desc;
#define PS_BATCH_NO_GRAPH 1u
ps_status ps_engine_extract_batch_u8(ps_engine *e, const uint8_t *d_pixels, const ps_image_desc *images,
                                     uint32_t n_images, const ps_sift_config *cfg, ps_feature *d_feat,
                                     float *d_desc, size_t cap, uint32_t *d_counts, uint32_t flags);

/* Blur tile autotuner: for the pyramid levels of the given image sizes (hw = height,
 * width pairs), times each tile height (16, 32, 64 rows) on the device and keeps the
 * fastest per (radius, source kind, level size class) for all later pyramid launches in
 * the process. *n_tuned (may be null) gets the number of level classes tuned. */
ps_status ps_engine_autotune_blur(ps_engine *e, const uint32_t *hw, uint32_t n_sizes, const ps_sift_config *cfg,
                                  uint32_t *n_tuned);
/* Grouped-launch autotuner for a batch (images as for ps_engine_extract_batch_u8): times
 * each launch shape of every grouped blur launch of the batch (tile heights 16 / 32 / 64,
 * sliding-window chunks of 64 / 128 / 192 rows) and both extrema tile heights (16, 32 rows),
 * in dependency order on the real images, and keeps the fastest per grouped-launch class
 * (radius, source kind, mean level size class, image count class) for all later batches in
 * the process. *n_tuned (may be null) gets the number of launch classes tuned. */
ps_status ps_engine_autotune_batch(ps_engine *e, const uint8_t *d_pixels, const ps_image_desc *images,
                                   uint32_t n_images, const ps_sift_config *cfg, uint32_t *n_tuned);
/* Forgets both tuned tables (single-level and grouped). */
void ps_autotune_reset(void);
uint32_t ps_autotune_entries(void);
/* Entry i of the tuned table: radius, source mode, height class, width class (floor(2 log2 of
 * the level's height / width)) and the launch shape chosen (tile rows, or PS_BLUR_SLIDE | chunk
 * rows for the sliding-window kernel), e.g. to persist a table or inspect it. */
#define PS_BLUR_SLIDE 65536
ps_status ps_autotune_entry(uint32_t i, int32_t out[5]);

/* ---- homography RANSAC ---------------------------------------------------------------
 * pipeline.rs:179-213 RansacOptions: reproj_threshold in [0.5, 100]; seed used when
 * has_seed (else the reference's default seeds, pipeline.rs:173-175). */
typedef struct {
    float reproj_threshold;
    uint8_t has_seed;
    uint8_t _pad[3];
    uint64_t seed;
} ps_ransac_options;

/* pipeline.rs:225-235 Homography::estimate -> estimate_homography_ransac(_with_seed)
 * (:2635-2755): similarity-seeded LO-RANSAC. out_inliers (capacity n) receives the inlier
 * indices into m in increasing order, *out_n their count; out_h the row-major homography
 * when *out_has_model. Same inliers and model as the sequential algorithm. */
ps_status ps_homography_estimate(const ps_feature_match *m, size_t n, const ps_ransac_options *opt,
                                 uint32_t *out_inliers, size_t *out_n, float out_h[9], int32_t *out_has_model);

/* Generalised batch form: many problems per launch (one workgroup each), a choice of
 * sampler and termination rule. The reference configuration is PS_SAMPLER_REFERENCE,
 * confidence 0.999, min 64, max 2000 iterations, local_opt = sim_seed_model = adaptive = 1.
 * PS_SAMPLER_WEIGHTS samples proportionally to caller weights (e.g. a learned inlier score)
 * and, when adaptive, stops when the sampled inlier mass reaches the confidence. A problem
 * whose weights have no usable mass (all zero, a total that is not finite and positive, or a
 * weight that is negative or not finite) is sampled uniformly instead, with the uniform
 * sampler's stop rule: uniform sampling is the guided sampler's fallback. The host-pointer
 * form ps_ransac_batch rejects negative or non-finite weights with PS_E_INVALID_ARG. */
#define PS_SAMPLER_REFERENCE 0   /* 1 / clamp(ratio, 0.05, 1), pipeline.rs:2006-2028 */
#define PS_SAMPLER_UNIFORM 1
#define PS_SAMPLER_WEIGHTS 2
typedef struct {
    float reproj_threshold;
    uint8_t sampler;
    uint8_t local_opt;
    uint8_t sim_seed_model;
    uint8_t adaptive;
    double confidence;
    uint32_t min_iters;
    uint32_t max_iters;
} ps_ransac_params;
typedef struct {
    uint32_t offset;    /* first match of the problem in the batch array */
    uint32_t n;         /* its match count */
    uint64_t sim_seed;  /* similarity-stage stream */
    uint64_t hom_seed;  /* homography-stage stream */
} ps_ransac_problem;
typedef struct {
    float h[9];
    uint32_t has_model, n_inliers, hyps;   /* hyps: hypotheses drawn before termination */
    float sim[4];                          /* similarity seed model (a, b, tx, ty) */
    uint32_t has_sim, n_sim;
} ps_ransac_result;
/* Device-resident batch on an engine's stream: d_inliers[offset .. offset + n_inliers) of
 * each problem gets its inlier indices; d_weights (per match) is read for the weighted
 * sampler only. */
ps_status ps_engine_ransac_batch(ps_engine *e, const ps_feature_match *d_matches, const float *d_weights,
                                 const ps_ransac_problem *d_problems, uint32_t n_problems, size_t total_matches,
                                 const ps_ransac_params *params, uint32_t *d_inliers, ps_ransac_result *d_results);
/* Host-pointer form of the batch call. */
ps_status ps_ransac_batch(const ps_feature_match *matches, const float *weights, const ps_ransac_problem *problems,
                          uint32_t n_problems, size_t total_matches, const ps_ransac_params *params,
                          uint32_t *inliers, ps_ransac_result *results);

/* pipeline.rs:238-246 Homography::estimate_similarity -> filter_matches_ransac(_with_seed)
 * (:2043-2191): similarity RANSAC with Procrustes refinement. out_inliers (capacity n) gets
 * the inlier indices in increasing order; out_model = a, b, tx, ty, scale, rotation_deg
 * (SimilarityTransform2D) when *out_has_model. */
ps_status ps_similarity_estimate(const ps_feature_match *m, size_t n, const ps_ransac_options *opt,
                                 uint32_t *out_inliers, size_t *out_n, float out_model[6], int32_t *out_has_model);

/* ---- two-view geometry: fundamental matrix and GRIC model selection (geometry.rs) ------- */
/* geometry.rs:108-128 FundamentalRansacOptions (default: 1.5, 2000, 0.999, 0xF00DCAFE12345678) */
typedef struct {
    double sampson_threshold;
    uint64_t max_iters;
    double confidence;
    uint64_t seed;
} ps_fundamental_options;
/* geometry.rs:857-1016 estimate_fundamental_ransac_with_options: normalised 8-point LO-RANSAC
 * with Sampson scoring (the 7-point solver when n == 7; no model below 7 matches). out_f is
 * the row-major FundamentalMatrix (f64); out_inliers (capacity n) gets the inlier indices in
 * increasing order. max_iters above 1,000,000 is rejected (the device runs the bound as given). */
ps_status ps_fundamental_estimate(const ps_feature_match *m, size_t n, const ps_fundamental_options *opt,
                                  uint32_t *out_inliers, size_t *out_n, double out_f[9], int32_t *out_has_model);
/* Batch form (one workgroup per problem); no counterpart in the reference. */
typedef struct {
    uint32_t offset, n;   /* matches [offset, offset + n) of the batch array */
    uint64_t seed;
} ps_fundamental_problem;
typedef struct {
    double f[9];
    uint32_t has_model, n_inliers, iters, _pad;   /* iters: RANSAC iterations run */
} ps_fundamental_result;
ps_status ps_engine_fundamental_batch(ps_engine *e, const ps_feature_match *d_matches,
                                      const ps_fundamental_problem *d_problems, uint32_t n_problems,
                                      size_t total_matches, double sampson_threshold, uint32_t max_iters,
                                      double confidence, uint32_t *d_inliers, ps_fundamental_result *d_results);
ps_status ps_fundamental_batch(const ps_feature_match *matches, const ps_fundamental_problem *problems,
                               uint32_t n_problems, size_t total_matches, double sampson_threshold, uint32_t max_iters,
                               double confidence, uint32_t *inliers, ps_fundamental_result *results);
/* geometry.rs:1233-1341 select_model_gric (Torr's GRIC; sigma floored at 0.25): h is the
 * row-major f32 homography or null, f the row-major f64 fundamental matrix or null.
 * has_comparison = 0 is the reference's None; selected: 0 Homography, 1 Fundamental. */
typedef struct {
    int32_t has_comparison, selected;
    double homography_gric, fundamental_gric;
} ps_gric_result;
ps_status ps_select_model_gric(const ps_feature_match *m, size_t n, const float *h, const double *f, double sigma,
                               ps_gric_result *out);
/* Batch form: per problem its match range and optional models. */
typedef struct {
    uint32_t offset, n;
    uint8_t has_h, has_f;
    uint8_t _pad[6];
    float h[9];
    uint32_t _pad2;
    double f[9];
} ps_gric_problem;
ps_status ps_engine_gric_batch(ps_engine *e, const ps_feature_match *d_matches, const ps_gric_problem *d_problems,
                               uint32_t n_problems, double sigma, ps_gric_result *d_results);

/* ---- per-pair match post-processing (cli.rs:290-318 run_match_from_bytes) --------------
 * For each pair's raw matches: summarize_tracking (pipeline.rs:2210-2272), the similarity
 * (Homography::estimate_similarity) and the homography (Homography::estimate) at the RANSAC
 * options, the fundamental matrix (estimate_fundamental_ransac_with_options with the default
 * options and Sampson threshold min(threshold, 3); only for 8 or more matches), GRIC (sigma
 * max(threshold, 1)) and the kept inliers (the homography's when they outnumber the
 * similarity's, else the similarity's). */
typedef struct {
    uint32_t offset, n;                    /* raw matches [offset, offset + n) */
    uint32_t qfeat_offset, tfeat_offset;   /* the pair's features in the query / target arrays */
} ps_pair_problem;
typedef struct {
    uint32_t raw_match_count, track_match_count;   /* TrackingSummary */
    float inlier_ratio, mean_dx, mean_dy, median_dx, median_dy, mean_scale_ratio, mean_rotation_deg;
    uint32_t has_transform, n_similarity;
    float transform[6];                            /* a, b, tx, ty, scale, rotation_deg */
    uint32_t has_homography, n_homography;
    float homography[9];
    uint32_t has_fundamental, n_fundamental;
    double fundamental[9];
    int32_t has_gric, gric_selected;               /* gric_selected: 0 Homography, 1 Fundamental */
    double homography_gric, fundamental_gric;
    uint32_t best_is_homography, n_best;
} ps_pair_geometry;
/* Device-resident batch on the engine's stream: d_best_inliers[offset .. offset + n_best) of
 * each pair receives the kept inlier indices. opt as for ps_homography_estimate. */
ps_status ps_engine_pair_geometry_batch(ps_engine *e, const ps_feature_match *d_matches, const ps_feature *d_qfeat,
                                        const ps_feature *d_tfeat, const ps_pair_problem *d_pairs, uint32_t n_pairs,
                                        size_t total_matches, const ps_ransac_options *opt, uint32_t *d_best_inliers,
                                        ps_pair_geometry *d_out);
/* Host form for one pair (q / t: the pair's features; m: its raw matches). */
ps_status ps_pair_geometry_estimate(const ps_feature *q, size_t nq, const ps_feature *t, size_t nt,
                                    const ps_feature_match *m, size_t n, const ps_ransac_options *opt,
                                    uint32_t *out_best_inliers, ps_pair_geometry *out);

/* ---- panorama pairwise stage (stitch.rs:478-571) ------------------------------------------
 * stitch.rs:45-69 PanoramaOptions: the fields this stage reads (projection, blending and
 * canvas limits belong to the rendering that follows it). */
typedef struct {
    uint32_t octaves, scales;
    float threshold, ratio, ransac_threshold;
    uint32_t _pad;
} ps_panorama_options;
/* stitch.rs:104-112 PairEdge (row-major h_from_to) */
typedef struct {
    uint64_t from, to;
    double h_from_to[9];
    uint64_t inliers;
    double mean_i_from, mean_i_to;
} ps_pair_edge;
/* One record per image pair i < j: has_edge = 0 where the reference makes no edge (fewer
 * than 6 raw matches or kept inliers, no alignment, a singular alignment); otherwise the
 * edges i -> j (h_from_to, mean_i_from = image i's mean) and j -> i (h_to_from, the means
 * swapped). raw_matches is the pair's ratio-match count. */
typedef struct {
    uint32_t from, to;
    uint32_t raw_matches, has_edge;
    uint64_t inliers;
    double h_from_to[9];
    double h_to_from[9];
    double mean_i_from, mean_i_to;
} ps_panorama_pair;
/* stitch.rs:478-571: every image's features (extract_sift_features: SIFT descriptors, first
 * octave upsampled, at most 4096 per image) from its luma plane (what to_luma8 yields), then
 * for every pair i < j the ratio matches (match_descriptors), similarity and homography RANSAC
 * at ransac_threshold with the default seeds, the kept inliers (the homography's when at least
 * as many), the safe alignment homography (resolve_safe_alignment_homography,
 * visualize.rs:361-374), its inverse (visualize.rs:202-232) and the mean patch intensities of
 * the kept inliers (stitch.rs:203-226). out gets the pair edges in the reference's order (per
 * pair with an edge: i -> j, then j -> i); *out_count their number (PS_E_CAPACITY when above
 * cap). Fewer than 2 images is the reference's error. */
ps_status ps_panorama_pair_edges(const uint8_t *const *luma, const uint32_t *heights, const uint32_t *widths,
                                 const uint32_t *row_strides, uint32_t n_images, const ps_panorama_options *opt,
                                 ps_pair_edge *out, size_t cap, size_t *out_count);
/* Device-resident form on the engine's stream: images in d_pixels as described by the host
 * array images (as for ps_engine_extract_batch_u8); d_pairs gets n (n - 1) / 2 records in pair
 * order ((0,1), (0,2), ..., (1,2), ...). The engine keeps the features until its next call. */
ps_status ps_engine_panorama_pairs(ps_engine *e, const uint8_t *d_pixels, const ps_image_desc *images,
                                   uint32_t n_images, const ps_panorama_options *opt, ps_panorama_pair *d_pairs);

/* ---- learned proposal ordering for RANSAC: the match-scoring MLP ----------------------
 * No counterpart in the reference (it samples by 1 / ratio, pipeline.rs:2006-2028). An
 * 8 -> 64 -> 64 -> 1 network scores each putative match; its sigmoid output is the sampling
 * weight of PS_SAMPLER_WEIGHTS. Parameters: 4801 f32 masters laid out as W1[64][8], b1[64],
 * W2[64][64], b2[64], w3[64], b3; the network computes with their BF16 values and FP32
 * accumulation. Inputs (8 per match): ratio (NaN -> 1, capped at 2), log(scale_t / scale_q),
 * sin and cos of (orientation_t - orientation_q), descriptor distance, query x / width,
 * query y / height, rank of the match by (ratio, index) within its pair / match count. */
#define PS_MLP_PARAMS 4801
#define PS_MLP_FEATURES 8
typedef struct {
    uint32_t match_offset, n;        /* the pair's matches in the batch array */
    uint32_t qfeat_offset, tfeat_offset;   /* its query / target features (indexed by the matches) */
    float width, height;             /* query image size */
} ps_mlp_pair;
ps_status ps_engine_mlp_features(ps_engine *e, const ps_feature_match *d_matches, const ps_feature *d_qfeat,
                                 const ps_feature *d_tfeat, const ps_mlp_pair *d_pairs, uint32_t n_pairs,
                                 float *d_x);
/* logits (and probabilities, when d_prob is not null) of n matches */
ps_status ps_engine_mlp_forward(ps_engine *e, const float *d_params, const float *d_x, size_t n, float *d_logit,
                                float *d_prob);
/* mean binary cross-entropy (with logits) over n labelled matches and its gradient */
ps_status ps_engine_mlp_grad(ps_engine *e, const float *d_params, const float *d_x, const float *d_y, size_t n,
                             float *d_grad, float *d_loss);
/* Adam step (beta 0.9 / 0.999, eps 1e-8) on the masters; step counts from 1 */
ps_status ps_engine_mlp_adam(ps_engine *e, float *d_params, float *d_m, float *d_v, const float *d_grad,
                             uint32_t step, float lr);
/* host-pointer forms for checks */
ps_status ps_mlp_forward(const float *params, const float *x, size_t n, float *logit);
ps_status ps_mlp_loss_grad(const float *params, const float *x, const float *y, size_t n, float *grad, float *loss);
/* Checkpoint / resume of the network (host buffers of PS_MLP_PARAMS floats). No counterpart in
 * the reference, whose only persisted state is the catalog. The file: magic "PSMLPCKP",
 * version 1, the layer sizes 8-64-64-1, the parameter count, the Adam step, flags (bit 0: the
 * Adam moments m, v follow the parameters), the f32 values and an FNV-1a 64 checksum, all
 * little-endian; written to `path`.tmp and renamed over `path`, as the catalog's files are
 * (catalog.rs:203-245). m and v are both null (parameters only) or both given. Loading
 * checks every header field, the size and the checksum (PS_E_FORMAT, PS_E_IO for file
 * errors); moments present in the file are copied only when m and v are given, *flags (may be
 * null) tells whether the file has them. */
#define PS_MLP_CKPT_MOMENTS 1u
ps_status ps_mlp_checkpoint_save(const char *path, const float *params, const float *m, const float *v,
                                 uint32_t adam_step);
ps_status ps_mlp_checkpoint_load(const char *path, float *params, float *m, float *v, uint32_t *adam_step,
                                 uint32_t *flags);

/* ---- 128-bit descriptor codes and the catalog shortlist ------------------------------
 * A code is two little-endian u64 words per descriptor: bit i (0..127) of the code is bit
 * i % 64 of word i / 64, the layout of the reference's u128 signature (catalog.rs:436). */
#define PS_HAMMING_MAX_K 1024

/* catalog.rs:37-45 quantize_descriptor_u128: bit i = desc[i] > 0.065. out_codes holds
 * 2 * n words. */
ps_status ps_quantize_u128(const float *desc, size_t n, uint64_t *out_codes);

/* catalog.rs:707-728 count_coarse_hamming_votes: the number of query codes whose minimum
 * Hamming distance over the entry codes is <= 24 (0 when ne == 0). */
ps_status ps_count_coarse_hamming_votes(const uint64_t *q_codes, size_t nq, const uint64_t *e_codes,
                                        size_t ne, size_t *out_votes);

/* Exact top-k Hamming shortlist of each query code over a catalog of n codes, ordered by
 * (distance asc, catalog index asc); 1 <= k <= PS_HAMMING_MAX_K, n < 2^32. Slots past
 * min(k, n) hold index UINT32_MAX and distance 255. out_idx / out_dist hold nq * k entries.
 * The reference votes instead of ranking (catalog.rs:596-620, :707-728); this is the
 * shortlist scan that replaces that pass in a GPU-resident catalog search. */
ps_status ps_hamming_topk(const uint64_t *catalog, size_t n, const uint64_t *q_codes, size_t nq,
                          uint32_t k, uint32_t *out_idx, uint8_t *out_dist);

/* FP32 rerank of a shortlist: for query q, the k indices shortlist[q*k ..] into cat_desc
 * (n x 128 f32) are reordered by (exact sequential squared L2 distance, index), the distance
 * of ps_match (pipeline.rs:1764-1786). Slots with index >= n go last as (UINT32_MAX, +inf). */
ps_status ps_rerank_l2(const float *q_desc, size_t nq, const float *cat_desc, size_t n,
                       const uint32_t *shortlist, uint32_t k, uint32_t *out_idx, float *out_d2);

/* Device-resident forms of the three calls above, asynchronous on the engine's stream. */
ps_status ps_engine_quantize_u128(ps_engine *e, const float *d_desc, size_t n, uint64_t *d_codes);
ps_status ps_engine_hamming_topk(ps_engine *e, const uint64_t *d_catalog, size_t n,
                                 const uint64_t *d_q_codes, size_t nq, uint32_t k,
                                 uint32_t *d_idx, uint8_t *d_dist);
ps_status ps_engine_rerank_l2(ps_engine *e, const float *d_q_desc, size_t nq, const float *d_cat_desc,
                              size_t n, const uint32_t *d_shortlist, uint32_t k, uint32_t *d_idx,
                              float *d_d2);

/* ---- GPU-resident feature catalog (catalog.rs) ----------------------------------------
 * FeatureCatalog (catalog.rs:179-197) with every entry's features, stored 128-bit signatures
 * and f32 descriptors resident in device memory of the library's device. Entry order, id
 * lookup and removal follow the reference (upsert replaces in place, remove swaps the last
 * entry in). Calls that change a catalog need exclusive access to it; searches may run
 * concurrently. */
typedef struct ps_catalog ps_catalog;
typedef struct {
    uint64_t n_entries, n_features;
    uint32_t octaves, scales;
    float contrast_threshold;
    uint8_t descriptor_kind;
    uint8_t _pad[3];
} ps_catalog_info;
typedef struct {
    const char *id;          /* UTF-8, not NUL-terminated; valid until the catalog changes */
    uint64_t id_len;
    uint32_t width, height;
    uint64_t feature_count;
} ps_catalog_entry_info;
/* catalog.rs:158-177 CatalogSearchHit with its TrackingSummary (pipeline.rs:1137-1153);
 * the entry's id, width, height and feature count come from ps_catalog_entry. */
typedef struct {
    uint32_t entry;          /* catalog index */
    uint32_t coarse_votes;
    uint64_t raw_match_count, match_count;
    float inlier_ratio;
    float mean_dx, mean_dy, median_dx, median_dy;
    float mean_scale_ratio, mean_rotation_deg;
    uint32_t has_transform, has_homography;
    float transform[6];      /* SimilarityTransform2D a, b, tx, ty, scale, rotation_deg */
    float homography[9];     /* Homography2D, row-major */
} ps_catalog_hit;

/* catalog.rs:272-282 FeatureCatalog::new (descriptor_kind as the catalog field, :186-188) */
ps_status ps_catalog_create(uint32_t octaves, uint32_t scales, float contrast_threshold, uint8_t descriptor_kind,
                            ps_catalog **out);
/* catalog.rs:447-540 load_binary from the bytes of a PSIFTCAT file: signatures kept as
 * stored, u8 descriptors dequantised on the device (pipeline.rs:1034-1044). PS_E_FORMAT
 * with the reference's message for a malformed file. */
ps_status ps_catalog_load_binary(const uint8_t *bytes, size_t len, ps_catalog **out);
/* catalog.rs:447 load_binary(path) */
ps_status ps_catalog_load_file(const char *path, ps_catalog **out);
/* catalog.rs:400-445 save_binary: u8-quantised descriptors (pipeline.rs:1024-1030), written
 * to a temporary file then renamed (catalog.rs:204-244). */
ps_status ps_catalog_save_binary(const ps_catalog *c, const char *path);
ps_status ps_catalog_destroy(ps_catalog *c);
ps_status ps_catalog_get_info(const ps_catalog *c, ps_catalog_info *out);
/* catalog.rs:347-357 list_summaries, one entry */
ps_status ps_catalog_entry(const ps_catalog *c, size_t idx, ps_catalog_entry_info *out);
/* An entry's features as stored on the device (any output may be null; cap rows each). */
ps_status ps_catalog_entry_features(const ps_catalog *c, size_t idx, ps_feature *out_feat, float *out_desc,
                                    uint64_t *out_sig, size_t cap);
/* catalog.rs:309-328 upsert_entry: signatures quantised from desc on the device
 * (catalog.rs:37-45); id must be valid UTF-8. */
ps_status ps_catalog_upsert(ps_catalog *c, const char *id, size_t id_len, uint32_t width, uint32_t height,
                            const ps_feature *feat, const float *desc, size_t n);
/* catalog.rs:331-339 remove; *removed = 1 when the id existed */
ps_status ps_catalog_remove(ps_catalog *c, const char *id, size_t id_len, int32_t *removed);
/* catalog.rs:595-633: the coarse vote of every entry (out_votes holds n_entries) for the
 * query descriptors, by the reference's path choice: more than 32 entries -> a vote needs a
 * signature within 24 bits that also shares a byte position with the query (the inverted
 * index), else within 24 bits (count_coarse_hamming_votes). At most 65535 * 256 queries. */
ps_status ps_catalog_coarse_votes(const ps_catalog *c, const float *q_desc, size_t nq, uint32_t *out_votes);
/* catalog.rs:576-702 search_features: coarse votes, shortlist by (votes desc, id asc) cut to
 * max(shortlist_k, 1), then per shortlisted entry the exact ratio match (pipeline.rs:1845-1858),
 * summarize_tracking (:2210-2272), similarity and homography RANSAC at ransac_threshold; hits
 * ordered by (match_count desc, inlier_ratio desc, coarse_votes desc, id asc), cut to top_k.
 * *out_count = the hit count (PS_E_CAPACITY when above cap). At most 262144 queries. */
ps_status ps_catalog_search(const ps_catalog *c, const ps_feature *q, const float *q_desc, size_t nq,
                            size_t top_k, size_t shortlist_k, float ratio, float ransac_threshold,
                            ps_catalog_hit *out, size_t cap, size_t *out_count);
/* The same with device-resident query features (e.g. from ps_engine_extract_u8), run on the
 * engine's stream; returns after the hits are on the host. */
ps_status ps_engine_catalog_search(ps_engine *e, const ps_catalog *c, const ps_feature *d_q, const float *d_q_desc,
                                   size_t nq, size_t top_k, size_t shortlist_k, float ratio,
                                   float ransac_threshold, ps_catalog_hit *out, size_t cap, size_t *out_count);

/* Tests only: the next compute call on this thread fails with PS_E_INJECTED before
 * touching the device. */
void ps_debug_fail_next(void);

#ifdef __cplusplus
}
#endif
#endif
