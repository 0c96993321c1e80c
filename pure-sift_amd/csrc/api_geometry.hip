// C ABI of the two-view geometry the per-pair match post-processing runs: similarity RANSAC
// (Homography::estimate_similarity), fundamental-matrix LO-RANSAC and GRIC model selection.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "host_common.h"

using namespace pshost;

namespace {

constexpr uint64_t kDefaultSimSeed = 0xD1B54A32D192ED03ull;   // pipeline.rs:173
constexpr uint64_t kDefaultHomSeed = 0x517CC1B727220A95ull;   // pipeline.rs:175
constexpr uint64_t kSeedMix = 0x9E3779B97F4A7C15ull;          // pipeline.rs:2650-2661
constexpr uint64_t kFundSeed = 0xF00DCAFE12345678ull;         // geometry.rs:125

ps_status check_reproj(float t) {   // RansacOptions::validate, pipeline.rs:203-212
    if (!std::isfinite(t) || !(t >= 0.5f && t <= 100.0f)) {
        char buf[160];
        snprintf(buf, sizeof buf, "Invalid `reproj_threshold`=%g: must be between 0.5 and 100.0", (double)t);
        return set_err(PS_E_CONFIG, buf);
    }
    return PS_OK;
}

ps_status check_ranges(const ps_fundamental_problem *p, uint32_t np, size_t total) {
    for (uint32_t i = 0; i < np; i++)
        if ((size_t)p[i].offset + p[i].n > total) return set_err(PS_E_INVALID_ARG, "problem range outside the match array");
    return PS_OK;
}

// The device loop runs the reference's iteration bound as given; without a good model it runs
// every iteration, so the bound is capped to keep one launch short.
constexpr uint64_t kMaxFundIters = 1000000;

ps_status check_iters(uint64_t v) {
    if (v > kMaxFundIters) return set_err(PS_E_INVALID_ARG, "max_iters above 1000000");
    return PS_OK;
}

ps_status enqueue_fundamental(ps_engine *e, const ps_feature_match *m, const ps_fundamental_problem *probs,
                              uint32_t np, size_t total, double thr, uint32_t iters, double conf, uint32_t *inl,
                              ps_fundamental_result *res, hipStream_t st = nullptr) {
    PS_TRY(ensure(e->aux[AUX_RS_7], ps::fundamental_work_bytes(std::max<size_t>(total, 1))));
    PS_HIP(ps::launch_fundamental(m, probs, np, thr, iters, conf, total, e->aux[AUX_RS_7].p, inl, res,
                                  st ? st : e->stream));
    return PS_OK;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// cli.rs:290-318 for n_pairs raw match sets, all on the engine stream (no host sync)
ps_status enqueue_pair_geometry(ps_engine *e, const ps_feature_match *dm, const ps_feature *dq, const ps_feature *dt,
                                const ps_pair_problem *dp, uint32_t np, size_t total, const ps_ransac_options &opt,
                                uint32_t *dbest, ps_pair_geometry *dout) {
    hipStream_t st = e->stream;
    const size_t tm = std::max<size_t>(total, 1);
    PS_TRY(ensure(e->aux[AUX_PG_0], np * sizeof(ps_ransac_problem)));
    PS_TRY(ensure(e->aux[AUX_PG_1], np * sizeof(ps_ransac_problem)));
    PS_TRY(ensure(e->aux[AUX_PG_2], np * sizeof(ps_fundamental_problem)));
    PS_TRY(ensure(e->aux[AUX_PG_3], np * sizeof(ps::TrackJob)));
    PS_TRY(ensure(e->aux[AUX_PG_4], np * sizeof(ps_ransac_result)));
    PS_TRY(ensure(e->aux[AUX_PG_5], np * sizeof(ps_ransac_result)));
    PS_TRY(ensure(e->aux[AUX_PG_6], ps::ransac_work_bytes(tm)));
    PS_TRY(ensure(e->aux[AUX_PG_8], tm * sizeof(uint32_t)));
    PS_TRY(ensure(e->aux[AUX_PG_9], np * sizeof(ps_fundamental_result)));
    PS_TRY(ensure(e->aux[AUX_PG_10], tm * sizeof(uint32_t)));
    const size_t o_gr = align256(np * sizeof(ps_gric_problem)), o_tr = o_gr + align256(np * sizeof(ps_gric_result));
    PS_TRY(ensure(e->aux[AUX_PG_11], o_tr + np * sizeof(ps::TrackOut)));
    char *small = static_cast<char *>(e->aux[AUX_PG_11].p);
    ps_gric_problem *gp = reinterpret_cast<ps_gric_problem *>(small);
    ps_gric_result *gr = reinterpret_cast<ps_gric_result *>(small + o_gr);
    ps::TrackOut *tr = reinterpret_cast<ps::TrackOut *>(small + o_tr);
    ps_ransac_problem *rp = as<ps_ransac_problem>(e->aux[AUX_PG_0]), *rp3 = as<ps_ransac_problem>(e->aux[AUX_PG_1]);
    ps_fundamental_problem *fp = as<ps_fundamental_problem>(e->aux[AUX_PG_2]);
    ps::TrackJob *jobs = as<ps::TrackJob>(e->aux[AUX_PG_3]);
    ps_ransac_result *rr = as<ps_ransac_result>(e->aux[AUX_PG_4]), *rr3 = rr;
    ps_fundamental_result *fr = as<ps_fundamental_result>(e->aux[AUX_PG_9]);
    const uint64_t sim_seed = opt.has_seed ? opt.seed : kDefaultSimSeed;
    const uint64_t hom_seed = opt.has_seed ? (opt.seed ^ kSeedMix) : kDefaultHomSeed;
    PS_HIP(ps::launch_pair_setup(dp, np, sim_seed, hom_seed, kFundSeed, rp, rp3, fp, jobs, st));
    // the fundamental RANSAC is independent of the similarity / homography chain: it runs on a
    // side stream beside it (not under stream capture, where the engine keeps one stream)
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    PS_HIP(hipStreamIsCapturing(st, &cap));
    const bool fork = cap == hipStreamCaptureStatusNone;
    hipStream_t fs = fork ? e->side[0] : st;
    if (fork) {
        PS_HIP(hipEventRecord(e->fork_ev[0], st));
        PS_HIP(hipStreamWaitEvent(fs, e->fork_ev[0], 0));
    }
    const double sampson = std::min<double>((double)opt.reproj_threshold, 3.0);   // cli.rs:302
    PS_TRY(enqueue_fundamental(e, dm, fp, np, total, sampson, 2000, 0.999, as<uint32_t>(e->aux[AUX_PG_10]), fr, fs));
    if (fork) PS_HIP(hipEventRecord(e->join_ev[0], fs));
    ps_ransac_params p;
    std::memset(&p, 0, sizeof p);
    p.reproj_threshold = opt.reproj_threshold;
    p.sampler = PS_SAMPLER_REFERENCE;
    p.local_opt = 1;
    p.sim_seed_model = 1;
    p.adaptive = 1;
    p.confidence = 0.999;   // pipeline.rs:2689
    p.min_iters = 64;
    p.max_iters = 2000;
    // similarity + homography at the options (their similarity stages coincide: same seed)
    PS_HIP(ps::launch_ransac(dm, nullptr, rp, np, p, total, e->aux[AUX_PG_6].p, as<uint32_t>(e->aux[AUX_PG_8]), rr,
                             st));
    const void *work3 = e->aux[AUX_PG_6].p;
    if (!(opt.reproj_threshold == 3.0f && !opt.has_seed)) {
        // summarize_tracking's own similarity RANSAC: 3 px, default seed (pipeline.rs:2227)
        PS_TRY(ensure(e->aux[AUX_PG_7], ps::ransac_work_bytes(tm)));
        ps_ransac_params p3 = p;
        p3.reproj_threshold = 3.0f;
        rr3 = as<ps_ransac_result>(e->aux[AUX_PG_5]);
        PS_TRY(ensure(e->aux[AUX_CS_11], tm * sizeof(uint32_t)));   // inlier list of this run (unused)
        PS_HIP(ps::launch_ransac(dm, nullptr, rp3, np, p3, total, e->aux[AUX_PG_7].p, as<uint32_t>(e->aux[AUX_CS_11]),
                                 rr3, st, true));
        work3 = e->aux[AUX_PG_7].p;
    }
    PS_HIP(ps::launch_tracking(dm, jobs, np, rr3, work3, total, dq, dt, tr, st));
    if (fork) PS_HIP(hipStreamWaitEvent(st, e->join_ev[0], 0));
    PS_HIP(ps::launch_pair_gric_problems(dp, np, rr, fr, gp, st));
    PS_HIP(ps::launch_gric(dm, gp, np, std::max<double>((double)opt.reproj_threshold, 1.0), gr, st));   // cli.rs:309
    PS_HIP(ps::launch_pair_finish(dp, np, rr, rr3, e->aux[AUX_PG_6].p, total, as<uint32_t>(e->aux[AUX_PG_8]), fr, gr,
                                  tr, dbest, dout, st));
    return PS_OK;
}

}  // namespace

extern "C" {

ps_status ps_engine_pair_geometry_batch(ps_engine *e, const ps_feature_match *d_matches, const ps_feature *d_qfeat,
                                        const ps_feature *d_tfeat, const ps_pair_problem *d_pairs, uint32_t n_pairs,
                                        size_t total_matches, const ps_ransac_options *opt, uint32_t *d_best_inliers,
                                        ps_pair_geometry *d_out) {
    PS_TRY(check_fail_injection());
    if (!e || !opt) return set_err(PS_E_INVALID_ARG, "null argument");
    PS_TRY(check_reproj(opt->reproj_threshold));
    if (n_pairs == 0) return PS_OK;
    if (!d_pairs || !d_out || (total_matches && (!d_matches || !d_qfeat || !d_tfeat || !d_best_inliers)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (total_matches >= 0x7fffffffull) return set_err(PS_E_INVALID_ARG, "too many matches");
    PS_HIP(hipSetDevice(e->device));
    return enqueue_pair_geometry(e, d_matches, d_qfeat, d_tfeat, d_pairs, n_pairs, total_matches, *opt,
                                 d_best_inliers, d_out);
}

ps_status ps_pair_geometry_estimate(const ps_feature *q, size_t nq, const ps_feature *t, size_t nt,
                                    const ps_feature_match *m, size_t n, const ps_ransac_options *opt,
                                    uint32_t *out_best_inliers, ps_pair_geometry *out) {
    TraceRange tr("ps.pair_geometry");
    PS_TRY(check_fail_injection());
    if (!opt || !out || (n && (!m || !q || !t || !out_best_inliers))) return set_err(PS_E_INVALID_ARG, "null argument");
    PS_TRY(check_reproj(opt->reproj_threshold));
    if (n >= 0x7fffffffull || nq >= 0x7fffffffull || nt >= 0x7fffffffull)
        return set_err(PS_E_INVALID_ARG, "too many matches or features");
    for (size_t i = 0; i < n; i++)
        if (m[i].query_index >= nq || m[i].target_index >= nt)
            return set_err(PS_E_INVALID_ARG, "a match indexes past the feature arrays");
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    hipStream_t st = e->stream;
    PS_TRY(ensure(e->aux[AUX_RS_1], std::max<size_t>(n, 1) * sizeof(ps_feature_match)));
    PS_TRY(ensure(e->aux[AUX_RS_2], std::max<size_t>(nq, 1) * sizeof(ps_feature)));
    PS_TRY(ensure(e->aux[AUX_RS_3], std::max<size_t>(nt, 1) * sizeof(ps_feature)));
    PS_TRY(ensure(e->aux[AUX_RS_4], sizeof(ps_pair_problem)));
    PS_TRY(ensure(e->aux[AUX_RS_5], std::max<size_t>(n, 1) * sizeof(uint32_t)));
    PS_TRY(ensure(e->aux[AUX_RS_6], sizeof(ps_pair_geometry)));
    const ps_pair_problem pp{0, (uint32_t)n, 0, 0};
    if (n) PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_1].p, m, n * sizeof(ps_feature_match), hipMemcpyHostToDevice, st));
    if (nq) PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_2].p, q, nq * sizeof(ps_feature), hipMemcpyHostToDevice, st));
    if (nt) PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_3].p, t, nt * sizeof(ps_feature), hipMemcpyHostToDevice, st));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_4].p, &pp, sizeof pp, hipMemcpyHostToDevice, st));
    PS_TRY(enqueue_pair_geometry(e, as<ps_feature_match>(e->aux[AUX_RS_1]), as<ps_feature>(e->aux[AUX_RS_2]),
                                 as<ps_feature>(e->aux[AUX_RS_3]), as<ps_pair_problem>(e->aux[AUX_RS_4]), 1, n, *opt,
                                 as<uint32_t>(e->aux[AUX_RS_5]), as<ps_pair_geometry>(e->aux[AUX_RS_6])));
    PS_HIP(hipMemcpyAsync(out, e->aux[AUX_RS_6].p, sizeof *out, hipMemcpyDeviceToHost, st));
    PS_HIP(hipStreamSynchronize(st));
    if (out->n_best)
        PS_HIP(hipMemcpyAsync(out_best_inliers, e->aux[AUX_RS_5].p, out->n_best * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, st));
    PS_HIP(hipStreamSynchronize(st));
    return PS_OK;
}


ps_status ps_similarity_estimate(const ps_feature_match *m, size_t n, const ps_ransac_options *opt,
                                 uint32_t *out_inliers, size_t *out_n, float out_model[6], int32_t *out_has_model) {
    PS_TRY(check_fail_injection());
    if (!opt || !out_n || !out_model || !out_has_model || (n && (!m || !out_inliers)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    PS_TRY(check_reproj(opt->reproj_threshold));
    *out_n = 0;
    *out_has_model = 0;
    for (int i = 0; i < 6; i++) out_model[i] = 0.0f;
    if (n == 0) return PS_OK;   // pipeline.rs:2060-2062
    if (n >= 0x7fffffffull) return set_err(PS_E_INVALID_ARG, "too many matches");
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->aux[AUX_RS_0], ps::ransac_work_bytes(n)));
    PS_TRY(ensure(e->aux[AUX_RS_1], n * sizeof(ps_feature_match)));
    PS_TRY(ensure(e->aux[AUX_RS_3], sizeof(ps_ransac_problem)));
    PS_TRY(ensure(e->aux[AUX_RS_4], n * sizeof(uint32_t)));
    PS_TRY(ensure(e->aux[AUX_RS_5], sizeof(ps_ransac_result)));
    ps_ransac_problem prob{0, (uint32_t)n, opt->has_seed ? opt->seed : kDefaultSimSeed, 0};
    ps_ransac_params p;
    std::memset(&p, 0, sizeof p);
    p.reproj_threshold = opt->reproj_threshold;
    p.sampler = PS_SAMPLER_REFERENCE;
    p.sim_seed_model = 1;
    p.confidence = 0.999;
    p.min_iters = 64;
    p.max_iters = 2000;
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_1].p, m, n * sizeof(ps_feature_match), hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_3].p, &prob, sizeof prob, hipMemcpyHostToDevice, e->stream));
    PS_HIP(ps::launch_ransac(as<ps_feature_match>(e->aux[AUX_RS_1]), nullptr, as<ps_ransac_problem>(e->aux[AUX_RS_3]),
                             1, p, n, e->aux[AUX_RS_0].p, as<uint32_t>(e->aux[AUX_RS_4]),
                             as<ps_ransac_result>(e->aux[AUX_RS_5]), e->stream, true));
    ps_ransac_result res;
    PS_HIP(hipMemcpyAsync(&res, e->aux[AUX_RS_5].p, sizeof res, hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipMemcpyAsync(out_inliers, e->aux[AUX_RS_4].p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    *out_n = res.n_sim;
    *out_has_model = res.has_sim ? 1 : 0;
    if (res.has_sim) {   // pipeline.rs:1962-1963 (f32 libm, as the reference's f32 calls)
        const float a = res.sim[0], bb = res.sim[1];
        const float rot = ::atan2f(bb, a);
        out_model[0] = a;
        out_model[1] = bb;
        out_model[2] = res.sim[2];
        out_model[3] = res.sim[3];
        out_model[4] = ::sqrtf(a * a + bb * bb);
        out_model[5] = rot * 57.2957795130823208767981548141051703f;
    }
    return PS_OK;
}

ps_status ps_fundamental_batch(const ps_feature_match *matches, const ps_fundamental_problem *problems,
                               uint32_t n_problems, size_t total_matches, double sampson_threshold, uint32_t max_iters,
                               double confidence, uint32_t *inliers, ps_fundamental_result *results) {
    PS_TRY(check_fail_injection());
    if (n_problems == 0) return PS_OK;
    if (!problems || !results || (total_matches && (!matches || !inliers)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (total_matches >= 0x7fffffffull) return set_err(PS_E_INVALID_ARG, "too many matches");
    PS_TRY(check_iters(max_iters));
    PS_TRY(check_ranges(problems, n_problems, total_matches));
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    const size_t tm = std::max<size_t>(total_matches, 1);
    PS_TRY(ensure(e->aux[AUX_RS_1], tm * sizeof(ps_feature_match)));
    PS_TRY(ensure(e->aux[AUX_RS_3], n_problems * sizeof(ps_fundamental_problem)));
    PS_TRY(ensure(e->aux[AUX_RS_4], tm * sizeof(uint32_t)));
    PS_TRY(ensure(e->aux[AUX_RS_5], n_problems * sizeof(ps_fundamental_result)));
    if (total_matches)
        PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_1].p, matches, total_matches * sizeof(ps_feature_match),
                              hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_3].p, problems, n_problems * sizeof(ps_fundamental_problem),
                          hipMemcpyHostToDevice, e->stream));
    PS_TRY(enqueue_fundamental(e, as<ps_feature_match>(e->aux[AUX_RS_1]), as<ps_fundamental_problem>(e->aux[AUX_RS_3]),
                               n_problems, total_matches, sampson_threshold, max_iters, confidence,
                               as<uint32_t>(e->aux[AUX_RS_4]), as<ps_fundamental_result>(e->aux[AUX_RS_5])));
    PS_HIP(hipMemcpyAsync(results, e->aux[AUX_RS_5].p, n_problems * sizeof(ps_fundamental_result),
                          hipMemcpyDeviceToHost, e->stream));
    if (total_matches)
        PS_HIP(hipMemcpyAsync(inliers, e->aux[AUX_RS_4].p, total_matches * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

ps_status ps_engine_fundamental_batch(ps_engine *e, const ps_feature_match *d_matches,
                                      const ps_fundamental_problem *d_problems, uint32_t n_problems,
                                      size_t total_matches, double sampson_threshold, uint32_t max_iters,
                                      double confidence, uint32_t *d_inliers, ps_fundamental_result *d_results) {
    PS_TRY(check_fail_injection());
    if (!e) return set_err(PS_E_INVALID_ARG, "null engine");
    if (n_problems == 0) return PS_OK;
    if (!d_problems || !d_results || (total_matches && (!d_matches || !d_inliers)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (total_matches >= 0x7fffffffull) return set_err(PS_E_INVALID_ARG, "too many matches");
    PS_TRY(check_iters(max_iters));
    PS_HIP(hipSetDevice(e->device));
    return enqueue_fundamental(e, d_matches, d_problems, n_problems, total_matches, sampson_threshold, max_iters,
                               confidence, d_inliers, d_results);
}

ps_status ps_fundamental_estimate(const ps_feature_match *m, size_t n, const ps_fundamental_options *opt,
                                  uint32_t *out_inliers, size_t *out_n, double out_f[9], int32_t *out_has_model) {
    PS_TRY(check_fail_injection());
    if (!opt || !out_n || !out_f || !out_has_model || (n && (!m || !out_inliers)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    *out_n = 0;
    *out_has_model = 0;
    for (int i = 0; i < 9; i++) out_f[i] = 0.0;
    PS_TRY(check_iters(opt->max_iters));
    if (n < 7) return PS_OK;   // geometry.rs:872-874
    ps_fundamental_problem prob{0, (uint32_t)std::min<size_t>(n, 0xffffffffu), opt->seed};
    ps_fundamental_result res;
    PS_TRY(ps_fundamental_batch(m, &prob, 1, n, opt->sampson_threshold, (uint32_t)opt->max_iters, opt->confidence,
                                out_inliers, &res));
    *out_n = res.n_inliers;
    *out_has_model = res.has_model ? 1 : 0;
    for (int i = 0; i < 9; i++) out_f[i] = res.f[i];
    return PS_OK;
}

ps_status ps_engine_gric_batch(ps_engine *e, const ps_feature_match *d_matches, const ps_gric_problem *d_problems,
                               uint32_t n_problems, double sigma, ps_gric_result *d_results) {
    PS_TRY(check_fail_injection());
    if (!e) return set_err(PS_E_INVALID_ARG, "null engine");
    if (n_problems == 0) return PS_OK;
    if (!d_problems || !d_results || !d_matches) return set_err(PS_E_INVALID_ARG, "null argument");
    PS_HIP(hipSetDevice(e->device));
    PS_HIP(ps::launch_gric(d_matches, d_problems, n_problems, sigma, d_results, e->stream));
    return PS_OK;
}

ps_status ps_select_model_gric(const ps_feature_match *m, size_t n, const float *h, const double *f, double sigma,
                               ps_gric_result *out) {
    PS_TRY(check_fail_injection());
    if (!out || (n && !m)) return set_err(PS_E_INVALID_ARG, "null argument");
    std::memset(out, 0, sizeof *out);
    if (n == 0) return PS_OK;   // geometry.rs:1239-1241
    if (!h && !f) return PS_OK;  // both scores infinite: None
    if (n >= 0x7fffffffull) return set_err(PS_E_INVALID_ARG, "too many matches");
    ps_gric_problem prob;
    std::memset(&prob, 0, sizeof prob);
    prob.n = (uint32_t)n;
    prob.has_h = h ? 1 : 0;
    prob.has_f = f ? 1 : 0;
    for (int i = 0; i < 9; i++) {
        prob.h[i] = h ? h[i] : 0.0f;
        prob.f[i] = f ? f[i] : 0.0;
    }
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->aux[AUX_RS_1], n * sizeof(ps_feature_match)));
    PS_TRY(ensure(e->aux[AUX_RS_6], sizeof(ps_gric_problem) + 256));
    PS_TRY(ensure(e->aux[AUX_RS_5], sizeof(ps_gric_result)));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_1].p, m, n * sizeof(ps_feature_match), hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_6].p, &prob, sizeof prob, hipMemcpyHostToDevice, e->stream));
    PS_HIP(ps::launch_gric(as<ps_feature_match>(e->aux[AUX_RS_1]), as<ps_gric_problem>(e->aux[AUX_RS_6]), 1, sigma,
                           as<ps_gric_result>(e->aux[AUX_RS_5]), e->stream));
    PS_HIP(hipMemcpyAsync(out, e->aux[AUX_RS_5].p, sizeof *out, hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

}  // extern "C"
