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
                              ps_fundamental_result *res) {
    PS_TRY(ensure(e->aux[AUX_RS_7], ps::fundamental_work_bytes(std::max<size_t>(total, 1))));
    PS_HIP(ps::launch_fundamental(m, probs, np, thr, iters, conf, total, e->aux[AUX_RS_7].p, inl, res, e->stream));
    return PS_OK;
}

}  // namespace

extern "C" {

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
