// C ABI of the homography RANSAC: the reference entry point (Homography::estimate) and the
// batched form with a selectable sampler.
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

// RansacOptions::validate, pipeline.rs:203-212
ps_status check_threshold(float t) {
    if (!std::isfinite(t) || !(t >= 0.5f && t <= 100.0f)) {
        char buf[160];
        snprintf(buf, sizeof buf, "Invalid `reproj_threshold`=%g: must be between 0.5 and 100.0", (double)t);
        return set_err(PS_E_CONFIG, buf);
    }
    return PS_OK;
}

ps_status check_params(const ps_ransac_params *p) {
    if (!p) return set_err(PS_E_INVALID_ARG, "null params");
    PS_TRY(check_threshold(p->reproj_threshold));
    if (p->sampler > PS_SAMPLER_WEIGHTS) return set_err(PS_E_INVALID_ARG, "unknown sampler");
    if (!(p->confidence > 0.0 && p->confidence < 1.0)) return set_err(PS_E_INVALID_ARG, "confidence must be in (0, 1)");
    if (p->max_iters == 0 || p->min_iters > p->max_iters)
        return set_err(PS_E_INVALID_ARG, "need 1 <= max_iters and min_iters <= max_iters");
    return PS_OK;
}

ps_status enqueue_batch(ps_engine *e, const ps_feature_match *m, const float *w, const ps_ransac_problem *probs,
                        uint32_t np, size_t total, const ps_ransac_params &p, uint32_t *inl, ps_ransac_result *res) {
    PS_TRY(ensure(e->aux[AUX_RS_0], ps::ransac_work_bytes(std::max<size_t>(total, 1))));
    PS_HIP(ps::launch_ransac(m, w, probs, np, p, total, e->aux[AUX_RS_0].p, inl, res, e->stream));
    return PS_OK;
}

}  // namespace

extern "C" {

ps_status ps_homography_estimate(const ps_feature_match *m, size_t n, const ps_ransac_options *opt,
                                 uint32_t *out_inliers, size_t *out_n, float out_h[9], int32_t *out_has_model) {
    PS_TRY(check_fail_injection());
    if (!opt || !out_n || !out_has_model || !out_h || (n && (!m || !out_inliers)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    PS_TRY(check_threshold(opt->reproj_threshold));
    *out_n = 0;
    *out_has_model = 0;
    if (n == 0) return PS_OK;
    if (n >= 0x7fffffffull) return set_err(PS_E_INVALID_ARG, "too many matches");
    ps_ransac_params p;
    std::memset(&p, 0, sizeof p);
    p.reproj_threshold = opt->reproj_threshold;
    p.sampler = PS_SAMPLER_REFERENCE;
    p.local_opt = 1;
    p.sim_seed_model = 1;
    p.adaptive = 1;
    p.confidence = 0.999;   // pipeline.rs:2689
    p.min_iters = 64;
    p.max_iters = 2000;
    ps_ransac_problem prob;
    prob.offset = 0;
    prob.n = (uint32_t)n;
    prob.sim_seed = opt->has_seed ? opt->seed : kDefaultSimSeed;
    prob.hom_seed = opt->has_seed ? (opt->seed ^ kSeedMix) : kDefaultHomSeed;
    ps_ransac_result res;
    size_t nout = 0;
    ps_status s = ps_ransac_batch(m, nullptr, &prob, 1, n, &p, out_inliers, &res);
    if (s != PS_OK) return s;
    nout = res.n_inliers;
    *out_n = nout;
    *out_has_model = res.has_model ? 1 : 0;
    for (int i = 0; i < 9; i++) out_h[i] = res.has_model ? res.h[i] : 0.0f;
    return PS_OK;
}

ps_status ps_ransac_batch(const ps_feature_match *matches, const float *weights, const ps_ransac_problem *problems,
                          uint32_t n_problems, size_t total_matches, const ps_ransac_params *params,
                          uint32_t *inliers, ps_ransac_result *results) {
    TraceRange tr("ps.ransac_batch");
    PS_TRY(check_fail_injection());
    PS_TRY(check_params(params));
    if (n_problems == 0) return PS_OK;
    if (!problems || !results || (total_matches && (!matches || !inliers)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (params->sampler == PS_SAMPLER_WEIGHTS && total_matches && !weights)
        return set_err(PS_E_INVALID_ARG, "the weighted sampler needs weights");
    if (params->sampler == PS_SAMPLER_WEIGHTS)
        for (size_t i = 0; i < total_matches; i++)
            if (!(weights[i] >= 0.0f) || !std::isfinite(weights[i]))
                return set_err(PS_E_INVALID_ARG, "sampling weights must be finite and non-negative");
    if (total_matches >= 0x7fffffffull) return set_err(PS_E_INVALID_ARG, "too many matches");
    for (uint32_t i = 0; i < n_problems; i++)
        if ((size_t)problems[i].offset + problems[i].n > total_matches)
            return set_err(PS_E_INVALID_ARG, "problem range outside the match array");
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    const size_t tm = std::max<size_t>(total_matches, 1);
    PS_TRY(ensure(e->aux[AUX_RS_1], tm * sizeof(ps_feature_match)));
    PS_TRY(ensure(e->aux[AUX_RS_2], tm * sizeof(float)));
    PS_TRY(ensure(e->aux[AUX_RS_3], n_problems * sizeof(ps_ransac_problem)));
    PS_TRY(ensure(e->aux[AUX_RS_4], tm * sizeof(uint32_t)));
    PS_TRY(ensure(e->aux[AUX_RS_5], n_problems * sizeof(ps_ransac_result)));
    if (total_matches)
        PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_1].p, matches, total_matches * sizeof(ps_feature_match),
                              hipMemcpyHostToDevice, e->stream));
    if (weights && total_matches)
        PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_2].p, weights, total_matches * sizeof(float), hipMemcpyHostToDevice,
                              e->stream));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_3].p, problems, n_problems * sizeof(ps_ransac_problem),
                          hipMemcpyHostToDevice, e->stream));
    PS_TRY(enqueue_batch(e, as<ps_feature_match>(e->aux[AUX_RS_1]), as<float>(e->aux[AUX_RS_2]),
                         as<ps_ransac_problem>(e->aux[AUX_RS_3]), n_problems, total_matches, *params,
                         as<uint32_t>(e->aux[AUX_RS_4]), as<ps_ransac_result>(e->aux[AUX_RS_5])));
    PS_HIP(hipMemcpyAsync(results, e->aux[AUX_RS_5].p, n_problems * sizeof(ps_ransac_result), hipMemcpyDeviceToHost,
                          e->stream));
    if (total_matches)
        PS_HIP(hipMemcpyAsync(inliers, e->aux[AUX_RS_4].p, total_matches * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

ps_status ps_engine_ransac_batch(ps_engine *e, const ps_feature_match *d_matches, const float *d_weights,
                                 const ps_ransac_problem *d_problems, uint32_t n_problems, size_t total_matches,
                                 const ps_ransac_params *params, uint32_t *d_inliers, ps_ransac_result *d_results) {
    PS_TRY(check_fail_injection());
    if (!e) return set_err(PS_E_INVALID_ARG, "null engine");
    PS_TRY(check_params(params));
    if (n_problems == 0) return PS_OK;
    if (!d_problems || !d_results || (total_matches && (!d_matches || !d_inliers)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (params->sampler == PS_SAMPLER_WEIGHTS && total_matches && !d_weights)
        return set_err(PS_E_INVALID_ARG, "the weighted sampler needs weights");
    if (total_matches >= 0x7fffffffull) return set_err(PS_E_INVALID_ARG, "too many matches");
    PS_HIP(hipSetDevice(e->device));
    return enqueue_batch(e, d_matches, d_weights, d_problems, n_problems, total_matches, *params, d_inliers,
                         d_results);
}

}  // extern "C"

// ------------------------------------------------------------------ match-scoring MLP
extern "C" {

ps_status ps_engine_mlp_features(ps_engine *e, const ps_feature_match *d_matches, const ps_feature *d_qfeat,
                                 const ps_feature *d_tfeat, const ps_mlp_pair *d_pairs, uint32_t n_pairs,
                                 float *d_x) {
    PS_TRY(check_fail_injection());
    if (!e || (n_pairs && (!d_matches || !d_qfeat || !d_tfeat || !d_pairs || !d_x)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    PS_HIP(hipSetDevice(e->device));
    PS_HIP(ps::launch_mlp_features(d_matches, d_qfeat, d_tfeat, d_pairs, n_pairs, d_x, e->stream));
    return PS_OK;
}

ps_status ps_engine_mlp_forward(ps_engine *e, const float *d_params, const float *d_x, size_t n, float *d_logit,
                                float *d_prob) {
    PS_TRY(check_fail_injection());
    if (!e || !d_params || (n && (!d_x || !d_logit))) return set_err(PS_E_INVALID_ARG, "null argument");
    if (n >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "too many matches");
    PS_HIP(hipSetDevice(e->device));
    PS_HIP(ps::launch_mlp_forward(d_params, d_x, (uint32_t)n, d_logit, d_prob, e->stream));
    return PS_OK;
}

ps_status ps_engine_mlp_grad(ps_engine *e, const float *d_params, const float *d_x, const float *d_y, size_t n,
                             float *d_grad, float *d_loss) {
    PS_TRY(check_fail_injection());
    if (!e || !d_params || !d_grad || (n && (!d_x || !d_y))) return set_err(PS_E_INVALID_ARG, "null argument");
    if (n >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "too many matches");
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->aux[AUX_RS_6], ps::mlp_grad_scratch_bytes()));
    PS_HIP(ps::launch_mlp_grad(d_params, d_x, d_y, (uint32_t)n, e->aux[AUX_RS_6].p, d_grad, d_loss, e->stream));
    return PS_OK;
}

ps_status ps_engine_mlp_adam(ps_engine *e, float *d_params, float *d_m, float *d_v, const float *d_grad,
                             uint32_t step, float lr) {
    PS_TRY(check_fail_injection());
    if (!e || !d_params || !d_m || !d_v || !d_grad) return set_err(PS_E_INVALID_ARG, "null argument");
    if (step == 0) return set_err(PS_E_INVALID_ARG, "Adam steps count from 1");
    PS_HIP(hipSetDevice(e->device));
    PS_HIP(ps::launch_mlp_adam(d_params, d_m, d_v, d_grad, step, lr, e->stream));
    return PS_OK;
}

ps_status ps_mlp_forward(const float *params, const float *x, size_t n, float *logit) {
    PS_TRY(check_fail_injection());
    if (!params || (n && (!x || !logit))) return set_err(PS_E_INVALID_ARG, "null argument");
    if (n == 0) return PS_OK;
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->aux[AUX_RS_7], PS_MLP_PARAMS * sizeof(float)));
    PS_TRY(ensure(e->aux[AUX_PM_0], n * 8 * sizeof(float)));
    PS_TRY(ensure(e->aux[AUX_PM_1], n * sizeof(float)));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_7].p, params, PS_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_PM_0].p, x, n * 8 * sizeof(float), hipMemcpyHostToDevice, e->stream));
    PS_TRY(ps_engine_mlp_forward(e, as<float>(e->aux[AUX_RS_7]), as<float>(e->aux[AUX_PM_0]), n,
                                 as<float>(e->aux[AUX_PM_1]), nullptr));
    PS_HIP(hipMemcpyAsync(logit, e->aux[AUX_PM_1].p, n * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

ps_status ps_mlp_loss_grad(const float *params, const float *x, const float *y, size_t n, float *grad, float *loss) {
    PS_TRY(check_fail_injection());
    if (!params || !grad || (n && (!x || !y))) return set_err(PS_E_INVALID_ARG, "null argument");
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    const size_t nn = n ? n : 1;
    PS_TRY(ensure(e->aux[AUX_RS_7], PS_MLP_PARAMS * sizeof(float)));
    PS_TRY(ensure(e->aux[AUX_PM_0], nn * 8 * sizeof(float)));
    PS_TRY(ensure(e->aux[AUX_PM_1], nn * sizeof(float)));
    PS_TRY(ensure(e->aux[AUX_PM_2], (PS_MLP_PARAMS + 1) * sizeof(float)));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RS_7].p, params, PS_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, e->stream));
    if (n) {
        PS_HIP(hipMemcpyAsync(e->aux[AUX_PM_0].p, x, n * 8 * sizeof(float), hipMemcpyHostToDevice, e->stream));
        PS_HIP(hipMemcpyAsync(e->aux[AUX_PM_1].p, y, n * sizeof(float), hipMemcpyHostToDevice, e->stream));
    }
    float *g = as<float>(e->aux[AUX_PM_2]);
    PS_TRY(ps_engine_mlp_grad(e, as<float>(e->aux[AUX_RS_7]), as<float>(e->aux[AUX_PM_0]), as<float>(e->aux[AUX_PM_1]),
                              n, g, g + PS_MLP_PARAMS));
    PS_HIP(hipMemcpyAsync(grad, g, PS_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    float l = 0.0f;
    PS_HIP(hipMemcpyAsync(&l, g + PS_MLP_PARAMS, sizeof(float), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    if (loss) *loss = l;
    return PS_OK;
}

}  // extern "C"
