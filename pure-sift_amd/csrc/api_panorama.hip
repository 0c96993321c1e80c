// C ABI of the grouped matcher (many set-against-set ratio matches in one launch chain) and of
// the panorama's pairwise stage (stitch.rs:478-571): batched extraction of every image, one
// grouped match over all image pairs, one batched RANSAC launch, one pair-edge launch.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host_common.h"

using namespace pshost;

namespace pshost {

// grouped exact ratio match on the engine stream (scratch: the caller's workspace slot)
ps_status enqueue_match_sets(ps_engine *e, const ps::MatchSetsArgs &a, DevBuf &scratch, ps_feature_match *out,
                             uint32_t *counts) {
    TraceRange tr("ps.match_sets");
    bool fresh = false;
    PS_TRY(ensure(scratch, ps::match_sets_scratch_bytes(a), &fresh));
    if (++e->nn_gen == 0) e->nn_gen = 1;
    // fresh memory: clear the two norm-bound words at its head (later calls clear them in turn)
    if (fresh) PS_HIP(ps::launch_fill_u32(static_cast<uint32_t *>(scratch.p), 0u, 4, e->stream));
    PS_HIP(ps::launch_match_sets(a, scratch.p, e->nn_gen, out, counts, e->stream));
    return PS_OK;
}

}  // namespace pshost

namespace {

constexpr uint64_t kDefaultSimSeed = 0xD1B54A32D192ED03ull;   // pipeline.rs:173
constexpr uint64_t kDefaultHomSeed = 0x517CC1B727220A95ull;   // pipeline.rs:175
constexpr uint32_t kPanoCap = PS_MAX_FEATURES;                // extract_sift_features' cap
constexpr uint32_t kMaxProblems = 65535;                      // one grid dimension

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

ps_status check_pano_options(const ps_panorama_options *o) {
    ps_sift_config cfg{};
    cfg.octaves = o->octaves;
    cfg.scales = o->scales;
    cfg.contrast_threshold = o->threshold;
    cfg.upsample_first_octave = 1;
    cfg.descriptor_kind = PS_DESCRIPTOR_SIFT;
    PS_TRY(validate_cfg(&cfg, nullptr, 0));
    if (!std::isfinite(o->ratio) || !(o->ratio > 0.0f))
        return set_err(PS_E_CONFIG, "panorama ratio must be a positive finite number");
    if (!std::isfinite(o->ransac_threshold) || !(o->ransac_threshold > 0.0f))
        return set_err(PS_E_CONFIG, "panorama ransac_threshold must be a positive finite number");
    return PS_OK;
}

// stitch.rs:493-571 on the engine's stream: features of every image into slabs of kPanoCap
// rows, the ratio matches of every pair i < j into slabs of kPanoCap matches, similarity +
// homography RANSAC for pairs with 6 or more matches, then the pair records.
ps_status enqueue_panorama(ps_engine *e, const uint8_t *d_pixels, const ps_image_desc *images, uint32_t n,
                           const ps_panorama_options &o, ps_panorama_pair *d_pairs) {
    hipStream_t st = e->stream;
    const uint32_t np = n * (n - 1) / 2;
    const size_t rows = (size_t)n * kPanoCap, total = (size_t)np * kPanoCap;
    ps_sift_config cfg{};
    cfg.octaves = o.octaves;
    cfg.scales = o.scales;
    cfg.contrast_threshold = o.threshold;
    cfg.upsample_first_octave = 1;   // DEFAULT_UPSAMPLE_FIRST_OCTAVE, pipeline.rs:1693
    cfg.descriptor_kind = PS_DESCRIPTOR_SIFT;   // extract_sift_features_with_options, pipeline.rs:1674
    PS_TRY(ensure(e->aux[AUX_PN_1], rows * sizeof(ps_feature)));
    PS_TRY(ensure(e->aux[AUX_PN_2], rows * 128 * sizeof(float)));
    // small tables: counts, sets, image descriptors, problems, match counts, RANSAC problems / results
    const size_t o_sets = al256(n * sizeof(uint32_t)), o_img = o_sets + al256(n * sizeof(ps_feature_set));
    const size_t o_prob = o_img + al256(n * sizeof(ps_image_desc));
    const size_t o_mc = o_prob + al256(np * sizeof(ps_match_problem));
    const size_t o_rp = o_mc + al256(np * sizeof(uint32_t)), o_rr = o_rp + al256(np * sizeof(ps_ransac_problem));
    PS_TRY(ensure(e->aux[AUX_PN_3], o_rr + np * sizeof(ps_ransac_result)));
    PS_TRY(ensure(e->aux[AUX_PN_5], total * sizeof(ps_feature_match)));
    PS_TRY(ensure(e->aux[AUX_PN_6], ps::ransac_work_bytes(total)));
    PS_TRY(ensure(e->aux[AUX_PN_7], total * sizeof(uint32_t)));
    char *small = static_cast<char *>(e->aux[AUX_PN_3].p);
    uint32_t *counts = reinterpret_cast<uint32_t *>(small);
    ps_feature_set *sets = reinterpret_cast<ps_feature_set *>(small + o_sets);
    ps_image_desc *d_img = reinterpret_cast<ps_image_desc *>(small + o_img);
    ps_match_problem *probs = reinterpret_cast<ps_match_problem *>(small + o_prob);
    uint32_t *mcount = reinterpret_cast<uint32_t *>(small + o_mc);
    ps_ransac_problem *rp = reinterpret_cast<ps_ransac_problem *>(small + o_rp);
    ps_ransac_result *rr = reinterpret_cast<ps_ransac_result *>(small + o_rr);
    ps_feature *feat = as<ps_feature>(e->aux[AUX_PN_1]);
    float *desc = as<float>(e->aux[AUX_PN_2]);
    ps_feature_match *matches = as<ps_feature_match>(e->aux[AUX_PN_5]);

    // host tables: the image descriptors and the pair list in the reference's order
    // (stitch.rs:500-505); uploaded again only when they differ from the last call's (the
    // copies read pageable host memory, so an upload waits for them)
    const bool same = e->pano_tables == small && e->pano_images.size() == n &&
                      std::memcmp(e->pano_images.data(), images, n * sizeof(ps_image_desc)) == 0;
    if (!same) {
        std::vector<ps_match_problem> hp(np);
        for (uint32_t i = 0, k = 0; i < n; i++)
            for (uint32_t j = i + 1; j < n; j++, k++) hp[k] = ps_match_problem{i, j, k * kPanoCap, 0u};
        PS_HIP(hipMemcpyAsync(d_img, images, n * sizeof(ps_image_desc), hipMemcpyHostToDevice, st));
        PS_HIP(hipMemcpyAsync(probs, hp.data(), np * sizeof(ps_match_problem), hipMemcpyHostToDevice, st));
        PS_HIP(hipStreamSynchronize(st));
        e->pano_images.assign(images, images + n);
        e->pano_tables = small;
    }

    PS_TRY(ps_engine_extract_batch_u8(e, d_pixels, images, n, &cfg, feat, desc, kPanoCap, counts, 0));
    PS_HIP(ps::launch_pano_sets(counts, n, kPanoCap, sets, st));
    ps::MatchSetsArgs a{};
    a.qf = a.tf = feat;
    a.qd = a.td = desc;
    a.q_rows = a.t_rows = rows;
    a.qsets = a.tsets = sets;
    a.n_qsets = a.n_tsets = n;
    a.cap = kPanoCap;
    a.probs = probs;
    a.n_probs = np;
    a.ratio = o.ratio;
    a.shared = true;
    PS_TRY(enqueue_match_sets(e, a, e->aux[AUX_PN_0], matches, mcount));
    PS_HIP(ps::launch_pano_ransac_problems(probs, mcount, np, kDefaultSimSeed, kDefaultHomSeed, rp, st));
    ps_ransac_params p;
    std::memset(&p, 0, sizeof p);
    p.reproj_threshold = o.ransac_threshold;
    p.sampler = PS_SAMPLER_REFERENCE;
    p.local_opt = 1;
    p.sim_seed_model = 1;
    p.adaptive = 1;
    p.confidence = 0.999;   // pipeline.rs:2689
    p.min_iters = 64;
    p.max_iters = 2000;
    // filter_matches_ransac and estimate_homography_ransac at the same threshold with the
    // default seeds: the homography's similarity stage is filter_matches_ransac's run
    PS_HIP(ps::launch_ransac(matches, nullptr, rp, np, p, total, e->aux[AUX_PN_6].p, as<uint32_t>(e->aux[AUX_PN_7]),
                             rr, st));
    PS_HIP(ps::launch_pano_edges(d_pixels, d_img, probs, mcount, np, matches, rr, as<uint32_t>(e->aux[AUX_PN_7]),
                                 e->aux[AUX_PN_6].p, total, d_pairs, st));
    return PS_OK;
}

ps_status check_images(const ps_image_desc *images, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) {
        if (images[i].row_stride < images[i].width) return set_err(PS_E_INVALID_ARG, "row_stride < width");
        if (images[i].height == 0 || images[i].width == 0) return set_err(PS_E_INVALID_ARG, "empty image");
    }
    return PS_OK;
}

}  // namespace

extern "C" {

ps_status ps_engine_match_sets(ps_engine *e, const ps_feature *d_qf, const float *d_qd, size_t q_rows,
                               const ps_feature_set *d_qsets, uint32_t n_qsets, const ps_feature *d_tf,
                               const float *d_td, size_t t_rows, const ps_feature_set *d_tsets, uint32_t n_tsets,
                               uint32_t set_cap, const ps_match_problem *d_problems, uint32_t n_problems,
                               float ratio, ps_feature_match *d_out, uint32_t *d_out_counts) {
    PS_TRY(check_fail_injection());
    if (!e) return set_err(PS_E_INVALID_ARG, "null argument");
    if (n_problems == 0) return PS_OK;
    if (!d_problems || !d_out || !d_out_counts || !d_qsets || !d_tsets || (q_rows && (!d_qf || !d_qd)) ||
        (t_rows && (!d_tf || !d_td)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (n_problems > kMaxProblems) return set_err(PS_E_INVALID_ARG, "at most 65535 problems per call");
    if (set_cap > 262144) return set_err(PS_E_INVALID_ARG, "set_cap above 262144");
    if (q_rows >= 0xffffffffull || t_rows >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "too many rows");
    if (!std::isfinite(ratio) || !(ratio > 0.0f)) return set_err(PS_E_CONFIG, "ratio must be a positive finite number");
    PS_HIP(hipSetDevice(e->device));
    ps::MatchSetsArgs a{};
    a.qf = d_qf;
    a.tf = d_tf;
    a.qd = d_qd;
    a.td = d_td;
    a.q_rows = q_rows;
    a.t_rows = t_rows;
    a.qsets = d_qsets;
    a.tsets = d_tsets;
    a.n_qsets = n_qsets;
    a.n_tsets = n_tsets;
    a.cap = set_cap;
    a.probs = d_problems;
    a.n_probs = n_problems;
    a.ratio = ratio;
    a.shared = d_qf == d_tf && d_qd == d_td && q_rows == t_rows && d_qsets == d_tsets && n_qsets == n_tsets;
    return enqueue_match_sets(e, a, e->aux[AUX_PN_0], d_out, d_out_counts);
}

ps_status ps_engine_panorama_pairs(ps_engine *e, const uint8_t *d_pixels, const ps_image_desc *images,
                                   uint32_t n_images, const ps_panorama_options *opt, ps_panorama_pair *d_pairs) {
    PS_TRY(check_fail_injection());
    if (!e || !opt) return set_err(PS_E_INVALID_ARG, "null argument");
    if (n_images < 2) return set_err(PS_E_INVALID_ARG, "Panorama stitching requires at least 2 input images");
    if (!d_pixels || !images || !d_pairs) return set_err(PS_E_INVALID_ARG, "null argument");
    if ((size_t)n_images * (n_images - 1) / 2 > kMaxProblems) return set_err(PS_E_INVALID_ARG, "at most 362 images");
    PS_TRY(check_pano_options(opt));
    PS_TRY(check_images(images, n_images));
    PS_HIP(hipSetDevice(e->device));
    return enqueue_panorama(e, d_pixels, images, n_images, *opt, d_pairs);
}

ps_status ps_panorama_pair_edges(const uint8_t *const *luma, const uint32_t *heights, const uint32_t *widths,
                                 const uint32_t *row_strides, uint32_t n_images, const ps_panorama_options *opt,
                                 ps_pair_edge *out, size_t cap, size_t *out_count) {
    TraceRange tr("ps.panorama_pair_edges");
    PS_TRY(check_fail_injection());
    if (!opt || !out_count) return set_err(PS_E_INVALID_ARG, "null argument");
    *out_count = 0;
    if (n_images < 2) return set_err(PS_E_INVALID_ARG, "Panorama stitching requires at least 2 input images");
    if (!luma || !heights || !widths || (cap && !out)) return set_err(PS_E_INVALID_ARG, "null argument");
    if ((size_t)n_images * (n_images - 1) / 2 > kMaxProblems) return set_err(PS_E_INVALID_ARG, "at most 362 images");
    PS_TRY(check_pano_options(opt));
    std::vector<ps_image_desc> im(n_images);
    size_t bytes = 0;
    for (uint32_t i = 0; i < n_images; i++) {
        if (!luma[i]) return set_err(PS_E_INVALID_ARG, "null argument");
        // packed on the device: row stride = width
        im[i] = ps_image_desc{bytes, heights[i], widths[i], widths[i], 0u};
        bytes += al256((size_t)heights[i] * widths[i]);
    }
    for (uint32_t i = 0; i < n_images; i++)
        if (row_strides && row_strides[i] < widths[i]) return set_err(PS_E_INVALID_ARG, "row_stride < width");
    PS_TRY(check_images(im.data(), n_images));
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    hipStream_t st = e->stream;
    const uint32_t np = n_images * (n_images - 1) / 2;
    PS_TRY(ensure(e->aux[AUX_PN_8], bytes));
    PS_TRY(ensure(e->aux[AUX_PN_9], np * sizeof(ps_panorama_pair)));
    uint8_t *d_pix = as<uint8_t>(e->aux[AUX_PN_8]);
    for (uint32_t i = 0; i < n_images; i++) {
        const size_t stride = row_strides ? row_strides[i] : widths[i];
        PS_HIP(hipMemcpy2DAsync(d_pix + im[i].offset, widths[i], luma[i], stride, widths[i], heights[i],
                                hipMemcpyHostToDevice, st));
    }
    PS_TRY(enqueue_panorama(e, d_pix, im.data(), n_images, *opt, as<ps_panorama_pair>(e->aux[AUX_PN_9])));
    std::vector<ps_panorama_pair> pairs(np);
    PS_HIP(hipMemcpyAsync(pairs.data(), e->aux[AUX_PN_9].p, np * sizeof(ps_panorama_pair), hipMemcpyDeviceToHost, st));
    PS_HIP(hipStreamSynchronize(st));
    // an image whose extraction overflowed the candidate capacity has incomplete features:
    // report it instead of edges built from them (the single-image path fails the same way)
    PS_TRY(ps_engine_check(e));
    size_t ne = 0;
    for (const ps_panorama_pair &p : pairs) {
        if (!p.has_edge) continue;
        if (ne + 2 <= cap) {
            ps_pair_edge &f = out[ne], &r = out[ne + 1];
            f.from = p.from;
            f.to = p.to;
            f.inliers = p.inliers;
            f.mean_i_from = p.mean_i_from;
            f.mean_i_to = p.mean_i_to;
            r.from = p.to;
            r.to = p.from;
            r.inliers = p.inliers;
            r.mean_i_from = p.mean_i_to;
            r.mean_i_to = p.mean_i_from;
            for (int c = 0; c < 9; c++) {
                f.h_from_to[c] = p.h_from_to[c];
                r.h_from_to[c] = p.h_to_from[c];
            }
        }
        ne += 2;
    }
    *out_count = ne;
    if (ne > cap) return set_err(PS_E_CAPACITY, "edge capacity too small");
    return PS_OK;
}

}  // extern "C"
