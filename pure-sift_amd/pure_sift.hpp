// pure_sift.hpp -- header-only C++ host mirror of pure-sift 0.2.0's hot-path API
// (src/pipeline.rs:22-170) over the C ABI in include/ps_sift.h.
//
//   auto sift = pure_sift::Sift::create(pure_sift::SiftConfig().octaves(4));   // throws ConfigError
//   std::vector<pure_sift::SiftFeature> f = sift.detect_and_compute(img, h, w, w);
//   auto m = pure_sift::Matcher::create({0.8f, false}).match_(fq, ft);
//
// Validation happens at construction (Result<_, String> in the reference -> ConfigError);
// compute calls never fall back to the CPU: a device failure throws DeviceError (the
// reference's infallible signatures become a panic in a Rust wrapper over the same C ABI).
#pragma once

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../include/ps_sift.h"

namespace pure_sift {

constexpr float BASE_SIGMA = 1.6f;                 // pipeline.rs:10
constexpr float ASSUMED_INPUT_SIGMA = 0.5f;        // pipeline.rs:13
constexpr bool DEFAULT_UPSAMPLE_FIRST_OCTAVE = true;
constexpr float DEFAULT_CONTRAST_THRESHOLD = 0.015f;
constexpr std::size_t SIFT_DESCRIPTOR_LEN = 128;   // pipeline.rs:998
constexpr float DEFAULT_MATCH_RATIO = 0.75f;       // pipeline.rs:1047
constexpr std::size_t MAX_FEATURES_PER_IMAGE = 4096;

struct ConfigError : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct DeviceError : std::runtime_error {
    ps_status status;
    DeviceError(ps_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void check(ps_status s) {
    if (s == PS_OK) return;
    std::string msg = ps_last_error();
    if (msg.empty()) msg = ps_status_string(s);
    if (s == PS_E_CONFIG || s == PS_E_FORMAT) throw ConfigError(msg);
    throw DeviceError(s, msg);
}

enum class DescriptorKind : uint8_t { Sift = PS_DESCRIPTOR_SIFT, RootSift = PS_DESCRIPTOR_ROOTSIFT };

// pipeline.rs:22-104
struct SiftConfig {
    std::size_t octaves_ = 3, scales_ = 3;
    float contrast_threshold_ = DEFAULT_CONTRAST_THRESHOLD;
    bool upsample_first_octave_ = DEFAULT_UPSAMPLE_FIRST_OCTAVE;
    DescriptorKind descriptor_kind_ = DescriptorKind::Sift;

    SiftConfig octaves(std::size_t v) const { auto c = *this; c.octaves_ = v; return c; }
    SiftConfig scales(std::size_t v) const { auto c = *this; c.scales_ = v; return c; }
    SiftConfig contrast_threshold(float v) const { auto c = *this; c.contrast_threshold_ = v; return c; }
    SiftConfig upsample_first_octave(bool v) const { auto c = *this; c.upsample_first_octave_ = v; return c; }
    SiftConfig descriptor_kind(DescriptorKind v) const { auto c = *this; c.descriptor_kind_ = v; return c; }

    ps_sift_config c_config() const {
        ps_sift_config c{};
        c.octaves = octaves_ > 0xffffffffu ? 0xffffffffu : (uint32_t)octaves_;
        c.scales = scales_ > 0xffffffffu ? 0xffffffffu : (uint32_t)scales_;
        c.contrast_threshold = contrast_threshold_;
        c.upsample_first_octave = upsample_first_octave_ ? 1 : 0;
        c.descriptor_kind = (uint8_t)descriptor_kind_;
        return c;
    }
    // Returns the reference's error message, or "" when valid.
    std::string validate() const {
        ps_sift_config c = c_config();
        char msg[256];
        return ps_validate_sift_config(&c, msg, sizeof msg) == PS_OK ? std::string() : std::string(msg);
    }
};

struct Keypoint {            // pipeline.rs:251-267
    float x, y, scale, response;
    std::size_t octave, layer;
};

struct SiftFeature {         // pipeline.rs:1058-1072
    float x, y, scale, response, orientation;
    std::vector<float> descriptor;
};

struct FeatureMatch {        // pipeline.rs:1075-1091
    std::size_t query_index, target_index;
    std::pair<float, float> query_point, target_point;
    float distance, ratio;
    std::pair<float, float> displacement;
};

struct MatchOptions {        // pipeline.rs:1720-1736
    float ratio = DEFAULT_MATCH_RATIO;
    bool mutual = false;
    uint8_t precision = PS_PRECISION_FP32;   // distance engine (PS_PRECISION_*)
};

// pipeline.rs:106-146
class Sift {
public:
    static Sift create(const SiftConfig &config) {
        std::string err = config.validate();
        if (!err.empty()) throw ConfigError(err);
        return Sift(config);
    }
    SiftConfig config() const { return config_; }

    // img: h x w f32 luma in [0, 1], row stride in elements
    std::vector<Keypoint> detect_keypoints(const float *img, std::size_t h, std::size_t w, std::size_t stride) const {
        ps_sift_config c = config_.c_config();
        std::size_t n = 0;
        std::vector<ps_keypoint> buf(1 << 14);
        ps_status s = ps_detect_keypoints(img, h, w, stride, &c, buf.data(), buf.size(), &n);
        if (s == PS_E_CAPACITY && n > buf.size()) {
            buf.resize(n);
            s = ps_detect_keypoints(img, h, w, stride, &c, buf.data(), buf.size(), &n);
        }
        check(s);
        std::vector<Keypoint> out(n);
        for (std::size_t i = 0; i < n; i++)
            out[i] = {buf[i].x, buf[i].y, buf[i].scale, buf[i].response, buf[i].octave, buf[i].layer};
        return out;
    }

    std::vector<SiftFeature> detect_and_compute(const float *img, std::size_t h, std::size_t w,
                                                std::size_t stride) const {
        return run(img, h, w, stride, false);
    }

    // pipeline.rs:1581-1620 compute_sift_descriptors_with_kind (this config's descriptor kind):
    // uncapped features for the given keypoints, pyramid parameters inferred from them
    std::vector<SiftFeature> compute_descriptors(const float *img, std::size_t h, std::size_t w, std::size_t stride,
                                                 const std::vector<Keypoint> &kps) const {
        std::vector<ps_keypoint> k(kps.size());
        for (std::size_t i = 0; i < kps.size(); i++)
            k[i] = {kps[i].x, kps[i].y, kps[i].scale, kps[i].response, (uint32_t)kps[i].octave, (uint32_t)kps[i].layer};
        const std::size_t cap = 4 * kps.size();
        std::vector<ps_feature> hdr(cap ? cap : 1);
        std::vector<float> desc(hdr.size() * SIFT_DESCRIPTOR_LEN);
        std::size_t n = 0;
        check(ps_compute_descriptors(img, h, w, stride, k.data(), k.size(), (uint8_t)config_.descriptor_kind_,
                                     hdr.data(), desc.data(), cap, &n));
        std::vector<SiftFeature> out(n);
        for (std::size_t i = 0; i < n; i++) {
            out[i] = {hdr[i].x, hdr[i].y, hdr[i].scale, hdr[i].response, hdr[i].orientation,
                      std::vector<float>(desc.begin() + i * SIFT_DESCRIPTOR_LEN,
                                         desc.begin() + (i + 1) * SIFT_DESCRIPTOR_LEN)};
        }
        return out;
    }
    std::vector<SiftFeature> detect_and_compute_u8(const uint8_t *img, std::size_t h, std::size_t w,
                                                   std::size_t stride) const {
        return run(img, h, w, stride, true);
    }

private:
    explicit Sift(const SiftConfig &c) : config_(c) {}
    std::vector<SiftFeature> run(const void *img, std::size_t h, std::size_t w, std::size_t stride, bool u8) const {
        ps_sift_config c = config_.c_config();
        std::vector<ps_feature> hdr(MAX_FEATURES_PER_IMAGE);
        std::vector<float> desc(MAX_FEATURES_PER_IMAGE * SIFT_DESCRIPTOR_LEN);
        std::size_t n = 0;
        check(u8 ? ps_detect_and_compute_u8(static_cast<const uint8_t *>(img), h, w, stride, &c, hdr.data(),
                                            desc.data(), hdr.size(), &n)
                 : ps_detect_and_compute(static_cast<const float *>(img), h, w, stride, &c, hdr.data(), desc.data(),
                                         hdr.size(), &n));
        std::vector<SiftFeature> out(n);
        for (std::size_t i = 0; i < n; i++) {
            out[i] = {hdr[i].x, hdr[i].y, hdr[i].scale, hdr[i].response, hdr[i].orientation,
                      std::vector<float>(desc.begin() + i * SIFT_DESCRIPTOR_LEN,
                                         desc.begin() + (i + 1) * SIFT_DESCRIPTOR_LEN)};
        }
        return out;
    }
    SiftConfig config_;
};

// pipeline.rs:148-170
class Matcher {
public:
    static Matcher create(const MatchOptions &o) {
        ps_match_options c{o.ratio, (uint8_t)(o.mutual ? 1 : 0), o.precision, {0, 0}};
        char msg[160];
        if (ps_validate_match_options(&c, msg, sizeof msg) != PS_OK) throw ConfigError(msg);
        return Matcher(o);
    }

    std::vector<FeatureMatch> match_(const std::vector<SiftFeature> &query,
                                     const std::vector<SiftFeature> &target) const {
        std::vector<ps_feature> qh, th;
        std::vector<float> qd, td;
        pack(query, qh, qd);
        pack(target, th, td);
        ps_match_options c{opt_.ratio, (uint8_t)(opt_.mutual ? 1 : 0), opt_.precision, {0, 0}};
        std::vector<ps_feature_match> buf(query.size() ? query.size() : 1);
        std::size_t n = 0;
        check(ps_match(qh.data(), qd.data(), qh.size(), th.data(), td.data(), th.size(), &c, buf.data(), buf.size(),
                       &n));
        std::vector<FeatureMatch> out(n);
        for (std::size_t i = 0; i < n; i++) {
            const ps_feature_match &m = buf[i];
            out[i] = {(std::size_t)m.query_index, (std::size_t)m.target_index, {m.qx, m.qy}, {m.tx, m.ty},
                      m.distance, m.ratio, {m.dx, m.dy}};
        }
        return out;
    }

private:
    explicit Matcher(const MatchOptions &o) : opt_(o) {}
    static void pack(const std::vector<SiftFeature> &f, std::vector<ps_feature> &h, std::vector<float> &d) {
        h.resize(f.size());
        d.resize(f.size() * SIFT_DESCRIPTOR_LEN);
        for (std::size_t i = 0; i < f.size(); i++) {
            if (f[i].descriptor.size() != SIFT_DESCRIPTOR_LEN)
                throw ConfigError("descriptor length must be 128");
            h[i] = {f[i].x, f[i].y, f[i].scale, f[i].response, f[i].orientation};
            std::copy(f[i].descriptor.begin(), f[i].descriptor.end(), d.begin() + i * SIFT_DESCRIPTOR_LEN);
        }
    }
    MatchOptions opt_;
};

// ---- feature catalog (src/catalog.rs) over ps_catalog_* --------------------------------

constexpr std::size_t DEFAULT_CATALOG_SHORTLIST = 16;   // catalog.rs:21
constexpr std::size_t DEFAULT_CATALOG_TOP_K = 5;        // catalog.rs:24
constexpr float DEFAULT_RANSAC_REPROJ_THRESHOLD = 3.0f; // pipeline.rs:1050

struct TrackingSummary {       // pipeline.rs:1137-1153
    std::size_t raw_match_count, match_count;
    float inlier_ratio;
    std::pair<float, float> mean_displacement, median_displacement;
    float mean_scale_ratio, mean_rotation_deg;
};
struct SimilarityTransform2D { float a, b, tx, ty, scale, rotation_deg; };   // pipeline.rs:1096-1109
struct CatalogEntrySummary {   // catalog.rs:47-57
    std::string id;
    std::size_t width, height, feature_count;
};
struct CatalogSearchHit {      // catalog.rs:158-177
    std::string id;
    std::size_t width, height, feature_count, coarse_votes;
    TrackingSummary tracking;
    std::optional<SimilarityTransform2D> transform;
    std::optional<std::array<std::array<float, 3>, 3>> homography;
};

// catalog.rs:179-197, resident in device memory. Mutation needs exclusive access.
class FeatureCatalog {
public:
    FeatureCatalog(std::size_t octaves, std::size_t scales, float threshold) {   // catalog.rs:272-282
        ps_catalog *c = nullptr;
        check(ps_catalog_create((uint32_t)octaves, (uint32_t)scales, threshold, PS_DESCRIPTOR_SIFT, &c));
        h_.reset(c);
    }
    static FeatureCatalog load_binary(const std::string &path) {              // catalog.rs:447
        ps_catalog *c = nullptr;
        check(ps_catalog_load_file(path.c_str(), &c));
        return FeatureCatalog(c);
    }
    static FeatureCatalog from_binary_bytes(const std::vector<uint8_t> &bytes) {
        ps_catalog *c = nullptr;
        check(ps_catalog_load_binary(bytes.data(), bytes.size(), &c));
        return FeatureCatalog(c);
    }
    void save_binary(const std::string &path) const { check(ps_catalog_save_binary(h_.get(), path.c_str())); }

    std::size_t size() const { return info().n_entries; }
    std::vector<CatalogEntrySummary> list_summaries() const {                    // catalog.rs:347-357
        std::vector<CatalogEntrySummary> out;
        for (std::size_t i = 0; i < size(); i++) out.push_back(summary(i));
        return out;
    }
    void upsert_entry(const std::string &id, std::size_t width, std::size_t height,
                      const std::vector<SiftFeature> &features) {                 // catalog.rs:309-328
        std::vector<ps_feature> h;
        std::vector<float> d;
        pack(features, h, d);
        check(ps_catalog_upsert(h_.get(), id.data(), id.size(), (uint32_t)width, (uint32_t)height, h.data(),
                                d.data(), h.size()));
    }
    bool remove(const std::string &id) {                                          // catalog.rs:331-339
        int32_t r = 0;
        check(ps_catalog_remove(h_.get(), id.data(), id.size(), &r));
        return r != 0;
    }
    std::vector<CatalogSearchHit> search_features(const std::vector<SiftFeature> &query,
                                                  std::size_t top_k = DEFAULT_CATALOG_TOP_K,
                                                  std::size_t shortlist_k = DEFAULT_CATALOG_SHORTLIST,
                                                  float ratio_threshold = DEFAULT_MATCH_RATIO,
                                                  float ransac_threshold = DEFAULT_RANSAC_REPROJ_THRESHOLD) const {
        std::vector<ps_feature> h;
        std::vector<float> d;
        pack(query, h, d);
        std::vector<ps_catalog_hit> buf(std::max<std::size_t>(std::min(top_k, size()), 1));
        std::size_t n = 0;
        check(ps_catalog_search(h_.get(), h.data(), d.data(), h.size(), top_k, shortlist_k, ratio_threshold,
                                ransac_threshold, buf.data(), buf.size(), &n));
        std::vector<CatalogSearchHit> out;
        for (std::size_t i = 0; i < n; i++) {
            const ps_catalog_hit &r = buf[i];
            const CatalogEntrySummary s = summary(r.entry);
            CatalogSearchHit hit{s.id, s.width, s.height, s.feature_count, r.coarse_votes,
                                 {(std::size_t)r.raw_match_count, (std::size_t)r.match_count, r.inlier_ratio,
                                  {r.mean_dx, r.mean_dy}, {r.median_dx, r.median_dy}, r.mean_scale_ratio,
                                  r.mean_rotation_deg},
                                 std::nullopt, std::nullopt};
            if (r.has_transform)
                hit.transform = SimilarityTransform2D{r.transform[0], r.transform[1], r.transform[2],
                                                      r.transform[3], r.transform[4], r.transform[5]};
            if (r.has_homography) {
                std::array<std::array<float, 3>, 3> m{};
                for (int k = 0; k < 9; k++) m[k / 3][k % 3] = r.homography[k];
                hit.homography = m;
            }
            out.push_back(std::move(hit));
        }
        return out;
    }

private:
    struct Del {
        void operator()(ps_catalog *c) const { ps_catalog_destroy(c); }
    };
    explicit FeatureCatalog(ps_catalog *c) : h_(c) {}
    ps_catalog_info info() const {
        ps_catalog_info i;
        check(ps_catalog_get_info(h_.get(), &i));
        return i;
    }
    CatalogEntrySummary summary(std::size_t idx) const {
        ps_catalog_entry_info e;
        check(ps_catalog_entry(h_.get(), idx, &e));
        return {std::string(e.id, e.id_len), e.width, e.height, (std::size_t)e.feature_count};
    }
    static void pack(const std::vector<SiftFeature> &f, std::vector<ps_feature> &h, std::vector<float> &d) {
        h.resize(f.size());
        d.resize(f.size() * SIFT_DESCRIPTOR_LEN);
        for (std::size_t i = 0; i < f.size(); i++) {
            if (f[i].descriptor.size() != SIFT_DESCRIPTOR_LEN) throw ConfigError("descriptor length must be 128");
            h[i] = {f[i].x, f[i].y, f[i].scale, f[i].response, f[i].orientation};
            std::copy(f[i].descriptor.begin(), f[i].descriptor.end(), d.begin() + i * SIFT_DESCRIPTOR_LEN);
        }
    }
    std::unique_ptr<ps_catalog, Del> h_;
};

// stitch.rs:45-86 PanoramaOptions (the pairwise stage's fields, the reference's defaults)
struct PanoramaOptions {
    uint32_t octaves = 3, scales = 3;
    float threshold = 0.015f, ratio = 0.75f, ransac_threshold = 3.0f;
};
// stitch.rs:104-112 PairEdge
struct PairEdge {
    std::size_t from, to;
    double h_from_to[3][3];
    std::size_t inliers;
    double mean_i_from, mean_i_to;
};
struct LumaImage {
    const uint8_t *pixels;
    uint32_t height, width, row_stride;
};
// stitch.rs:478-571: the pair edges of all image pairs, in the reference's order
inline std::vector<PairEdge> panorama_pair_edges(const std::vector<LumaImage> &images,
                                                 const PanoramaOptions &o = PanoramaOptions()) {
    std::vector<const uint8_t *> px;
    std::vector<uint32_t> hs, ws, ss;
    for (const LumaImage &im : images) {
        px.push_back(im.pixels);
        hs.push_back(im.height);
        ws.push_back(im.width);
        ss.push_back(im.row_stride);
    }
    const ps_panorama_options c{o.octaves, o.scales, o.threshold, o.ratio, o.ransac_threshold, 0};
    const std::size_t n = images.size();
    std::vector<ps_pair_edge> raw(n > 1 ? n * (n - 1) : 1);
    std::size_t cnt = 0;
    check(ps_panorama_pair_edges(px.data(), hs.data(), ws.data(), ss.data(), (uint32_t)n, &c, raw.data(), raw.size(),
                                 &cnt));
    std::vector<PairEdge> out(cnt);
    for (std::size_t k = 0; k < cnt; k++) {
        out[k].from = raw[k].from;
        out[k].to = raw[k].to;
        for (int r = 0; r < 3; r++)
            for (int col = 0; col < 3; col++) out[k].h_from_to[r][col] = raw[k].h_from_to[r * 3 + col];
        out[k].inliers = raw[k].inliers;
        out[k].mean_i_from = raw[k].mean_i_from;
        out[k].mean_i_to = raw[k].mean_i_to;
    }
    return out;
}

}  // namespace pure_sift
