// Exercises the C++ host mirror (pure-sift_amd/pure_sift.hpp) end to end: validation
// errors, detect_keypoints, detect_and_compute (f32 and u8), compute_descriptors, match_ (plain
// and mutual), the feature catalog and the panorama pair edges.
// Reads a raw u8 image (h, w from argv) and prints a one-line summary the GPU test checks.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "pure_sift.hpp"

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s query.u8 target.u8 h w\n", argv[0]);
        return 2;
    }
    std::size_t h = std::stoul(argv[3]), w = std::stoul(argv[4]);
    auto load = [&](const char *p) {
        std::vector<uint8_t> v(h * w);
        std::ifstream f(p, std::ios::binary);
        f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)v.size());
        return v;
    };
    std::vector<uint8_t> q8 = load(argv[1]), t8 = load(argv[2]);
    int bad = 0;
    try {
        pure_sift::Sift::create(pure_sift::SiftConfig().scales(2));
        bad++;
    } catch (const pure_sift::ConfigError &e) {
        if (std::string(e.what()).find("Invalid `scales`=2") == std::string::npos) bad++;
    }
    try {
        pure_sift::Matcher::create({1.5f, false});
        bad++;
    } catch (const pure_sift::ConfigError &) {
    }
    auto sift = pure_sift::Sift::create(pure_sift::SiftConfig().octaves(4).scales(3));
    std::vector<float> qf(h * w);
    for (std::size_t i = 0; i < h * w; i++) qf[i] = (float)q8[i] * (1.0f / 255.0f);
    auto kps = sift.detect_keypoints(qf.data(), h, w, w);
    auto fq = sift.detect_and_compute_u8(q8.data(), h, w, w);
    auto fq32 = sift.detect_and_compute(qf.data(), h, w, w);
    auto ft = sift.detect_and_compute_u8(t8.data(), h, w, w);
    if (fq.size() != fq32.size()) bad++;
    for (std::size_t i = 0; i < fq.size() && i < fq32.size(); i++)
        if (fq[i].descriptor != fq32[i].descriptor || fq[i].orientation != fq32[i].orientation) bad++;
    // uncapped CLI extract path: detect_keypoints + compute_descriptors on the same keypoints
    auto fk = sift.compute_descriptors(qf.data(), h, w, w, kps);
    if (fk.size() < kps.size()) bad++;
    auto m = pure_sift::Matcher::create({0.8f, false}).match_(fq, ft);
    auto mm = pure_sift::Matcher::create({0.8f, true}).match_(fq, ft);
    // feature catalog: the target under two ids, then search with the query; save / load round trip
    pure_sift::FeatureCatalog cat(4, 3, 0.015f);
    cat.upsert_entry("decoy", w, h, std::vector<pure_sift::SiftFeature>(ft.begin(), ft.begin() + ft.size() / 10));
    cat.upsert_entry("target", w, h, ft);
    auto hits = cat.search_features(fq);
    if (hits.empty() || hits[0].id != "target" || hits[0].tracking.match_count == 0) bad++;
    const char *dir = std::getenv("MIRROR_DEMO_TMP");
    if (dir) {
        const std::string path = std::string(dir) + "/demo.psift";
        cat.save_binary(path);
        auto back = pure_sift::FeatureCatalog::load_binary(path);
        auto names = back.list_summaries();
        if (names.size() != 2 || names[1].id != "target" || names[1].feature_count != ft.size()) bad++;
    }
    if (!cat.remove("decoy") || cat.remove("decoy") || cat.size() != 1) bad++;
    // panorama pairwise stage over (query, target, query): the two identical images align by the identity
    std::vector<pure_sift::LumaImage> pano{{q8.data(), (uint32_t)h, (uint32_t)w, (uint32_t)w},
                                           {t8.data(), (uint32_t)h, (uint32_t)w, (uint32_t)w},
                                           {q8.data(), (uint32_t)h, (uint32_t)w, (uint32_t)w}};
    auto edges = pure_sift::panorama_pair_edges(pano);
    bool ident = false;
    for (const auto &e : edges)
        if (e.from == 0 && e.to == 2)
            ident = std::abs(e.h_from_to[0][2]) < 0.5 && std::abs(e.h_from_to[1][2]) < 0.5 &&
                    e.mean_i_from == e.mean_i_to;
    if (edges.size() % 2 != 0 || !ident) bad++;
    try {
        pure_sift::panorama_pair_edges({pano[0]});
        bad++;
    } catch (const pure_sift::DeviceError &) {
    }
    std::printf("keypoints=%zu features=%zu,%zu described=%zu matches=%zu mutual=%zu catalog_hits=%zu pano_edges=%zu "
                "bad=%d\n",
                kps.size(), fq.size(), ft.size(), fk.size(), m.size(), mm.size(), hits.size(), edges.size(), bad);
    return bad ? 1 : 0;
}
