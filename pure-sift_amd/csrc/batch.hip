// Batched extraction of many uneven-size images as grouped persistent launches, and the
// launch-shape autotuners.
//
// Batching: every stage of the extraction runs once for the whole batch. One blur launch takes
// a pyramid level (octave o, level s) of every image that has octave o, one extrema launch an
// octave index of every image, and the ordering, orientation, feature-flattening and descriptor
// launches every image at once -- about 30 launches per batch whatever the number of images, each
// walking the units (tiles, strips, keypoints, features) of all its images with a resident grid
// (the grouped kernels of pyramid.hip and describe.hip, which run the single-image kernels'
// bodies, so every image gets exactly its single-image result). As in the single-image path,
// the dependent chain of levels 1..S across octaves stays on the engine stream and each octave's
// leaf work (level S+1 and its extrema) forks to a side stream. The sequence is captured once
// into a HIP graph and replayed while images, buffers, config and tuned shapes are unchanged.
//
// Autotuners: ps_engine_autotune_blur times every launch shape (tile kernel with 16, 32 or 64
// rows; sliding-window kernel with 64-, 128- or 192-row chunks) of single-level blur launches and
// keeps the fastest per (radius, source mode, level size class); ps_engine_autotune_batch times
// every shape of each grouped blur launch of a batch (and 16- or 32-row extrema tiles) and keeps
// the fastest per grouped-launch class. Both tables are process-wide.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "host_common.h"

using namespace pshost;

namespace {

using GroupClass = std::tuple<int, int, int, int, int>;   // kind, radius, mode, mean px class, jobs class

std::mutex g_tile_mu;
std::map<std::tuple<int, int, int, int>, int> g_tiles;   // (radius, mode, h class, w class) -> shape
std::map<GroupClass, int> g_group_tiles;                  // grouped launch class -> shape / rows
uint64_t g_tile_gen = 0;                                  // bumped on every table change

int size_class(int v) { return (int)std::floor(2.0 * std::log2((double)std::max(v, 1))); }
int size_class_d(double v) { return (int)std::floor(2.0 * std::log2(std::max(v, 1.0))); }

void key_push(std::vector<uint64_t> &k, uint64_t v) { k.push_back(v); }

constexpr int kKindBlur = 0, kKindExtrema = 1;
constexpr int kBlurShapes[] = {16, 32, 64, ps::kBlurSlide | 64, ps::kBlurSlide | 128, ps::kBlurSlide | 192};
constexpr uint32_t kCntWords = 16;    // counter words per image, as an engine's `cnt`

}  // namespace

namespace pshost {

// Untuned default launch shape of a blur level, from device timings of every shape on the
// levels of a 1920x1080 pyramid. Levels of >= 3M pixels: radius 5..8 take the sliding-window
// kernel with 128-row chunks, radius <= 4 32-row tiles, larger radii and the upsampling first
// level 64-row tiles; levels of >= 1M pixels take 32-row tiles, smaller ones 16-row tiles
// (more workgroups for a launch that cannot fill the device).
static int default_shape(int radius, int src_mode, int h, int w) {
    const long px = (long)h * w;
    const bool up = src_mode == ps::SRC_U8_UP || src_mode == ps::SRC_F32_UP;
    if (px >= (3L << 20)) {
        if (up || radius > 8) return 64;
        return radius <= 4 ? 32 : (ps::kBlurSlide | 128);
    }
    if (px >= (1L << 20)) return 32;
    return 16;
}

int blur_tile(int radius, int src_mode, int h, int w) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    auto it = g_tiles.find(std::make_tuple(radius, src_mode, size_class(h), size_class(w)));
    return it == g_tiles.end() ? default_shape(radius, src_mode, h, w) : it->second;
}

}  // namespace pshost

// ------------------------------------------------------------------ batch state
struct GBlur {                        // one grouped blur launch
    int o = 0, s = 0;                 // octave, level (s = 0: the first level, from the input)
    int radius = 0, mode = 0;
    bool center = false;              // every job also writes its source plane (level 0 of octave o)
    bool side = false;                // on octave o's side stream
    const ps::Taps *taps = nullptr;
    GroupClass cls;
    double mean_px = 0;
    int shape = 0;
    std::vector<ps::BlurJob> jobs;
    uint32_t n_units = 0;
    size_t table = 0;                 // byte offset of the jobs in the device table buffer
};
struct GExtrema {                     // one grouped extrema launch (octave o)
    int o = 0;
    GroupClass cls;
    double mean_px = 0;
    int rows = 16;
    std::vector<ps::ExtremaJob> jobs;
    uint32_t n_units = 0;
    size_t table = 0;
};
struct BatchState {
    std::vector<uint64_t> key;        // images, buffers, config
    uint64_t shapes_gen = ~0ull;      // table generation the launch shapes were taken from
    uint32_t n = 0;
    int n_oct = 0;                    // octaves of the image with the most
    ps_sift_config cfg{};
    ps::PyrPlan ref;                  // taps, radii and per-octave scales (size-independent)
    std::vector<GBlur> blur;          // launch order: first level, then per octave levels 1..S+1
    std::vector<GExtrema> ex;         // one per octave
    bool full = false;                // ordering and description included (not for the autotuner)
    size_t cap = 0;
    uint32_t *d_counts = nullptr;
    size_t t_rank = 0, t_desc = 0;    // byte offsets of the ordering / description tables
    std::vector<unsigned char> fixed; // those two tables (the shaped tables follow them)
    DevBuf pyr, cand, rank, sorted, cnt, ori, nori, fkp, fj, off, tables;
};

namespace {

__global__ void clear_counters_kernel(uint32_t *cnt, uint32_t n_words, uint32_t keep) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_words && i % kCntWords != keep) cnt[i] = 0u;
}

// every counter of every image but the sticky error words (keep = kErrWord), or all of them
// (keep = kCntWords)
hipError_t launch_clear_counters(uint32_t *cnt, uint32_t n, uint32_t keep, hipStream_t st) {
    const uint32_t words = n * kCntWords;
    if (!words) return hipSuccess;
    clear_counters_kernel<<<(words + 255) / 256, 256, 0, st>>>(cnt, words, keep);
    return hipGetLastError();
}

int blur_shape_for(const GBlur &g) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    auto it = g_group_tiles.find(g.cls);
    if (it != g_group_tiles.end()) return it->second;
    const int side = (int)std::sqrt(std::max(g.mean_px, 1.0));   // a square level of the mean size
    return default_shape(g.radius, g.mode, side, side);
}

int extrema_rows_for(const GExtrema &x) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    auto it = g_group_tiles.find(x.cls);
    if (it != g_group_tiles.end()) return it->second;
    return x.mean_px >= (double)(1L << 20) ? 32 : 16;
}

void number_blur(GBlur &g) {
    uint32_t u = 0;
    for (ps::BlurJob &j : g.jobs) {
        j.unit0 = u;
        j.units_x = (uint32_t)((j.a.w + 63) / 64);
        u += ps::group_units(j.a.h, j.a.w, g.shape);
    }
    g.n_units = u;
}

void number_extrema(GExtrema &x) {
    uint32_t u = 0;
    for (ps::ExtremaJob &j : x.jobs) {
        j.unit0 = u;
        j.units_x = (uint32_t)((j.w - 2 + 63) / 64);
        u += ps::extrema_group_units(j.h, j.w, x.rows);
    }
    x.n_units = u;
}

template <typename T>
size_t put_table(std::vector<unsigned char> &buf, const std::vector<T> &v) {
    const size_t at = (buf.size() + 255) & ~(size_t)255;
    buf.resize(at + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(buf.data() + at, v.data(), v.size() * sizeof(T));
    return at;
}

// Plans, job tables and workspace of a batch. Without d_counts (the autotuner) only the
// pyramid and extrema launches are set up.
ps_status build_batch(ps_engine *e, BatchState &B, const uint8_t *d_pixels, const ps_image_desc *im, uint32_t n,
                      const ps_sift_config &cfg, ps_feature *d_feat, float *d_desc, size_t cap,
                      uint32_t *d_counts) {
    const uint32_t kMax = PS_MAX_FEATURES;
    const int S = (int)cfg.scales;
    B.n = n;
    B.cfg = cfg;
    B.cap = cap;
    B.d_counts = d_counts;
    B.full = d_counts != nullptr;
    B.blur.clear();
    B.ex.clear();
    B.fixed.clear();
    B.shapes_gen = ~0ull;
    std::vector<ps::PyrPlan> plans(n);
    std::vector<std::array<size_t, ps::kMaxOct>> ooff(n);   // float offset of octave o of image i
    size_t total = 0;
    int ref_i = -1;
    B.n_oct = 0;
    for (uint32_t i = 0; i < n; i++) {
        ps::PyrPlan &p = plans[i];
        if (!ps::make_plan((int)im[i].height, (int)im[i].width, cfg, p))
            return set_err(PS_E_CONFIG, "unsupported blur radius");
        for (int o = 0; o < p.n_oct; o++) {
            ooff[i][o] = total;
            total += (size_t)(S + 2) * p.h[o] * p.w[o];   // levels 0..S+1 (S+2 is never stored)
        }
        total = (total + 63) & ~(size_t)63;
        if (p.n_oct > B.n_oct) {
            B.n_oct = p.n_oct;
            ref_i = (int)i;
        }
    }
    // (no image with an octave: no scale-space launches, every image orders and describes nothing)
    if (ref_i >= 0) B.ref = plans[ref_i];
    PS_TRY(ensure(B.pyr, total * sizeof(float)));
    PS_TRY(ensure(B.cand, (size_t)n * ps::kCandCap * sizeof(ps::Cand)));
    float *P = as<float>(B.pyr);
    ps::Cand *cand = as<ps::Cand>(B.cand);
    uint32_t *cnt = as<uint32_t>(B.cnt);
    auto cls_of = [&](int kind, int radius, int mode, double mean_px, size_t jobs) {
        return std::make_tuple(kind, radius, mode, size_class_d(mean_px), size_class((int)jobs));
    };
    if (B.n_oct > 0) {   // the first level, from the u8 input (fused luma conversion and 2x upsample)
        GBlur g;
        g.radius = B.ref.init_r;
        g.mode = B.ref.upsample ? ps::SRC_U8_UP : ps::SRC_U8;
        g.taps = &B.ref.init_taps;
        for (uint32_t i = 0; i < n; i++) {
            const ps::PyrPlan &p = plans[i];
            if (p.n_oct == 0) continue;
            ps::BlurJob j{};
            j.a.src = d_pixels + im[i].offset;
            j.a.src_stride = im[i].row_stride;
            j.a.src_h = (int)im[i].height;
            j.a.src_w = (int)im[i].width;
            j.a.h = p.h[0];
            j.a.w = p.w[0];
            j.a.dst_g = P + ooff[i][0];
            g.jobs.push_back(j);
        }
        B.blur.push_back(std::move(g));
    }
    for (int o = 0; o < B.n_oct; o++) {
        for (int s = 1; s <= S + 1; s++) {
            GBlur g;
            g.o = o;
            g.s = s;
            g.radius = B.ref.lvl_r[s];
            g.mode = (s == 1 && o > 0) ? ps::SRC_DOWN2 : ps::SRC_F32;
            g.center = g.mode == ps::SRC_DOWN2;
            g.side = s == S + 1 && o + 1 < B.n_oct;
            g.taps = &B.ref.lvl_taps[s];
            for (uint32_t i = 0; i < n; i++) {
                const ps::PyrPlan &p = plans[i];
                if (p.n_oct <= o) continue;
                const size_t plane = (size_t)p.h[o] * p.w[o];
                ps::BlurJob j{};
                j.a.h = p.h[o];
                j.a.w = p.w[o];
                j.a.dst_g = P + ooff[i][o] + (size_t)s * plane;
                if (g.mode == ps::SRC_DOWN2) {   // octave base = downsample_2x(gaussians[S]) (pipeline.rs:624-626)
                    j.a.src = P + ooff[i][o - 1] + (size_t)S * p.h[o - 1] * p.w[o - 1];
                    j.a.src_stride = (size_t)p.w[o - 1];
                    j.a.src_h = p.h[o - 1];
                    j.a.src_w = p.w[o - 1];
                    j.a.dst_src = P + ooff[i][o];
                } else {
                    j.a.src = P + ooff[i][o] + (size_t)(s - 1) * plane;
                    j.a.src_stride = (size_t)p.w[o];
                    j.a.src_h = p.h[o];
                    j.a.src_w = p.w[o];
                }
                g.jobs.push_back(j);
            }
            B.blur.push_back(std::move(g));
        }
        GExtrema x;
        x.o = o;
        for (uint32_t i = 0; i < n; i++) {
            const ps::PyrPlan &p = plans[i];
            if (p.n_oct <= o || p.h[o] < 3 || p.w[o] < 3) continue;
            ps::ExtremaJob j{};
            j.gauss = P + ooff[i][o];
            j.out = cand + (size_t)i * ps::kCandCap;
            j.count = cnt + (size_t)i * kCntWords + 0;
            j.err = cnt + (size_t)i * kCntWords + kErrWord;
            j.h = p.h[o];
            j.w = p.w[o];
            x.jobs.push_back(j);
        }
        B.ex.push_back(std::move(x));
    }
    for (GBlur &g : B.blur) {
        double a = 0;
        for (const ps::BlurJob &j : g.jobs) a += (double)j.a.h * j.a.w;
        g.mean_px = g.jobs.empty() ? 0.0 : a / (double)g.jobs.size();
        g.cls = cls_of(kKindBlur, g.radius, g.mode, g.mean_px, g.jobs.size());
    }
    for (GExtrema &x : B.ex) {
        double a = 0;
        for (const ps::ExtremaJob &j : x.jobs) a += (double)j.h * j.w;
        x.mean_px = x.jobs.empty() ? 0.0 : a / (double)x.jobs.size();
        x.cls = cls_of(kKindExtrema, S, 0, x.mean_px, x.jobs.size());
    }
    if (!B.full) return PS_OK;
    // ordering and description
    const size_t rank_bytes = ps::rank_scratch_bytes(ps::kCandCap);
    PS_TRY(ensure(B.rank, (size_t)n * rank_bytes));
    PS_TRY(ensure(B.off, (size_t)2 * (n + 1) * sizeof(uint32_t)));
    PS_TRY(ensure(B.sorted, (size_t)n * kMax * sizeof(ps::Cand)));
    PS_TRY(ensure(B.ori, (size_t)n * kMax * 4 * sizeof(float)));
    PS_TRY(ensure(B.nori, (size_t)n * kMax * sizeof(uint32_t)));
    PS_TRY(ensure(B.fkp, (size_t)n * kMax * sizeof(uint32_t)));
    PS_TRY(ensure(B.fj, (size_t)n * kMax * sizeof(uint32_t)));
    std::vector<ps::RankJob> rj(n);
    std::vector<ps::DescJob> dj(n);
    for (uint32_t i = 0; i < n; i++) {
        const ps::PyrPlan &p = plans[i];
        uint32_t *ci = cnt + (size_t)i * kCntWords;
        ps::Cand *sorted = as<ps::Cand>(B.sorted) + (size_t)i * kMax;
        ps::RankJob &r = rj[i];
        std::memset(&r, 0, sizeof r);
        r.cand = cand + (size_t)i * ps::kCandCap;
        r.count = ci + 0;
        r.out = sorted;
        r.out_count = ci + 2;
        r.scratch = static_cast<unsigned char *>(B.rank.p) + (size_t)i * rank_bytes;
        ps::DescJob &d = dj[i];
        std::memset(&d, 0, sizeof d);
        d.pv.base = P;
        d.pv.n_oct = p.n_oct;
        d.pv.S = S;
        for (int o = 0; o < p.n_oct; o++) {
            d.pv.off[o] = ooff[i][o];
            d.pv.h[o] = p.h[o];
            d.pv.w[o] = p.w[o];
            d.pv.scale_factor[o] = p.scale_factor[o];
        }
        d.kps = sorted;
        d.nkp = ci + 2;
        d.ori = as<float>(B.ori) + (size_t)i * kMax * 4;
        d.nori = as<uint32_t>(B.nori) + (size_t)i * kMax;
        d.feat_kp = as<uint32_t>(B.fkp) + (size_t)i * kMax;
        d.feat_j = as<uint32_t>(B.fj) + (size_t)i * kMax;
        d.nfeat = d_counts + i;
        d.out_f = d_feat + (size_t)i * cap;
        d.out_d = d_desc + (size_t)i * cap * 128;
    }
    B.t_rank = put_table(B.fixed, rj);
    B.t_desc = put_table(B.fixed, dj);
    return PS_OK;
}

// Launch shapes from the tables (or the given ones), unit numbering, and every job table on the
// device. An earlier batch on the stream may still read the tables: wait for it first.
ps_status upload_tables(ps_engine *e, BatchState &B) {
    std::vector<unsigned char> buf(B.fixed);
    for (GBlur &g : B.blur) {
        number_blur(g);
        g.table = put_table(buf, g.jobs);
    }
    for (GExtrema &x : B.ex) {
        number_extrema(x);
        x.table = put_table(buf, x.jobs);
    }
    PS_HIP(hipStreamSynchronize(e->stream));
    PS_TRY(ensure(B.tables, buf.size()));
    PS_HIP(hipMemcpy(B.tables.p, buf.data(), buf.size(), hipMemcpyHostToDevice));
    return PS_OK;
}

ps_status shape_batch(ps_engine *e, BatchState &B, uint64_t gen) {
    for (GBlur &g : B.blur) g.shape = blur_shape_for(g);
    for (GExtrema &x : B.ex) x.rows = extrema_rows_for(x);
    PS_TRY(upload_tables(e, B));
    B.shapes_gen = gen;
    return PS_OK;
}

template <typename T>
const T *table_at(const BatchState &B, size_t at) {
    return reinterpret_cast<const T *>(static_cast<const unsigned char *>(B.tables.p) + at);
}

hipError_t launch_gblur(const BatchState &B, const GBlur &g, hipStream_t st) {
    return ps::launch_blur_group(g.radius, g.mode, table_at<ps::BlurJob>(B, g.table), (uint32_t)g.jobs.size(),
                                 g.n_units, *g.taps, g.shape, g.center, st);
}

hipError_t launch_gextrema(const BatchState &B, const GExtrema &x, hipStream_t st) {
    const ps::PyrPlan &r = B.ref;
    return ps::launch_extrema_group(table_at<ps::ExtremaJob>(B, x.table), (uint32_t)x.jobs.size(), x.n_units, r.S,
                                    x.rows, B.cfg.contrast_threshold, r.scale_factor[x.o], r.eff_scale[x.o], r.log2k,
                                    x.o, ps::kCandCap, r.lvl_taps[r.S + 2], r.lvl_r[r.S + 2], st);
}

// The pyramid and extrema of the whole batch: the chain (first level, levels 1..S of every
// octave) on the engine stream; level S+1 and the extrema of octave o on a side stream forked
// once level S of octave o is written (the last octave's on the engine stream); all joined back.
ps_status enqueue_scale_space_group(ps_engine *e, BatchState &B) {
    bool forked[ps_engine::kSides] = {};
    size_t bi = 0;
    if (!B.blur.empty()) PS_HIP(launch_gblur(B, B.blur[bi++], e->stream));
    for (int o = 0; o < B.n_oct; o++) {
        hipStream_t leaf = e->stream;
        for (; bi < B.blur.size() && B.blur[bi].o == o; bi++) {
            const GBlur &g = B.blur[bi];
            if (g.side) {
                const int k = o % ps_engine::kSides;
                PS_HIP(hipEventRecord(e->fork_ev[o], e->stream));
                PS_HIP(hipStreamWaitEvent(e->side[k], e->fork_ev[o], 0));
                forked[k] = true;
                leaf = e->side[k];
            }
            PS_HIP(launch_gblur(B, g, g.s == B.ref.S + 1 ? leaf : e->stream));
        }
        PS_HIP(launch_gextrema(B, B.ex[o], leaf));
    }
    for (int k = 0; k < ps_engine::kSides; k++)
        if (forked[k]) {
            PS_HIP(hipEventRecord(e->join_ev[k], e->side[k]));
            PS_HIP(hipStreamWaitEvent(e->stream, e->join_ev[k], 0));
        }
    return PS_OK;
}

// The whole batched extraction: counters, scale space, ordering (4096 cap), orientations,
// feature flattening and descriptors.
ps_status enqueue_group(ps_engine *e, BatchState &B) {
    TraceRange tr("ps.extract_batch");
    const uint32_t kMax = PS_MAX_FEATURES;
    PS_HIP(launch_clear_counters(as<uint32_t>(B.cnt), B.n, kErrWord, e->stream));
    PS_TRY(enqueue_scale_space_group(e, B));
    if (!B.full) return PS_OK;
    const uint32_t fcap = (uint32_t)std::min<size_t>(B.cap, kMax);
    PS_HIP(ps::launch_rank_group(table_at<ps::RankJob>(B, B.t_rank), B.n, ps::kCandCap, kMax, e->stream));
    uint32_t *off = as<uint32_t>(B.off);
    // keypoint counts: word 2 of each image's counters; feature counts: the caller's d_counts
    PS_HIP(ps::launch_describe_group(table_at<ps::DescJob>(B, B.t_desc), B.n, kMax, fcap, B.cfg.descriptor_kind, off,
                                     off + B.n + 1, as<uint32_t>(B.cnt) + 2, kCntWords, B.d_counts, e->stream));
    return PS_OK;
}

// The engine's batch state for these images and buffers (rebuilt when they change) with its
// launch shapes current.
ps_status prepare_batch(ps_engine *e, const uint8_t *d_pixels, const ps_image_desc *im, uint32_t n,
                        const ps_sift_config &cfg, ps_feature *d_feat, float *d_desc, size_t cap, uint32_t *d_counts,
                        std::vector<uint64_t> &key, uint64_t gen) {
    key.clear();
    key_push(key, (uint64_t)(uintptr_t)d_pixels);
    key_push(key, (uint64_t)(uintptr_t)d_feat);
    key_push(key, (uint64_t)(uintptr_t)d_desc);
    key_push(key, (uint64_t)(uintptr_t)d_counts);
    key_push(key, cap);
    key_push(key, n);
    uint64_t c0 = 0, c1 = 0;
    static_assert(sizeof(ps_sift_config) <= 16, "config key");
    std::memcpy(&c0, &cfg, std::min<size_t>(8, sizeof cfg));
    if (sizeof cfg > 8) std::memcpy(&c1, reinterpret_cast<const char *>(&cfg) + 8, sizeof cfg - 8);
    key_push(key, c0);
    key_push(key, c1);
    for (uint32_t i = 0; i < n; i++) {
        key_push(key, im[i].offset);
        key_push(key, ((uint64_t)im[i].height << 32) | im[i].width);
        key_push(key, im[i].row_stride);
    }
    if (!e->batch) e->batch = new BatchState();
    BatchState &B = *e->batch;
    if (B.key != key) {
        B.key.clear();
        PS_HIP(hipStreamSynchronize(e->stream));   // an earlier batch may still use the workspace
        // every counter word of the new images starts at zero, the error words included: the
        // words of an earlier batch's images (or fresh memory) mean nothing for these images
        PS_TRY(ensure(B.cnt, (size_t)std::max(n, 1u) * kCntWords * sizeof(uint32_t)));
        PS_HIP(launch_clear_counters(as<uint32_t>(B.cnt), std::max(n, 1u), kCntWords, e->stream));
        PS_TRY(build_batch(e, B, d_pixels, im, n, cfg, d_feat, d_desc, cap, d_counts));
        B.key = key;
    }
    if (B.shapes_gen != gen) PS_TRY(shape_batch(e, B, gen));
    key_push(key, gen);
    return PS_OK;
}

}  // namespace

namespace pshost {

void release_batch(ps_engine *e) {
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
    e->graph_key.clear();
    if (BatchState *B = e->batch) {
        DevBuf *bufs[] = {&B->pyr, &B->cand, &B->rank, &B->sorted, &B->cnt, &B->ori, &B->nori,
                          &B->fkp, &B->fj, &B->off, &B->tables};
        for (DevBuf *b : bufs)
            if (b->p) (void)hipFree(b->p);
        delete B;
        e->batch = nullptr;
    }
}

ps_status check_batch_err(ps_engine *e) {
    BatchState *B = e->batch;
    if (!B || !B->cnt.p || !B->n) return PS_OK;
    std::vector<uint32_t> words((size_t)B->n * kCntWords);
    PS_HIP(hipMemcpyAsync(words.data(), B->cnt.p, words.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    uint32_t err = 0;
    for (uint32_t i = 0; i < B->n; i++) err |= words[(size_t)i * kCntWords + kErrWord];
    if (err) {
        PS_HIP(launch_clear_counters(as<uint32_t>(B->cnt), B->n, kCntWords, e->stream));
        PS_HIP(hipStreamSynchronize(e->stream));
    }
    for (uint32_t i = 0; i < B->n; i++) {
        const std::string where = " (image " + std::to_string(i) + " of the batch)";
        PS_TRY(err_status(&words[(size_t)i * kCntWords], where.c_str()));
    }
    return PS_OK;
}

}  // namespace pshost

extern "C" {

ps_status ps_engine_extract_batch_u8(ps_engine *e, const uint8_t *d_pixels, const ps_image_desc *images,
                                     uint32_t n_images, const ps_sift_config *cfg, ps_feature *d_feat,
                                     float *d_desc, size_t cap, uint32_t *d_counts, uint32_t flags) {
    PS_TRY(check_fail_injection());
    if (!e || !cfg || (n_images && (!d_pixels || !images || !d_counts || (cap && (!d_feat || !d_desc)))))
        return set_err(PS_E_INVALID_ARG, "null argument");
    PS_TRY(validate_cfg(cfg, nullptr, 0));
    if (n_images == 0) return PS_OK;
    for (uint32_t i = 0; i < n_images; i++)
        if (images[i].row_stride < images[i].width) return set_err(PS_E_INVALID_ARG, "row_stride < width");
    PS_HIP(hipSetDevice(e->device));
    uint64_t gen;
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        gen = g_tile_gen;
    }
    std::vector<uint64_t> key;
    PS_TRY(prepare_batch(e, d_pixels, images, n_images, *cfg, d_feat, d_desc, cap, d_counts, key, gen));
    BatchState &B = *e->batch;
    const bool use_graph = !(flags & PS_BATCH_NO_GRAPH);
    if (!use_graph) return enqueue_group(e, B);
    if (e->graph_exec && key == e->graph_key) {
        PS_HIP(hipGraphLaunch(e->graph_exec, e->stream));
        return PS_OK;
    }
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
    e->graph_key.clear();
    PS_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    ps_status st = enqueue_group(e, B);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (st != PS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (ce != hipSuccess) return set_err(PS_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
    hipGraphExec_t ex = nullptr;
    hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        return set_err(PS_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
    }
    e->graph = g;
    e->graph_exec = ex;
    e->graph_key = key;
    PS_HIP(hipGraphLaunch(ex, e->stream));
    return PS_OK;
}

ps_status ps_engine_autotune_batch(ps_engine *e, const uint8_t *d_pixels, const ps_image_desc *images,
                                   uint32_t n_images, const ps_sift_config *cfg, uint32_t *n_tuned) {
    PS_TRY(check_fail_injection());
    if (!e || !cfg || (n_images && (!d_pixels || !images))) return set_err(PS_E_INVALID_ARG, "null argument");
    PS_TRY(validate_cfg(cfg, nullptr, 0));
    if (n_tuned) *n_tuned = 0;
    if (n_images == 0) return PS_OK;
    for (uint32_t i = 0; i < n_images; i++)
        if (images[i].row_stride < images[i].width) return set_err(PS_E_INVALID_ARG, "row_stride < width");
    PS_HIP(hipSetDevice(e->device));
    // a state of its own (no outputs): the engine's batch and graph are rebuilt by the next call
    release_batch(e);
    std::vector<uint64_t> key;
    PS_TRY(prepare_batch(e, d_pixels, images, n_images, *cfg, nullptr, nullptr, 0, nullptr, key, ~0ull - 1));
    BatchState &B = *e->batch;
    PS_HIP(launch_clear_counters(as<uint32_t>(B.cnt), B.n, kCntWords, e->stream));
    hipEvent_t a, b;
    PS_HIP(hipEventCreate(&a));
    PS_HIP(hipEventCreate(&b));
    std::map<GroupClass, int> result;
    // time every candidate of one launch (warm-up + 3 runs each, its inputs already produced by
    // the launches before it) and keep the fastest; the launch is left run with that shape
    auto tune = [&](auto &launch, const int *cands, int n_cands, int &chosen) -> ps_status {
        float best_ms = 1e30f;
        int best = cands[0];
        for (int c = 0; c < n_cands; c++) {
            chosen = cands[c];
            PS_TRY(upload_tables(e, B));
            PS_HIP(launch(e->stream));
            PS_HIP(hipEventRecord(a, e->stream));
            for (int rep = 0; rep < 3; rep++) PS_HIP(launch(e->stream));
            PS_HIP(hipEventRecord(b, e->stream));
            PS_HIP(hipEventSynchronize(b));
            float ms = 0.0f;
            PS_HIP(hipEventElapsedTime(&ms, a, b));
            if (ms < best_ms) {
                best_ms = ms;
                best = cands[c];
            }
        }
        chosen = best;
        PS_TRY(upload_tables(e, B));
        PS_HIP(launch(e->stream));
        return PS_OK;
    };
    ps_status st = PS_OK;
    size_t bi = 0;
    for (int o = 0; o < B.n_oct && st == PS_OK; o++) {
        // octave o's launches in dependency order (the first level belongs to octave 0)
        for (; bi < B.blur.size() && st == PS_OK && B.blur[bi].o == o; bi++) {
            GBlur &g = B.blur[bi];
            auto launch = [&](hipStream_t s) { return launch_gblur(B, g, s); };
            st = tune(launch, kBlurShapes, (int)(sizeof kBlurShapes / sizeof kBlurShapes[0]), g.shape);
            if (st == PS_OK) result[g.cls] = g.shape;
        }
        if (st != PS_OK) break;
        GExtrema &x = B.ex[o];
        const int rows[] = {16, 32};
        auto launch = [&](hipStream_t s) { return launch_gextrema(B, x, s); };
        st = tune(launch, rows, 2, x.rows);
        if (st == PS_OK) result[x.cls] = x.rows;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    release_batch(e);
    PS_TRY(st);
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        for (auto &kv : result) g_group_tiles[kv.first] = kv.second;
        g_tile_gen++;
    }
    if (n_tuned) *n_tuned = (uint32_t)result.size();
    return PS_OK;
}

ps_status ps_engine_autotune_blur(ps_engine *e, const uint32_t *hw, uint32_t n_sizes, const ps_sift_config *cfg,
                                  uint32_t *n_tuned) {
    PS_TRY(check_fail_injection());
    if (!e || !cfg || (n_sizes && !hw)) return set_err(PS_E_INVALID_ARG, "null argument");
    PS_TRY(validate_cfg(cfg, nullptr, 0));
    PS_HIP(hipSetDevice(e->device));
    struct Lvl {
        int r, mode, h, w, sh, sw;
        const ps::Taps *taps;
    };
    std::vector<Lvl> todo;
    std::vector<ps::PyrPlan> plans(n_sizes);
    std::map<std::tuple<int, int, int, int>, int> seen;
    size_t max_src = 16, max_dst = 16;
    for (uint32_t i = 0; i < n_sizes; i++) {
        ps::PyrPlan &p = plans[i];
        const int h = (int)hw[2 * i], w = (int)hw[2 * i + 1];
        if (!ps::make_plan(h, w, *cfg, p)) return set_err(PS_E_CONFIG, "unsupported blur radius");
        if (p.n_oct == 0) continue;
        auto add = [&](int r, int mode, int lh, int lw, int sh, int sw, const ps::Taps *t) {
            auto key = std::make_tuple(r, mode, size_class(lh), size_class(lw));
            if (seen.count(key)) return;
            seen[key] = 1;
            todo.push_back(Lvl{r, mode, lh, lw, sh, sw, t});
            max_src = std::max(max_src, (size_t)sh * sw * sizeof(float));
            max_dst = std::max(max_dst, (size_t)lh * lw * sizeof(float));
        };
        add(p.init_r, p.upsample ? ps::SRC_U8_UP : ps::SRC_U8, p.h[0], p.w[0], h, w, &p.init_taps);
        for (int o = 0; o < p.n_oct; o++)
            for (int s = 1; s <= p.S + 2; s++) {
                if (s == 1 && o > 0)
                    add(p.lvl_r[s], ps::SRC_DOWN2, p.h[o], p.w[o], p.h[o - 1], p.w[o - 1], &p.lvl_taps[s]);
                else
                    add(p.lvl_r[s], ps::SRC_F32, p.h[o], p.w[o], p.h[o], p.w[o], &p.lvl_taps[s]);
            }
    }
    PS_TRY(ensure(e->aux[AUX_PM_6], max_src));
    PS_TRY(ensure(e->aux[AUX_PM_7], 3 * max_dst));
    PS_HIP(hipMemsetAsync(e->aux[AUX_PM_6].p, 0, max_src, e->stream));
    hipEvent_t a, b;
    PS_HIP(hipEventCreate(&a));
    PS_HIP(hipEventCreate(&b));
    std::map<std::tuple<int, int, int, int>, int> result;
    for (const Lvl &L : todo) {
        ps::BlurArgs args;
        args.src = e->aux[AUX_PM_6].p;
        args.src_stride = (size_t)L.sw;
        args.src_h = L.sh;
        args.src_w = L.sw;
        args.h = L.h;
        args.w = L.w;
        float *d = as<float>(e->aux[AUX_PM_7]);
        args.dst_g = d;
        args.dst_d = nullptr;   // extraction writes no DoG planes
        args.dst_src = L.mode == ps::SRC_DOWN2 ? d + 2 * (max_dst / 4) : nullptr;
        float best_ms = 1e30f;
        int best_th = 32;
        for (int th : {16, 32, 64, ps::kBlurSlide | 64, ps::kBlurSlide | 128, ps::kBlurSlide | 192}) {
            PS_HIP(ps::launch_blur(L.r, L.mode, args, *L.taps, th, e->stream));   // warm-up
            PS_HIP(hipEventRecord(a, e->stream));
            for (int rep = 0; rep < 3; rep++) PS_HIP(ps::launch_blur(L.r, L.mode, args, *L.taps, th, e->stream));
            PS_HIP(hipEventRecord(b, e->stream));
            PS_HIP(hipEventSynchronize(b));
            float ms = 0.0f;
            PS_HIP(hipEventElapsedTime(&ms, a, b));
            if (ms < best_ms) {
                best_ms = ms;
                best_th = th;
            }
        }
        result[std::make_tuple(L.r, L.mode, size_class(L.h), size_class(L.w))] = best_th;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        for (auto &kv : result) g_tiles[kv.first] = kv.second;
        g_tile_gen++;
    }
    if (n_tuned) *n_tuned = (uint32_t)result.size();
    return PS_OK;
}

void ps_autotune_reset(void) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tiles.clear();
    g_group_tiles.clear();
    g_tile_gen++;
}

uint32_t ps_autotune_entries(void) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    return (uint32_t)g_tiles.size();
}

ps_status ps_autotune_entry(uint32_t i, int32_t out[5]) {
    if (!out) return set_err(PS_E_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(g_tile_mu);
    if (i >= g_tiles.size()) return set_err(PS_E_INVALID_ARG, "autotune entry index out of range");
    auto it = g_tiles.begin();
    std::advance(it, i);
    out[0] = std::get<0>(it->first);
    out[1] = std::get<1>(it->first);
    out[2] = std::get<2>(it->first);
    out[3] = std::get<3>(it->first);
    out[4] = it->second;
    return PS_OK;
}

}  // extern "C"
