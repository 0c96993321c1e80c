// Batched extraction of many uneven-size images and the blur-tile autotuner.
//
// Batching: the images of a batch are spread round-robin over up to kSlots child engines,
// each with its own stream and workspace, so small images run concurrently. The whole
// fork / per-image pipelines / join sequence is captured once into a HIP graph and
// replayed while the batch (image shapes, buffers, config) stays the same, which removes
// the host launch cost of the ~25 kernels per image.
//
// Autotuner: for every (radius, source mode, level size class) of the given image sizes,
// each blur launch shape (tile kernel with 16, 32 or 64 rows; sliding-window kernel with
// 64-, 128- or 192-row chunks) is timed on the device and the fastest is kept in a
// process-wide table that the pyramid launcher consults.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "host_common.h"

using namespace pshost;

namespace {

constexpr uint32_t kSlots = 8;

std::mutex g_tile_mu;
std::map<std::tuple<int, int, int, int>, int> g_tiles;   // (radius, mode, h class, w class) -> rows
uint64_t g_tile_gen = 0;                                  // bumped on every table change

int size_class(int v) { return (int)std::floor(2.0 * std::log2((double)std::max(v, 1))); }

void key_push(std::vector<uint64_t> &k, uint64_t v) { k.push_back(v); }

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

void release_batch(ps_engine *e) {
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
    e->graph_key.clear();
    for (hipEvent_t ev : e->batch_events) (void)hipEventDestroy(ev);
    e->batch_events.clear();
    for (ps_engine *c : e->children) engine_free(c);
    e->children.clear();
}

}  // namespace pshost

namespace {

ps_status ensure_children(ps_engine *e, uint32_t k) {
    while (e->children.size() < k) {
        ps_engine *c = nullptr;
        PS_TRY(engine_new(e->device, &c));
        e->children.push_back(c);
    }
    while (e->batch_events.size() < k + 1) {
        hipEvent_t ev;
        PS_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        e->batch_events.push_back(ev);
    }
    return PS_OK;
}

ps_status enqueue_batch(ps_engine *e, const uint8_t *d_pixels, const ps_image_desc *im, uint32_t n, uint32_t k,
                        const ps_sift_config &cfg, ps_feature *d_feat, float *d_desc, size_t cap,
                        uint32_t *d_counts) {
    PS_HIP(hipEventRecord(e->batch_events[0], e->stream));
    for (uint32_t c = 0; c < k; c++) PS_HIP(hipStreamWaitEvent(e->children[c]->stream, e->batch_events[0], 0));
    const uint32_t fcap = (uint32_t)std::min<size_t>(cap, PS_MAX_FEATURES);
    for (uint32_t i = 0; i < n; i++) {
        ps_engine *c = e->children[i % k];
        PS_TRY(enqueue_extract(c, d_pixels + im[i].offset, true, im[i].height, im[i].width, im[i].row_stride, cfg,
                               d_feat + (size_t)i * cap, d_desc + (size_t)i * cap * 128, fcap, d_counts + i));
    }
    for (uint32_t c = 0; c < k; c++) {
        PS_HIP(hipEventRecord(e->batch_events[c + 1], e->children[c]->stream));
        PS_HIP(hipStreamWaitEvent(e->stream, e->batch_events[c + 1], 0));
    }
    return PS_OK;
}

}  // namespace

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
    const uint32_t k = std::min(n_images, kSlots);
    PS_TRY(ensure_children(e, k));
    // size every child's workspace for its largest image before any capture
    for (uint32_t c = 0; c < k; c++) {
        size_t mh = 0, mw = 0, best = 0;
        for (uint32_t i = c; i < n_images; i += k) {
            size_t a = (size_t)images[i].height * images[i].width;
            if (a >= best) {
                best = a;
                mh = images[i].height;
                mw = images[i].width;
            }
        }
        // the pyramid size grows with both sides: take the per-side maxima too
        for (uint32_t i = c; i < n_images; i += k) {
            mh = std::max<size_t>(mh, images[i].height);
            mw = std::max<size_t>(mw, images[i].width);
        }
        PS_TRY(prepare_extract(e->children[c], mh, mw, *cfg));
    }
    const bool use_graph = !(flags & PS_BATCH_NO_GRAPH);
    if (!use_graph) {
        if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
        if (e->graph) (void)hipGraphDestroy(e->graph);
        e->graph_exec = nullptr;
        e->graph = nullptr;
        e->graph_key.clear();
        return enqueue_batch(e, d_pixels, images, n_images, k, *cfg, d_feat, d_desc, cap, d_counts);
    }
    std::vector<uint64_t> key;
    key_push(key, (uint64_t)(uintptr_t)d_pixels);
    key_push(key, (uint64_t)(uintptr_t)d_feat);
    key_push(key, (uint64_t)(uintptr_t)d_desc);
    key_push(key, (uint64_t)(uintptr_t)d_counts);
    key_push(key, cap);
    key_push(key, n_images);
    uint64_t c0 = 0, c1 = 0;
    std::memcpy(&c0, cfg, 8);
    std::memcpy(&c1, reinterpret_cast<const char *>(cfg) + 8, 8);
    key_push(key, c0);
    key_push(key, c1);
    for (uint32_t i = 0; i < n_images; i++) {
        key_push(key, images[i].offset);
        key_push(key, ((uint64_t)images[i].height << 32) | images[i].width);
        key_push(key, images[i].row_stride);
    }
    // tile choices are baked into the graph: include them
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        key_push(key, g_tile_gen);
    }
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
    ps_status st = enqueue_batch(e, d_pixels, images, n_images, k, *cfg, d_feat, d_desc, cap, d_counts);
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
