// Engine (stream + device workspace), pyramid planning and the C ABI of
// libps_sift.so (include/ps_sift.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "host_common.h"

#pragma clang fp contract(off)

namespace pshost {

thread_local std::string t_last_error;
thread_local bool t_fail_next = false;

ps_status set_err(ps_status s, const std::string &msg) {
    t_last_error = msg;
    return s;
}

}  // namespace pshost

using namespace pshost;

namespace {

// ------------------------------------------------------------------ host numerics
constexpr float kBaseSigma = 1.6f;          // pipeline.rs:10
constexpr float kAssumedInputSigma = 0.5f;  // pipeline.rs:13
constexpr float kPiF = 3.14159265358979323846f;

// f32::powi (compiler-rt __powisf2 binary exponentiation)
float powi_rs(float a, int b) {
    bool recip = b < 0;
    float r = 1.0f;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0f / r : r;
}

// gaussian_kernel_1d, pipeline.rs:275-296 (host, glibc expf -> bit-identical weights)
int gaussian_taps(float sigma, ps::Taps &t) {
    float cs = std::fmax(sigma, 1e-3f);
    long radius = (long)std::ceil(3.0f * cs);
    if (radius < 1) radius = 1;
    if (2 * radius + 1 > ps::kMaxTaps) return -1;
    float two_sigma_sq = 2.0f * cs * cs;
    float norm = 1.0f / (std::sqrt(2.0f * kPiF) * cs);
    int n = (int)(2 * radius + 1);
    std::memset(&t, 0, sizeof(t));
    for (int i = 0; i < n; i++) {
        float x = (float)(i - radius);
        t.w[i] = norm * expf(-((x * x) / two_sigma_sq));
    }
    float sum = 0.0f;
    for (int i = 0; i < n; i++) sum += t.w[i];
    if (sum > 0.0f) {
        float inv = 1.0f / sum;
        for (int i = 0; i < n; i++) t.w[i] *= inv;
    }
    return (int)radius;
}

}  // namespace

namespace ps {

// build_scale_space_pyramid_with_options geometry and sigma schedule (pipeline.rs:550-638)
bool make_plan(int h, int w, const ps_sift_config &cfg, PyrPlan &p) {
    p = PyrPlan();
    int S = (int)cfg.scales;
    p.S = S;
    p.upsample = cfg.upsample_first_octave ? 1 : 0;
    p.in_h = h;
    p.in_w = w;
    if (h == 0 || w == 0 || cfg.octaves == 0 || S == 0) return true;
    p.k = powf(2.0f, 1.0f / (float)S);
    p.log2k = std::log2((double)p.k);
    int bh = p.upsample ? 2 * h : h, bw = p.upsample ? 2 * w : w;
    float sigma_in = p.upsample ? 2.0f * kAssumedInputSigma : kAssumedInputSigma;
    float init_sigma = std::sqrt(std::fmax(kBaseSigma * kBaseSigma - sigma_in * sigma_in, 1e-4f));
    p.init_r = gaussian_taps(init_sigma, p.init_taps);
    if (p.init_r < 0) return false;
    float prev = kBaseSigma;
    for (int s = 1; s <= S + 2; s++) {
        float target = kBaseSigma * powi_rs(p.k, s);
        float delta = std::sqrt(std::fmax(target * target - prev * prev, 1e-4f));
        p.lvl_r[s] = gaussian_taps(delta, p.lvl_taps[s]);
        if (p.lvl_r[s] < 0 || p.lvl_r[s] > 12) return false;
        prev = target;
    }
    float base_mult = p.upsample ? 0.5f : 1.0f;
    size_t off = 0;
    int noct = std::min((int)cfg.octaves, kMaxOct);
    for (int o = 0; o < noct; o++) {
        if (bh < 3 || bw < 3) break;
        p.h[o] = bh;
        p.w[o] = bw;
        p.off[o] = off;
        off += (size_t)(2 * S + 5) * bh * bw;
        p.scale_factor[o] = base_mult * (float)(1u << o);
        for (int s = 1; s <= S; s++) p.eff_scale[o][s] = kBaseSigma * p.scale_factor[o] * powi_rs(p.k, s);
        p.n_oct = o + 1;
        bh /= 2;
        bw /= 2;
    }
    p.total_floats = off;
    return true;
}

}  // namespace ps

namespace pshost {

ps_status ensure(DevBuf &b, size_t bytes, bool *fresh) {
    if (fresh) *fresh = false;
    if (bytes == 0) bytes = 16;
    if (b.n >= bytes) return PS_OK;
    if (b.p) PS_HIP(hipFree(b.p));
    b.p = nullptr;
    b.n = 0;
    bytes = (bytes + 255) & ~(size_t)255;
    PS_HIP(hipMalloc(&b.p, bytes));
    b.n = bytes;
    // (a new allocation can have the address of the one just freed: callers that must clear
    // fresh memory ask for this flag instead of comparing pointers)
    if (fresh) *fresh = true;
    return PS_OK;
}

// The engine's counter words; a fresh buffer is cleared once so that the error word, which
// only ps_engine_check / clear_err_word reset, starts at zero.
ps_status ensure_cnt(ps_engine *e) {
    bool fresh = false;
    PS_TRY(ensure(e->cnt, 64, &fresh));
    if (fresh) PS_HIP(ps::launch_fill_u32(as<uint32_t>(e->cnt), 0u, 16, e->stream));
    return PS_OK;
}

std::mutex g_mu;
bool g_inited = false;
int g_device = 0;
std::vector<ps_engine *> g_pool;

ps_status check_fail_injection() {
    if (t_fail_next) {
        t_fail_next = false;
        return set_err(PS_E_INJECTED, "failure injected by ps_debug_fail_next");
    }
    return PS_OK;
}

ps_status init_locked(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return set_err(PS_E_NO_DEVICE, "no HIP device available");
    if (device < 0) device = 0;
    if (device >= n) return set_err(PS_E_INVALID_ARG, "device index out of range");
    hipDeviceProp_t prop;
    PS_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_err(PS_E_NO_DEVICE, std::string("libps_sift.so is built for gfx950, device is ") + prop.gcnArchName);
    PS_HIP(hipSetDevice(device));
    PS_TRY(static_scratch_setup(device, nullptr));
    g_device = device;
    g_inited = true;
    return PS_OK;
}

static void release_streams(ps_engine *e) {
    for (hipEvent_t ev : e->fork_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : e->join_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t s : e->side)
        if (s) (void)hipStreamDestroy(s);
    if (e->stream) (void)hipStreamDestroy(e->stream);
}

ps_status engine_new(int device, ps_engine **out) {
    auto *e = new ps_engine();
    e->device = device;
    hipError_t err = hipSetDevice(device);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    for (int i = 0; i < ps_engine::kSides && err == hipSuccess; i++) {
        err = hipStreamCreateWithFlags(&e->side[i], hipStreamNonBlocking);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&e->join_ev[i], hipEventDisableTiming);
    }
    for (int i = 0; i < ps::kMaxOct && err == hipSuccess; i++)
        err = hipEventCreateWithFlags(&e->fork_ev[i], hipEventDisableTiming);
    if (err != hipSuccess) {
        release_streams(e);
        delete e;
        return set_err(PS_E_DEVICE, std::string("engine stream: ") + hipGetErrorString(err));
    }
    ps_status st = static_scratch_setup(device, nullptr);
    if (st != PS_OK) {
        release_streams(e);
        delete e;
        return st;
    }
    *out = e;
    return PS_OK;
}

constexpr int kEvPerCall = 7;

void prof_release(ps_engine *e) {
    for (hipEvent_t ev : e->prof_ev) (void)hipEventDestroy(ev);
    e->prof_ev.clear();
    e->prof_max = e->prof_ext = e->prof_mat = 0;
    e->prof_on = false;
}

// records stage boundary `slot` of the current extract (0..4) or match (5..6) call on `st`
// (the engine stream unless given)
void prof_mark(ps_engine *e, int slot, hipStream_t st = nullptr) {
    if (!e->prof_on) return;
    uint32_t call = slot < 5 ? e->prof_ext : e->prof_mat;
    if (call >= e->prof_max) return;
    (void)hipEventRecord(e->prof_ev[(size_t)call * kEvPerCall + slot], st ? st : e->stream);
}

void engine_free(ps_engine *e) {
    if (!e) return;
    prof_release(e);
    (void)hipSetDevice(e->device);
    DevBuf *bufs[] = {&e->pyr, &e->cand, &e->rank, &e->sorted, &e->ori, &e->nori, &e->feat_kp, &e->feat_j, &e->cnt,
                      &e->img, &e->feat, &e->desc, &e->kpout, &e->qf, &e->qd, &e->tf, &e->td,
                      &e->mcnt, &e->mout, &e->part, &e->part2, &e->bidx, &e->bdist, &e->sdist, &e->ridx, &e->ridx_d, &e->fin};
    for (DevBuf *b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (DevBuf &b : e->aux)
        if (b.p) (void)hipFree(b.p);
    release_batch(e);
    release_streams(e);
    delete e;
}

Borrow::Borrow() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_inited) {
        st = init_locked(-1);
        if (st != PS_OK) return;
    }
    if (!g_pool.empty()) {
        e = g_pool.back();
        g_pool.pop_back();
    } else {
        st = engine_new(g_device, &e);
    }
}

Borrow::~Borrow() {
    if (!e) return;
    std::lock_guard<std::mutex> lk(g_mu);
    g_pool.push_back(e);
}

ps_status validate_cfg(const ps_sift_config *c, char *msg, size_t cap) {
    char buf[256];
    ps_status s = PS_OK;
    if (!c) return set_err(PS_E_INVALID_ARG, "null config");
    if (c->octaves < 1 || c->octaves > 8) {   // pipeline.rs:81-86
        snprintf(buf, sizeof buf, "Invalid `octaves`=%u: must be between 1 and 8", c->octaves);
        s = PS_E_CONFIG;
    } else if (c->scales < 3 || c->scales > 10) {   // pipeline.rs:87-92
        snprintf(buf, sizeof buf, "Invalid `scales`=%u: must be between 3 and 10 for 3D extrema extraction",
                 c->scales);
        s = PS_E_CONFIG;
    } else if (!std::isfinite(c->contrast_threshold) || c->contrast_threshold <= 0.0f ||
               c->contrast_threshold > 1.0f) {   // pipeline.rs:93-101
        snprintf(buf, sizeof buf, "Invalid `contrast_threshold`=%g: must be finite and in (0.0, 1.0]",
                 (double)c->contrast_threshold);
        s = PS_E_CONFIG;
    } else if (c->descriptor_kind > 1) {
        snprintf(buf, sizeof buf, "Invalid `descriptor_kind`=%u", (unsigned)c->descriptor_kind);
        s = PS_E_CONFIG;
    }
    if (s != PS_OK) {
        if (msg && cap) {
            std::strncpy(msg, buf, cap - 1);
            msg[cap - 1] = 0;
        }
        return set_err(s, buf);
    }
    if (msg && cap) msg[0] = 0;
    return PS_OK;
}

ps_status validate_opt(const ps_match_options *o, char *msg, size_t cap) {
    if (!o) return set_err(PS_E_INVALID_ARG, "null match options");
    if (!std::isfinite(o->ratio) || o->ratio <= 0.0f || o->ratio > 1.0f) {   // pipeline.rs:157-162
        char buf[160];
        snprintf(buf, sizeof buf, "Invalid `ratio`=%g: must be finite and in (0.0, 1.0]", (double)o->ratio);
        if (msg && cap) {
            std::strncpy(msg, buf, cap - 1);
            msg[cap - 1] = 0;
        }
        return set_err(PS_E_CONFIG, buf);
    }
    if (o->precision > PS_PRECISION_FP8) {
        char buf[160];
        snprintf(buf, sizeof buf, "Invalid `precision`=%u: must be 0 (FP32), 1 (BF16) or 2 (FP8)",
                 (unsigned)o->precision);
        if (msg && cap) {
            std::strncpy(msg, buf, cap - 1);
            msg[cap - 1] = 0;
        }
        return set_err(PS_E_CONFIG, buf);
    }
    if (msg && cap) msg[0] = 0;
    return PS_OK;
}

ps::PyrView pyr_view(ps_engine *e) {
    ps::PyrView v;
    std::memset(&v, 0, sizeof(v));
    v.base = as<float>(e->pyr);
    v.n_oct = e->plan.n_oct;
    v.S = e->plan.S;
    for (int o = 0; o < ps::kMaxOct; o++) {
        v.off[o] = e->plan.off[o];
        v.h[o] = e->plan.h[o];
        v.w[o] = e->plan.w[o];
        v.scale_factor[o] = e->plan.scale_factor[o];
    }
    return v;
}

// Gaussian + DoG pyramid (pipeline.rs:550-638) into e->pyr (src is device memory) and, with
// `extrema`, every octave's extrema appended to e->cand (cnt[0] candidates, cnt[kErrWord] error bits).
// Octave o+1 depends only on level S of octave o, so the levels 1..S of all octaves form one
// dependent chain: it stays on the engine stream (a cross-stream event costs ~10 us of
// latency), and each octave's leaf work -- the levels above S and the extrema pass -- is forked
// to a side stream once level S is written (octave 0's one octave later when octave 1 is small,
// leaf_delay); every side stream is joined back into the engine stream before this returns.
// Candidate order is fixed later by the ordering stage, so it does not matter which stream
// appends first. No fork happens while the engine stream is being
// captured: a capturing caller (the batched path) already overlaps whole images on separate
// streams.

ps_status enqueue_scale_space(ps_engine *e, const void *src, bool src_u8, size_t h, size_t w, size_t stride,
                              const ps_sift_config &cfg, bool extrema, uint32_t *clear = nullptr, int n_clear = 0) {
    TraceRange tr("ps.scale_space");
    ps::PyrPlan &p = e->plan;
    if (!ps::make_plan((int)h, (int)w, cfg, p)) return set_err(PS_E_CONFIG, "unsupported blur radius");
    PS_TRY(ensure(e->pyr, p.total_floats * sizeof(float)));
    if (extrema) PS_TRY(ensure(e->cand, (size_t)ps::kCandCap * sizeof(ps::Cand)));
    if (p.n_oct == 0) {
        if (clear) PS_HIP(ps::launch_fill_u32(clear, 0u, (size_t)n_clear, e->stream));
        prof_mark(e, 1);
        return PS_OK;
    }
    float *P = as<float>(e->pyr);
    uint32_t *cnt = as<uint32_t>(e->cnt);
    auto level = [&](int o, int s, hipStream_t st) -> hipError_t {
        ps::BlurArgs b;
        b.h = p.h[o];
        b.w = p.w[o];
        b.dst_g = P + p.gauss(o, s);
        // extraction forms DoG from the Gaussian planes inside the extrema kernel; only the
        // pyramid dump stores the DoG planes
        b.dst_d = extrema ? nullptr : P + p.dog(o, s - 1);
        b.dst_src = nullptr;
        int m = ps::SRC_F32;
        if (s == 1 && o > 0) {   // octave base = downsample_2x(gaussians[S]) (pipeline.rs:624-626)
            b.src = P + p.gauss(o - 1, p.S);
            b.src_stride = (size_t)p.w[o - 1];
            b.src_h = p.h[o - 1];
            b.src_w = p.w[o - 1];
            b.dst_src = P + p.gauss(o, 0);
            m = ps::SRC_DOWN2;
        } else {
            b.src = P + p.gauss(o, s - 1);
            b.src_stride = (size_t)p.w[o];
            b.src_h = p.h[o];
            b.src_w = p.w[o];
        }
        return ps::launch_blur(p.lvl_r[s], m, b, p.lvl_taps[s], blur_tile(p.lvl_r[s], m, b.h, b.w), st);
    };
    auto octave_extrema = [&](int o, hipStream_t st) -> hipError_t {
        return ps::launch_extrema(P + p.gauss(o, 0), p.h[o], p.w[o], p.S, cfg.contrast_threshold, p.scale_factor[o],
                                  p.eff_scale[o], p.log2k, o, as<ps::Cand>(e->cand), cnt + 0, ps::kCandCap, cnt + kErrWord,
                                  p.lvl_taps[p.S + 2], p.lvl_r[p.S + 2], st);
    };
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    PS_HIP(hipStreamIsCapturing(e->stream, &cap));
    const bool fork = cap == hipStreamCaptureStatusNone;
    ps::BlurArgs a;
    a.src = src;
    a.src_stride = stride;
    a.src_h = (int)h;
    a.src_w = (int)w;
    a.h = p.h[0];
    a.w = p.w[0];
    a.dst_g = P + p.gauss(0, 0);
    a.dst_d = nullptr;
    a.dst_src = nullptr;
    a.clear = clear;
    a.n_clear = n_clear;
    int mode = src_u8 ? (p.upsample ? ps::SRC_U8_UP : ps::SRC_U8) : (p.upsample ? ps::SRC_F32_UP : ps::SRC_F32);
    PS_HIP(ps::launch_blur(p.init_r, mode, a, p.init_taps, blur_tile(p.init_r, mode, a.h, a.w), e->stream));
    // octave o's leaf work -- its remaining levels and its extrema -- on a side stream beside the
    // chain (the last octave's on the engine stream: nothing follows it there)
    auto leaf = [&](int o) -> ps_status {
        const bool side = fork && o + 1 < p.n_oct;
        hipStream_t st = side ? e->side[o % ps_engine::kSides] : e->stream;
        if (side) {
            PS_HIP(hipEventRecord(e->fork_ev[o], e->stream));
            PS_HIP(hipStreamWaitEvent(st, e->fork_ev[o], 0));
        }
        // level S+2 feeds only the upper DoG plane of slice-S candidates: the extraction path
        // leaves it to the extrema kernel, which forms its few needed values from level S+1
        for (int s = p.S + 1; s <= (extrema ? p.S + 1 : p.S + 2); s++) PS_HIP(level(o, s, st));
        if (extrema) PS_HIP(octave_extrema(o, st));
        return PS_OK;
    };
    bool leaf_done[ps::kMaxOct] = {};
    for (int o = 0; o < p.n_oct; o++) {
        // the chain: levels 1..S of every octave on the engine stream (octave o+1 starts from
        // level S of octave o), so it crosses no stream
        for (int s = 1; s <= p.S; s++) PS_HIP(level(o, s, e->stream));
        if (o == 0) prof_mark(e, 1);
        for (int j = 0; j <= o; j++)
            if (!leaf_done[j] && std::min(j + leaf_delay(j, (long)p.h[1] * p.w[1]), p.n_oct - 1) <= o) {
                PS_TRY(leaf(j));
                leaf_done[j] = true;
            }
    }
    if (fork)
        for (int i = 0; i < ps_engine::kSides && i + 1 < p.n_oct; i++) {
            PS_HIP(hipEventRecord(e->join_ev[i], e->side[i]));
            PS_HIP(hipStreamWaitEvent(e->stream, e->join_ev[i], 0));
        }
    return PS_OK;
}

// pyramid -> extrema -> ordered (and, for extraction, capped) candidates in e->sorted;
// counters: cnt[0] candidates, cnt[2] ordered keypoints, cnt[3] features, cnt[kErrWord] error bits
ps_status enqueue_keypoints(ps_engine *e, const void *src, bool src_u8, size_t h, size_t w, size_t stride,
                            const ps_sift_config &cfg, uint32_t limit) {
    PS_TRY(ensure_cnt(e));
    uint32_t *cnt = as<uint32_t>(e->cnt);
    PS_TRY(ensure(e->rank, ps::rank_scratch_bytes(ps::kCandCap)));
    PS_TRY(ensure(e->sorted, (size_t)limit * sizeof(ps::Cand)));
    prof_mark(e, 0);
    // every counter but the error word is zeroed by the first blur launch (or here when the image
    // is too small for an octave)
    PS_TRY(enqueue_scale_space(e, src, src_u8, h, w, stride, cfg, true, cnt, (int)kErrWord));
    prof_mark(e, 2);
    const ps::PyrPlan &p = e->plan;
    const bool packed = p.n_oct <= 8 && p.S < 16 && (p.n_oct == 0 || (p.h[0] <= 4096 && p.w[0] <= 4096));
    PS_HIP(ps::launch_rank(as<ps::Cand>(e->cand), cnt + 0, ps::kCandCap, limit, as<ps::Cand>(e->sorted), cnt + 2,
                           e->rank.p, packed, e->stream));
    prof_mark(e, 3);
    return PS_OK;
}

ps_status enqueue_extract(ps_engine *e, const void *src, bool src_u8, size_t h, size_t w, size_t stride,
                          const ps_sift_config &cfg, ps_feature *d_feat, float *d_desc, uint32_t cap,
                          uint32_t *d_count) {
    TraceRange tr("ps.extract");
    const uint32_t kMax = PS_MAX_FEATURES;
    PS_TRY(enqueue_keypoints(e, src, src_u8, h, w, stride, cfg, kMax));
    PS_TRY(ensure(e->ori, kMax * 4 * sizeof(float)));
    PS_TRY(ensure(e->nori, kMax * sizeof(uint32_t)));
    PS_TRY(ensure(e->feat_kp, kMax * sizeof(uint32_t)));
    PS_TRY(ensure(e->feat_j, kMax * sizeof(uint32_t)));
    uint32_t *cnt = as<uint32_t>(e->cnt);
    ps::PyrView pv = pyr_view(e);
    uint32_t fcap = std::min(cap, kMax);
    PS_HIP(ps::launch_orientations(pv, as<ps::Cand>(e->sorted), cnt + 2, kMax, as<float>(e->ori),
                                   as<uint32_t>(e->nori), e->stream));
    PS_HIP(ps::launch_feature_scan(cnt + 2, as<uint32_t>(e->nori), kMax, fcap, as<uint32_t>(e->feat_kp),
                                   as<uint32_t>(e->feat_j), d_count, e->stream));
    PS_HIP(ps::launch_descriptors(pv, as<ps::Cand>(e->sorted), as<float>(e->ori), as<uint32_t>(e->feat_kp),
                                  as<uint32_t>(e->feat_j), d_count, fcap, cfg.descriptor_kind, d_feat, d_desc,
                                  e->stream));
    prof_mark(e, 4);
    if (e->prof_on) e->prof_ext++;
    return PS_OK;
}

// best/second/index per query into e->bidx / bdist / sdist (or the given outputs)
// Grows every extraction buffer to what an h x w image needs (no allocation happens in a
// later enqueue_extract of that size, so the sequence can be captured into a graph).
ps_status prepare_extract(ps_engine *e, size_t h, size_t w, const ps_sift_config &cfg) {
    ps::PyrPlan p;
    if (!ps::make_plan((int)h, (int)w, cfg, p)) return set_err(PS_E_CONFIG, "unsupported blur radius");
    const uint32_t kMax = PS_MAX_FEATURES;
    PS_TRY(ensure(e->pyr, p.total_floats * sizeof(float)));
    PS_TRY(ensure_cnt(e));
    PS_TRY(ensure(e->cand, (size_t)ps::kCandCap * sizeof(ps::Cand)));
    PS_TRY(ensure(e->rank, ps::rank_scratch_bytes(ps::kCandCap)));
    PS_TRY(ensure(e->sorted, (size_t)kMax * sizeof(ps::Cand)));
    PS_TRY(ensure(e->ori, kMax * 4 * sizeof(float)));
    PS_TRY(ensure(e->nori, kMax * sizeof(uint32_t)));
    PS_TRY(ensure(e->feat_kp, kMax * sizeof(uint32_t)));
    PS_TRY(ensure(e->feat_j, kMax * sizeof(uint32_t)));
    return PS_OK;
}

ps_status enqueue_nn(ps_engine *e, const float *qd, const uint32_t *nq, size_t nqm, const float *td,
                     const uint32_t *nt, size_t ntm, DevBuf &scratch, uint32_t *bidx, float *bd, float *sd,
                     int precision = PS_PRECISION_FP32, float *approx = nullptr) {
    if (precision == PS_PRECISION_BF16 || precision == PS_PRECISION_FP8) {
        PS_TRY(ensure(scratch, ps::nn_precision_scratch_bytes(precision, nqm, ntm)));
        PS_HIP(ps::launch_nn_precision(precision, qd, nq, nqm, td, nt, ntm, scratch.p, bidx, bd, sd, approx,
                                       e->stream));
        return PS_OK;
    }
    bool fresh = false;
    PS_TRY(ensure(scratch, ps::nn_mfma_scratch_bytes(nqm, ntm), &fresh));
    if (++e->nn_gen == 0) e->nn_gen = 1;
    // fresh memory: clear the two norm-bound words at its head (later calls clear them in turn)
    if (fresh) PS_HIP(ps::launch_fill_u32(static_cast<uint32_t *>(scratch.p), 0u, 4, e->stream));
    PS_HIP(ps::launch_nn_mfma(qd, nq, nqm, td, nt, ntm, scratch.p, bidx, bd, sd, e->nn_gen, e->stream));
    return PS_OK;
}

ps_status enqueue_match(ps_engine *e, const ps_feature *qf, const float *qd, const uint32_t *nq, size_t nq_max,
                        const ps_feature *tf, const float *td, const uint32_t *nt, size_t nt_max,
                        const ps_match_options &opt, ps_feature_match *out, uint32_t *out_count) {
    TraceRange tr("ps.match");
    size_t nqm = std::max<size_t>(nq_max, 1), ntm = std::max<size_t>(nt_max, 1);
    prof_mark(e, 5);
    PS_TRY(ensure(e->bidx, nqm * sizeof(uint32_t)));
    PS_TRY(ensure(e->bdist, nqm * sizeof(float)));
    PS_TRY(ensure(e->sdist, nqm * sizeof(float)));
    if (nqm > 262144) return set_err(PS_E_INVALID_ARG, "at most 262144 query features per match call");
    PS_TRY(enqueue_nn(e, qd, nq, nqm, td, nt, ntm, e->part, as<uint32_t>(e->bidx), as<float>(e->bdist),
                      as<float>(e->sdist), opt.precision));
    uint32_t *rev = nullptr;
    if (opt.mutual) {
        // reverse pass (pipeline.rs:1815-1840): nearest query of every target, first index on ties
        PS_TRY(ensure(e->ridx, ntm * sizeof(uint32_t)));
        PS_TRY(ensure(e->ridx_d, 2 * ntm * sizeof(float)));
        PS_TRY(enqueue_nn(e, td, nt, ntm, qd, nq, nqm, e->part2, as<uint32_t>(e->ridx), as<float>(e->ridx_d),
                          as<float>(e->ridx_d) + ntm, opt.precision));
        rev = as<uint32_t>(e->ridx);
    }
    PS_TRY(ensure(e->fin, ps::match_finish_scratch_bytes(nqm)));
    PS_HIP(ps::launch_match_finish(qf, tf, nq, nqm, nt, as<uint32_t>(e->bidx), as<float>(e->bdist),
                                   as<float>(e->sdist), rev, opt.ratio, opt.mutual ? 1 : 0, out, out_count, e->fin.p,
                                   e->stream));
    prof_mark(e, 6);
    if (e->prof_on) e->prof_mat++;
    return PS_OK;
}

// The error word (counter word kErrWord) accumulates the error bits of every extraction on the
// engine since the last check; a check reads and clears it.
ps_status clear_err_word(ps_engine *e) {
    PS_TRY(ensure_cnt(e));
    PS_HIP(ps::launch_fill_u32(as<uint32_t>(e->cnt) + kErrWord, 0u, 1, e->stream));
    return PS_OK;
}

ps_status err_status(const uint32_t *words, const char *where) {
    if (!(words[kErrWord] & ps::kErrCandOverflow)) return PS_OK;
    char buf[160];
    snprintf(buf, sizeof buf, "more than 2^20 extremum candidates in one image%s (%u found)", where, words[0]);
    return set_err(PS_E_CAPACITY, buf);
}

ps_status check_err_word(ps_engine *e) {
    uint32_t words[16] = {};
    PS_HIP(hipMemcpyAsync(words, as<uint32_t>(e->cnt), sizeof words, hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    if (words[kErrWord]) {
        PS_TRY(clear_err_word(e));
        PS_HIP(hipStreamSynchronize(e->stream));
    }
    return err_status(words, "");
}

template <typename T>
ps_status extract_host(const T *img, size_t h, size_t w, size_t row_stride, const ps_sift_config *cfg,
                       ps_feature *out_feat, float *out_desc, size_t cap, size_t *out_count) {
    PS_TRY(check_fail_injection());
    if (!img || !cfg || !out_count || (cap && (!out_feat || !out_desc)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (row_stride < w) return set_err(PS_E_INVALID_ARG, "row_stride < width");
    PS_TRY(validate_cfg(cfg, nullptr, 0));
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    const bool u8 = sizeof(T) == 1;
    PS_TRY(ensure(e->img, h * w * sizeof(T)));
    if (h && w)
        PS_HIP(hipMemcpy2DAsync(e->img.p, w * sizeof(T), img, row_stride * sizeof(T), w * sizeof(T), h,
                                hipMemcpyHostToDevice, e->stream));
    PS_TRY(ensure(e->feat, PS_MAX_FEATURES * sizeof(ps_feature)));
    PS_TRY(ensure(e->desc, PS_MAX_FEATURES * 128 * sizeof(float)));
    PS_TRY(ensure(e->mcnt, 64));
    uint32_t *dcount = as<uint32_t>(e->mcnt);
    PS_TRY(clear_err_word(e));
    PS_TRY(enqueue_extract(e, e->img.p, u8, h, w, w, *cfg, as<ps_feature>(e->feat), as<float>(e->desc),
                           PS_MAX_FEATURES, dcount));
    PS_HIP(hipMemcpyAsync(&e->host_counts[0], dcount, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    PS_TRY(check_err_word(e));
    size_t n = e->host_counts[0];
    *out_count = n;
    if (n > cap) return set_err(PS_E_CAPACITY, "output capacity smaller than the feature count");
    if (n) {
        PS_HIP(hipMemcpyAsync(out_feat, e->feat.p, n * sizeof(ps_feature), hipMemcpyDeviceToHost, e->stream));
        PS_HIP(hipMemcpyAsync(out_desc, e->desc.p, n * 128 * sizeof(float), hipMemcpyDeviceToHost, e->stream));
        PS_HIP(hipStreamSynchronize(e->stream));
    }
    return PS_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

ps_status ps_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_inited && (device < 0 || device == g_device)) return PS_OK;
    return init_locked(device);
}

const char *ps_status_string(ps_status s) {
    switch (s) {
    case PS_OK: return "ok";
    case PS_E_INVALID_ARG: return "invalid argument";
    case PS_E_CONFIG: return "invalid configuration";
    case PS_E_DEVICE: return "device error";
    case PS_E_CAPACITY: return "insufficient capacity";
    case PS_E_NO_DEVICE: return "no usable device";
    case PS_E_OUT_OF_MEMORY: return "device out of memory";
    case PS_E_INJECTED: return "injected failure";
    case PS_E_IO: return "catalog file error";
    case PS_E_FORMAT: return "malformed catalog";
    default: return "unknown status";
    }
}

const char *ps_last_error(void) { return t_last_error.c_str(); }

ps_status ps_validate_sift_config(const ps_sift_config *cfg, char *msg, size_t cap) {
    return validate_cfg(cfg, msg, cap);
}

ps_status ps_validate_match_options(const ps_match_options *opt, char *msg, size_t cap) {
    return validate_opt(opt, msg, cap);
}

void ps_debug_fail_next(void) { t_fail_next = true; }

ps_status ps_detect_and_compute(const float *img, size_t h, size_t w, size_t row_stride, const ps_sift_config *cfg,
                                ps_feature *out_feat, float *out_desc, size_t cap, size_t *out_count) {
    return extract_host<float>(img, h, w, row_stride, cfg, out_feat, out_desc, cap, out_count);
}

ps_status ps_detect_and_compute_u8(const uint8_t *img, size_t h, size_t w, size_t row_stride,
                                   const ps_sift_config *cfg, ps_feature *out_feat, float *out_desc, size_t cap,
                                   size_t *out_count) {
    return extract_host<uint8_t>(img, h, w, row_stride, cfg, out_feat, out_desc, cap, out_count);
}

ps_status ps_detect_keypoints(const float *img, size_t h, size_t w, size_t row_stride, const ps_sift_config *cfg,
                              ps_keypoint *out, size_t cap, size_t *out_count) {
    PS_TRY(check_fail_injection());
    if (!img || !cfg || !out_count || (cap && !out)) return set_err(PS_E_INVALID_ARG, "null argument");
    if (row_stride < w) return set_err(PS_E_INVALID_ARG, "row_stride < width");
    PS_TRY(validate_cfg(cfg, nullptr, 0));
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->img, h * w * sizeof(float)));
    if (h && w)
        PS_HIP(hipMemcpy2DAsync(e->img.p, w * sizeof(float), img, row_stride * sizeof(float), w * sizeof(float), h,
                                hipMemcpyHostToDevice, e->stream));
    PS_TRY(clear_err_word(e));
    PS_TRY(enqueue_keypoints(e, e->img.p, false, h, w, w, *cfg, ps::kCandCap));
    PS_HIP(hipMemcpyAsync(&e->host_counts[0], as<uint32_t>(e->cnt) + 2, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          e->stream));
    PS_TRY(check_err_word(e));
    size_t n = e->host_counts[0];
    *out_count = n;
    if (n > cap) return set_err(PS_E_CAPACITY, "output capacity smaller than the keypoint count");
    if (n) {
        PS_TRY(ensure(e->kpout, n * sizeof(ps_keypoint)));
        PS_HIP(ps::launch_cand_to_keypoints(as<ps::Cand>(e->sorted), as<uint32_t>(e->cnt) + 2, (uint32_t)n,
                                            as<ps_keypoint>(e->kpout), e->stream));
        PS_HIP(hipMemcpyAsync(out, e->kpout.p, n * sizeof(ps_keypoint), hipMemcpyDeviceToHost, e->stream));
        PS_HIP(hipStreamSynchronize(e->stream));
    }
    return PS_OK;
}

// Rust `f32 as usize`: saturating, NaN -> 0
static size_t f32_to_usize(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551615.0f) return SIZE_MAX;
    return (size_t)v;
}

ps_status ps_compute_descriptors(const float *img, size_t h, size_t w, size_t row_stride, const ps_keypoint *kps,
                                 size_t n, uint8_t kind, ps_feature *out_feat, float *out_desc, size_t cap,
                                 size_t *out_count) {
    PS_TRY(check_fail_injection());
    if (!img || !out_count || (n && !kps) || (cap && (!out_feat || !out_desc)))
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (row_stride < w) return set_err(PS_E_INVALID_ARG, "row_stride < width");
    if (kind != PS_DESCRIPTOR_SIFT && kind != PS_DESCRIPTOR_ROOTSIFT)
        return set_err(PS_E_INVALID_ARG, "descriptor kind must be PS_DESCRIPTOR_SIFT or PS_DESCRIPTOR_ROOTSIFT");
    *out_count = 0;
    if (h < 3 || w < 3 || n == 0) return PS_OK;   // pipeline.rs:1587-1589
    // pyramid parameters (pipeline.rs:1591-1617), host f32 with the reference's libm calls
    uint32_t max_layer = 0;
    for (size_t i = 0; i < n; i++) max_layer = std::max(max_layer, kps[i].layer);
    const uint32_t scales = std::max<uint32_t>(max_layer, 3);
    if (scales > (uint32_t)ps::kMaxScales) return set_err(PS_E_CONFIG, "keypoint layers need more than 10 scales");
    const float k = std::pow(2.0f, 1.0f / (float)scales);
    bool upsample = false;
    size_t max_oct = 0;
    bool have_layer = false;
    for (size_t i = 0; i < n; i++) {
        const ps_keypoint &kp = kps[i];
        size_t o;
        if (kp.layer > 0) {
            o = kp.octave;
            if (!have_layer) {
                if (kp.octave >= 31) return set_err(PS_E_CONFIG, "keypoint octave out of range");
                const float expected = kBaseSigma * (float)(1u << kp.octave) * powi_rs(k, (int)kp.layer);
                upsample = (kp.scale / expected) < 0.707f;
                have_layer = true;
            }
        } else {
            o = f32_to_usize(std::floor(std::log2(std::max(kp.scale / kBaseSigma, 1.0f))));
        }
        max_oct = std::max(max_oct, o);
    }
    if (max_oct + 1 > (size_t)ps::kMaxOct) return set_err(PS_E_CONFIG, "keypoints need more than 8 octaves");
    ps_sift_config cfg{};
    cfg.octaves = (uint32_t)(max_oct + 1);
    cfg.scales = scales;
    cfg.contrast_threshold = 0.015f;   // unused: no extremum search on this path
    cfg.upsample_first_octave = upsample ? 1 : 0;
    cfg.descriptor_kind = kind;
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->img, h * w * sizeof(float)));
    PS_HIP(hipMemcpy2DAsync(e->img.p, w * sizeof(float), img, row_stride * sizeof(float), w * sizeof(float), h,
                            hipMemcpyHostToDevice, e->stream));
    PS_TRY(ensure_cnt(e));
    PS_TRY(enqueue_scale_space(e, e->img.p, false, h, w, w, cfg, false));
    const ps::PyrPlan &p = e->plan;
    if (p.n_oct == 0) return PS_OK;
    // per-keypoint level, resolve_octave_and_layer (pipeline.rs:1206-1227)
    const int g_len = p.S + 3;
    const int searchable = std::max(g_len - 3, 1);
    std::vector<ps::Cand> c(n);
    for (size_t i = 0; i < n; i++) {
        const ps_keypoint &kp = kps[i];
        size_t oct, layer;
        if (kp.layer > 0) {
            oct = std::min<size_t>(kp.octave, (size_t)p.n_oct - 1);
            layer = std::min<size_t>(kp.layer, (size_t)g_len - 1);
        } else {
            oct = std::min(f32_to_usize(std::floor(std::log2(std::max(kp.scale / kBaseSigma, 1.0f)))),
                           (size_t)p.n_oct - 1);
            const float local_sigma = std::max(kp.scale / (kBaseSigma * (float)(1u << oct)), 1.0f);
            const size_t v = f32_to_usize(std::round(std::log2(local_sigma) * (float)searchable));
            layer = std::min(std::max<size_t>(v, 1), (size_t)std::max(g_len - 2, 1));
        }
        c[i] = ps::Cand{kp.x, kp.y, kp.scale, kp.response, (uint32_t)oct, (uint32_t)layer, 0};
    }
    const size_t fmax = 4 * n;
    if (fmax > UINT32_MAX / 2) return set_err(PS_E_INVALID_ARG, "too many keypoints");
    PS_TRY(ensure(e->sorted, n * sizeof(ps::Cand)));
    PS_TRY(ensure(e->ori, n * 4 * sizeof(float)));
    PS_TRY(ensure(e->nori, n * sizeof(uint32_t)));
    PS_TRY(ensure(e->feat_kp, fmax * sizeof(uint32_t)));
    PS_TRY(ensure(e->feat_j, fmax * sizeof(uint32_t)));
    PS_TRY(ensure(e->feat, fmax * sizeof(ps_feature)));
    PS_TRY(ensure(e->desc, fmax * 128 * sizeof(float)));
    PS_TRY(ensure(e->mcnt, 64));
    uint32_t *cnt = as<uint32_t>(e->cnt), *dcount = as<uint32_t>(e->mcnt);
    PS_HIP(hipMemcpyAsync(e->sorted.p, c.data(), n * sizeof(ps::Cand), hipMemcpyHostToDevice, e->stream));
    PS_HIP(ps::launch_fill_u32(cnt + 2, (uint32_t)n, 1, e->stream));
    ps::PyrView pv = pyr_view(e);
    PS_HIP(ps::launch_orientations(pv, as<ps::Cand>(e->sorted), cnt + 2, (uint32_t)n, as<float>(e->ori),
                                   as<uint32_t>(e->nori), e->stream));
    PS_HIP(ps::launch_feature_scan(cnt + 2, as<uint32_t>(e->nori), (uint32_t)n, (uint32_t)fmax,
                                   as<uint32_t>(e->feat_kp), as<uint32_t>(e->feat_j), dcount, e->stream));
    PS_HIP(ps::launch_descriptors(pv, as<ps::Cand>(e->sorted), as<float>(e->ori), as<uint32_t>(e->feat_kp),
                                  as<uint32_t>(e->feat_j), dcount, (uint32_t)fmax, kind, as<ps_feature>(e->feat),
                                  as<float>(e->desc), e->stream));
    PS_HIP(hipMemcpyAsync(&e->host_counts[0], dcount, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    const size_t nf = e->host_counts[0];
    *out_count = nf;
    if (nf > cap) return set_err(PS_E_CAPACITY, "output capacity smaller than the feature count");
    if (nf) {
        PS_HIP(hipMemcpyAsync(out_feat, e->feat.p, nf * sizeof(ps_feature), hipMemcpyDeviceToHost, e->stream));
        PS_HIP(hipMemcpyAsync(out_desc, e->desc.p, nf * 128 * sizeof(float), hipMemcpyDeviceToHost, e->stream));
        PS_HIP(hipStreamSynchronize(e->stream));
    }
    return PS_OK;
}

ps_status ps_build_pyramid(const float *img, size_t h, size_t w, size_t row_stride, const ps_sift_config *cfg,
                           float *out, size_t out_cap, size_t *out_len, int32_t *dims, int32_t *n_octaves) {
    PS_TRY(check_fail_injection());
    if (!img || !cfg || !out_len) return set_err(PS_E_INVALID_ARG, "null argument");
    if (row_stride < w) return set_err(PS_E_INVALID_ARG, "row_stride < width");
    PS_TRY(validate_cfg(cfg, nullptr, 0));
    ps::PyrPlan p;
    if (!ps::make_plan((int)h, (int)w, *cfg, p)) return set_err(PS_E_CONFIG, "unsupported blur radius");
    *out_len = p.total_floats;
    if (n_octaves) *n_octaves = p.n_oct;
    if (dims)
        for (int o = 0; o < ps::kMaxOct; o++) {
            dims[2 * o] = p.h[o];
            dims[2 * o + 1] = p.w[o];
        }
    if (!out) return PS_OK;
    if (out_cap < p.total_floats) return set_err(PS_E_CAPACITY, "pyramid output too small");
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->img, h * w * sizeof(float)));
    if (h && w)
        PS_HIP(hipMemcpy2DAsync(e->img.p, w * sizeof(float), img, row_stride * sizeof(float), w * sizeof(float), h,
                                hipMemcpyHostToDevice, e->stream));
    PS_TRY(ensure_cnt(e));
    PS_TRY(enqueue_scale_space(e, e->img.p, false, h, w, w, *cfg, false));
    if (p.total_floats)
        PS_HIP(hipMemcpyAsync(out, e->pyr.p, p.total_floats * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

ps_status ps_nn2(const float *q_desc, size_t nq, const float *t_desc, size_t nt, uint32_t *best_idx, float *best_d2,
                 float *second_d2) {
    PS_TRY(check_fail_injection());
    if ((nq && (!q_desc || !best_idx || !best_d2 || !second_d2)) || (nt && !t_desc))
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (nq == 0) return PS_OK;
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->qd, nq * 128 * sizeof(float)));
    PS_TRY(ensure(e->td, std::max<size_t>(nt, 1) * 128 * sizeof(float)));
    PS_TRY(ensure(e->mcnt, 64));
    e->host_counts[0] = (uint32_t)nq;
    e->host_counts[1] = (uint32_t)nt;
    uint32_t *dc = as<uint32_t>(e->mcnt);
    PS_HIP(hipMemcpyAsync(dc, e->host_counts, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->qd.p, q_desc, nq * 128 * sizeof(float), hipMemcpyHostToDevice, e->stream));
    if (nt) PS_HIP(hipMemcpyAsync(e->td.p, t_desc, nt * 128 * sizeof(float), hipMemcpyHostToDevice, e->stream));
    PS_TRY(ensure(e->bidx, nq * sizeof(uint32_t)));
    PS_TRY(ensure(e->bdist, nq * sizeof(float)));
    PS_TRY(ensure(e->sdist, nq * sizeof(float)));
    PS_TRY(enqueue_nn(e, as<float>(e->qd), dc, nq, as<float>(e->td), dc + 1, std::max<size_t>(nt, 1), e->part,
                      as<uint32_t>(e->bidx), as<float>(e->bdist), as<float>(e->sdist)));
    PS_HIP(hipMemcpyAsync(best_idx, e->bidx.p, nq * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipMemcpyAsync(best_d2, e->bdist.p, nq * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipMemcpyAsync(second_d2, e->sdist.p, nq * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

ps_status ps_nn2_precision(const float *q_desc, size_t nq, const float *t_desc, size_t nt, int precision,
                           uint32_t *best_idx, float *best_d2, float *second_d2, float *approx_best_d2) {
    PS_TRY(check_fail_injection());
    if (precision < PS_PRECISION_FP32 || precision > PS_PRECISION_FP8)
        return set_err(PS_E_INVALID_ARG, "precision must be PS_PRECISION_FP32, _BF16 or _FP8");
    if ((nq && (!q_desc || !best_idx || !best_d2 || !second_d2)) || (nt && !t_desc))
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (nq == 0) return PS_OK;
    if (nq >= 0xffffffffull || nt >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "too many descriptors");
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->qd, nq * 128 * sizeof(float)));
    PS_TRY(ensure(e->td, std::max<size_t>(nt, 1) * 128 * sizeof(float)));
    PS_TRY(ensure(e->mcnt, 64));
    e->host_counts[0] = (uint32_t)nq;
    e->host_counts[1] = (uint32_t)nt;
    uint32_t *dc = as<uint32_t>(e->mcnt);
    PS_HIP(hipMemcpyAsync(dc, e->host_counts, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->qd.p, q_desc, nq * 128 * sizeof(float), hipMemcpyHostToDevice, e->stream));
    if (nt) PS_HIP(hipMemcpyAsync(e->td.p, t_desc, nt * 128 * sizeof(float), hipMemcpyHostToDevice, e->stream));
    PS_TRY(ensure(e->bidx, nq * sizeof(uint32_t)));
    PS_TRY(ensure(e->bdist, 2 * nq * sizeof(float)));
    PS_TRY(ensure(e->sdist, nq * sizeof(float)));
    float *approx = approx_best_d2 ? as<float>(e->bdist) + nq : nullptr;
    PS_TRY(enqueue_nn(e, as<float>(e->qd), dc, nq, as<float>(e->td), dc + 1, std::max<size_t>(nt, 1), e->part,
                      as<uint32_t>(e->bidx), as<float>(e->bdist), as<float>(e->sdist), precision, approx));
    PS_HIP(hipMemcpyAsync(best_idx, e->bidx.p, nq * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipMemcpyAsync(best_d2, e->bdist.p, nq * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipMemcpyAsync(second_d2, e->sdist.p, nq * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    if (approx) PS_HIP(hipMemcpyAsync(approx_best_d2, approx, nq * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

ps_status ps_match(const ps_feature *q, const float *q_desc, size_t nq, const ps_feature *t, const float *t_desc,
                   size_t nt, const ps_match_options *opt, ps_feature_match *out, size_t cap, size_t *out_count) {
    PS_TRY(check_fail_injection());
    if (!opt || !out_count) return set_err(PS_E_INVALID_ARG, "null argument");
    if ((nq && (!q || !q_desc)) || (nt && (!t || !t_desc))) return set_err(PS_E_INVALID_ARG, "null argument");
    PS_TRY(validate_opt(opt, nullptr, 0));
    *out_count = 0;
    if (nq == 0 || nt < 2) return PS_OK;   // pipeline.rs:1749-1751
    if (cap < nq && !out) return set_err(PS_E_INVALID_ARG, "null output");
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->qf, nq * sizeof(ps_feature)));
    PS_TRY(ensure(e->tf, nt * sizeof(ps_feature)));
    PS_TRY(ensure(e->qd, nq * 128 * sizeof(float)));
    PS_TRY(ensure(e->td, nt * 128 * sizeof(float)));
    PS_TRY(ensure(e->mout, nq * sizeof(ps_feature_match)));
    PS_TRY(ensure(e->mcnt, 64));
    e->host_counts[0] = (uint32_t)nq;
    e->host_counts[1] = (uint32_t)nt;
    uint32_t *dc = as<uint32_t>(e->mcnt);
    PS_HIP(hipMemcpyAsync(dc, e->host_counts, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->qf.p, q, nq * sizeof(ps_feature), hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->tf.p, t, nt * sizeof(ps_feature), hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->qd.p, q_desc, nq * 128 * sizeof(float), hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->td.p, t_desc, nt * 128 * sizeof(float), hipMemcpyHostToDevice, e->stream));
    PS_TRY(enqueue_match(e, as<ps_feature>(e->qf), as<float>(e->qd), dc, nq, as<ps_feature>(e->tf), as<float>(e->td),
                         dc + 1, nt, *opt, as<ps_feature_match>(e->mout), dc + 2));
    PS_HIP(hipMemcpyAsync(&e->host_counts[2], dc + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    size_t n = e->host_counts[2];
    *out_count = n;
    if (n > cap) return set_err(PS_E_CAPACITY, "output capacity smaller than the match count");
    if (n) {
        PS_HIP(hipMemcpyAsync(out, e->mout.p, n * sizeof(ps_feature_match), hipMemcpyDeviceToHost, e->stream));
        PS_HIP(hipStreamSynchronize(e->stream));
    }
    return PS_OK;
}

ps_status ps_engine_create(int device, ps_engine **out) {
    if (!out) return set_err(PS_E_INVALID_ARG, "null argument");
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_inited || (device >= 0 && device != g_device)) PS_TRY(init_locked(device));
        if (device < 0) device = g_device;
    }
    return engine_new(device, out);
}

ps_status ps_engine_destroy(ps_engine *e) {
    if (!e) return PS_OK;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    engine_free(e);
    return PS_OK;
}

void *ps_engine_stream(ps_engine *e) { return e ? (void *)e->stream : nullptr; }

ps_status ps_engine_sync(ps_engine *e) {
    if (!e) return set_err(PS_E_INVALID_ARG, "null engine");
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

ps_status ps_engine_extract_u8(ps_engine *e, const uint8_t *d_img, size_t h, size_t w, size_t row_stride,
                               const ps_sift_config *cfg, ps_feature *d_feat, float *d_desc, size_t cap,
                               uint32_t *d_count) {
    PS_TRY(check_fail_injection());
    if (!e || !cfg || !d_count || (h && w && !d_img)) return set_err(PS_E_INVALID_ARG, "null argument");
    if (row_stride < w) return set_err(PS_E_INVALID_ARG, "row_stride < width");
    PS_TRY(validate_cfg(cfg, nullptr, 0));
    PS_HIP(hipSetDevice(e->device));
    return enqueue_extract(e, d_img, true, h, w, row_stride, *cfg, d_feat, d_desc,
                           (uint32_t)std::min<size_t>(cap, PS_MAX_FEATURES), d_count);
}

ps_status ps_engine_match(ps_engine *e, const ps_feature *d_qf, const float *d_qd, const uint32_t *d_nq, size_t nq_max,
                          const ps_feature *d_tf, const float *d_td, const uint32_t *d_nt, size_t nt_max,
                          const ps_match_options *opt, ps_feature_match *d_out, uint32_t *d_out_count) {
    PS_TRY(check_fail_injection());
    if (!e || !opt || !d_nq || !d_nt || !d_out_count) return set_err(PS_E_INVALID_ARG, "null argument");
    PS_TRY(validate_opt(opt, nullptr, 0));
    PS_HIP(hipSetDevice(e->device));
    return enqueue_match(e, d_qf, d_qd, d_nq, nq_max, d_tf, d_td, d_nt, nt_max, *opt, d_out, d_out_count);
}

ps_status ps_engine_nn2(ps_engine *e, const float *d_q_desc, const uint32_t *d_nq, size_t nq_max,
                        const float *d_t_desc, const uint32_t *d_nt, size_t nt_max, int precision,
                        uint32_t *d_best_idx, float *d_best_d2, float *d_second_d2, float *d_approx) {
    PS_TRY(check_fail_injection());
    if (!e || !d_nq || !d_nt || !d_best_idx || !d_best_d2 || !d_second_d2)
        return set_err(PS_E_INVALID_ARG, "null argument");
    if (precision < PS_PRECISION_FP32 || precision > PS_PRECISION_FP8)
        return set_err(PS_E_INVALID_ARG, "precision must be PS_PRECISION_FP32, _BF16 or _FP8");
    if (nq_max == 0) return PS_OK;
    PS_HIP(hipSetDevice(e->device));
    return enqueue_nn(e, d_q_desc, d_nq, nq_max, d_t_desc, d_nt, std::max<size_t>(nt_max, 1), e->part, d_best_idx,
                      d_best_d2, d_second_d2, precision, d_approx);
}

ps_status ps_engine_profile_begin(ps_engine *e, uint32_t max_calls) {
    if (!e) return set_err(PS_E_INVALID_ARG, "null engine");
    PS_HIP(hipSetDevice(e->device));
    prof_release(e);
    e->prof_ev.resize((size_t)max_calls * kEvPerCall);
    for (auto &ev : e->prof_ev) PS_HIP(hipEventCreate(&ev));
    e->prof_max = max_calls;
    e->prof_on = true;
    return PS_OK;
}

ps_status ps_engine_profile_end(ps_engine *e, float stage_ms[PS_NUM_STAGES], uint32_t *calls) {
    if (!e || !stage_ms) return set_err(PS_E_INVALID_ARG, "null argument");
    PS_HIP(hipSetDevice(e->device));
    PS_HIP(hipStreamSynchronize(e->stream));
    for (int i = 0; i < PS_NUM_STAGES; i++) stage_ms[i] = 0.0f;
    uint32_t ne = std::min(e->prof_ext, e->prof_max), nm = std::min(e->prof_mat, e->prof_max);
    for (uint32_t c = 0; c < ne; c++) {
        hipEvent_t *ev = &e->prof_ev[(size_t)c * kEvPerCall];
        for (int st = 0; st < 4; st++) {
            float ms = 0.0f;
            PS_HIP(hipEventElapsedTime(&ms, ev[st], ev[st + 1]));
            stage_ms[st] += ms;
        }
    }
    for (uint32_t c = 0; c < nm; c++) {
        hipEvent_t *ev = &e->prof_ev[(size_t)c * kEvPerCall];
        float ms = 0.0f;
        PS_HIP(hipEventElapsedTime(&ms, ev[5], ev[6]));
        stage_ms[4] += ms;
    }
    if (calls) *calls = ne;
    prof_release(e);
    return PS_OK;
}

ps_status ps_engine_check(ps_engine *e) {
    if (!e) return set_err(PS_E_INVALID_ARG, "null engine");
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(check_batch_err(e));
    if (!e->cnt.p) {
        PS_HIP(hipStreamSynchronize(e->stream));
        return PS_OK;
    }
    return check_err_word(e);
}

}  // extern "C"
