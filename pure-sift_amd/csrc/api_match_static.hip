// No-allocation form of the device-resident exact matcher (ps_engine_match_static).
//
// Every scratch buffer of the FP32 ratio matcher -- the split-BF16 distance pass's candidate
// lists, norms and splits for the forward pass and for the mutual check's reverse pass, the
// nearest / second-nearest / index arrays and the ratio-test compaction words -- lives in one
// file-scope device array, `scratch`, sized for 4096 query and 4096 target features (the
// per-image feature cap). Nothing here calls hipMalloc, hipFree or an allocating library
// routine on the call path, so callers that cannot allocate (code under HIP graph capture,
// tight serving loops) can match. Each GPU holds its own copy of the array for the life of
// the process; every call on a GPU shares it, so calls that overlap on one GPU race.
#include <atomic>
#include <mutex>

#include "host_common.h"

using namespace pshost;

namespace {

constexpr size_t kStaticRows = PS_MAX_FEATURES;           // features per side
constexpr size_t kStaticBytes = (size_t)17 << 20;         // 17 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets inside the array. They are fixed (laid out for the maximum size) so that the
// norm-bound words at the head of each distance-pass region keep their place across calls of
// any size: the distance pass clears the word of the next generation for the next call.
struct Layout {
    size_t fwd, rev, bidx, bdist, sdist, ridx, ridx_d, fin, end;
};

Layout layout() {
    Layout L;
    const size_t nn = al256(ps::nn_mfma_scratch_bytes(kStaticRows, kStaticRows));
    L.fwd = 0;
    L.rev = L.fwd + nn;
    L.bidx = L.rev + nn;
    L.bdist = L.bidx + al256(kStaticRows * sizeof(uint32_t));
    L.sdist = L.bdist + al256(kStaticRows * sizeof(float));
    L.ridx = L.sdist + al256(kStaticRows * sizeof(float));
    L.ridx_d = L.ridx + al256(kStaticRows * sizeof(uint32_t));
    L.fin = L.ridx_d + al256(2 * kStaticRows * sizeof(float));
    L.end = L.fin + al256(ps::match_finish_scratch_bytes(kStaticRows));
    return L;
}

struct DeviceSlot {
    std::once_flag once;
    ps_status status = PS_E_DEVICE;
    std::string msg;
    unsigned char *ptr = nullptr;           // device address of `scratch` on this GPU
    std::atomic<uint32_t> gen{0};           // generation tag of the distance passes
};

DeviceSlot g_slots[kMaxDevices];

// Once per device, with that device current: resolve the array's device address (the host
// symbol `scratch` is only a stand-in) and clear the norm-bound words of both distance-pass
// regions so no stale word can carry a generation tag of a later call.
void setup_device(int device, DeviceSlot &s) {
    const Layout L = layout();
    if (L.end > kStaticBytes) {
        s.status = PS_E_DEVICE;
        s.msg = "static matcher scratch smaller than its layout";
        return;
    }
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        s.status = PS_E_DEVICE;
        s.msg = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    unsigned char *base = static_cast<unsigned char *>(p);
    e = ps::launch_fill_u32(reinterpret_cast<uint32_t *>(base + L.fwd), 0u, 64, nullptr);
    if (e == hipSuccess) e = ps::launch_fill_u32(reinterpret_cast<uint32_t *>(base + L.rev), 0u, 64, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        s.status = PS_E_DEVICE;
        s.msg = std::string("static matcher scratch set-up: ") + hipGetErrorString(e);
        return;
    }
    (void)device;
    s.ptr = base;
    s.status = PS_OK;
}

}  // namespace

namespace pshost {

// Called by ps_init and by engine creation with `device` current; the entry point calls it
// too (its first call on a device does the set-up if neither ran there).
ps_status static_scratch_setup(int device, unsigned char **out) {
    if (device < 0 || device >= kMaxDevices) return set_err(PS_E_INVALID_ARG, "device index out of range");
    DeviceSlot &s = g_slots[device];
    std::call_once(s.once, [&] { setup_device(device, s); });
    if (s.status != PS_OK) return set_err(s.status, s.msg);
    if (out) *out = s.ptr;
    return PS_OK;
}

}  // namespace pshost

extern "C" {

ps_status ps_engine_match_static(ps_engine *e, const ps_feature *d_qf, const float *d_qd, const uint32_t *d_nq,
                                 size_t nq_max, const ps_feature *d_tf, const float *d_td, const uint32_t *d_nt,
                                 size_t nt_max, const ps_match_options *opt, ps_feature_match *d_out,
                                 uint32_t *d_out_count) {
    PS_TRY(check_fail_injection());
    if (!e || !opt || !d_nq || !d_nt || !d_out_count) return set_err(PS_E_INVALID_ARG, "null argument");
    PS_TRY(validate_opt(opt, nullptr, 0));
    if (opt->precision != PS_PRECISION_FP32)
        return set_err(PS_E_INVALID_ARG, "the no-allocation matcher runs the exact FP32 path only");
    if (nq_max > kStaticRows || nt_max > kStaticRows)
        return set_err(PS_E_INVALID_ARG, "the no-allocation matcher takes at most 4096 query and 4096 target features");
    PS_HIP(hipSetDevice(e->device));
    unsigned char *base = nullptr;
    PS_TRY(static_scratch_setup(e->device, &base));
    const Layout L = layout();
    const size_t nqm = nq_max ? nq_max : 1, ntm = nt_max ? nt_max : 1;
    DeviceSlot &s = g_slots[e->device];
    auto next_gen = [&s] {
        uint32_t g = s.gen.fetch_add(1, std::memory_order_relaxed) + 1;
        if (g == 0) g = s.gen.fetch_add(1, std::memory_order_relaxed) + 1;
        return g;
    };
    uint32_t *bidx = reinterpret_cast<uint32_t *>(base + L.bidx);
    float *bdist = reinterpret_cast<float *>(base + L.bdist);
    float *sdist = reinterpret_cast<float *>(base + L.sdist);
    PS_HIP(ps::launch_nn_mfma(d_qd, d_nq, nqm, d_td, d_nt, ntm, base + L.fwd, bidx, bdist, sdist, next_gen(),
                              e->stream));
    uint32_t *rev = nullptr;
    if (opt->mutual) {
        // reverse pass (pipeline.rs:1815-1840): nearest query of every target
        rev = reinterpret_cast<uint32_t *>(base + L.ridx);
        float *rd = reinterpret_cast<float *>(base + L.ridx_d);
        PS_HIP(ps::launch_nn_mfma(d_td, d_nt, ntm, d_qd, d_nq, nqm, base + L.rev, rev, rd, rd + ntm, next_gen(),
                                  e->stream));
    }
    PS_HIP(ps::launch_match_finish(d_qf, d_tf, d_nq, nqm, d_nt, bidx, bdist, sdist, rev, opt->ratio,
                                   opt->mutual ? 1 : 0, d_out, d_out_count, base + L.fin, e->stream));
    return PS_OK;
}

}  // extern "C"
