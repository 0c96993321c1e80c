// Host-side state shared by the C ABI translation units: the engine (stream +
// device workspace), the per-thread error message and the pooled borrow used by
// host-pointer entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ps_internal.h"
#include <rocprofiler-sdk-roctx/roctx.h>

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

// Workspace slots of the catalog / precision-matcher / RANSAC entry points.
enum AuxSlot {
    AUX_HAM_CAT = 0, AUX_HAM_Q, AUX_HAM_BOUND, AUX_HAM_CNT, AUX_HAM_KEYS, AUX_HAM_IDX, AUX_HAM_DIST, AUX_HAM_FLAG,
    AUX_RR_QD, AUX_RR_CD, AUX_RR_IDX, AUX_RR_DIST,
    AUX_PM_0, AUX_PM_1, AUX_PM_2, AUX_PM_3, AUX_PM_4, AUX_PM_5, AUX_PM_6, AUX_PM_7,
    AUX_RS_0, AUX_RS_1, AUX_RS_2, AUX_RS_3, AUX_RS_4, AUX_RS_5, AUX_RS_6, AUX_RS_7,
    AUX_CS_0, AUX_CS_1, AUX_CS_2, AUX_CS_3, AUX_CS_4, AUX_CS_5, AUX_CS_6, AUX_CS_7, AUX_CS_8, AUX_CS_9,
    AUX_CS_10, AUX_CS_11, AUX_CS_12, AUX_CS_13,
    AUX_PG_0, AUX_PG_1, AUX_PG_2, AUX_PG_3, AUX_PG_4, AUX_PG_5, AUX_PG_6, AUX_PG_7, AUX_PG_8, AUX_PG_9,
    AUX_PG_10, AUX_PG_11,
    AUX_PN_0, AUX_PN_1, AUX_PN_2, AUX_PN_3, AUX_PN_4, AUX_PN_5, AUX_PN_6, AUX_PN_7, AUX_PN_8, AUX_PN_9,
    AUX_COUNT
};

struct BatchState;   // batch.hip

// A named host-side trace range (roctx) around an enqueue sequence: rocprofv3 --marker-trace
// shows it beside the kernels the sequence launched (host time of the call; with no tracer
// attached the calls return at once).
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

// Chains of later octaves that the leaf work of octave o waits for before it is forked, for an
// image whose octave 1 has oct1_px pixels. Octave 0's leaf (its level S+1 and its extrema, the
// largest work of the scale space) would otherwise fill the device while octave 1's chain runs,
// and every later octave waits for that chain: where that chain is small enough to be
// latency-bound it goes first, octave 0's leaf forked after octave 1's levels 1..S (1080p frame
// 0.455 -> 0.448 ms, 720p 0.376 -> 0.350 ms; delaying every leaf by one octave, or octave 0's by
// two, measured slower at 1080p: 0.464 / 0.469 ms). Larger octaves 1 keep the fork at octave 0
// (a 3840x2160 frame, octave 1 8.3M px, and the batched scale space, whose grouped launches are
// throughput-bound, measured slower with the delay).
inline int leaf_delay(int o, long oct1_px) { return o == 0 && oct1_px <= (4L << 20) ? 1 : 0; }

struct ps_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    // the leaf work of octave o of an extraction (levels above S, extrema) runs on
    // side[o % kSides], forked once level S of octave o (of octave 1 for o = 0) is written and
    // joined back into `stream` at the end of the scale space
    static constexpr int kSides = 3;
    hipStream_t side[kSides] = {};
    hipEvent_t fork_ev[ps::kMaxOct] = {};
    hipEvent_t join_ev[kSides] = {};
    DevBuf pyr, cand, rank, sorted, ori, nori, feat_kp, feat_j, cnt;
    DevBuf img, feat, desc, kpout;                 // host-API staging
    DevBuf qf, qd, tf, td, mcnt, mout;             // host-API match staging
    DevBuf part, part2, bidx, bdist, sdist, ridx, ridx_d, fin;  // matcher scratch
    DevBuf aux[AUX_COUNT];
    uint32_t nn_gen = 0;                           // per-call tag of the exact matcher's scratch words
    // batched extraction: the grouped launches of the last batch (their job tables and
    // workspace), kept with an instantiated graph of the sequence
    BatchState *batch = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;
    ps::PyrPlan plan;
    uint32_t host_counts[4] = {};
    // panorama tables last uploaded (image descriptors, then the pair list follows from their
    // number): an unchanged call re-uses them without a host round trip
    std::vector<ps_image_desc> pano_images;
    const void *pano_tables = nullptr;
    // stage profiling: per recorded call, events [extract: 0..4][match: 5..6]
    std::vector<hipEvent_t> prof_ev;
    uint32_t prof_max = 0, prof_ext = 0, prof_mat = 0;
    bool prof_on = false;
};

namespace pshost {

// Counter word of an engine's `cnt` buffer that accumulates the error bits of its extractions
// (ps::kErrCandOverflow) until ps_engine_check reads and clears it; words 0..kErrWord-1 are
// the per-extraction counters.
constexpr int kErrWord = 15;
ps_status ensure_cnt(ps_engine *e);
ps_status clear_err_word(ps_engine *e);
ps_status check_err_word(ps_engine *e);
// the error of a 16-word counter block (PS_OK without overflow bits); `where` names the image
ps_status err_status(const uint32_t *words, const char *where);

ps_status set_err(ps_status s, const std::string &msg);
// grows b to at least `bytes` (contents are not kept); *fresh (optional) tells whether memory
// was allocated by this call
ps_status ensure(DevBuf &b, size_t bytes, bool *fresh = nullptr);
ps_status check_fail_injection();

template <typename T>
T *as(DevBuf &b) {
    return static_cast<T *>(b.p);
}

ps_status engine_new(int device, ps_engine **out);
void engine_free(ps_engine *e);
void release_batch(ps_engine *e);
// error words of the last batch's images (checked and cleared like check_err_word)
ps_status check_batch_err(ps_engine *e);
ps_status validate_cfg(const ps_sift_config *c, char *msg, size_t cap);
ps_status validate_opt(const ps_match_options *o, char *msg, size_t cap);
// device address of the no-allocation matcher's static scratch on `device` (set up once per
// device, with that device current)
ps_status static_scratch_setup(int device, unsigned char **out);
ps_status prepare_extract(ps_engine *e, size_t h, size_t w, const ps_sift_config &cfg);
// exact ratio match (pipeline.rs:1744-1841) of device features, counts as device u32s
ps_status enqueue_match(ps_engine *e, const ps_feature *qf, const float *qd, const uint32_t *nq, size_t nq_max,
                        const ps_feature *tf, const float *td, const uint32_t *nt, size_t nt_max,
                        const ps_match_options &opt, ps_feature_match *out, uint32_t *out_count);
// many set-against-set ratio matches in one launch chain (ps_engine_match_sets)
ps_status enqueue_match_sets(ps_engine *e, const ps::MatchSetsArgs &a, DevBuf &scratch, ps_feature_match *out,
                             uint32_t *counts);
ps_status enqueue_extract(ps_engine *e, const void *src, bool src_u8, size_t h, size_t w, size_t stride,
                          const ps_sift_config &cfg, ps_feature *d_feat, float *d_desc, uint32_t cap,
                          uint32_t *d_count);
// blur launch shape for a level (autotuned table, else a size-based default; see launch_blur)
int blur_tile(int radius, int src_mode, int h, int w);

// RAII borrow of a pooled engine for host-pointer calls
struct Borrow {
    ps_engine *e = nullptr;
    ps_status st = PS_OK;
    Borrow();
    ~Borrow();
    Borrow(const Borrow &) = delete;
    Borrow &operator=(const Borrow &) = delete;
};

}  // namespace pshost

#define PS_HIP(call)                                                                                       \
    do {                                                                                                   \
        hipError_t _e = (call);                                                                            \
        if (_e != hipSuccess)                                                                              \
            return pshost::set_err(_e == hipErrorOutOfMemory ? PS_E_OUT_OF_MEMORY : PS_E_DEVICE,           \
                                   std::string(#call) + ": " + hipGetErrorString(_e));                     \
    } while (0)

#define PS_TRY(expr)                                                                                       \
    do {                                                                                                   \
        ps_status _s = (expr);                                                                             \
        if (_s != PS_OK) return _s;                                                                        \
    } while (0)
