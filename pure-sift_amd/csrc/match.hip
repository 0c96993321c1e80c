// Ratio test, mutual check and ordered compaction of the Lowe-ratio matcher, bit-exact with
// pure-sift 0.2.0's match_descriptors_with_options (src/pipeline.rs:1744-1841). The nearest /
// second-nearest distances come from the split-BF16 MFMA pass with exact f32 rescoring
// (match_mfma.hip); this file turns them into FeatureMatch records in query order.
#include "ps_internal.h"

#pragma clang fp contract(off)

namespace ps {

// Ratio test (pipeline.rs:1788-1805), optional mutual check (pipeline.rs:1815-1840) and
// ordered compaction of the accepted matches in query order: match_flag_kernel decides every
// query in parallel and counts per 256-query block; match_compact_kernel writes the records
// (up to 16 blocks of queries: match_finish_fused_kernel alone).

// the ratio (and mutual) decision of a valid query q with distances bd, sd
__device__ __forceinline__ bool match_accept_vals(uint32_t q, float bd, float sd, const uint32_t *best_idx,
                                                  const uint32_t *rev_idx, float ratio_thr, int mutual) {
    if (!(sd > 1e-12f && isfinite(bd))) return false;
    const float r = sqrtf(bd) / sqrtf(sd);
    return r < ratio_thr && (!mutual || rev_idx[best_idx[q]] == q);
}

// the ratio (and mutual) decision of query q
__device__ __forceinline__ bool match_accept(uint32_t q, uint32_t nq, uint32_t nt, const uint32_t *best_idx,
                                             const float *best, const float *second, const uint32_t *rev_idx,
                                             float ratio_thr, int mutual) {
    if (q >= nq || nt < 2) return false;
    return match_accept_vals(q, best[q], second[q], best_idx, rev_idx, ratio_thr, mutual);
}

__global__ __launch_bounds__(256) void match_flag_kernel(const uint32_t *nq_p, uint32_t nq_max, const uint32_t *nt_p,
                                                         const uint32_t *best_idx, const float *best,
                                                         const float *second, const uint32_t *rev_idx,
                                                         float ratio_thr, int mutual, uint8_t *flag,
                                                         uint32_t *block_count) {
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    const bool acc = match_accept(q, min(*nq_p, nq_max), *nt_p, best_idx, best, second, rev_idx, ratio_thr, mutual);
    if (q < nq_max) flag[q] = acc ? 1 : 0;
    if (acc) atomicAdd(&s_cnt, 1u);
    __syncthreads();
    if (threadIdx.x == 0) block_count[blockIdx.x] = s_cnt;
}

// the output record of accepted query q at position pos
__device__ __forceinline__ void write_match(const ps_feature *qf, const ps_feature *tf, const uint32_t *best_idx,
                                            const float *best, const float *second, uint32_t q, uint32_t pos,
                                            ps_feature_match *out) {
    const float d1 = sqrtf(best[q]), d2 = sqrtf(second[q]);
    const uint32_t t = best_idx[q];
    const ps_feature a = qf[q], c = tf[t];
    ps_feature_match r;
    r.query_index = q; r.target_index = t;
    r.qx = a.x; r.qy = a.y; r.tx = c.x; r.ty = c.y;
    r.distance = d1; r.ratio = d1 / d2;
    r.dx = c.x - a.x; r.dy = c.y - a.y;
    out[pos] = r;
}

// One workgroup per 256-query block: its output offset is the sum of the earlier blocks'
// counts, the position inside the block a ballot prefix over the flags.
__global__ __launch_bounds__(256) void match_compact_kernel(const ps_feature *qf, const ps_feature *tf,
                                                            const uint32_t *nq_p, uint32_t nq_max,
                                                            const uint32_t *best_idx, const float *best,
                                                            const float *second, const uint8_t *flag,
                                                            const uint32_t *block_count, uint32_t nblocks,
                                                            ps_feature_match *out, uint32_t *out_count) {
    __shared__ uint32_t s_red[256];
    __shared__ uint32_t s_wave[4];
    const uint32_t b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nq = min(*nq_p, nq_max);
    // offset of this block (and, in block 0, the total)
    uint32_t acc = 0;
    for (uint32_t k = tid; k < nblocks; k += 256)
        if (k < b || b == 0) acc += block_count[k];
    s_red[tid] = acc;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (tid < o) s_red[tid] += s_red[tid + o];
        __syncthreads();
    }
    const uint32_t base = b == 0 ? 0u : s_red[0];
    if (b == 0 && tid == 0) *out_count = s_red[0];
    const uint32_t q = b * 256 + tid;
    const bool f = q < nq && flag[q];
    const uint64_t m = __ballot(f);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = base;
    for (uint32_t k = 0; k < wave; k++) before += s_wave[k];
    if (f) write_match(qf, tf, best_idx, best, second, q, before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)), out);
}

// Up to kFusedBlocks blocks of queries in one launch: block b decides the queries of the blocks
// before it once more (at most 15 x 256, a few per thread) to find its output offset instead of
// reading a separate flag pass's counts; block 0 decides every query for the total.
constexpr uint32_t kFusedBlocks = 16;
__global__ __launch_bounds__(256) void match_finish_fused_kernel(const ps_feature *qf, const ps_feature *tf,
                                                                 const uint32_t *nq_p, uint32_t nq_max,
                                                                 const uint32_t *nt_p, const uint32_t *best_idx,
                                                                 const float *best, const float *second,
                                                                 const uint32_t *rev_idx, float ratio_thr, int mutual,
                                                                 ps_feature_match *out, uint32_t *out_count) {
    __shared__ uint32_t s_red[256];
    __shared__ uint32_t s_wave[4];
    const uint32_t b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nq = min(*nq_p, nq_max), nt = *nt_p;
    const uint32_t lim = min((b == 0 ? gridDim.x : b) * 256, nt < 2 ? 0u : nq);
    // every distance load of the thread issued before any decision
    float bd[kFusedBlocks], sd[kFusedBlocks];
#pragma unroll
    for (uint32_t k = 0; k < kFusedBlocks; k++) {
        const uint32_t q = tid + 256 * k;
        bd[k] = q < lim ? best[q] : 0.0f;
        sd[k] = q < lim ? second[q] : 0.0f;
    }
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < kFusedBlocks; k++) {
        const uint32_t q = tid + 256 * k;
        if (q < lim && match_accept_vals(q, bd[k], sd[k], best_idx, rev_idx, ratio_thr, mutual)) cnt++;
    }
    s_red[tid] = cnt;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (tid < o) s_red[tid] += s_red[tid + o];
        __syncthreads();
    }
    if (b == 0 && tid == 0) *out_count = s_red[0];
    const uint32_t base = b == 0 ? 0u : s_red[0];
    const uint32_t q = b * 256 + tid;
    const bool f = match_accept(q, nq, nt, best_idx, best, second, rev_idx, ratio_thr, mutual);
    const uint64_t m = __ballot(f);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = base;
    for (uint32_t k = 0; k < wave; k++) before += s_wave[k];
    if (f) write_match(qf, tf, best_idx, best, second, q, before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)), out);
}

size_t match_finish_scratch_bytes(size_t nq_max) { return nq_max + ((nq_max + 255) / 256) * 4 + 64; }

hipError_t launch_match_finish(const ps_feature *qf, const ps_feature *tf, const uint32_t *nq, size_t nq_max,
                               const uint32_t *nt, const uint32_t *best_idx, const float *best, const float *second,
                               const uint32_t *rev_idx, float ratio, int mutual, ps_feature_match *out,
                               uint32_t *out_count, void *scratch, hipStream_t st) {
    const uint32_t nblocks = (uint32_t)((nq_max + 255) / 256);
    if (nblocks > 1024) return hipErrorInvalidValue;
    if (nblocks <= kFusedBlocks) {
        match_finish_fused_kernel<<<nblocks, 256, 0, st>>>(qf, tf, nq, (uint32_t)nq_max, nt, best_idx, best, second,
                                                           rev_idx, ratio, mutual, out, out_count);
        return hipGetLastError();
    }
    uint32_t *block_count = static_cast<uint32_t *>(scratch);
    uint8_t *flag = reinterpret_cast<uint8_t *>(block_count + nblocks);
    match_flag_kernel<<<nblocks, 256, 0, st>>>(nq, (uint32_t)nq_max, nt, best_idx, best, second, rev_idx, ratio, mutual,
                                               flag, block_count);
    match_compact_kernel<<<nblocks, 256, 0, st>>>(qf, tf, nq, (uint32_t)nq_max, best_idx, best, second, flag,
                                                   block_count, nblocks, out, out_count);
    return hipGetLastError();
}

}  // namespace ps
