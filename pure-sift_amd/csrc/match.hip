// Brute-force Lowe-ratio matcher, bit-exact with pure-sift 0.2.0's
// match_descriptors_with_options (src/pipeline.rs:1700-1841).
//
// Every (query, target) distance is the reference's sequential f32 sum
// sum += (q_i - t_i)^2 over i = 0..127 with separate roundings. The early exit of
// descriptor_dist_sq_early_exit (pipeline.rs:1712) never changes a result
// (partial sums are monotone and it only fires when the sum already reaches the
// running second-best), so it is not needed for exactness. The per-query state
// (best, second, best index) is merged across threads and workgroups with an
// order-independent rule -- best = min distance, lowest index on ties; second =
// min(second_a, second_b, max(best_a, best_b)) -- which equals the reference's
// sequential scan in index order.
#include "ps_internal.h"

#pragma clang fp contract(off)

namespace ps {

constexpr int kQT = 64;       // queries per workgroup
constexpr int kTT = 64;       // targets per tile
constexpr int kDP = 132;      // padded row stride in LDS (floats)

__device__ __forceinline__ void nn_push(float d, uint32_t t, float &b, float &s, uint32_t &bi) {
    if (d < b) { s = b; b = d; bi = t; }          // pipeline.rs:1779-1785
    else if (d < s) s = d;
}

__device__ __forceinline__ void nn_combine(float &b, float &s, uint32_t &bi, float b2, float s2, uint32_t bi2) {
    float other;
    if (b2 < b || (b2 == b && bi2 < bi)) { other = b; b = b2; bi = bi2; }
    else other = b2;
    s = fminf(fminf(s, s2), other);
}

__global__ __launch_bounds__(256) void nn_partial_kernel(const float *qd, const uint32_t *nq_p, uint32_t nq_max,
                                                         const float *td, const uint32_t *nt_p, uint32_t nt_max,
                                                         int nsplit, NNPart *part) {
    __shared__ __attribute__((aligned(16))) float sq[kQT * kDP];
    __shared__ __attribute__((aligned(16))) float st[kTT * kDP];
    const uint32_t nq = min(*nq_p, nq_max), nt = min(*nt_p, nt_max);
    const uint32_t q0 = blockIdx.x * kQT;
    if (q0 >= nq) return;
    const int tid = threadIdx.x;
    const uint32_t ntiles = (nt + kTT - 1) / kTT;
    const uint32_t per = (ntiles + nsplit - 1) / nsplit;
    const uint32_t tb = blockIdx.y * per, te = min(tb + per, ntiles);

    for (int i = tid; i < kQT * 32; i += 256) {
        int r = i >> 5, c = (i & 31) * 4;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (q0 + r < nq) v = *reinterpret_cast<const float4 *>(qd + (size_t)(q0 + r) * 128 + c);
        *reinterpret_cast<float4 *>(sq + r * kDP + c) = v;
    }
    const int tx = tid & 15, ty = tid >> 4;
    float best[4], second[4];
    uint32_t bidx[4];
#pragma unroll
    for (int i = 0; i < 4; i++) { best[i] = INFINITY; second[i] = INFINITY; bidx[i] = 0; }

    for (uint32_t tile = tb; tile < te; tile++) {
        const uint32_t t0 = tile * kTT;
        float4 v[kTT * 32 / 256];
#pragma unroll
        for (int j = 0; j < kTT * 32 / 256; j++) {
            int i = tid + j * 256, r = i >> 5, c = (i & 31) * 4;
            v[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (t0 + r < nt) v[j] = *reinterpret_cast<const float4 *>(td + (size_t)(t0 + r) * 128 + c);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kTT * 32 / 256; j++) {
            int i = tid + j * 256, r = i >> 5, c = (i & 31) * 4;
            *reinterpret_cast<float4 *>(st + r * kDP + c) = v[j];
        }
        __syncthreads();
        float acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = 0.0f;
#pragma unroll 2
        for (int d = 0; d < 128; d += 4) {
            float4 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = *reinterpret_cast<const float4 *>(sq + (ty * 4 + i) * kDP + d);
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = *reinterpret_cast<const float4 *>(st + (tx * 4 + j) * kDP + d);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    float e;
                    e = a[i].x - b[j].x; acc[i][j] += e * e;
                    e = a[i].y - b[j].y; acc[i][j] += e * e;
                    e = a[i].z - b[j].z; acc[i][j] += e * e;
                    e = a[i].w - b[j].w; acc[i][j] += e * e;
                }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            uint32_t t = t0 + tx * 4 + j;
            if (t < nt) {
#pragma unroll
                for (int i = 0; i < 4; i++) nn_push(acc[i][j], t, best[i], second[i], bidx[i]);
            }
        }
    }
    // merge the 16 partial states of each query (threads with equal ty)
    __syncthreads();
    float *sb = sq;                                   // reuse LDS: [64 queries][16] x 3
    float *ss = sq + kQT * 16;
    uint32_t *si = reinterpret_cast<uint32_t *>(sq + 2 * kQT * 16);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int q = ty * 4 + i;
        sb[q * 16 + tx] = best[i];
        ss[q * 16 + tx] = second[i];
        si[q * 16 + tx] = bidx[i];
    }
    __syncthreads();
    if (tid < kQT && q0 + tid < nq) {
        float b = sb[tid * 16], s = ss[tid * 16];
        uint32_t bi = si[tid * 16];
        for (int k = 1; k < 16; k++) nn_combine(b, s, bi, sb[tid * 16 + k], ss[tid * 16 + k], si[tid * 16 + k]);
        NNPart p;
        p.best = b; p.second = s; p.idx = bi; p.pad = 0;
        part[(size_t)blockIdx.y * nq_max + q0 + tid] = p;
    }
}

int nn_splits(size_t nq_max, size_t nt_max) {
    size_t qtiles = (nq_max + kQT - 1) / kQT, ttiles = (nt_max + kTT - 1) / kTT;
    if (qtiles == 0 || ttiles == 0) return 1;
    size_t ns = (2048 + qtiles - 1) / qtiles;
    if (ns > ttiles) ns = ttiles;
    if (ns > 256) ns = 256;
    return (int)(ns < 1 ? 1 : ns);
}

hipError_t launch_nn_partial(const float *qd, const uint32_t *nq, size_t nq_max, const float *td, const uint32_t *nt,
                             size_t nt_max, int nsplit, NNPart *part, hipStream_t st) {
    if (nq_max == 0) return hipSuccess;
    dim3 grid((unsigned)((nq_max + kQT - 1) / kQT), (unsigned)nsplit);
    nn_partial_kernel<<<grid, 256, 0, st>>>(qd, nq, (uint32_t)nq_max, td, nt, (uint32_t)nt_max, nsplit, part);
    return hipGetLastError();
}

__global__ void nn_merge_kernel(const NNPart *part, int nsplit, const uint32_t *nq_p, uint32_t nq_max,
                                uint32_t *best_idx, float *best, float *second) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= min(*nq_p, nq_max)) return;
    NNPart p = part[q];
    float b = p.best, s = p.second;
    uint32_t bi = p.idx;
    for (int k = 1; k < nsplit; k++) {
        NNPart o = part[(size_t)k * nq_max + q];
        nn_combine(b, s, bi, o.best, o.second, o.idx);
    }
    best_idx[q] = bi;
    if (best) best[q] = b;
    if (second) second[q] = s;
}

hipError_t launch_nn_merge(const NNPart *part, int nsplit, const uint32_t *nq, size_t nq_max, uint32_t *best_idx,
                           float *best, float *second, hipStream_t st) {
    if (nq_max == 0) return hipSuccess;
    nn_merge_kernel<<<(unsigned)((nq_max + 255) / 256), 256, 0, st>>>(part, nsplit, nq, (uint32_t)nq_max, best_idx,
                                                                      best, second);
    return hipGetLastError();
}

// Ratio test (pipeline.rs:1788-1805), optional mutual check (pipeline.rs:1815-1840) and
// ordered compaction of the accepted matches in query order: match_flag_kernel decides every
// query in parallel and counts per 256-query block; match_compact_kernel writes the records.
__global__ __launch_bounds__(256) void match_flag_kernel(const uint32_t *nq_p, uint32_t nq_max, const uint32_t *nt_p,
                                                         const uint32_t *best_idx, const float *best,
                                                         const float *second, const uint32_t *rev_idx,
                                                         float ratio_thr, int mutual, uint8_t *flag,
                                                         uint32_t *block_count) {
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    const uint32_t nq = min(*nq_p, nq_max);
    bool acc = false;
    if (q < nq && *nt_p >= 2) {
        float bd = best[q], sd = second[q];
        if (sd > 1e-12f && isfinite(bd)) {
            float r = sqrtf(bd) / sqrtf(sd);
            acc = r < ratio_thr && (!mutual || rev_idx[best_idx[q]] == q);
        }
    }
    if (q < nq_max) flag[q] = acc ? 1 : 0;
    if (acc) atomicAdd(&s_cnt, 1u);
    __syncthreads();
    if (threadIdx.x == 0) block_count[blockIdx.x] = s_cnt;
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
    if (f) {
        const uint32_t pos = before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        float d1 = sqrtf(best[q]), d2 = sqrtf(second[q]);
        uint32_t t = best_idx[q];
        ps_feature a = qf[q], c = tf[t];
        ps_feature_match r;
        r.query_index = q; r.target_index = t;
        r.qx = a.x; r.qy = a.y; r.tx = c.x; r.ty = c.y;
        r.distance = d1; r.ratio = d1 / d2;
        r.dx = c.x - a.x; r.dy = c.y - a.y;
        out[pos] = r;
    }
}

size_t match_finish_scratch_bytes(size_t nq_max) { return nq_max + ((nq_max + 255) / 256) * 4 + 64; }

hipError_t launch_match_finish(const ps_feature *qf, const ps_feature *tf, const uint32_t *nq, size_t nq_max,
                               const uint32_t *nt, const uint32_t *best_idx, const float *best, const float *second,
                               const uint32_t *rev_idx, float ratio, int mutual, ps_feature_match *out,
                               uint32_t *out_count, void *scratch, hipStream_t st) {
    const uint32_t nblocks = (uint32_t)((nq_max + 255) / 256);
    if (nblocks > 1024) return hipErrorInvalidValue;
    uint32_t *block_count = static_cast<uint32_t *>(scratch);
    uint8_t *flag = reinterpret_cast<uint8_t *>(block_count + nblocks);
    match_flag_kernel<<<nblocks, 256, 0, st>>>(nq, (uint32_t)nq_max, nt, best_idx, best, second, rev_idx, ratio, mutual,
                                               flag, block_count);
    match_compact_kernel<<<nblocks, 256, 0, st>>>(qf, tf, nq, (uint32_t)nq_max, best_idx, best, second, flag,
                                                   block_count, nblocks, out, out_count);
    return hipGetLastError();
}

}  // namespace ps
