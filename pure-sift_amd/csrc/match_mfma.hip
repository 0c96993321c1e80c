// Exact Lowe-ratio matching through an FP32 MFMA distance GEMM.
//
// Pass 1 (nn_mfma_kernel): d'(q,t) = |t|^2 - 2 q.t on v_mfma_f32_32x32x2f32 (the per-query
// constant |q|^2 does not change the ranking; |t|^2 comes from the LDS tile); every lane keeps
// the 4 smallest (d', index) of one query over its targets in registers.
// Pass 2 (nn_rescore_kernel): the per-split lists are merged into the global top-4, each
// candidate is rescored with the reference's sequential f32 sum (pipeline.rs:1700-1717),
// and best/second/index follow the reference's rule. Exactness guard: with E a bound on
// |d'+|q|^2 - exact| for any pair, every non-candidate t has exact(t) >= d'_4 + |q|^2 - E;
// if that exceeds the exact second-best the result equals the full scan's. Otherwise the
// query is flagged and pass 3 (nn_fallback_*) scans all targets exactly.
#include "ps_internal.h"

#pragma clang fp contract(off)

namespace ps {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kMQ = 128;   // queries per workgroup (4 waves x 32 query columns)
constexpr int kMT = 64;    // targets per LDS tile (2 MFMA row blocks)
constexpr int kMP = 132;   // LDS row stride in floats
constexpr uint32_t kNoIdx = 0xffffffffu;

__device__ __forceinline__ void top4_insert(float d, uint32_t t, float (&k)[4], uint32_t (&ix)[4]) {
    if (!(d < k[3])) return;
    if (d < k[2]) {
        k[3] = k[2]; ix[3] = ix[2];
        if (d < k[1]) {
            k[2] = k[1]; ix[2] = ix[1];
            if (d < k[0]) { k[1] = k[0]; ix[1] = ix[0]; k[0] = d; ix[0] = t; }
            else { k[1] = d; ix[1] = t; }
        } else { k[2] = d; ix[2] = t; }
    } else { k[3] = d; ix[3] = t; }
}

// (d, idx) lexicographic insertion for merging lists whose indices are not monotone
// (fully unrolled so the lists stay in registers)
__device__ __forceinline__ bool lex_less(float a, uint32_t ia, float b, uint32_t ib) {
    return a < b || (a == b && ia < ib);
}
__device__ __forceinline__ void top4_insert_lex(float d, uint32_t t, float (&k)[4], uint32_t (&ix)[4]) {
    if (!lex_less(d, t, k[3], ix[3])) return;
    k[3] = d;
    ix[3] = t;
#pragma unroll
    for (int p = 3; p > 0; p--) {
        if (lex_less(k[p], ix[p], k[p - 1], ix[p - 1])) {
            float td = k[p]; k[p] = k[p - 1]; k[p - 1] = td;
            uint32_t ti = ix[p]; ix[p] = ix[p - 1]; ix[p - 1] = ti;
        }
    }
}

struct Cand4 {
    float d[4];
    uint32_t i[4];
};

__global__ __launch_bounds__(256) void nn_mfma_kernel(const float *qd, const uint32_t *nq_p, uint32_t nq_max,
                                                      const float *td, const uint32_t *nt_p, uint32_t nt_max,
                                                      int nsplit, Cand4 *part, float *qn, uint32_t *tmax_bits) {
    __shared__ __attribute__((aligned(16))) float s_t[kMT * kMP];
    __shared__ float s_tn[kMT];
    const uint32_t nq = min(*nq_p, nq_max), nt = min(*nt_p, nt_max);
    const uint32_t qb = blockIdx.x * kMQ;
    if (qb >= nq) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    const uint32_t q = qb + wave * 32 + col;

    // B operand: this lane's query, dims half*64 .. half*64+63 (k order is free for a dot product)
    float b[64];
    if (q < nq) {
        const float4 *p = reinterpret_cast<const float4 *>(qd + (size_t)q * 128 + half * 64);
#pragma unroll
        for (int m = 0; m < 16; m++) {
            float4 v = p[m];
            b[4 * m] = v.x; b[4 * m + 1] = v.y; b[4 * m + 2] = v.z; b[4 * m + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int s = 0; s < 64; s++) b[s] = 0.0f;
    }
    // |q|^2 (any order: it only feeds the error bound) for the rescoring pass
    if (blockIdx.y == 0) {
        float qq = 0.0f;
#pragma unroll
        for (int s = 0; s < 64; s++) qq += b[s] * b[s];
        qq += __shfl_xor(qq, 32);
        if (half == 0 && q < nq) qn[q] = qq;
    }
    float kd[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
    uint32_t ki[4] = {kNoIdx, kNoIdx, kNoIdx, kNoIdx};

    const uint32_t ntiles = (nt + kMT - 1) / kMT;
    const uint32_t per = (ntiles + nsplit - 1) / nsplit;
    const uint32_t tb = blockIdx.y * per, te = min(tb + per, ntiles);
    for (uint32_t tile = tb; tile < te; tile++) {
        const uint32_t t0 = tile * kMT;
        float4 v[kMT * 32 / 256];
#pragma unroll
        for (int j = 0; j < kMT * 32 / 256; j++) {
            int i = tid + j * 256, r = i >> 5, c = (i & 31) * 4;
            v[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (t0 + r < nt) v[j] = *reinterpret_cast<const float4 *>(td + (size_t)(t0 + r) * 128 + c);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kMT * 32 / 256; j++) {
            int i = tid + j * 256, r = i >> 5, c = (i & 31) * 4;
            *reinterpret_cast<float4 *>(s_t + r * kMP + c) = v[j];
        }
        __syncthreads();
        {   // |t|^2 of the tile rows: 4 threads per row, combined by shuffles
            const int r = tid >> 2, part = tid & 3;
            const float *row = s_t + r * kMP + part * 32;
            float tt = 0.0f;
#pragma unroll
            for (int c = 0; c < 32; c += 4) {
                float4 x = *reinterpret_cast<const float4 *>(row + c);
                tt += x.x * x.x + x.y * x.y + x.z * x.z + x.w * x.w;
            }
            tt += __shfl_xor(tt, 1);
            tt += __shfl_xor(tt, 2);
            if (part == 0) {
                s_tn[r] = tt;
                if (blockIdx.x == 0 && t0 + r < nt)
                    atomicMax(tmax_bits, isfinite(tt) ? __float_as_uint(tt) : 0x7f800000u);
            }
        }
        __syncthreads();
        floatx16 acc0 = {}, acc1 = {};
        const float *r0 = s_t + col * kMP + half * 64;
        const float *r1 = s_t + (32 + col) * kMP + half * 64;
#pragma unroll
        for (int m = 0; m < 16; m++) {
            float4 a0 = *reinterpret_cast<const float4 *>(r0 + 4 * m);
            float4 a1 = *reinterpret_cast<const float4 *>(r1 + 4 * m);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b[4 * m], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b[4 * m], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b[4 * m + 1], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b[4 * m + 1], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, b[4 * m + 2], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, b[4 * m + 2], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, b[4 * m + 3], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, b[4 * m + 3], acc1, 0, 0, 0);
        }
        // D layout: lane holds column (query) lane%32, rows 8*(r/4) + 4*(lane/32) + r%4
#pragma unroll
        for (int r = 0; r < 16; r++) {
            int i = 8 * (r >> 2) + 4 * half + (r & 3);
            uint32_t t = t0 + i;
            if (t < nt) top4_insert(s_tn[i] - 2.0f * acc0[r], t, kd, ki);
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            int i = 32 + 8 * (r >> 2) + 4 * half + (r & 3);
            uint32_t t = t0 + i;
            if (t < nt) top4_insert(s_tn[i] - 2.0f * acc1[r], t, kd, ki);
        }
    }
    // merge the two lane halves that hold the same query
    float od[4];
    uint32_t oi[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        od[j] = __shfl_xor(kd[j], 32);
        oi[j] = __shfl_xor(ki[j], 32);
    }
    if (half == 0 && q < nq) {
#pragma unroll
        for (int j = 0; j < 4; j++) top4_insert_lex(od[j], oi[j], kd, ki);
        Cand4 c;
#pragma unroll
        for (int j = 0; j < 4; j++) { c.d[j] = kd[j]; c.i[j] = ki[j]; }
        part[(size_t)blockIdx.y * nq_max + q] = c;
    }
}

// sequential f32 distance of pipeline.rs:1700-1717 (without the exactness-neutral early exit)
// One thread per (query, candidate): the 4 exact distances of a query are computed by 4
// adjacent lanes (each a 128-step sequential sum) and combined with shuffles.
__global__ __launch_bounds__(256) void nn_rescore_kernel(const Cand4 *part, int nsplit, const float *qd, const float *qn,
                                                         const uint32_t *nq_p, uint32_t nq_max, const float *td,
                                                         const uint32_t *nt_p, const uint32_t *tmax_bits,
                                                         uint32_t *best_idx, float *best, float *second,
                                                         uint32_t *flag_list, uint32_t *flag_count) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t q = gid >> 2;
    const int j = gid & 3;
    const bool active = q < min(*nq_p, nq_max);
    const uint32_t nt = *nt_p;
    float kd[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
    uint32_t ki[4] = {kNoIdx, kNoIdx, kNoIdx, kNoIdx};
    if (active) {
        // the per-split lists are loaded 8 at a time so their loads overlap
        for (int s0 = 0; s0 < nsplit; s0 += 8) {
            Cand4 c[8];
#pragma unroll
            for (int v = 0; v < 8; v++)
                if (s0 + v < nsplit) c[v] = part[(size_t)(s0 + v) * nq_max + q];
#pragma unroll
            for (int v = 0; v < 8; v++) {
                if (s0 + v >= nsplit) break;
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (c[v].i[u] != kNoIdx) top4_insert_lex(c[v].d[u], c[v].i[u], kd, ki);
            }
        }
    }
    uint32_t my = kNoIdx;
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (u == j) my = ki[u];
    float d = INFINITY;
    if (active && my != kNoIdx) d = seq_dist128(qd + (size_t)q * 128, td + (size_t)my * 128);
    // gather the 4 (d, idx) of this query into every one of its lanes
    float dv[4];
    uint32_t iv[4];
    const int base = (threadIdx.x & 63) & ~3;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        dv[u] = __shfl(d, base + u);
        iv[u] = __shfl(my, base + u);
    }
    if (!active || j != 0) return;
    float b = INFINITY, sc = INFINITY;
    uint32_t bi = 0;
    int ncand = 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        if (iv[u] == kNoIdx) continue;
        ncand++;
        if (dv[u] < b || (dv[u] == b && iv[u] < bi)) { sc = b; b = dv[u]; bi = iv[u]; }
        else if (dv[u] < sc) sc = dv[u];
    }
    // exactness guard (see header comment); all targets are candidates when nt <= 4
    bool ok = ncand >= (int)min(nt, 4u);
    if (ok && nt > 4) {
        const float qq = qn[q], tt = __uint_as_float(*tmax_bits);
        const float E = 1.0e-4f * (qq + tt + 2.0f * sqrtf(qq * tt)) + 1e-30f;
        ok = (kd[3] + qq) - E > sc;
    }
    best_idx[q] = bi;
    best[q] = b;
    second[q] = sc;
    if (!ok) flag_list[atomicAdd(flag_count, 1u)] = q;
}

// Exact scan for flagged queries over a fixed 64 x 16 grid: workgroup (a, b) takes flagged
// queries a, a+64, ... and targets b*256 + tid (+ 4096 strides), one target per thread, then
// reduces its 256 partial states; nn_fallback_merge_kernel combines the 16 splits in order.
constexpr int kFbQ = 64, kFbT = 16;
__global__ __launch_bounds__(256) void nn_fallback_partial_kernel(const float *qd, const uint32_t *nq_p,
                                                                  uint32_t nq_max, const float *td,
                                                                  const uint32_t *nt_p, uint32_t nt_max,
                                                                  const uint32_t *flag_list,
                                                                  const uint32_t *flag_count, NNPart *fb) {
    __shared__ float sb[256], ss[256];
    __shared__ uint32_t si[256];
    const uint32_t nflag = min(*flag_count, min(*nq_p, nq_max));
    const uint32_t nt = min(*nt_p, nt_max);
    for (uint32_t f = blockIdx.x; f < nflag; f += kFbQ) {
        const uint32_t q = flag_list[f];
        float b = INFINITY, s = INFINITY;
        uint32_t bi = 0;
        for (uint32_t t = blockIdx.y * 256 + threadIdx.x; t < nt; t += kFbT * 256) {
            float d = seq_dist128(qd + (size_t)q * 128, td + (size_t)t * 128);
            if (d < b) { s = b; b = d; bi = t; }
            else if (d < s) s = d;
        }
        __syncthreads();
        sb[threadIdx.x] = b; ss[threadIdx.x] = s; si[threadIdx.x] = bi;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {   // tree reduction with the order-independent rule
            if (threadIdx.x < (unsigned)w) {
                float B = sb[threadIdx.x], S = ss[threadIdx.x];
                uint32_t BI = si[threadIdx.x];
                float b2 = sb[threadIdx.x + w], s2 = ss[threadIdx.x + w];
                uint32_t i2 = si[threadIdx.x + w];
                float other;
                if (b2 < B || (b2 == B && i2 < BI)) { other = B; B = b2; BI = i2; }
                else other = b2;
                sb[threadIdx.x] = B; ss[threadIdx.x] = fminf(fminf(S, s2), other); si[threadIdx.x] = BI;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            NNPart p;
            p.best = sb[0]; p.second = ss[0]; p.idx = si[0]; p.pad = 0;
            fb[(size_t)f * kFbT + blockIdx.y] = p;
        }
    }
}

__global__ void nn_fallback_merge_kernel(const NNPart *fb, const uint32_t *flag_list, const uint32_t *flag_count,
                                         const uint32_t *nq_p, uint32_t nq_max, uint32_t *best_idx, float *best,
                                         float *second) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= min(*flag_count, min(*nq_p, nq_max))) return;
    NNPart p = fb[(size_t)f * kFbT];
    float B = p.best, S = p.second;
    uint32_t BI = p.idx;
    for (int k = 1; k < kFbT; k++) {
        NNPart o = fb[(size_t)f * kFbT + k];
        float other;
        if (o.best < B || (o.best == B && o.idx < BI)) { other = B; B = o.best; BI = o.idx; }
        else other = o.best;
        S = fminf(fminf(S, o.second), other);
    }
    const uint32_t q = flag_list[f];
    best_idx[q] = BI;
    best[q] = B;
    second[q] = S;
}

int nn_mfma_splits(size_t nq_max, size_t nt_max) {
    size_t qb = (nq_max + kMQ - 1) / kMQ, tt = (nt_max + kMT - 1) / kMT;
    if (qb == 0 || tt == 0) return 1;
    size_t ns = (768 + qb - 1) / qb;
    if (ns > tt) ns = tt;
    if (ns > 1024) ns = 1024;
    return (int)(ns < 1 ? 1 : ns);
}

size_t nn_mfma_scratch_bytes(size_t nq_max, size_t nt_max) {
    size_t ns = (size_t)nn_mfma_splits(nq_max, nt_max);
    return ns * nq_max * sizeof(Cand4) + nq_max * kFbT * sizeof(NNPart) + nq_max * (sizeof(float) + sizeof(uint32_t)) +
           256;
}

// Scratch layout: [part: ns*nq_max Cand4][fb: nq_max*16 NNPart][qn: nq_max][flag_list: nq_max]
// [counters: tmax, flag_count]
hipError_t launch_nn_mfma(const float *qd, const uint32_t *nq, size_t nq_max, const float *td, const uint32_t *nt,
                          size_t nt_max, void *scratch, uint32_t *best_idx, float *best, float *second,
                          hipStream_t st) {
    if (nq_max == 0) return hipSuccess;
    const int ns = nn_mfma_splits(nq_max, nt_max);
    char *p = static_cast<char *>(scratch);
    Cand4 *part = reinterpret_cast<Cand4 *>(p);
    p += (size_t)ns * nq_max * sizeof(Cand4);
    NNPart *fb = reinterpret_cast<NNPart *>(p);
    p += nq_max * kFbT * sizeof(NNPart);
    float *qn = reinterpret_cast<float *>(p);
    p += nq_max * sizeof(float);
    uint32_t *flag_list = reinterpret_cast<uint32_t *>(p);
    p += nq_max * sizeof(uint32_t);
    uint32_t *tmax = reinterpret_cast<uint32_t *>(p);
    uint32_t *flag_count = tmax + 1;
    hipError_t e = launch_fill_u32(tmax, 0u, 2, st);
    if (e != hipSuccess) return e;
    dim3 grid((unsigned)((nq_max + kMQ - 1) / kMQ), (unsigned)ns);
    nn_mfma_kernel<<<grid, 256, 0, st>>>(qd, nq, (uint32_t)nq_max, td, nt, (uint32_t)nt_max, ns, part, qn, tmax);
    nn_rescore_kernel<<<(unsigned)((4 * nq_max + 255) / 256), 256, 0, st>>>(part, ns, qd, qn, nq, (uint32_t)nq_max, td,
                                                                            nt, tmax, best_idx, best, second,
                                                                            flag_list, flag_count);
    nn_fallback_partial_kernel<<<dim3(kFbQ, kFbT), 256, 0, st>>>(qd, nq, (uint32_t)nq_max, td, nt, (uint32_t)nt_max,
                                                                 flag_list, flag_count, fb);
    nn_fallback_merge_kernel<<<(unsigned)((nq_max + 255) / 256), 256, 0, st>>>(fb, flag_list, flag_count, nq,
                                                                               (uint32_t)nq_max, best_idx, best,
                                                                               second);
    return hipGetLastError();
}

}  // namespace ps
