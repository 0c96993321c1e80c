// Exact Lowe-ratio matching through a split-BF16 MFMA distance GEMM.
//
// Pass 1 (nn_mfma_kernel): d'(q,t) = |t|^2 - 2 q.t with q.t from three bf16 MFMA products:
// every f32 element x is split as x = hi + lo + e with hi = bf16(x), lo = bf16(x - hi)
// (round to nearest even; x - hi is exact in f32), |e| <= 2^-16 |x|, and
// q.t ~ qh.th + qh.tl + ql.th on v_mfma_f32_32x32x16_bf16 (products exact in f32, f32
// accumulation). The dropped terms are below 3.1 * 2^-16 |q||t| and the accumulation of 384
// terms below 384 * 2^-24 |q||t|, so |error of q.t| < 7e-5 |q||t| -- about 5x the cheapest
// exact alternative (v_mfma_f32_32x32x2f32) in matrix throughput. The per-query constant
// |q|^2 does not change the ranking; |t|^2 comes from the LDS tile; every lane keeps the 4
// smallest (d', index) of one query over its targets in registers.
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
constexpr int kMaxNNSplit = 16;   // target splits (the rescore holds one entry per split per lane)
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

__device__ __forceinline__ uint32_t bf16_bits_rne(float x) {
    const uint32_t b = __float_as_uint(x);
    return (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;
}

// Matcher prep, one launch for both sides: |x|^2 of each row (any summation order: it only
// ranks and bounds) and the row's bf16 split (128 words: hi pairs, then lo pairs; elements 2j
// and 2j+1 in word j). Workgroups [0, tblocks) take target rows, the rest query rows,
// 4 x kPrepRows rows each (each wave kPrepRows, their loads issued together); rows in [n, rows)
// get norm +inf, so they never rank, and a zero split. The largest target norm goes to tmax[gen & 1] as (gen << 32 | ordered
// bits) by one 64-bit atomicMax per workgroup, and workgroup 0 zeroes tmax[(gen + 1) & 1],
// the word the next call (next generation) will take: stream order puts that store after
// this generation's readers of the other word. A word of another generation (fresh memory
// before its first clearing, a wrapped counter) reads as +inf, which is safe.
constexpr int kPrepRows = 4;   // rows per wave (16 per workgroup: 512 workgroups for 4096 + 4096 rows)
struct PrepSide {
    const float *x;
    const uint32_t *n_p;
    uint32_t n_max, rows;
    float *norm;
    uint32_t *split;
};
__global__ __launch_bounds__(256) void nn_prep_kernel(PrepSide t, PrepSide q, uint32_t tblocks,
                                                      unsigned long long *tmax, uint32_t gen) {
    __shared__ float s_max[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool tside = blockIdx.x < tblocks;
    const PrepSide &sd = tside ? t : q;
    const uint32_t blk = tside ? blockIdx.x : blockIdx.x - tblocks;
    const uint32_t n = min(*sd.n_p, sd.n_max);
    float wmax = 0.0f;
    float2 v[kPrepRows];
#pragma unroll
    for (int k = 0; k < kPrepRows; k++) {   // all loads first so their latencies overlap
        const uint32_t r = (blk * 4 + w) * kPrepRows + k;
        v[k] = r < n ? reinterpret_cast<const float2 *>(sd.x + (size_t)r * 128)[lane] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < kPrepRows; k++) {
        const uint32_t r = (blk * 4 + w) * kPrepRows + k;
        float s = v[k].x * v[k].x + v[k].y * v[k].y;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0 && r < sd.rows) sd.norm[r] = r < n ? s : INFINITY;
        if (r < n) wmax = fmaxf(wmax, isfinite(s) ? s : INFINITY);
        if (r < sd.rows) {
            const uint32_t hx = bf16_bits_rne(v[k].x), hy = bf16_bits_rne(v[k].y);
            const float rx = v[k].x - __uint_as_float(hx << 16), ry = v[k].y - __uint_as_float(hy << 16);
            sd.split[(size_t)r * 128 + lane] = hx | (hy << 16);
            sd.split[(size_t)r * 128 + 64 + lane] = bf16_bits_rne(rx) | (bf16_bits_rne(ry) << 16);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) tmax[(gen + 1) & 1] = 0ull;
    if (!tside) return;
    if (lane == 0) s_max[w] = wmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float m = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
        // the word only grows during a generation: a workgroup whose maximum does not exceed
        // the value it reads skips the atomic (hundreds of workgroups, one address)
        const unsigned long long v = ((unsigned long long)gen << 32) | __float_as_uint(m);
        if (__hip_atomic_load(tmax + (gen & 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v)
            atomicMax(tmax + (gen & 1), v);
    }
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Pass 1. Workgroup: 128 queries (4 waves x 32 B-operand columns; lane (col, kh) holds the
// hi and lo words of its query for k = 16s + 8kh .. +7, s = 0..7, in VGPRs) against one split of
// the targets, 64 split target rows per LDS tile; the next tile is loaded into registers while
// the current one is multiplied. Every lane keeps the 4 smallest (|t|^2 - 2 q.t, index) of
// the targets it sees; both halves' lists are written (8 candidates per query and split).
// nn_mfma_scan is one workgroup's share: queries qb .. qb + 127 of the split rows qs (nq valid)
// against target tiles [tb, te) of ts (nt valid rows, norms tn); rows at or past nt load as
// zero with norm +inf, so they never rank.
__device__ __forceinline__ void nn_mfma_scan(const uint32_t *qs, uint32_t nq, uint32_t qb, const uint32_t *ts,
                                             uint32_t nt, const float *tn, uint32_t tb, uint32_t te,
                                             float (&kd)[4], uint32_t (&ki)[4]) {
    __shared__ __attribute__((aligned(16))) uint32_t s_t[kMT * kMP];
    __shared__ __attribute__((aligned(16))) float s_tn[kMT];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    const uint32_t q = qb + wave * 32 + col;

    uint4 bh[8], bl[8];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        if (q < nq) {
            bh[s] = *reinterpret_cast<const uint4 *>(qs + (size_t)q * 128 + 8 * s + 4 * half);
            bl[s] = *reinterpret_cast<const uint4 *>(qs + (size_t)q * 128 + 64 + 8 * s + 4 * half);
        } else {
            bh[s] = make_uint4(0u, 0u, 0u, 0u);
            bl[s] = make_uint4(0u, 0u, 0u, 0u);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        kd[j] = INFINITY;
        ki[j] = kNoIdx;
    }
    constexpr int kLd = kMT * 32 / 256;   // 16-byte loads per thread per tile
    uint4 v[kLd];
    float vn = 0.0f;
#define NN_LOAD(TILE)                                                                                 \
    do {                                                                                              \
        const uint32_t t0_ = (TILE) * kMT;                                                            \
        _Pragma("unroll") for (int j = 0; j < kLd; j++) {                                             \
            int i = tid + j * 256, r = i >> 5, c = (i & 31) * 4;                                     \
            v[j] = t0_ + r < nt ? *reinterpret_cast<const uint4 *>(ts + (size_t)(t0_ + r) * 128 + c)   \
                                : make_uint4(0u, 0u, 0u, 0u);                                         \
        }                                                                                             \
        if (tid < kMT) vn = t0_ + tid < nt ? tn[t0_ + tid] : INFINITY;                                \
    } while (0)
    if (tb < te) NN_LOAD(tb);
    for (uint32_t tile = tb; tile < te; tile++) {
        __syncthreads();   // previous tile's readers are done
#pragma unroll
        for (int j = 0; j < kLd; j++) {
            int i = tid + j * 256, r = i >> 5, c = (i & 31) * 4;
            *reinterpret_cast<uint4 *>(s_t + r * kMP + c) = v[j];
        }
        if (tid < kMT) s_tn[tid] = vn;
        __syncthreads();
        if (tile + 1 < te) NN_LOAD(tile + 1);
        floatx16 acc0 = {}, acc1 = {};
        // A operand: target row col (block 0) / 32 + col (block 1), words 8s + 4kh (hi) and
        // 64 + 8s + 4kh (lo)
        const uint32_t *r0 = s_t + col * kMP + 4 * half;
        const uint32_t *r1 = s_t + (32 + col) * kMP + 4 * half;
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const bf16x8 ah0 = as_bf16x8(*reinterpret_cast<const uint4 *>(r0 + 8 * s));
            const bf16x8 al0 = as_bf16x8(*reinterpret_cast<const uint4 *>(r0 + 64 + 8 * s));
            const bf16x8 ah1 = as_bf16x8(*reinterpret_cast<const uint4 *>(r1 + 8 * s));
            const bf16x8 al1 = as_bf16x8(*reinterpret_cast<const uint4 *>(r1 + 64 + 8 * s));
            const bf16x8 qh = as_bf16x8(bh[s]), ql = as_bf16x8(bl[s]);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0, qh, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1, qh, acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0, ql, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1, ql, acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al0, qh, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al1, qh, acc1, 0, 0, 0);
        }
        // D layout: lane holds column (query) lane%32, rows 8*(r/4) + 4*(lane/32) + r%4;
        // padding rows have |t|^2 = +inf and never rank
        const uint32_t t0 = tile * kMT;
        float vmin = INFINITY;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const float4 n0 = *reinterpret_cast<const float4 *>(&s_tn[8 * rr + 4 * half]);
            const float4 n1 = *reinterpret_cast<const float4 *>(&s_tn[32 + 8 * rr + 4 * half]);
            const float a[4] = {n0.x, n0.y, n0.z, n0.w}, c[4] = {n1.x, n1.y, n1.z, n1.w};
#pragma unroll
            for (int u = 0; u < 4; u++) {
                acc0[rr * 4 + u] = a[u] - 2.0f * acc0[rr * 4 + u];
                acc1[rr * 4 + u] = c[u] - 2.0f * acc1[rr * 4 + u];
                vmin = fminf(vmin, fminf(acc0[rr * 4 + u], acc1[rr * 4 + u]));
            }
        }
        if (vmin < kd[3]) {
#pragma unroll
            for (int r = 0; r < 16; r++) top4_insert(acc0[r], t0 + 8 * (r >> 2) + 4 * half + (r & 3), kd, ki);
#pragma unroll
            for (int r = 0; r < 16; r++) top4_insert(acc1[r], t0 + 32 + 8 * (r >> 2) + 4 * half + (r & 3), kd, ki);
        }
    }
#undef NN_LOAD
}

__device__ __forceinline__ uint32_t nn_tiles_begin(uint32_t nt, int nsplit, uint32_t split, uint32_t &te) {
    const uint32_t ntiles = (nt + kMT - 1) / kMT;
    const uint32_t per = (ntiles + nsplit - 1) / nsplit;
    const uint32_t tb = split * per;
    te = min(tb + per, ntiles);
    return tb;
}

__global__ __launch_bounds__(256) void nn_mfma_kernel(const uint32_t *qs, const uint32_t *nq_p, uint32_t nq_max,
                                                      const uint32_t *ts, const uint32_t *nt_p, uint32_t nt_max,
                                                      const float *tn, int nsplit, Cand4 *part) {
    const uint32_t nq = min(*nq_p, nq_max), nt = min(*nt_p, nt_max);
    const uint32_t qb = blockIdx.x * kMQ;
    if (qb >= nq) return;
    uint32_t te;
    const uint32_t tb = nn_tiles_begin(nt, nsplit, blockIdx.y, te);
    float kd[4];
    uint32_t ki[4];
    nn_mfma_scan(qs, nq, qb, ts, nt, tn, tb, te, kd, ki);
    const int lane = threadIdx.x & 63;
    const uint32_t q = qb + (threadIdx.x >> 6) * 32 + (lane & 31);
    if (q < nq) {
        Cand4 c;
#pragma unroll
        for (int j = 0; j < 4; j++) { c.d[j] = kd[j]; c.i[j] = ki[j]; }
        part[((size_t)blockIdx.y * nq_max + q) * 2 + (lane >> 5)] = c;
    }
}

// Pass 2. Eight lanes per query: the 8 smallest (value, index) over the 2 * nsplit lists
// (lane j holds entry j of every list), each rescored with the reference's sequential f32
// sum (pipeline.rs:1700-1717), then best / second in (distance, index) order. Exactness
// guard: with E a bound on |value + |q|^2 - exact| for any pair, every target that is not a
// candidate has exact distance >= tau + |q|^2 - E, where tau is the smaller of the 8th
// selected value and the 4th (last) entry of every list (a list drops only targets above
// its own 4th); if that exceeds the exact second-best the result equals the full scan's,
// otherwise the query is flagged and the workgroup scans all targets for it exactly
// (pipeline.rs:1700-1740 order rules).
// nn_rescore_block is one workgroup's share (kRsThreads threads, kRsQ = kRsThreads / 8 queries):
// queries kRsQ qblk .. + kRsQ - 1 (nq valid) of lists
// part[(split * pstride + q) * 2 + half], targets td (nt of them, nt_scan scanned exactly).
constexpr int kRsThreads = 128;   // 16 queries per workgroup: 256 workgroups for 4096 queries
constexpr int kRsQ = kRsThreads / 8;
__device__ __forceinline__ void nn_rescore_block(const Cand4 *part, int nsplit, size_t pstride, uint32_t qblk,
                                                 const float *qd, const float *qn, uint32_t nq, const float *td,
                                                 uint32_t nt, uint32_t nt_scan, const unsigned long long *tmax,
                                                 uint32_t gen, uint32_t *best_idx, float *best, float *second) {
    __shared__ uint32_t s_flag[kRsQ];
    __shared__ uint32_t s_nflag;
    __shared__ float sb[kRsThreads], ss[kRsThreads];
    __shared__ uint32_t si[kRsThreads];
    if (threadIdx.x == 0) s_nflag = 0;
    __syncthreads();
    const uint32_t q = qblk * kRsQ + (threadIdx.x >> 3);
    const int j = threadIdx.x & 7;
    const bool active = q < nq;
    // lane j: entry j % 4 of the half-(j / 4) list of every split
    float md[kMaxNNSplit];
    uint32_t mi[kMaxNNSplit];
    float last4 = INFINITY;
    {
        const int u = j & 3;
        const uint32_t qs = active ? q : 0u;
#pragma unroll
        for (int s = 0; s < kMaxNNSplit; s++) {   // loads first, then the bookkeeping
            const int ss = s < nsplit ? s : 0;
            const float *e = reinterpret_cast<const float *>(part + ((size_t)ss * pstride + qs) * 2 + (j >> 2));
            md[s] = e[u];
            mi[s] = reinterpret_cast<const uint32_t *>(e)[4 + u];
        }
#pragma unroll
        for (int s = 0; s < kMaxNNSplit; s++) {
            if (!active || s >= nsplit) {
                md[s] = INFINITY;
                mi[s] = kNoIdx;
            }
            if (u == 3) last4 = fminf(last4, md[s]);
        }
    }
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) last4 = fminf(last4, __shfl_xor(last4, o));
    float kth = INFINITY;
    uint32_t sel_i = kNoIdx;
    for (int round = 0; round < 8; round++) {
        float m = INFINITY;
        uint32_t mix = kNoIdx;
        int ms = -1;
#pragma unroll
        for (int s = 0; s < kMaxNNSplit; s++)
            if (lex_less(md[s], mi[s], m, mix)) { m = md[s]; mix = mi[s]; ms = s; }
        float g = m;
        uint32_t gi = mix;
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            float ov = __shfl_xor(g, o);
            uint32_t oi = __shfl_xor(gi, o);
            if (lex_less(ov, oi, g, gi)) { g = ov; gi = oi; }
        }
        if (j == round) sel_i = gi;
        if (round == 7) kth = g;
        if (ms >= 0 && mix == gi && m == g) {
#pragma unroll
            for (int s = 0; s < kMaxNNSplit; s++)
                if (s == ms) { md[s] = INFINITY; mi[s] = kNoIdx; }
        }
    }
    float d = INFINITY;
    if (active && sel_i != kNoIdx) d = seq_dist128(qd + (size_t)q * 128, td + (size_t)sel_i * 128);
    float bd = d;
    uint32_t bi = sel_i;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        float od = __shfl_xor(bd, o);
        uint32_t oi = __shfl_xor(bi, o);
        if (lex_less(od, oi, bd, bi)) { bd = od; bi = oi; }
    }
    float sd = (sel_i == bi) ? INFINITY : d;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) sd = fminf(sd, __shfl_xor(sd, o));
    int ncand = sel_i != kNoIdx;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) ncand += __shfl_xor(ncand, o);
    if (active && j == 0) {
    // all targets are candidates when nt <= 8
    bool ok = ncand >= (int)min(nt, 8u);
    if (ok && nt > 8) {
        // the largest target norm of this call (a word of another generation: +inf, which
        // flags the query -- always safe)
        const unsigned long long tw = tmax[gen & 1];
        const float qq = qn[q], tt = (uint32_t)(tw >> 32) == gen ? __uint_as_float((uint32_t)tw) : INFINITY;
        // E bounds |value + |q|^2 - exact| for every pair: twice the pass-1 dot-product error
        // (< 7e-5 |q||t|) plus the f32 roundings of |q|^2, |t|^2 and the sequential exact
        // distance (each < 128 * 2^-24 of its magnitude) stay below 5e-5 (|q| + |t|)^2; the
        // factor 2e-4 leaves a 4x margin (a flagged query only costs the exact scan)
        const float E = 2.0e-4f * (qq + tt + 2.0f * sqrtf(qq * tt)) + 1e-30f;
        ok = (fminf(kth, last4) + qq) - E > sd;
    }
    best_idx[q] = bi == kNoIdx ? 0u : bi;
    best[q] = bd;
    second[q] = sd;
    if (!ok) s_flag[atomicAdd(&s_nflag, 1u)] = q;
    }
    __syncthreads();
    // exact scan of every target for this workgroup's flagged queries: per thread the
    // reference's sequential rule over its targets, then a tree reduction with the
    // order-independent rule ((value, index) best; second = smallest of the rest)
    const uint32_t nflag = s_nflag;
    for (uint32_t f = 0; f < nflag; f++) {
        const uint32_t fq = s_flag[f];
        float b = INFINITY, s2 = INFINITY;
        uint32_t bix = 0;
        for (uint32_t t = threadIdx.x; t < nt_scan; t += kRsThreads) {
            const float d2 = seq_dist128(qd + (size_t)fq * 128, td + (size_t)t * 128);
            if (d2 < b) { s2 = b; b = d2; bix = t; }
            else if (d2 < s2) s2 = d2;
        }
        sb[threadIdx.x] = b; ss[threadIdx.x] = s2; si[threadIdx.x] = bix;
        __syncthreads();
        for (int w = kRsThreads / 2; w > 0; w >>= 1) {
            if (threadIdx.x < (unsigned)w) {
                float B = sb[threadIdx.x], S = ss[threadIdx.x];
                uint32_t BI = si[threadIdx.x];
                const float b2 = sb[threadIdx.x + w], sv2 = ss[threadIdx.x + w];
                const uint32_t i2 = si[threadIdx.x + w];
                float other;
                if (b2 < B || (b2 == B && i2 < BI)) { other = B; B = b2; BI = i2; }
                else other = b2;
                sb[threadIdx.x] = B; ss[threadIdx.x] = fminf(fminf(S, sv2), other); si[threadIdx.x] = BI;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            best_idx[fq] = si[0];
            best[fq] = sb[0];
            second[fq] = ss[0];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kRsThreads) void nn_rescore_kernel(const Cand4 *part, int nsplit, const float *qd, const float *qn,
                                                         const uint32_t *nq_p, uint32_t nq_max, const float *td,
                                                         const uint32_t *nt_p, uint32_t nt_max,
                                                         const unsigned long long *tmax, uint32_t gen,
                                                         uint32_t *best_idx, float *best, float *second) {
    const uint32_t nt = *nt_p;
    nn_rescore_block(part, nsplit, nq_max, blockIdx.x, qd, qn, min(*nq_p, nq_max), td, nt, min(nt, nt_max), tmax,
                     gen, best_idx, best, second);
}

int nn_mfma_splits(size_t nq_max, size_t nt_max) {
    size_t qb = (nq_max + kMQ - 1) / kMQ, tt = (nt_max + kMT - 1) / kMT;
    if (qb == 0 || tt == 0) return 1;
    size_t ns = (512 + qb - 1) / qb;
    if (ns > tt) ns = tt;
    if (ns > (size_t)kMaxNNSplit) ns = kMaxNNSplit;
    return (int)(ns < 1 ? 1 : ns);
}

static size_t tn_rows(size_t nt_max) { return (nt_max + kMT - 1) / kMT * kMT; }

size_t nn_mfma_scratch_bytes(size_t nq_max, size_t nt_max) {
    size_t ns = (size_t)nn_mfma_splits(nq_max, nt_max);
    return ns * nq_max * 2 * sizeof(Cand4) + nq_max * sizeof(float) + tn_rows(nt_max) * sizeof(float) +
           (nq_max + tn_rows(nt_max)) * 128 * sizeof(uint32_t) + 512;
}

// Scratch layout: [tmax: 2 x u64, padded to 256 B][part: ns*nq_max*2 Cand4][qn: nq_max]
// [tn: padded nt_max][q split: nq_max*128 words][t split: padded nt_max*128 words]. The
// tmax words sit at a fixed offset so that calls of any size find the previous call's
// clearing there.
// Three launches: prep, MFMA distances, rescore (+ exact scans of flagged queries).
hipError_t launch_nn_mfma(const float *qd, const uint32_t *nq, size_t nq_max, const float *td, const uint32_t *nt,
                          size_t nt_max, void *scratch, uint32_t *best_idx, float *best, float *second, uint32_t gen,
                          hipStream_t st) {
    if (nq_max == 0) return hipSuccess;
    const int ns = nn_mfma_splits(nq_max, nt_max);
    char *p = static_cast<char *>(scratch);
    unsigned long long *tmax = reinterpret_cast<unsigned long long *>(p);
    p += 256;
    Cand4 *part = reinterpret_cast<Cand4 *>(p);
    p += (size_t)ns * nq_max * 2 * sizeof(Cand4);
    float *qn = reinterpret_cast<float *>(p);
    p += nq_max * sizeof(float);
    float *tn = reinterpret_cast<float *>(p);
    p += tn_rows(nt_max) * sizeof(float);
    uint32_t *qsplit = reinterpret_cast<uint32_t *>(p);
    p += nq_max * 128 * sizeof(uint32_t);
    uint32_t *tsplit = reinterpret_cast<uint32_t *>(p);
    const uint32_t trows = (uint32_t)tn_rows(nt_max);
    constexpr uint32_t kPrepBlock = 4 * kPrepRows;
    const uint32_t tblocks = (trows + kPrepBlock - 1) / kPrepBlock,
                   qblocks = (uint32_t)((nq_max + kPrepBlock - 1) / kPrepBlock);
    const PrepSide ts{td, nt, (uint32_t)nt_max, trows, tn, tsplit};
    const PrepSide qs{qd, nq, (uint32_t)nq_max, (uint32_t)nq_max, qn, qsplit};
    nn_prep_kernel<<<tblocks + qblocks, 256, 0, st>>>(ts, qs, tblocks, tmax, gen);
    dim3 grid((unsigned)((nq_max + kMQ - 1) / kMQ), (unsigned)ns);
    nn_mfma_kernel<<<grid, 256, 0, st>>>(qsplit, nq, (uint32_t)nq_max, tsplit, nt, (uint32_t)nt_max, tn, ns, part);
    nn_rescore_kernel<<<(unsigned)((nq_max + kRsQ - 1) / kRsQ), kRsThreads, 0, st>>>(part, ns, qd, qn, nq, (uint32_t)nq_max, td, nt,
                                                                      (uint32_t)nt_max, tmax, gen, best_idx, best,
                                                                      second);
    return hipGetLastError();
}

// ---- grouped form: many (query set, target set) problems in one chain -------------------

// The sets' tables and bounds: a problem whose set index is out of range, or the part of a set
// past its array's rows or past cap, counts as empty.
struct SetsView {
    const ps_feature_set *qsets, *tsets;
    uint32_t n_qsets, n_tsets, cap, q_rows, t_rows;
};
__device__ __forceinline__ uint32_t set_len(const ps_feature_set &S, uint32_t cap, uint32_t rows) {
    return S.offset >= rows ? 0u : min(min(S.count, cap), rows - S.offset);
}
__device__ __forceinline__ void resolve_sets(const ps_match_problem &P, const SetsView &V, ps_feature_set &Q,
                                             ps_feature_set &T, uint32_t &nq, uint32_t &nt) {
    Q = P.query_set < V.n_qsets ? V.qsets[P.query_set] : ps_feature_set{0u, 0u};
    T = P.target_set < V.n_tsets ? V.tsets[P.target_set] : ps_feature_set{0u, 0u};
    nq = set_len(Q, V.cap, V.q_rows);
    nt = set_len(T, V.cap, V.t_rows);
}

// Prep over feature sets: norms and bf16 splits of rows [offset, offset + min(count, cap)) of
// every set (rows outside the sets are never read by the later passes). On the target side
// the largest norm goes to tmax[gen & 1] and workgroup (0, 0) zeroes the other word, as in
// nn_prep_kernel. Grid (cap / 64 blocks, sets).
__global__ __launch_bounds__(256) void nn_prep_sets_kernel(const float *x, const ps_feature_set *sets, uint32_t cap,
                                                           uint32_t rows, float *norm, uint32_t *split,
                                                           int target_side, unsigned long long *tmax, uint32_t gen) {
    __shared__ float s_max[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const ps_feature_set S = sets[blockIdx.y];
    const uint32_t n = set_len(S, cap, rows);
    const uint32_t lb = blockIdx.x * 64;
    float wmax = 0.0f;
    if (lb < n) {
        float2 v[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint32_t l = lb + w * 16 + k, r = S.offset + l;
            v[k] = l < n && r < rows ? reinterpret_cast<const float2 *>(x + (size_t)r * 128)[lane]
                                     : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint32_t l = lb + w * 16 + k, r = S.offset + l;
            const bool ok = l < n && r < rows;
            float s = v[k].x * v[k].x + v[k].y * v[k].y;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
            if (ok) {
                if (lane == 0) norm[r] = s;
                wmax = fmaxf(wmax, isfinite(s) ? s : INFINITY);
                const uint32_t hx = bf16_bits_rne(v[k].x), hy = bf16_bits_rne(v[k].y);
                const float rx = v[k].x - __uint_as_float(hx << 16), ry = v[k].y - __uint_as_float(hy << 16);
                split[(size_t)r * 128 + lane] = hx | (hy << 16);
                split[(size_t)r * 128 + 64 + lane] = bf16_bits_rne(rx) | (bf16_bits_rne(ry) << 16);
            }
        }
    }
    if (!target_side) return;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) tmax[(gen + 1) & 1] = 0ull;
    if (lane == 0) s_max[w] = wmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float m = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
        // the word only grows during a generation: a workgroup whose maximum does not exceed
        // the value it reads skips the atomic (hundreds of workgroups, one address)
        const unsigned long long v = ((unsigned long long)gen << 32) | __float_as_uint(m);
        if (__hip_atomic_load(tmax + (gen & 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v)
            atomicMax(tmax + (gen & 1), v);
    }
}

// Pass 1 per problem: grid (query blocks, problems, target splits); lists of problem k at
// part[((k * nsplit + split) * cap + q) * 2 + half], indices relative to the target set.
__global__ __launch_bounds__(256) void nn_mfma_sets_kernel(const uint32_t *qsplit, const uint32_t *tsplit,
                                                           const float *tn, SetsView V, const ps_match_problem *probs,
                                                           int nsplit, Cand4 *part) {
    const uint32_t cap = V.cap;
    ps_feature_set Q, T;
    uint32_t nq, nt;
    resolve_sets(probs[blockIdx.y], V, Q, T, nq, nt);
    const uint32_t qb = blockIdx.x * kMQ;
    if (qb >= nq) return;
    uint32_t te;
    const uint32_t tb = nn_tiles_begin(nt, nsplit, blockIdx.z, te);
    float kd[4];
    uint32_t ki[4];
    nn_mfma_scan(qsplit + (size_t)Q.offset * 128, nq, qb, tsplit + (size_t)T.offset * 128, nt, tn + T.offset, tb, te,
                 kd, ki);
    const int lane = threadIdx.x & 63;
    const uint32_t q = qb + (threadIdx.x >> 6) * 32 + (lane & 31);
    if (q < nq) {
        Cand4 c;
#pragma unroll
        for (int j = 0; j < 4; j++) { c.d[j] = kd[j]; c.i[j] = ki[j]; }
        part[(((size_t)blockIdx.y * nsplit + blockIdx.z) * cap + q) * 2 + (lane >> 5)] = c;
    }
}

// Pass 2 per problem: grid (kRsQ-query blocks, problems); results at [k * cap + q].
__global__ __launch_bounds__(kRsThreads) void nn_rescore_sets_kernel(const Cand4 *part, int nsplit, const float *qd,
                                                                     const float *qn, const float *td, SetsView V,
                                                              const ps_match_problem *probs,
                                                              const unsigned long long *tmax, uint32_t gen,
                                                              uint32_t *best_idx, float *best, float *second) {
    const uint32_t k = blockIdx.y, cap = V.cap;
    ps_feature_set Q, T;
    uint32_t nq, nt;
    resolve_sets(probs[k], V, Q, T, nq, nt);
    if (blockIdx.x * kRsQ >= nq) return;
    const size_t o = (size_t)k * cap;
    nn_rescore_block(part + (size_t)k * nsplit * cap * 2, nsplit, cap, blockIdx.x, qd + (size_t)Q.offset * 128,
                     qn + Q.offset, nq, td + (size_t)T.offset * 128, nt, nt, tmax, gen, best_idx + o, best + o,
                     second + o);
}

// Ratio test (pipeline.rs:1788-1805) and ordered compaction per problem: one workgroup walks
// its queries 256 at a time, the output position a ballot prefix over the accepted ones.
__global__ __launch_bounds__(256) void match_sets_finish_kernel(const ps_feature *qf, const ps_feature *tf, SetsView V,
                                                                const ps_match_problem *probs,
                                                                const uint32_t *best_idx, const float *best,
                                                                const float *second, float ratio_thr,
                                                                ps_feature_match *out, uint32_t *out_counts) {
    __shared__ uint32_t s_wave[4];
    const uint32_t k = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, cap = V.cap;
    const ps_match_problem P = probs[k];
    ps_feature_set Q, T;
    uint32_t nq, nt;
    resolve_sets(P, V, Q, T, nq, nt);
    const size_t o = (size_t)k * cap;
    const ps_feature *QF = qf + Q.offset, *TF = tf + T.offset;
    ps_feature_match *dst = out + P.out_offset;
    uint32_t base = 0;
    for (uint32_t c = 0; c < nq; c += 256) {
        const uint32_t q = c + tid;
        bool acc = false;
        if (q < nq && nt >= 2) {
            const float bd = best[o + q], sd = second[o + q];
            if (sd > 1e-12f && isfinite(bd)) acc = sqrtf(bd) / sqrtf(sd) < ratio_thr;
        }
        const uint64_t m = __ballot(acc);
        if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = base, tot = 0;
#pragma unroll
        for (uint32_t w = 0; w < 4; w++) {
            if (w < wave) before += s_wave[w];
            tot += s_wave[w];
        }
        if (acc) {
            const uint32_t pos = before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            const float d1 = sqrtf(best[o + q]), d2 = sqrtf(second[o + q]);
            const uint32_t t = best_idx[o + q];
            const ps_feature a = QF[q], b = TF[t];
            ps_feature_match r;
            r.query_index = q;
            r.target_index = t;
            r.qx = a.x; r.qy = a.y; r.tx = b.x; r.ty = b.y;
            r.distance = d1;
            r.ratio = d1 / d2;
            r.dx = b.x - a.x;
            r.dy = b.y - a.y;
            dst[pos] = r;
        }
        base += tot;
        __syncthreads();
    }
    if (tid == 0) out_counts[k] = base;
}

namespace {

int sets_splits(const MatchSetsArgs &a) {
    const size_t qb = (a.cap + kMQ - 1) / kMQ, tt = (a.cap + kMT - 1) / kMT;
    const size_t blocks = qb * a.n_probs;
    if (blocks == 0 || tt == 0) return 1;
    size_t ns = (2048 + blocks - 1) / blocks;
    if (ns > tt) ns = tt;
    if (ns > (size_t)kMaxNNSplit) ns = kMaxNNSplit;
    return (int)(ns < 1 ? 1 : ns);
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

struct SetsLayout {
    size_t part, qn, tn, qs, ts, bidx, best, second, total;
};

SetsLayout sets_layout(const MatchSetsArgs &a) {
    SetsLayout L;
    const size_t ns = (size_t)sets_splits(a), pc = (size_t)a.n_probs * a.cap;
    size_t o = 256;   // the two norm-bound words
    L.part = o;
    o += al256(ns * pc * 2 * sizeof(Cand4));
    L.qn = o;
    o += al256(a.q_rows * sizeof(float));
    L.tn = a.shared ? L.qn : o;
    if (!a.shared) o += al256(a.t_rows * sizeof(float));
    L.qs = o;
    o += al256(a.q_rows * 128 * sizeof(uint32_t));
    L.ts = a.shared ? L.qs : o;
    if (!a.shared) o += al256(a.t_rows * 128 * sizeof(uint32_t));
    L.bidx = o;
    o += al256(pc * sizeof(uint32_t));
    L.best = o;
    o += al256(pc * sizeof(float));
    L.second = o;
    o += al256(pc * sizeof(float));
    L.total = o;
    return L;
}

}  // namespace

size_t match_sets_scratch_bytes(const MatchSetsArgs &a) { return sets_layout(a).total; }

// Four launches for all problems: prep (one per distinct side), MFMA distances, rescore (+
// exact scans of flagged queries), ratio test + compaction.
hipError_t launch_match_sets(const MatchSetsArgs &a, void *scratch, uint32_t gen, ps_feature_match *out,
                             uint32_t *out_counts, hipStream_t st) {
    if (a.n_probs == 0) return hipSuccess;
    const SetsLayout L = sets_layout(a);
    const int ns = sets_splits(a);
    char *b = static_cast<char *>(scratch);
    unsigned long long *tmax = reinterpret_cast<unsigned long long *>(b);
    Cand4 *part = reinterpret_cast<Cand4 *>(b + L.part);
    float *qn = reinterpret_cast<float *>(b + L.qn), *tn = reinterpret_cast<float *>(b + L.tn);
    uint32_t *qs = reinterpret_cast<uint32_t *>(b + L.qs), *ts = reinterpret_cast<uint32_t *>(b + L.ts);
    uint32_t *bidx = reinterpret_cast<uint32_t *>(b + L.bidx);
    float *best = reinterpret_cast<float *>(b + L.best), *second = reinterpret_cast<float *>(b + L.second);
    const uint32_t rb = (a.cap + 63) / 64;
    const SetsView V{a.qsets, a.tsets, a.n_qsets, a.n_tsets, a.cap, (uint32_t)a.q_rows, (uint32_t)a.t_rows};
    if (a.cap > 0) {
        if (!a.shared && a.n_qsets > 0)
            nn_prep_sets_kernel<<<dim3(rb, a.n_qsets), 256, 0, st>>>(a.qd, a.qsets, a.cap, (uint32_t)a.q_rows, qn, qs,
                                                                     0, tmax, gen);
        if (a.n_tsets > 0)
            nn_prep_sets_kernel<<<dim3(rb, a.n_tsets), 256, 0, st>>>(a.td, a.tsets, a.cap, (uint32_t)a.t_rows, tn, ts,
                                                                     1, tmax, gen);
        nn_mfma_sets_kernel<<<dim3((a.cap + kMQ - 1) / kMQ, a.n_probs, ns), 256, 0, st>>>(qs, ts, tn, V, a.probs, ns,
                                                                                          part);
        nn_rescore_sets_kernel<<<dim3((a.cap + kRsQ - 1) / kRsQ, a.n_probs), kRsThreads, 0, st>>>(part, ns, a.qd, qn, a.td, V, a.probs,
                                                                                   tmax, gen, bidx, best, second);
    }
    match_sets_finish_kernel<<<a.n_probs, 256, 0, st>>>(a.qf, a.tf, V, a.probs, bidx, best, second, a.ratio, out,
                                                        out_counts);
    return hipGetLastError();
}

}  // namespace ps
