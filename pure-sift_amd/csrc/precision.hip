// Precision form of the ratio matcher's nearest / second-nearest search
// (pipeline.rs:1764-1786 semantics for the final decision):
//
//   1. prep: each descriptor row is rounded to BF16 (round to nearest even) or to FP8
//      E4M3 with a per-row power-of-two scale (the row's max |x| mapped to <= 448), and its
//      f32 squared norm is stored (+inf for padding rows, so padding never ranks);
//   2. topk: an MFMA GEMM of targets (A, 32 rows per tile) against queries (B, 32 columns
//      per tile, held in VGPRs for the whole scan) with FP32 accumulation; the epilogue turns
//      each dot product into the ranking value |t|^2 - 2 q.t and keeps, per lane, the 4
//      smallest (value, index) of the targets it sees. Each query is seen by 2 lanes per
//      target split, so every split leaves 8 candidates per query;
//   3. rescore: the 8 smallest approximate candidates over all splits are rescored with the
//      exact sequential f32 distance of the FP32 path and the best / second are taken in
//      (distance, index) order, which is the FP32 path's decision whenever its best and
//      second-best targets are among the 8.
//
// MFMA forms: v_mfma_f32_32x32x16_bf16 (8 k-steps over 128) and the gfx950 scaled
// v_mfma_f32_32x32x64_f8f6f4 with E4M3 operands (2 k-steps) whose per-lane E8M0 scales are
// the rows' power-of-two scales, so the accumulator already holds the unscaled dot product.
// A and B lanes read the same k positions (lane half h takes k in [64h, 64h + 64) of the
// row), so the k order inside an instruction does not affect the sum.
#include <algorithm>
#include <atomic>
#include <type_traits>

#include "ps_internal.h"

namespace ps {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kPQB = 256;        // queries per workgroup (4 waves x 64)
constexpr int kPTT = 64;         // targets per LDS tile
constexpr int kPCand = 4;        // candidates kept per lane
constexpr int kPRescore = 8;     // candidates rescored per query
// Ranking keys are |t|^2 + kKeyBias - 2 q.t: positive for unit descriptors and far from the
// subnormal range, so replacing their low mantissa bits (see the top-k epilogue) keeps them
// ordinary floats near the value
constexpr float kKeyBias = 8.0f;

template <bool FP8>
struct Prec {
    // BF16 rows carry 144 values: the 128 components and the bias-shifted norm term (see
    // prep_bf16_kernel); FP8 rows the 128 components
    static constexpr int kRowBytes = FP8 ? 128 : 288;
    static constexpr int kLdsStride = kRowBytes + (FP8 ? 32 : 16);   // pad: rows start in different banks
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ uint32_t bf16_rne(float x) {
    uint32_t b = __float_as_uint(x);
    return (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;
}

// One wave per row of 144 BF16 values (round to nearest even); rows >= *n (padding) become
// zero with norm +inf. The MFMA's k = 144 folds the ranking value's norm term in: a query row
// is (q, 1, 1, 1, 0...), a target row (-2 t, n1, n2, n3, 0...) with n1 + n2 + n3 the
// target's |t|^2 + kKeyBias split into three BF16 terms (each the rounded remainder of the
// previous, so their sum is within 2^-24 relative of it), so the FP32 accumulator holds
// (|t|^2 + kKeyBias) - 2 q.t directly; scaling by -2 is exact in BF16. Padding targets get
// (0, +inf, 0, 0, ...), a +inf value.
__global__ __launch_bounds__(256) void prep_bf16_kernel(const float *x, const uint32_t *n_p, uint32_t rows,
                                                        bool target, uint32_t *out, float *norm) {
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    const bool valid = r < *n_p;
    float2 v = valid ? reinterpret_cast<const float2 *>(x + (size_t)r * 128)[lane] : make_float2(0.f, 0.f);
    const float sc = target ? -2.0f : 1.0f;
    uint32_t *row = out + (size_t)r * 72;
    row[lane] = bf16_rne(sc * v.x) | (bf16_rne(sc * v.y) << 16);
    const float s = wave_sum(v.x * v.x + v.y * v.y);
    if (lane == 0) norm[r] = valid ? s : INFINITY;
    if (lane < 8) {
        uint32_t e[2] = {0u, 0u};   // values 128 + 2 lane, 128 + 2 lane + 1
        if (lane < 2) {
            uint32_t n[3];
            if (!target) {
                n[0] = n[1] = n[2] = 0x3f80u;   // 1.0
            } else if (!valid) {
                n[0] = 0x7f80u;                 // +inf
                n[1] = n[2] = 0u;
            } else {
                const float t = s + kKeyBias;
                n[0] = bf16_rne(t);
                const float r1 = t - __uint_as_float(n[0] << 16);
                n[1] = bf16_rne(r1);
                n[2] = bf16_rne(r1 - __uint_as_float(n[1] << 16));
            }
            e[0] = n[2 * lane];
            e[1] = lane == 0 ? n[1] : 0u;
        }
        row[64 + lane] = e[0] | (e[1] << 16);
    }
}

__global__ __launch_bounds__(256) void prep_fp8_kernel(const float *x, const uint32_t *n_p, uint32_t rows,
                                                       uint32_t *out, float *norm, uint8_t *exps) {
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    const bool valid = r < *n_p;
    float2 v = valid ? reinterpret_cast<const float2 *>(x + (size_t)r * 128)[lane] : make_float2(0.f, 0.f);
    const float amax = wave_max(fmaxf(fabsf(v.x), fabsf(v.y)));
    // smallest e with amax * 2^-e <= 448
    int e = 0;
    if (amax > 0.0f) {
        int ex;
        float m = frexpf(amax / 448.0f, &ex);
        e = (m == 0.5f) ? ex - 1 : ex;
    }
    const float a = ldexpf(v.x, -e), b = ldexpf(v.y, -e);
    const uint32_t pk = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false) & 0xffffu;
    const uint32_t hi = __shfl_down(pk, 1);
    if ((lane & 1) == 0) out[(size_t)r * 32 + (lane >> 1)] = pk | (hi << 16);
    float s = wave_sum(v.x * v.x + v.y * v.y);
    if (lane == 0) {
        norm[r] = valid ? s : INFINITY;
        exps[r] = (uint8_t)(127 + e);
    }
}

__device__ __forceinline__ void cand_insert(float v, uint32_t idx, float (&bv)[kPCand], uint32_t (&bi)[kPCand]) {
#pragma unroll
    for (int j = 0; j < kPCand; j++) {
        bool lt = v < bv[j] || (v == bv[j] && idx < bi[j]);
        float tv = lt ? bv[j] : v;
        uint32_t ti = lt ? bi[j] : idx;
        bv[j] = lt ? v : bv[j];
        bi[j] = lt ? idx : bi[j];
        v = tv;
        idx = ti;
    }
}

struct PCand {
    float v;
    uint32_t idx;
};

template <bool FP8>
__global__ __launch_bounds__(256, 2) void prec_topk_kernel(const uint8_t *qlow, const uint8_t *qexp,
                                                           const uint8_t *tlow, const float *tn,
                                                           const uint8_t *texp, uint32_t nq_pad,
                                                           uint32_t tiles_per_split, uint32_t ntiles,
                                                           PCand *part) {
    using P = Prec<FP8>;
    __shared__ __attribute__((aligned(32))) uint8_t s_t[2][kPTT * P::kLdsStride];
    __shared__ __attribute__((aligned(16))) float s_tn[2][kPTT];
    __shared__ uint32_t s_te[2][kPTT];
    __shared__ int s_pad[2];   // the tile holds padding rows (a suffix of the rows, so: its last row does)
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, h = l >> 5, c = l & 31;
    const uint32_t tile0 = blockIdx.y * tiles_per_split;
    const uint32_t tile1 = min(ntiles, tile0 + tiles_per_split);
    PCand *out = part + ((size_t)blockIdx.y * nq_pad) * (2 * kPCand);
    const uint32_t qbase = blockIdx.x * kPQB + w * 64;

    // query operands (B), both 32-query sets of this wave
    using BT = typename std::conditional<FP8, i32x8, bf16x8>::type;
    constexpr int KS = P::kRowBytes / 2 / (int)sizeof(BT);   // MFMA k-steps per lane half (9 BF16, 2 FP8)
    BT b[2][KS];
    int qe[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const uint8_t *row = qlow + (size_t)(qbase + s * 32 + c) * P::kRowBytes + h * (P::kRowBytes / 2);
#pragma unroll
        for (int t = 0; t < KS; t++) b[s][t] = *reinterpret_cast<const BT *>(row + t * sizeof(BT));
        qe[s] = FP8 ? (int)qexp[qbase + s * 32 + c] : 127;
    }
    float bv[2][kPCand];
    uint32_t bi[2][kPCand];
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int j = 0; j < kPCand; j++) {
            bv[s][j] = INFINITY;
            bi[s][j] = 0xffffffffu;
        }

    // cooperative tile load: kPTT rows x kRowBytes, 16 B per load, staged through VGPRs
    // (kLoads per thread: 5 for BF16 rows, the last by half the threads, 2 for FP8)
    constexpr int kChunks = kPTT * P::kRowBytes / 16;
    constexpr int kLoads = (kChunks + 255) / 256;
    constexpr int kSegs = P::kRowBytes / 16;
    static_assert(kLoads <= 5, "tile load split");
    uint4 pre0 = {}, pre1 = {}, pre2 = {}, pre3 = {}, pre4 = {};
    float pre_tn = 0.0f;
    uint32_t pre_te = 127u;
// chunk I (16 bytes) of the tile for this thread, in the named register preI
#define PREC_LD1(I)                                                                                 \
    do {                                                                                            \
        if constexpr ((I) < kLoads)                                                                 \
            if ((I) * 256 + tid < kChunks) pre##I = src_[(I) * 256 + tid];                          \
    } while (0)
#define PREC_PUT1(I)                                                                                \
    do {                                                                                            \
        if constexpr ((I) < kLoads)                                                                 \
            if ((I) * 256 + tid < kChunks) PREC_PUT((I), pre##I);                                   \
    } while (0)
#define PREC_LOAD_TILE(TILE)                                                                        \
    do {                                                                                            \
        const uint4 *src_ = reinterpret_cast<const uint4 *>(tlow + (size_t)(TILE) * kPTT * P::kRowBytes); \
        PREC_LD1(0); PREC_LD1(1); PREC_LD1(2); PREC_LD1(3); PREC_LD1(4);                           \
        if (tid < kPTT) {                                                                           \
            pre_tn = tn[(size_t)(TILE) * kPTT + tid];                                               \
            pre_te = FP8 ? (uint32_t)texp[(size_t)(TILE) * kPTT + tid] : 127u;                      \
        }                                                                                           \
    } while (0)
#define PREC_PUT(I, V)                                                                              \
    *reinterpret_cast<uint4 *>(&s_t[buf_st][((tid + (I) * 256) / kSegs) * P::kLdsStride +             \
                                            ((tid + (I) * 256) % kSegs) * 16]) = (V)
#define PREC_STORE_TILE(BUF)                                                                        \
    do {                                                                                            \
        const int buf_st = (BUF);                                                                   \
        PREC_PUT1(0); PREC_PUT1(1); PREC_PUT1(2); PREC_PUT1(3); PREC_PUT1(4);                      \
        if (tid < kPTT) {                                                                           \
            s_tn[buf_st][tid] = pre_tn + kKeyBias;                                                  \
            s_te[buf_st][tid] = pre_te;                                                             \
            if (tid == kPTT - 1) s_pad[buf_st] = isinf(pre_tn) ? 1 : 0;                             \
        }                                                                                           \
    } while (0)

    if (tile0 < tile1) {
        PREC_LOAD_TILE(tile0);
        PREC_STORE_TILE(0);
    }
    __syncthreads();
    int buf = 0;
    for (uint32_t tile = tile0; tile < tile1; tile++) {
        const bool more = tile + 1 < tile1;
        if (more) PREC_LOAD_TILE(tile + 1);
        f32x16 acc[2][2];
        BT a_op[2][KS];
        int te_op[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const uint8_t *arow = &s_t[buf][(u * 32 + c) * P::kLdsStride + h * (P::kRowBytes / 2)];
            te_op[u] = (int)s_te[buf][u * 32 + c];
#pragma unroll
            for (int t = 0; t < KS; t++) a_op[u][t] = *reinterpret_cast<const BT *>(arow + t * sizeof(BT));
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
#pragma unroll
            for (int s = 0; s < 2; s++) acc[s][u] = f32x16{};
            const int te = te_op[u];
#pragma unroll
            for (int t = 0; t < KS; t++) {
                const BT a = a_op[u][t];
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    if constexpr (FP8)
                        acc[s][u] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b[s][t], acc[s][u], 0, 0, 0,
                                                                                    te, 0, qe[s]);
                    else
                        acc[s][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[s][t], acc[s][u], 0, 0, 0);
                }
            }
        }
        // epilogue: D lane l, register r -> query column c, target row 8(r/4) + 4h + r%4.
        // Each of the lane's 32 values per query set becomes a key: the ranking value offset by
        // kKeyBias, (|t|^2 + kKeyBias) - 2 q.t, with its 5 low mantissa bits replaced by the
        // value's position p = 16u + r (a change below 2^-18 relative, far inside the
        // approximation the FP32 rescore removes). A lane's keys are then distinct, so a tree of
        // fminf (which returns one of its operands) yields the smallest key and with it its
        // target; the same tournament (of 3-input min / median operations) carries the second
        // smallest. When the second smallest is not below the lane's 4th-best key, at most the
        // smallest enters the list (one insertion, no per-value branch: almost every (tile,
        // lane) after a split's first tiles); when exactly two keys are below it, those two
        // enter. Otherwise, and in the tile holding padding rows (+inf norms), every value is
        // tried in turn.
        const uint32_t tbase = tile * kPTT;
        const bool pad_tile = s_pad[buf] != 0;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            float key[32];
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const float4 t4 = *reinterpret_cast<const float4 *>(&s_tn[buf][u * 32 + 8 * rr + 4 * h]);
                    const float tv[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
                    for (int q4 = 0; q4 < 4; q4++) {
                        const int p = u * 16 + rr * 4 + q4;
                        // BF16: the MFMA already formed (|t|^2 + kKeyBias) - 2 q.t
                        const float v = FP8 ? __builtin_fmaf(-2.0f, acc[s][u][rr * 4 + q4], tv[q4])
                                            : acc[s][u][rr * 4 + q4];
                        key[p] = __uint_as_float((__float_as_uint(v) & ~31u) | (uint32_t)p);
                    }
                }
            const float thr = bv[s][kPCand - 1];
            bool slow = pad_tile;
            if (!pad_tile) {
                // 3-way tournament: triples give (min3, med3), merging sorted (first, second)
                // groups a, b, c gives min3(a1, b1, c1) and min(med3(a1, b1, c1), min3(a2, b2, c2))
                float f[11], g[11];
#pragma unroll
                for (int i = 0; i < 10; i++) {
                    f[i] = fminf(fminf(key[3 * i], key[3 * i + 1]), key[3 * i + 2]);
                    g[i] = __builtin_amdgcn_fmed3f(key[3 * i], key[3 * i + 1], key[3 * i + 2]);
                }
                f[10] = fminf(key[30], key[31]);
                g[10] = fmaxf(key[30], key[31]);
                auto merge3 = [](float a1, float a2, float b1, float b2, float c1, float c2, float &n1, float &n2) {
                    n1 = fminf(fminf(a1, b1), c1);
                    n2 = fminf(__builtin_amdgcn_fmed3f(a1, b1, c1), fminf(fminf(a2, b2), c2));
                };
                auto merge2 = [](float a1, float a2, float b1, float b2, float &n1, float &n2) {
                    n1 = fminf(a1, b1);
                    n2 = __builtin_amdgcn_fmed3f(a1, b1, fminf(a2, b2));
                };
                float f1[4], g1[4];
#pragma unroll
                for (int i = 0; i < 3; i++)
                    merge3(f[3 * i], g[3 * i], f[3 * i + 1], g[3 * i + 1], f[3 * i + 2], g[3 * i + 2], f1[i], g1[i]);
                merge2(f[9], g[9], f[10], g[10], f1[3], g1[3]);
                float f2, g2, lo0, hi0;
                merge3(f1[0], g1[0], f1[1], g1[1], f1[2], g1[2], f2, g2);
                merge2(f2, g2, f1[3], g1[3], lo0, hi0);
                auto target = [&](float k) {
                    const uint32_t p = __float_as_uint(k) & 31u;
                    return tbase + (p >> 4) * 32 + 8 * ((p & 15) >> 2) + 4 * h + (p & 3);
                };
                if (hi0 < thr) {
                    // two or more new candidates (mostly in a split's first tiles): exactly two
                    // are the two smallest keys; more take the per-value loop
                    int below = 0;
#pragma unroll
                    for (int p = 0; p < 32; p++) below += key[p] < thr ? 1 : 0;
                    if (below > 2) {
                        slow = true;
                    } else {
                        cand_insert(lo0, target(lo0), bv[s], bi[s]);
                        cand_insert(hi0, target(hi0), bv[s], bi[s]);
                    }
                } else if (lo0 < thr) {
                    cand_insert(lo0, target(lo0), bv[s], bi[s]);
                }
            }
            if (slow) {
#pragma unroll
                for (int p = 0; p < 32; p++) {
                    // padding rows: +inf norms give +inf (or NaN once marked) keys, which never rank
                    const float k = key[p];
                    if (k < bv[s][kPCand - 1])
                        cand_insert(k, tbase + (p >> 4) * 32 + 8 * ((p & 15) >> 2) + 4 * h + (p & 3), bv[s], bi[s]);
                }
            }
        }
        if (more) {
            __syncthreads();
            PREC_STORE_TILE(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }
#undef PREC_LD1
#undef PREC_PUT1
#undef PREC_LOAD_TILE
#undef PREC_PUT
#undef PREC_STORE_TILE
#pragma unroll
    for (int s = 0; s < 2; s++) {
        PCand *o = out + (size_t)(qbase + s * 32 + c) * (2 * kPCand) + h * kPCand;
#pragma unroll
        for (int j = 0; j < kPCand; j++) o[j] = PCand{bv[s][j], bi[s][j]};
    }
}

// 8 lanes per query: the 8 smallest (value, index) of the 8 * nsplit candidates, their exact
// sequential f32 distances, then best / second in (distance, index) order.
__global__ __launch_bounds__(256) void prec_rescore_kernel(const PCand *part, uint32_t nsplit, uint32_t nq_pad,
                                                           const uint32_t *nq_p, const float *qd, const float *td,
                                                           const float *qn, uint32_t *best_idx, float *best,
                                                           float *second, float *approx_best) {
    const uint32_t gq = blockIdx.x * 32 + (threadIdx.x >> 3);
    const int j = threadIdx.x & 7;
    const uint32_t nq = *nq_p;
    const bool active = gq < nq;
    // candidates of this query: lane j holds entries j, j + 8, ... (one per split)
    constexpr int kMaxSplit = 16;
    PCand mine[kMaxSplit];
#pragma unroll
    for (int s = 0; s < kMaxSplit; s++) {
        mine[s] = PCand{INFINITY, 0xffffffffu};
        if (active && (uint32_t)s < nsplit) mine[s] = part[((size_t)s * nq_pad + gq) * 8 + j];
    }
    // 8 rounds of group-wide minimum; the round-r winner is rescored by lane r
    PCand sel{INFINITY, 0xffffffffu};
    for (int round = 0; round < kPRescore; round++) {
        PCand m{INFINITY, 0xffffffffu};
        int ms = -1;
#pragma unroll
        for (int s = 0; s < kMaxSplit; s++) {
            bool lt = mine[s].v < m.v || (mine[s].v == m.v && mine[s].idx < m.idx);
            if (lt) {
                m = mine[s];
                ms = s;
            }
        }
        PCand g = m;
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            float ov = __shfl_xor(g.v, o);
            uint32_t oi = __shfl_xor(g.idx, o);
            if (ov < g.v || (ov == g.v && oi < g.idx)) {
                g.v = ov;
                g.idx = oi;
            }
        }
        if (j == round) sel = g;
        // the owner drops its entry (indices are unique among valid entries)
        if (ms >= 0 && m.idx == g.idx && m.v == g.v) {
#pragma unroll
            for (int s = 0; s < kMaxSplit; s++)
                if (s == ms) mine[s] = PCand{INFINITY, 0xffffffffu};
        }
    }
    float d = INFINITY;
    uint32_t di = 0xffffffffu;
    if (active && sel.idx != 0xffffffffu) {
        d = seq_dist128(qd + (size_t)gq * 128, td + (size_t)sel.idx * 128);
        di = sel.idx;
    }
    // best and second of the 8 in (distance, index) order
    float bd = d, bv = sel.v;
    uint32_t bidx = di;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        float od = __shfl_xor(bd, o), ov = __shfl_xor(bv, o);
        uint32_t oi = __shfl_xor(bidx, o);
        if (od < bd || (od == bd && oi < bidx)) {
            bd = od;
            bidx = oi;
            bv = ov;
        }
    }
    float sd = (di == bidx) ? INFINITY : d;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) sd = fminf(sd, __shfl_xor(sd, o));
    if (active && j == 0) {
        best_idx[gq] = bidx == 0xffffffffu ? 0u : bidx;
        best[gq] = bd;
        second[gq] = sd;
        if (approx_best) approx_best[gq] = (bv - kKeyBias) + qn[gq];
    }
}

struct PrecLayout {
    size_t nq_pad, nt_pad, row_bytes;
    size_t off_qlow, off_tlow, off_qn, off_tn, off_qe, off_te, off_part, total;
    uint32_t nsplit, tiles_per_split, ntiles;
};

// compute units of the current device (256 on MI355X), cached per device
int device_cus() {
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cached[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cached[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

PrecLayout prec_layout(int prec, size_t nqm, size_t ntm) {
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    PrecLayout L{};
    L.nq_pad = (nqm + kPQB - 1) / kPQB * kPQB;
    L.nt_pad = (ntm + kPTT - 1) / kPTT * kPTT;
    L.row_bytes = prec == 2 ? 128 : 288;
    L.ntiles = (uint32_t)(L.nt_pad / kPTT);
    const uint32_t qblocks = (uint32_t)(L.nq_pad / kPQB);
    // target splits: the smallest count whose workgroups fill their last round of resident
    // slots (two per CU) to at least 90%, else the best-filled count (at most 16 splits, at
    // least 4 tiles each)
    const uint32_t slots = 2u * (uint32_t)device_cus();
    const uint32_t ns_max = std::min<uint32_t>(16, std::max<uint32_t>(1, L.ntiles / 4));
    uint32_t ns = 1;
    double best_fill = -1.0;
    for (uint32_t k = 1; k <= ns_max; k++) {
        const uint64_t wgs = (uint64_t)qblocks * k;
        const uint64_t rounds = (wgs + slots - 1) / slots;
        const double fill = (double)wgs / (double)(rounds * slots);
        if (fill >= 0.9) {
            ns = k;
            break;
        }
        if (fill > best_fill) {
            best_fill = fill;
            ns = k;
        }
    }
    L.tiles_per_split = (L.ntiles + ns - 1) / ns;
    L.nsplit = (L.ntiles + L.tiles_per_split - 1) / L.tiles_per_split;
    size_t o = 0;
    L.off_qlow = o; o += al(L.nq_pad * L.row_bytes);
    L.off_tlow = o; o += al(L.nt_pad * L.row_bytes);
    L.off_qn = o; o += al(L.nq_pad * 4);
    L.off_tn = o; o += al(L.nt_pad * 4);
    L.off_qe = o; o += al(L.nq_pad);
    L.off_te = o; o += al(L.nt_pad);
    L.off_part = o; o += al((size_t)L.nsplit * L.nq_pad * 2 * kPCand * sizeof(PCand));
    L.total = o;
    return L;
}

}  // namespace

size_t nn_precision_scratch_bytes(int prec, size_t nq_max, size_t nt_max) {
    return prec_layout(prec, std::max<size_t>(nq_max, 1), std::max<size_t>(nt_max, 1)).total;
}

hipError_t launch_nn_precision(int prec, const float *qd, const uint32_t *nq, size_t nq_max, const float *td,
                               const uint32_t *nt, size_t nt_max, void *scratch, uint32_t *best_idx, float *best,
                               float *second, float *approx_best, hipStream_t st) {
    if (prec != 1 && prec != 2) return hipErrorInvalidValue;
    const size_t nqm = std::max<size_t>(nq_max, 1), ntm = std::max<size_t>(nt_max, 1);
    const PrecLayout L = prec_layout(prec, nqm, ntm);
    char *base = static_cast<char *>(scratch);
    uint8_t *qlow = reinterpret_cast<uint8_t *>(base + L.off_qlow);
    uint8_t *tlow = reinterpret_cast<uint8_t *>(base + L.off_tlow);
    float *qn = reinterpret_cast<float *>(base + L.off_qn);
    float *tn = reinterpret_cast<float *>(base + L.off_tn);
    uint8_t *qe = reinterpret_cast<uint8_t *>(base + L.off_qe);
    uint8_t *te = reinterpret_cast<uint8_t *>(base + L.off_te);
    PCand *part = reinterpret_cast<PCand *>(base + L.off_part);
    const uint32_t qrows = (uint32_t)L.nq_pad, trows = (uint32_t)L.nt_pad;
    if (prec == 1) {
        prep_bf16_kernel<<<(qrows + 3) / 4, 256, 0, st>>>(qd, nq, qrows, false, reinterpret_cast<uint32_t *>(qlow),
                                                          qn);
        prep_bf16_kernel<<<(trows + 3) / 4, 256, 0, st>>>(td, nt, trows, true, reinterpret_cast<uint32_t *>(tlow),
                                                          tn);
    } else {
        prep_fp8_kernel<<<(qrows + 3) / 4, 256, 0, st>>>(qd, nq, qrows, reinterpret_cast<uint32_t *>(qlow), qn, qe);
        prep_fp8_kernel<<<(trows + 3) / 4, 256, 0, st>>>(td, nt, trows, reinterpret_cast<uint32_t *>(tlow), tn, te);
    }
    dim3 grid(qrows / kPQB, L.nsplit);
    if (prec == 1)
        prec_topk_kernel<false><<<grid, 256, 0, st>>>(qlow, qe, tlow, tn, te, qrows, L.tiles_per_split, L.ntiles,
                                                     part);
    else
        prec_topk_kernel<true><<<grid, 256, 0, st>>>(qlow, qe, tlow, tn, te, qrows, L.tiles_per_split, L.ntiles,
                                                    part);
    prec_rescore_kernel<<<(qrows + 31) / 32, 256, 0, st>>>(part, L.nsplit, qrows, nq, qd, td, qn, best_idx, best,
                                                          second, approx_best);
    return hipGetLastError();
}

}  // namespace ps
