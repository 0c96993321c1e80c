// 128-bit descriptor codes: quantisation (catalog.rs:37-45), the coarse Hamming vote count
// (catalog.rs:707-728), an exact top-k Hamming shortlist over a code catalog, and the FP32
// L2 rerank of a shortlist.
//
// Codes are two little-endian u64 words (bit i of the code = bit i%64 of word i/64); on the
// device a code is read as one uint4 (16 B). Hamming distance = 4 XOR + 4 popcounts.
//
// Top-k order is (distance asc, catalog index asc). The fast path is a three-kernel
// bounded scan:
//   1. sample: exact histogram of one query against the first s = min(n, 16384) codes;
//      picks a per-query bound B so that about `target` catalog codes fall below it
//      (when s == n the histogram is exact and B = T + 1, T the k-th smallest distance);
//   2. collect: every (query, code) pair with d < B appends the key (d << 32 | index) to
//      the query's candidate list (at most kHamCandCap slots);
//   3. select: per query, histogram of the candidates gives the exact k-th distance T
//      (valid when the list holds at least k codes and did not overflow), the keys with
//      d <= T are sorted in LDS and the first k written out.
// A query whose list is short or overflowed is appended to a flag list and solved by an
// exact single-workgroup pass over the whole catalog (full histogram, then an
// index-ordered scan that keeps every d < T and the first k - count(d < T) codes at d == T).
// Both paths produce exactly the (distance, index) order, so the output never depends on
// which one ran.
#include <algorithm>

#include "ps_internal.h"

namespace ps {

namespace {

constexpr uint32_t kSample = 16384;
constexpr int kCollectCPT = 8;       // codes per thread in the collect kernel
constexpr int kCollectQPB = 128;     // queries per collect workgroup
constexpr uint32_t kSelectMax = 2048;

__device__ __forceinline__ uint32_t ham(uint4 a, uint4 b) {
    uint32_t d = __popc(a.x ^ b.x);
    d = __popc(a.y ^ b.y) + d;
    d = __popc(a.z ^ b.z) + d;
    return __popc(a.w ^ b.w) + d;
}

// LDS bitonic sort (ascending) of P = 2^m 64-bit keys by all threads of the workgroup.
__device__ void bitonic_sort_u64(uint64_t *a, uint32_t P, uint32_t tid, uint32_t nth) {
    for (uint32_t size = 2; size <= P; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = tid; i < P / 2; i += nth) {
                uint32_t lo = 2 * i - (i & (stride - 1));
                uint32_t hi = lo + stride;
                bool asc = (lo & size) == 0;
                uint64_t x = a[lo], y = a[hi];
                if ((x > y) == asc) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            __syncthreads();
        }
}

__device__ __forceinline__ uint32_t pow2_ceil(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// catalog.rs:37-45: bit i = desc[i] > 0.065. One wave per descriptor: lanes 0..63 test
// elements 0..63 (word 0) and 64..127 (word 1); the two ballots are the code.
__global__ __launch_bounds__(256) void quantize_kernel(const float *desc, uint32_t n, uint64_t *codes) {
    const uint32_t j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= n) return;
    const float *d = desc + (size_t)j * 128;
    uint64_t lo = __ballot(d[lane] > 0.065f);
    uint64_t hi = __ballot(d[64 + lane] > 0.065f);
    if (lane == 0) {
        codes[2 * (size_t)j] = lo;
        codes[2 * (size_t)j + 1] = hi;
    }
}

// Minimum distance of each query over one chunk of entries; entries are read with
// workgroup-uniform addresses.
__global__ __launch_bounds__(256) void ham_min_kernel(const uint4 *q, uint32_t nq, const uint4 *e, uint32_t ne,
                                                      uint32_t chunk, uint32_t *mind) {
    const uint32_t qi = blockIdx.x * 256 + threadIdx.x;
    const uint32_t j0 = blockIdx.y * chunk, j1 = min(j0 + chunk, ne);
    const bool valid = qi < nq;
    const uint4 qq = valid ? q[qi] : make_uint4(0, 0, 0, 0);
    uint32_t m = 255;
    for (uint32_t j = j0; j < j1; j++) m = min(m, ham(qq, e[j]));
    if (valid) atomicMin(&mind[qi], m);
}

// catalog.rs:725: a query votes when its minimum distance is <= 24.
__global__ __launch_bounds__(256) void ham_votes_kernel(const uint32_t *mind, uint32_t nq, uint32_t *votes) {
    const uint32_t qi = blockIdx.x * 256 + threadIdx.x;
    uint64_t m = __ballot(qi < nq && mind[qi] <= 24u);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(votes, (uint32_t)__popcll(m));
}

__global__ __launch_bounds__(256) void ham_sample_kernel(const uint4 *cat, uint32_t n, const uint4 *qs, uint32_t s,
                                                         uint32_t kk, uint32_t need, uint32_t *bound) {
    __shared__ uint32_t hist[4][132];
    const uint32_t q = blockIdx.x, tid = threadIdx.x, w = tid >> 6;
    for (uint32_t i = tid; i < 4 * 132; i += 256) (&hist[0][0])[i] = 0;
    __syncthreads();
    const uint4 qq = qs[q];
    for (uint32_t j = tid; j < s; j += 256) atomicAdd(&hist[w][ham(qq, cat[j])], 1u);
    __syncthreads();
    if (tid < 129) hist[0][tid] = hist[0][tid] + hist[1][tid] + hist[2][tid] + hist[3][tid];
    __syncthreads();
    if (tid == 0) {
        // exact sample (s == n): B = T + 1 with count(d <= T) >= kk; else the smallest B
        // with count(d < B) >= need in the sample
        const uint32_t want = s == n ? kk : need;
        uint32_t cum = 0, b = 129;
        for (uint32_t t = 0; t <= 128; t++) {
            cum += hist[0][t];
            if (cum >= want) {
                b = t + 1;
                break;
            }
        }
        bound[q] = b;
    }
}

// One workgroup: 256 * CPT consecutive codes (held in VGPRs) against kCollectQPB queries
// (staged in LDS, read with uniform addresses). Per (query, code): 4 XOR + 4 popcounts + 1
// min; the rare wave-wide hit (min distance below the query's bound) takes the append
// branch. Hits go to a small per-query LDS list (LDS atomics, no global round trip); the
// lists are flushed at the end with one global slot reservation per query, and a list that
// fills up spills further hits straight to the global list.
constexpr int kLocalSlots = 16;
template <int CPT>
__global__ __launch_bounds__(256) void ham_collect_kernel(const uint4 *cat, uint32_t n, const uint4 *qs,
                                                          const uint32_t *bound, uint32_t nq, uint64_t *keys,
                                                          uint32_t *cnt) {
    __shared__ uint4 s_q[kCollectQPB];
    __shared__ uint32_t s_b[kCollectQPB];
    __shared__ uint32_t s_n[kCollectQPB];
    __shared__ uint64_t s_keys[kCollectQPB][kLocalSlots];
    const uint32_t base = blockIdx.x * (256 * CPT) + threadIdx.x;
    const uint32_t qa = blockIdx.y * kCollectQPB, nqb = min((uint32_t)kCollectQPB, nq - qa);
    if (threadIdx.x < nqb) {
        s_q[threadIdx.x] = qs[qa + threadIdx.x];
        s_b[threadIdx.x] = bound[qa + threadIdx.x];
        s_n[threadIdx.x] = 0;
    }
    uint4 c[CPT];
#pragma unroll
    for (int i = 0; i < CPT; i++) {
        uint32_t j = base + i * 256;
        c[i] = j < n ? cat[j] : make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    for (uint32_t qi = 0; qi < nqb; qi++) {
        const uint4 qq = s_q[qi];
        const uint32_t bq = s_b[qi];
        uint32_t d[CPT], dmin = 0xffffffffu;
#pragma unroll
        for (int i = 0; i < CPT; i++) {
            d[i] = ham(qq, c[i]);
            dmin = min(dmin, d[i]);
        }
        if (dmin < bq) {
#pragma unroll
            for (int i = 0; i < CPT; i++) {
                const uint32_t j = base + i * 256;
                if (d[i] < bq && j < n) {
                    const uint64_t key = ((uint64_t)d[i] << 32) | j;
                    uint32_t ls = atomicAdd(&s_n[qi], 1u);
                    if (ls < kLocalSlots) {
                        s_keys[qi][ls] = key;
                    } else {
                        const uint32_t q = qa + qi;
                        uint32_t slot = atomicAdd(&cnt[q], 1u);
                        if (slot < kHamCandCap) keys[(size_t)q * kHamCandCap + slot] = key;
                    }
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < nqb) {
        const uint32_t m = min(s_n[threadIdx.x], (uint32_t)kLocalSlots);
        if (m) {
            const uint32_t q = qa + threadIdx.x;
            const uint32_t slot = atomicAdd(&cnt[q], m);
            for (uint32_t i = 0; i < m; i++)
                if (slot + i < kHamCandCap) keys[(size_t)q * kHamCandCap + slot + i] = s_keys[threadIdx.x][i];
        }
    }
}

__device__ __forceinline__ void flag_query(uint32_t *flags, uint32_t q) {
    uint32_t slot = atomicAdd(&flags[0], 1u);
    flags[1 + slot] = q;
}

__device__ __forceinline__ void emit_sorted(const uint64_t *s_keys, uint32_t kk, uint32_t k, uint32_t tid,
                                            uint32_t nth, uint32_t *out_idx, uint8_t *out_dist) {
    for (uint32_t p = tid; p < k; p += nth) {
        if (p < kk) {
            uint64_t key = s_keys[p];
            out_idx[p] = (uint32_t)key;
            out_dist[p] = (uint8_t)(key >> 32);
        } else {
            out_idx[p] = 0xffffffffu;
            out_dist[p] = 0xff;
        }
    }
}

__global__ __launch_bounds__(256) void ham_select_kernel(const uint64_t *keys, const uint32_t *cnt, uint32_t kk,
                                                         uint32_t k, uint32_t *out_idx, uint8_t *out_dist,
                                                         uint32_t *flags) {
    __shared__ uint32_t hist[132];
    __shared__ uint64_t s_keys[kSelectMax];
    __shared__ uint32_t s_n, s_T, s_total;
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const uint32_t m = cnt[q];
    if (m > kHamCandCap || m < kk) {
        if (tid == 0) flag_query(flags, q);
        return;
    }
    const uint64_t *kq = keys + (size_t)q * kHamCandCap;
    for (uint32_t i = tid; i < 132; i += 256) hist[i] = 0;
    if (tid == 0) s_n = 0;
    __syncthreads();
    for (uint32_t i = tid; i < m; i += 256) atomicAdd(&hist[(uint32_t)(kq[i] >> 32)], 1u);
    __syncthreads();
    if (tid == 0) {
        uint32_t cum = 0, t = 0;
        for (t = 0; t <= 128; t++) {
            if (cum + hist[t] >= kk) break;
            cum += hist[t];
        }
        s_T = t;
        s_total = cum + hist[t];
    }
    __syncthreads();
    const uint32_t T = s_T, total = s_total;
    if (total > kSelectMax) {
        if (tid == 0) flag_query(flags, q);
        return;
    }
    for (uint32_t i = tid; i < m; i += 256) {
        uint64_t key = kq[i];
        if ((uint32_t)(key >> 32) <= T) s_keys[atomicAdd(&s_n, 1u)] = key;
    }
    const uint32_t P = pow2_ceil(total);
    __syncthreads();
    for (uint32_t i = total + tid; i < P; i += 256) s_keys[i] = ~0ull;
    __syncthreads();
    bitonic_sort_u64(s_keys, P, tid, 256);
    emit_sorted(s_keys, kk, k, tid, 256, out_idx + (size_t)q * k, out_dist + (size_t)q * k);
}

// Exact path for flagged queries; the grid strides over the flag list.
__global__ __launch_bounds__(1024) void ham_fallback_kernel(const uint4 *cat, uint32_t n, const uint4 *qs,
                                                            const uint32_t *flags, uint32_t kk, uint32_t k,
                                                            uint32_t *out_idx, uint8_t *out_dist) {
    __shared__ uint32_t hist[132];
    __shared__ uint64_t s_keys[kHamMaxK];
    __shared__ uint32_t s_weq[16];
    __shared__ uint32_t s_lt, s_eqrun, s_T, s_below;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t nflag = flags[0];
    for (uint32_t f = blockIdx.x; f < nflag; f += gridDim.x) {
        const uint32_t q = flags[1 + f];
        const uint4 qq = qs[q];
        for (uint32_t i = tid; i < 132; i += 1024) hist[i] = 0;
        __syncthreads();
        for (uint32_t j = tid; j < n; j += 1024) atomicAdd(&hist[ham(qq, cat[j])], 1u);
        __syncthreads();
        if (tid == 0) {
            uint32_t cum = 0, t = 0;
            for (t = 0; t <= 128; t++) {
                if (cum + hist[t] >= kk) break;
                cum += hist[t];
            }
            s_T = t;
            s_below = cum;
            s_lt = 0;
            s_eqrun = 0;
        }
        __syncthreads();
        const uint32_t T = s_T, below = s_below, take = kk - below;
        for (uint32_t base = 0; base < n; base += 1024) {
            const uint32_t j = base + tid;
            const uint32_t d = j < n ? ham(qq, cat[j]) : 0x1000u;
            if (d < T) s_keys[atomicAdd(&s_lt, 1u)] = ((uint64_t)d << 32) | j;   // count(d < T) = below < kk
            const bool eq = d == T;
            const uint64_t mask = __ballot(eq);
            if (lane == 0) s_weq[w] = (uint32_t)__popcll(mask);
            __syncthreads();
            uint32_t off = s_eqrun;
            for (uint32_t v = 0; v < w; v++) off += s_weq[v];
            off += (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            if (eq && off < take) s_keys[below + off] = ((uint64_t)d << 32) | j;
            __syncthreads();
            if (tid == 0) {
                uint32_t tot = 0;
                for (uint32_t v = 0; v < 16; v++) tot += s_weq[v];
                s_eqrun += tot;
            }
            __syncthreads();
        }
        const uint32_t P = pow2_ceil(kk);
        for (uint32_t i = kk + tid; i < P; i += 1024) s_keys[i] = ~0ull;
        __syncthreads();
        bitonic_sort_u64(s_keys, P, tid, 1024);
        emit_sorted(s_keys, kk, k, tid, 1024, out_idx + (size_t)q * k, out_dist + (size_t)q * k);
        __syncthreads();
    }
}

// FP32 rerank of a shortlist: exact sequential squared L2 to each listed catalog row,
// then (distance asc, index asc). Invalid slots (index >= n) sort last as (UINT32_MAX, inf).
__global__ __launch_bounds__(256) void rerank_kernel(const float *qd, const float *cd, uint32_t n,
                                                     const uint32_t *shortlist, uint32_t k, uint32_t *out_idx,
                                                     float *out_d2) {
    __shared__ uint64_t s_keys[kHamMaxK];
    __shared__ float4 s_q[32];
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    if (tid < 32) s_q[tid] = reinterpret_cast<const float4 *>(qd + (size_t)q * 128)[tid];
    __syncthreads();
    const uint32_t P = pow2_ceil(k);
    for (uint32_t c = tid; c < P; c += 256) {
        uint64_t key = ~0ull;
        if (c < k) {
            uint32_t idx = shortlist[(size_t)q * k + c];
            if (idx < n) {
                float d = seq_dist128(reinterpret_cast<const float *>(s_q), cd + (size_t)idx * 128);
                key = ((uint64_t)__float_as_uint(d) << 32) | idx;
            }
        }
        s_keys[c] = key;
    }
    __syncthreads();
    bitonic_sort_u64(s_keys, P, tid, 256);
    for (uint32_t p = tid; p < k; p += 256) {
        uint64_t key = s_keys[p];
        bool ok = key != ~0ull;
        out_idx[(size_t)q * k + p] = ok ? (uint32_t)key : 0xffffffffu;
        out_d2[(size_t)q * k + p] = ok ? __uint_as_float((uint32_t)(key >> 32)) : INFINITY;
    }
}

struct TopkScratch {
    uint32_t *bound, *cnt, *flags;
    uint64_t *keys;
};

TopkScratch carve(void *scratch, uint32_t nq_chunk) {
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    char *p = static_cast<char *>(scratch);
    TopkScratch s;
    s.bound = reinterpret_cast<uint32_t *>(p);
    p += al((size_t)nq_chunk * 4);
    s.cnt = reinterpret_cast<uint32_t *>(p);
    p += al((size_t)nq_chunk * 4);
    s.flags = reinterpret_cast<uint32_t *>(p);
    p += al(((size_t)nq_chunk + 1) * 4);
    s.keys = reinterpret_cast<uint64_t *>(p);
    return s;
}

}  // namespace

hipError_t launch_quantize_u128(const float *desc, uint32_t n, uint64_t *codes, hipStream_t st) {
    if (n == 0) return hipSuccess;
    quantize_kernel<<<(n + 3) / 4, 256, 0, st>>>(desc, n, codes);
    return hipGetLastError();
}

// scratch: nq + 1 u32 (per-query minimum, then the vote count)
hipError_t launch_hamming_votes(const uint64_t *q, uint32_t nq, const uint64_t *e, uint32_t ne, uint32_t *scratch,
                                hipStream_t st) {
    hipError_t err = launch_fill_u32(scratch, 255u, nq, st);
    if (err == hipSuccess) err = launch_fill_u32(scratch + nq, 0u, 1, st);
    if (err != hipSuccess || nq == 0 || ne == 0) return err;
    const uint32_t chunk = std::max(512u, (ne + 65534u) / 65535u);
    dim3 grid((nq + 255) / 256, (ne + chunk - 1) / chunk);
    ham_min_kernel<<<grid, 256, 0, st>>>(reinterpret_cast<const uint4 *>(q), nq, reinterpret_cast<const uint4 *>(e),
                                         ne, chunk, scratch);
    ham_votes_kernel<<<(nq + 255) / 256, 256, 0, st>>>(scratch, nq, scratch + nq);
    return hipGetLastError();
}

size_t hamming_topk_scratch_bytes(uint32_t nq_chunk) {
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    return al((size_t)nq_chunk * 4) * 2 + al(((size_t)nq_chunk + 1) * 4) +
           (size_t)nq_chunk * kHamCandCap * sizeof(uint64_t);
}

// Queries are processed in chunks of kHamQueryChunk; scratch must hold
// hamming_topk_scratch_bytes(min(nq, kHamQueryChunk)). Requires 1 <= k <= kHamMaxK, n >= 1.
hipError_t launch_hamming_topk(const uint64_t *cat, uint32_t n, const uint64_t *qs, uint32_t nq, uint32_t k,
                               void *scratch, uint32_t *out_idx, uint8_t *out_dist, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    if (k == 0 || k > kHamMaxK || n == 0) return hipErrorInvalidValue;
    const uint4 *c4 = reinterpret_cast<const uint4 *>(cat);
    const uint32_t kk = std::min(k, n);
    const uint32_t s = std::min(n, kSample);
    const uint64_t target = std::min<uint64_t>(4ull * k, kHamCandCap / 2);
    const uint32_t need = (uint32_t)std::max<uint64_t>(1, (target * s + n - 1) / n);
    for (uint32_t q0 = 0; q0 < nq; q0 += kHamQueryChunk) {
        const uint32_t m = std::min(kHamQueryChunk, nq - q0);
        const uint4 *q4 = reinterpret_cast<const uint4 *>(qs) + q0;
        TopkScratch sc = carve(scratch, std::min(nq, kHamQueryChunk));
        hipError_t err = launch_fill_u32(sc.cnt, 0u, m, st);
        if (err == hipSuccess) err = launch_fill_u32(sc.flags, 0u, 1, st);
        if (err != hipSuccess) return err;
        ham_sample_kernel<<<m, 256, 0, st>>>(c4, n, q4, s, kk, need, sc.bound);
        dim3 cg((n + 256 * kCollectCPT - 1) / (256 * kCollectCPT), (m + kCollectQPB - 1) / kCollectQPB);
        ham_collect_kernel<kCollectCPT><<<cg, 256, 0, st>>>(c4, n, q4, sc.bound, m, sc.keys, sc.cnt);
        uint32_t *oi = out_idx + (size_t)q0 * k;
        uint8_t *od = out_dist + (size_t)q0 * k;
        ham_select_kernel<<<m, 256, 0, st>>>(sc.keys, sc.cnt, kk, k, oi, od, sc.flags);
        ham_fallback_kernel<<<std::min(m, 256u), 1024, 0, st>>>(c4, n, q4, sc.flags, kk, k, oi, od);
        err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

hipError_t launch_rerank_l2(const float *qd, uint32_t nq, const float *cd, uint32_t n, const uint32_t *shortlist,
                            uint32_t k, uint32_t *out_idx, float *out_d2, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    if (k == 0 || k > kHamMaxK) return hipErrorInvalidValue;
    rerank_kernel<<<nq, 256, 0, st>>>(qd, cd, n, shortlist, k, out_idx, out_d2);
    return hipGetLastError();
}

}  // namespace ps
