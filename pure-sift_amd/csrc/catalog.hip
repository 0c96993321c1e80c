// Device side of the GPU-resident feature catalog (catalog.rs): u8 descriptor
// dequantisation at load, the coarse vote pass over every catalog signature, and the
// per-hit tracking summary.
#include <hip/hip_runtime.h>

#include "ps_internal.h"

#pragma clang fp contract(off)

namespace ps {

namespace {

// pipeline.rs:1034-1044: b / 512 (exact), then a sequential f32 sum of squares in element
// order, sqrt, and one multiply by 1 / norm when norm > 1e-7. One row per thread.
__global__ __launch_bounds__(256) void dequant_u8_kernel(const uint8_t *q, uint32_t n, float *out) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const uint4 *src = reinterpret_cast<const uint4 *>(q + (size_t)r * 128);
    float4 *dst = reinterpret_cast<float4 *>(out + (size_t)r * 128);
    float acc = 0.0f;
    for (int c = 0; c < 8; c++) {
        const uint4 w = src[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
        for (int j = 0; j < 4; j++)
            for (int b = 0; b < 4; b++) {
                const float v = (float)((ws[j] >> (8 * b)) & 0xFFu) * (1.0f / 512.0f);
                acc += v * v;
            }
    }
    const float norm = sqrtf(acc);
    const bool scale = norm > 1e-7f;
    const float inv = scale ? 1.0f / norm : 1.0f;
    for (int c = 0; c < 8; c++) {
        const uint4 w = src[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
        for (int j = 0; j < 4; j++) {
            float v[4];
            for (int b = 0; b < 4; b++) {
                v[b] = (float)((ws[j] >> (8 * b)) & 0xFFu) * (1.0f / 512.0f);
                if (scale) v[b] *= inv;
            }
            dst[c * 4 + j] = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
}

// Nonzero iff some byte of v is zero: the borrow of (v - 0x01..) can only set bit 7 of a
// byte above a zero byte, so the "any" answer is exact.
__device__ __forceinline__ uint32_t zero_byte_bits(uint32_t v) { return (v - 0x01010101u) & ~v & 0x80808080u; }

// catalog.rs:595-633. Entry e gets one vote per query that has a catalog signature of e
// within 24 bits; on the inverted-index path (MIH, taken above 32 entries) the signature
// must also equal the query in at least one of its 16 bytes, because the reference only
// examines signatures found through the bucket of one of the query's bytes
// (catalog.rs:599-613). Grid (entry, 256-query block); each wave walks the entry's
// signatures with wave-uniform (scalar) loads and leaves once all its queries have voted.
template <bool MIH>
__global__ __launch_bounds__(256) void catalog_vote_kernel(const uint4 *sig, const uint32_t *off, const uint4 *q,
                                                           uint32_t nq, uint32_t *votes) {
    const uint32_t e = blockIdx.x;
    const uint32_t b = off[e], end = off[e + 1];
    const uint32_t qi = blockIdx.y * 256 + threadIdx.x;
    const bool live = qi < nq;
    const uint4 qs = live ? q[qi] : make_uint4(0u, 0u, 0u, 0u);
    bool found = !live;
    constexpr uint32_t G = 8;   // signatures per group: all loads issued before any test
    for (uint32_t i = b; i < end; i += G) {
        if (__all(found)) break;
        // past the end, re-test the entry's last signature: the vote is an OR, so a repeat
        // changes nothing and every load stays in bounds without a branch
        uint4 s[G];
#pragma unroll
        for (uint32_t k = 0; k < G; k++) s[k] = sig[i + k < end ? i + k : end - 1];
        bool near[G], any_near = false;
#pragma unroll
        for (uint32_t k = 0; k < G; k++) {
            const uint32_t d = __popc(qs.x ^ s[k].x) + __popc(qs.y ^ s[k].y) + __popc(qs.z ^ s[k].z) +
                               __popc(qs.w ^ s[k].w);
            near[k] = d <= 24u;
            any_near = any_near || near[k];
        }
        if (!MIH) {
            found = found || any_near;
        } else if (__any(any_near)) {   // the byte test only where some lane is within 24 bits
#pragma unroll
            for (uint32_t k = 0; k < G; k++) {
                const uint32_t z = zero_byte_bits(qs.x ^ s[k].x) | zero_byte_bits(qs.y ^ s[k].y) |
                                   zero_byte_bits(qs.z ^ s[k].z) | zero_byte_bits(qs.w ^ s[k].w);
                found = found || (near[k] && z != 0u);
            }
        }
    }
    const uint64_t bal = __ballot(found && live);
    if ((threadIdx.x & 63) == 0 && bal) atomicAdd(&votes[e], (uint32_t)__popcll(bal));
}

__device__ __forceinline__ uint32_t f32_order_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float f32_from_order_key(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// k-th smallest (0-based) of the na keys key(i), by a 4 x 8-bit radix select in one
// workgroup (the value of sorted[k] of the reference's sort; equal keys are equal values).
template <typename KeyFn>
__device__ uint32_t select_kth(uint32_t na, uint32_t k, KeyFn key, uint32_t *s_hist, uint32_t *s_state) {
    const int tid = threadIdx.x;
    uint32_t prefix = 0, mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        s_hist[tid] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < na; i += 256) {
            const uint32_t v = key(i);
            if ((v & mask) == prefix) atomicAdd(&s_hist[(v >> shift) & 255u], 1u);
        }
        __syncthreads();
        // inclusive prefix sum over the 256 bins (Hillis-Steele)
        uint32_t c = s_hist[tid];
        for (int o = 1; o < 256; o <<= 1) {
            __syncthreads();
            const uint32_t add = tid >= o ? s_hist[tid - o] : 0u;
            __syncthreads();
            c += add;
            s_hist[tid] = c;
        }
        __syncthreads();
        const uint32_t excl = tid ? s_hist[tid - 1] : 0u;
        if (excl <= k && k < c) {
            s_state[0] = (uint32_t)tid;
            s_state[1] = excl;
        }
        __syncthreads();
        prefix |= s_state[0] << shift;
        mask |= 255u << shift;
        k -= s_state[1];
        __syncthreads();
    }
    return prefix;
}

// pipeline.rs:2210-2272 for one problem per workgroup: the active set is the similarity
// inliers at the default 3 px threshold (or every match when there are none); the five
// sums run sequentially in the active order (one lane each, in parallel), the medians by
// radix selection.
__global__ __launch_bounds__(256) void tracking_kernel(const ps_feature_match *matches, const TrackJob *jobs,
                                                       const ps_ransac_result *sim_res, const void *sim_work,
                                                       size_t sim_total, const ps_feature *qf, const ps_feature *tf,
                                                       TrackOut *out) {
    __shared__ float s_v[5][256];
    __shared__ uint32_t s_hist[256];
    __shared__ uint32_t s_state[2];
    const int tid = threadIdx.x;
    const TrackJob J = jobs[blockIdx.x];
    const uint32_t n = J.n;
    const uint32_t nsim = sim_res[blockIdx.x].n_sim;
    const uint32_t *inl = ransac_sim_inliers(sim_work, sim_total, J.moff, n);
    const bool use_inl = nsim > 0;
    const uint32_t na = use_inl ? nsim : n;
    if (na == 0) {   // no matches: no summary (the caller drops the entry)
        if (tid == 0) out[blockIdx.x] = TrackOut{};
        return;
    }
    const ps_feature_match *mm = matches + J.moff;
    const ps_feature *T = tf + J.tfoff, *Qf = qf + J.qfoff;
    float sum = 0.0f;   // lane l < 5 owns sum l: dx, dy, scale ratio, sin, cos
    for (uint32_t base = 0; base < na; base += 256) {
        const uint32_t k = base + tid;
        if (k < na) {
            const ps_feature_match M = mm[use_inl ? inl[k] : k];
            const ps_feature Q = Qf[M.query_index], Tt = T[M.target_index];
            s_v[0][tid] = M.dx;
            s_v[1][tid] = M.dy;
            s_v[2][tid] = Q.scale > 1e-6f ? Tt.scale / Q.scale : 0.0f;   // + 0 leaves a sum >= +0 unchanged
            const float d = Tt.orientation - Q.orientation;
            s_v[3][tid] = sinf(d);
            s_v[4][tid] = cosf(d);
        }
        __syncthreads();
        if (tid < 5) {
            const uint32_t cnt = na - base < 256u ? na - base : 256u;
            for (uint32_t j = 0; j < cnt; j++) sum += s_v[tid][j];
        }
        __syncthreads();
    }
    __shared__ float s_sum[5];
    if (tid < 5) s_sum[tid] = sum;
    __syncthreads();
    float med[2];
    for (int c = 0; c < 2; c++) {
        auto key = [&](uint32_t i) {
            const ps_feature_match &M = mm[use_inl ? inl[i] : i];
            return f32_order_key(c == 0 ? M.dx : M.dy);
        };
        const uint32_t mid = na / 2;
        const float hi = f32_from_order_key(select_kth(na, mid, key, s_hist, s_state));
        if (na % 2 == 1) {
            med[c] = hi;
        } else {
            const float lo = f32_from_order_key(select_kth(na, mid - 1, key, s_hist, s_state));
            med[c] = 0.5f * (lo + hi);
        }
    }
    if (tid == 0) {
        const float inv_n = 1.0f / (float)na;
        TrackOut o;
        o.mean_dx = s_sum[0] * inv_n;
        o.mean_dy = s_sum[1] * inv_n;
        o.median_dx = med[0];
        o.median_dy = med[1];
        o.mean_scale_ratio = s_sum[2] * inv_n;
        o.mean_rotation_deg = atan2f(s_sum[3], s_sum[4]) * 57.2957795130823208767981548141051703f;
        o.active = na;
        o.pad = 0;
        out[blockIdx.x] = o;
    }
}

// catalog.rs:656-671 on the device: slot k's RANSAC problem and tracking job over its match
// block (k * nq, count k); a slot without matches gets an empty problem
__global__ void catalog_problems_kernel(const uint32_t *cnt, const ps_feature_set *tsets, uint32_t ns, uint32_t nq,
                                        uint64_t sim_seed, uint64_t hom_seed, ps_ransac_problem *rp, TrackJob *jobs) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    rp[k] = ps_ransac_problem{k * nq, cnt[k], sim_seed, hom_seed};
    jobs[k] = TrackJob{k * nq, cnt[k], tsets[k].offset, 0u};
}

}  // namespace

hipError_t launch_catalog_problems(const uint32_t *cnt, const ps_feature_set *tsets, uint32_t ns, uint32_t nq,
                                   uint64_t sim_seed, uint64_t hom_seed, ps_ransac_problem *rp, TrackJob *jobs,
                                   hipStream_t st) {
    if (ns == 0) return hipSuccess;
    catalog_problems_kernel<<<(ns + 255) / 256, 256, 0, st>>>(cnt, tsets, ns, nq, sim_seed, hom_seed, rp, jobs);
    return hipGetLastError();
}

hipError_t launch_dequant_u8(const uint8_t *q, uint32_t n, float *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    dequant_u8_kernel<<<(n + 255) / 256, 256, 0, st>>>(q, n, out);
    return hipGetLastError();
}

hipError_t launch_catalog_votes(const uint64_t *sig, const uint32_t *off, uint32_t n_entries, const uint64_t *q,
                                uint32_t nq, bool mih, uint32_t *votes, hipStream_t st) {
    hipError_t err = launch_fill_u32(votes, 0u, n_entries, st);
    if (err != hipSuccess || n_entries == 0 || nq == 0) return err;
    const dim3 grid(n_entries, (nq + 255) / 256);
    const uint4 *s4 = reinterpret_cast<const uint4 *>(sig), *q4 = reinterpret_cast<const uint4 *>(q);
    if (mih)
        catalog_vote_kernel<true><<<grid, 256, 0, st>>>(s4, off, q4, nq, votes);
    else
        catalog_vote_kernel<false><<<grid, 256, 0, st>>>(s4, off, q4, nq, votes);
    return hipGetLastError();
}

hipError_t launch_tracking(const ps_feature_match *matches, const TrackJob *jobs, uint32_t n_jobs,
                           const ps_ransac_result *sim_res, const void *sim_work, size_t sim_total,
                           const ps_feature *qf, const ps_feature *tf, TrackOut *out, hipStream_t st) {
    if (n_jobs == 0) return hipSuccess;
    tracking_kernel<<<n_jobs, 256, 0, st>>>(matches, jobs, sim_res, sim_work, sim_total, qf, tf, out);
    return hipGetLastError();
}

}  // namespace ps
