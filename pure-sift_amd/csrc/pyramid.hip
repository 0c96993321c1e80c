// Scale-space pyramid (fused upsample + separable Gaussian + DoG), extremum
// detection with Hessian edge rejection and 3-D Taylor refinement, and the
// keypoint ordering / 4096 cap -- gfx950 kernels.
//
// Bit-exactness with pure-sift 0.2.0 (src/pipeline.rs): every output pixel sums
// its taps in the reference's order with separate multiply and add roundings
// (no FMA: -ffp-contract=off plus the pragma below), Gaussian weights come from
// the host (glibc expf, exactly pipeline.rs:275-296), and the extremum tests /
// refinement repeat the reference's f32 expressions operation by operation.
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

#include "ps_internal.h"

#pragma clang fp contract(off)

namespace ps {

// ------------------------------------------------------------------ blur + DoG
// One workgroup produces a TW x TH tile of Gaussian level s (and DoG s) from level
// s-1: the (TH+2R) x (TW+2R) clamped source window is staged in LDS, the horizontal
// pass writes (TH+2R) x TW partial rows to LDS, the vertical pass reads them back.
// Clamping the staged window reproduces the reference's per-tap index clamp
// (pipeline.rs:339, :401) because every staged element is the clamped source pixel.
// Tile height TH in {16, 32, 64} is a template parameter (picked per level shape by the
// engine's autotuner); the vertical pass gives each thread TH / 4 rows of one column.
constexpr int kTW = 64;
constexpr int kBlurThreads = 256;

// Workgroups are dispatched to the 8 XCDs round-robin by linear workgroup ID, and each XCD has
// its own L2. xcd_tile maps the linear ID b (of nb) to a tile index so that every XCD owns one
// contiguous, row-major range of tiles: neighbouring tiles, whose halos overlap, then share an
// L2 instead of fetching the shared cache lines from HBM once per XCD. The map is a bijection
// on [0, nb) (XCD k owns nb/8 + (k < nb%8) consecutive tiles).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
    const uint32_t q = nb / kXcds, r = nb % kXcds, k = b % kXcds;
    return k * q + min(k, r) + b / kXcds;
}

template <int SRC>
__device__ __forceinline__ float load_src(const BlurArgs &a, int y, int x) {
    const float inv255 = 1.0f / 255.0f;   // luma_to_array2, pipeline.rs:961
    if constexpr (SRC == SRC_F32) {
        return static_cast<const float *>(a.src)[(size_t)y * a.src_stride + x];
    } else if constexpr (SRC == SRC_U8) {
        return (float)static_cast<const uint8_t *>(a.src)[(size_t)y * a.src_stride + x] * inv255;
    } else if constexpr (SRC == SRC_DOWN2) {   // downsample_2x, pipeline.rs:443-446
        return static_cast<const float *>(a.src)[(size_t)(2 * y) * a.src_stride + 2 * x];
    } else {                                   // upsample_2x_bilinear, pipeline.rs:469-479
        int y0 = y >> 1, x0 = x >> 1;
        int y1 = min(y0 + (y & 1), a.src_h - 1), x1 = min(x0 + (x & 1), a.src_w - 1);
        float p00, p01, p10, p11;
        if constexpr (SRC == SRC_F32_UP) {
            const float *s = static_cast<const float *>(a.src);
            p00 = s[(size_t)y0 * a.src_stride + x0]; p01 = s[(size_t)y0 * a.src_stride + x1];
            p10 = s[(size_t)y1 * a.src_stride + x0]; p11 = s[(size_t)y1 * a.src_stride + x1];
        } else {
            const uint8_t *s = static_cast<const uint8_t *>(a.src);
            p00 = (float)s[(size_t)y0 * a.src_stride + x0] * inv255;
            p01 = (float)s[(size_t)y0 * a.src_stride + x1] * inv255;
            p10 = (float)s[(size_t)y1 * a.src_stride + x0] * inv255;
            p11 = (float)s[(size_t)y1 * a.src_stride + x1] * inv255;
        }
        return 0.25f * (p00 + p01 + p10 + p11);
    }
}

// One output tile at (x0, y0): the body shared by the single-level launch and the grouped
// (many image levels per launch) persistent launch.
template <int R, int SRC, int kTH>
__device__ __forceinline__ void blur_tile(const BlurArgs &a, const Taps &taps, int x0, int y0, int tid) {
    constexpr int IW = kTW + 2 * R;
    constexpr int IWP = (IW + 3) & ~3;
    constexpr int IH = kTH + 2 * R;
    constexpr int NT = 2 * R + 1;
    __shared__ __attribute__((aligned(16))) float s_in[IH * IWP];
    __shared__ __attribute__((aligned(16))) float s_tmp[IH * kTW];
    const int h = a.h, w = a.w;
    if (a.clear && x0 == 0 && y0 == 0 && tid < a.n_clear) a.clear[tid] = 0u;

    // stage the clamped window: all loads of a thread are issued before its LDS stores so
    // they overlap instead of paying one memory latency each
    constexpr int NL = (IH * IW + kBlurThreads - 1) / kBlurThreads;
    {
        float v[NL];
#pragma unroll
        for (int j = 0; j < NL; j++) {
            int i = tid + j * kBlurThreads;
            int r = i / IW, c = i - r * IW;
            int gy = min(max(y0 - R + r, 0), h - 1);
            int gx = min(max(x0 - R + c, 0), w - 1);
            v[j] = (i < IH * IW) ? load_src<SRC>(a, gy, gx) : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < NL; j++) {
            int i = tid + j * kBlurThreads;
            int r = i / IW, c = i - r * IW;
            if (i < IH * IW) s_in[r * IWP + c] = v[j];
        }
    }
    __syncthreads();

    // horizontal: 4 consecutive outputs per thread, taps in k order (pipeline.rs:351-354)
    constexpr int HC = kTW / 4;
    for (int i = tid; i < IH * HC; i += kBlurThreads) {
        int r = i / HC, c = (i - r * HC) * 4;
        const float *src = s_in + r * IWP + c;
        float v[NT + 3];
#pragma unroll
        for (int j = 0; j < NT + 3; j++) v[j] = src[j];
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
#pragma unroll
        for (int k = 0; k < NT; k++) {
            float wk = taps.w[k];
            a0 += v[k] * wk;
            a1 += v[k + 1] * wk;
            a2 += v[k + 2] * wk;
            a3 += v[k + 3] * wk;
        }
        float4 o = make_float4(a0, a1, a2, a3);
        *reinterpret_cast<float4 *>(s_tmp + r * kTW + c) = o;
    }
    __syncthreads();

    // vertical: TH / 4 consecutive rows of one column per thread (pipeline.rs:400-406)
    constexpr int VR = kTH / 4;
    for (int i = tid; i < (kTH / VR) * kTW; i += kBlurThreads) {
        int c = i % kTW, r0 = (i / kTW) * VR;
        float v[VR + NT - 1];
#pragma unroll
        for (int j = 0; j < VR + NT - 1; j++) v[j] = s_tmp[(r0 + j) * kTW + c];
        float acc[VR];
#pragma unroll
        for (int j = 0; j < VR; j++) acc[j] = 0.0f;
#pragma unroll
        for (int k = 0; k < NT; k++) {
            float wk = taps.w[k];
#pragma unroll
            for (int j = 0; j < VR; j++) acc[j] += v[j + k] * wk;
        }
        int x = x0 + c;
        if (x < w) {
#pragma unroll
            for (int j = 0; j < VR; j++) {
                int y = y0 + r0 + j;
                if (y < h) {
                    size_t o = (size_t)y * w + x;
                    float center = s_in[(r0 + j + R) * IWP + c + R];
                    if (a.dst_g) a.dst_g[o] = acc[j];
                    if (a.dst_d) a.dst_d[o] = acc[j] - center;   // pipeline.rs:611
                    if (a.dst_src) a.dst_src[o] = center;
                }
            }
        }
    }
}

template <int R, int SRC, int kTH>
__global__ __launch_bounds__(kBlurThreads) void blur_kernel(BlurArgs a, Taps taps) {
    const uint32_t tile = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    blur_tile<R, SRC, kTH>(a, taps, (int)(tile % gridDim.x) * kTW, (int)(tile / gridDim.x) * kTH, threadIdx.x);
}

// Sliding-window form of the same blur for large levels. A workgroup owns a 64-column strip
// and `chunk` output rows of it and walks down in steps of kSlideTH rows: each step stages
// kSlideTH source rows (with the 2R-column halo), runs the horizontal pass once for them into
// a ring of horizontal-pass rows in LDS, and emits the kSlideTH output rows whose vertical
// window is complete. The horizontal pass therefore runs once per source row instead of once
// per tile (plus P steps of prologue per chunk), staging uses loop-invariant clamped columns
// and wave-uniform rows, and the next step's global loads are in flight during the vertical
// pass. The ring keeps RING = (P + 1) * kSlideTH rows and is written twice (at slot and at
// slot + RING) so any window of up to RING consecutive rows is contiguous in LDS. Per output
// the arithmetic is that of blur_kernel (same taps, same order, same clamping): the two
// kernels are bit-identical.
constexpr int kSlideTH = kBlurSlideStep;

// CENTER: the level also writes its DoG plane or its source values (the pyramid dump, and the
// first level of octaves >= 1, which materialises G0): only then are the ring rows' centre
// values kept, so the common case needs a third less LDS and fits more workgroups per CU.
template <int R, int SRC, bool CENTER>
__device__ __forceinline__ void blur_strip(const BlurArgs &a, const Taps &taps, int x0, int y_start, int chunk,
                                           int tid) {
    constexpr int NT = 2 * R + 1;
    constexpr int IW = kTW + 2 * R;
    constexpr int IWP = (IW + 3) & ~3;
    constexpr int TH = kSlideTH;
    constexpr int P = (2 * R + TH - 1) / TH;   // steps before the first output block
    constexpr int RING = (P + 1) * TH;         // horizontal-pass rows kept in LDS
    constexpr int VR = TH / 4;                 // output rows per thread per step
    static_assert(IW <= 2 * kTW, "two staged columns per lane");
    static_assert(VR + 2 * R <= RING, "vertical window fits the ring");
    static_assert(kBlurThreads == 4 * 64 && TH == 4 * VR && TH * (kTW / 4) == kBlurThreads, "step shape");
    __shared__ __attribute__((aligned(16))) float s_in[TH * IWP];
    __shared__ __attribute__((aligned(16))) float s_h[2 * RING * kTW];
    __shared__ __attribute__((aligned(16))) float s_c[CENTER ? RING * kTW : 4];
    // the wave index is uniform: readfirstlane lets row arithmetic stay scalar
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int h = a.h, w = a.w;
    if (a.clear && x0 == 0 && y_start == 0 && tid < a.n_clear) a.clear[tid] = 0u;
    const int y_end = min(y_start + chunk, h);
    const int n_steps = (y_end - y_start + TH - 1) / TH + P;
    const int hs = y_start - R;   // source row of horizontal-pass row 0
    // staged columns of this lane: lane and lane + 64 of the (kTW + 2R)-wide window
    const int gxa = min(max(x0 - R + lane, 0), w - 1);
    const int gxb = min(max(x0 - R + kTW + lane, 0), w - 1);
    const bool has_b = lane + kTW < IW;
    float ra[4], rb[4];
    auto fetch = [&](int k) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int gy = min(max(hs + k * TH + 4 * wv + i, 0), h - 1);
            ra[i] = load_src<SRC>(a, gy, gxa);
            rb[i] = has_b ? load_src<SRC>(a, gy, gxb) : 0.0f;
        }
    };
    fetch(0);
    const int hr = tid >> 4, hc = (tid & 15) * 4;   // this thread's horizontal-pass item
    for (int k = 0; k < n_steps; k++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            s_in[(4 * wv + i) * IWP + lane] = ra[i];
            if (has_b) s_in[(4 * wv + i) * IWP + kTW + lane] = rb[i];
        }
        __syncthreads();
        {   // horizontal pass of row k*TH + hr, taps in k order (pipeline.rs:351-354)
            const float *src = s_in + hr * IWP + hc;
            float v[NT + 3];
#pragma unroll
            for (int j = 0; j < NT + 3; j++) v[j] = src[j];
            float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
#pragma unroll
            for (int q = 0; q < NT; q++) {
                float wk = taps.w[q];
                a0 += v[q] * wk;
                a1 += v[q + 1] * wk;
                a2 += v[q + 2] * wk;
                a3 += v[q + 3] * wk;
            }
            const int slot = (k % (P + 1)) * TH + hr;
            const float4 o = make_float4(a0, a1, a2, a3);
            *reinterpret_cast<float4 *>(s_h + slot * kTW + hc) = o;
            *reinterpret_cast<float4 *>(s_h + (slot + RING) * kTW + hc) = o;
            if constexpr (CENTER)
                *reinterpret_cast<float4 *>(s_c + slot * kTW + hc) = make_float4(v[R], v[R + 1], v[R + 2], v[R + 3]);
        }
        if (k + 1 < n_steps) fetch(k + 1);
        __syncthreads();
        if (k >= P) {   // vertical pass of output rows (k-P)*TH + 4*wv + [0, VR) (pipeline.rs:400-406)
            const int r0 = (k - P) * TH + 4 * wv;   // relative to y_start
            const int base = r0 % RING;
            float v[VR + 2 * R];
#pragma unroll
            for (int t = 0; t < VR + 2 * R; t++) v[t] = s_h[(base + t) * kTW + lane];
            float acc[VR];
#pragma unroll
            for (int j = 0; j < VR; j++) acc[j] = 0.0f;
#pragma unroll
            for (int q = 0; q < NT; q++) {
                float wk = taps.w[q];
#pragma unroll
                for (int j = 0; j < VR; j++) acc[j] += v[j + q] * wk;
            }
            const int x = x0 + lane;
#pragma unroll
            for (int j = 0; j < VR; j++) {
                const int y = y_start + r0 + j;
                if (y < y_end && x < w) {
                    const size_t o = (size_t)y * w + x;
                    if (a.dst_g) a.dst_g[o] = acc[j];
                    if constexpr (CENTER) {
                        int cs = base + R + j;
                        if (cs >= RING) cs -= RING;
                        const float center = s_c[cs * kTW + lane];
                        if (a.dst_d) a.dst_d[o] = acc[j] - center;   // pipeline.rs:611
                        if (a.dst_src) a.dst_src[o] = center;
                    }
                }
            }
        }
    }
}

template <int R, int SRC, bool CENTER>
__global__ __launch_bounds__(kBlurThreads) void blur_slide_kernel(BlurArgs a, Taps taps, int chunk) {
    const uint32_t tile = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    blur_strip<R, SRC, CENTER>(a, taps, (int)(tile % gridDim.x) * kTW, (int)(tile / gridDim.x) * chunk, chunk,
                               threadIdx.x);
}

template <int R>
static hipError_t blur_slide_dispatch(int src_mode, const BlurArgs &a, const Taps &t, int chunk, hipStream_t st) {
    if (chunk <= 0 || chunk % kSlideTH) return hipErrorInvalidValue;
    dim3 grid((a.w + kTW - 1) / kTW, (a.h + chunk - 1) / chunk);
    const bool center = a.dst_d || a.dst_src;
#define PS_SLIDE(M)                                                                                  \
    (center ? (blur_slide_kernel<R, M, true><<<grid, kBlurThreads, 0, st>>>(a, t, chunk), 0)         \
            : (blur_slide_kernel<R, M, false><<<grid, kBlurThreads, 0, st>>>(a, t, chunk), 0))
    switch (src_mode) {
    case SRC_F32: PS_SLIDE(SRC_F32); break;
    case SRC_DOWN2: PS_SLIDE(SRC_DOWN2); break;
    case SRC_U8:
    case SRC_F32_UP:
    case SRC_U8_UP:
        if constexpr (R == 4 || R == 5) {
            if (src_mode == SRC_U8) PS_SLIDE(SRC_U8);
            else if (src_mode == SRC_F32_UP) PS_SLIDE(SRC_F32_UP);
            else PS_SLIDE(SRC_U8_UP);
            break;
        } else {
            return hipErrorInvalidValue;
        }
    default: return hipErrorInvalidValue;
    }
#undef PS_SLIDE
    return hipGetLastError();
}

template <int R, int TH>
static hipError_t blur_dispatch_th(int src_mode, const BlurArgs &a, const Taps &t, hipStream_t st) {
    dim3 grid((a.w + kTW - 1) / kTW, (a.h + TH - 1) / TH);
    switch (src_mode) {
    case SRC_F32: blur_kernel<R, SRC_F32, TH><<<grid, kBlurThreads, 0, st>>>(a, t); break;
    case SRC_DOWN2: blur_kernel<R, SRC_DOWN2, TH><<<grid, kBlurThreads, 0, st>>>(a, t); break;
    case SRC_U8:
    case SRC_F32_UP:
    case SRC_U8_UP:
        if constexpr (R == 4 || R == 5) {
            if (src_mode == SRC_U8) blur_kernel<R, SRC_U8, TH><<<grid, kBlurThreads, 0, st>>>(a, t);
            else if (src_mode == SRC_F32_UP) blur_kernel<R, SRC_F32_UP, TH><<<grid, kBlurThreads, 0, st>>>(a, t);
            else blur_kernel<R, SRC_U8_UP, TH><<<grid, kBlurThreads, 0, st>>>(a, t);
            break;
        } else {
            return hipErrorInvalidValue;
        }
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int R>
static hipError_t blur_dispatch_src(int src_mode, const BlurArgs &a, const Taps &t, int th, hipStream_t st) {
    if (th & kBlurSlide) return blur_slide_dispatch<R>(src_mode, a, t, th & (kBlurSlide - 1), st);
    switch (th) {
    case 16: return blur_dispatch_th<R, 16>(src_mode, a, t, st);
    case 32: return blur_dispatch_th<R, 32>(src_mode, a, t, st);
    case 64: return blur_dispatch_th<R, 64>(src_mode, a, t, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_blur(int radius, int src_mode, const BlurArgs &a, const Taps &t, int th, hipStream_t st) {
    if (a.h <= 0 || a.w <= 0) return hipSuccess;
    switch (radius) {
    case 1: return blur_dispatch_src<1>(src_mode, a, t, th, st);
    case 2: return blur_dispatch_src<2>(src_mode, a, t, th, st);
    case 3: return blur_dispatch_src<3>(src_mode, a, t, th, st);
    case 4: return blur_dispatch_src<4>(src_mode, a, t, th, st);
    case 5: return blur_dispatch_src<5>(src_mode, a, t, th, st);
    case 6: return blur_dispatch_src<6>(src_mode, a, t, th, st);
    case 7: return blur_dispatch_src<7>(src_mode, a, t, th, st);
    case 8: return blur_dispatch_src<8>(src_mode, a, t, th, st);
    case 9: return blur_dispatch_src<9>(src_mode, a, t, th, st);
    case 10: return blur_dispatch_src<10>(src_mode, a, t, th, st);
    case 11: return blur_dispatch_src<11>(src_mode, a, t, th, st);
    case 12: return blur_dispatch_src<12>(src_mode, a, t, th, st);
    default: return hipErrorInvalidValue;
    }
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    fill_u32_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, v, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ extrema
struct EffScale {
    float v[kMaxLevels];
};

// pipeline.rs:664-687
__device__ __forceinline__ bool hessian_edge_ok(const float *c, int w, int x, int y, float center) {
    const float *rp = c + (size_t)(y - 1) * w, *rc = c + (size_t)y * w, *rn = c + (size_t)(y + 1) * w;
    float dxx = rc[x + 1] - 2.0f * center + rc[x - 1];
    float dyy = rn[x] - 2.0f * center + rp[x];
    float dxy = 0.25f * ((rn[x + 1] - rn[x - 1]) - (rp[x + 1] - rp[x - 1]));
    float trace = dxx + dyy;
    float det = dxx * dyy - dxy * dxy;
    const float THRESHOLD_RATIO = ((10.0f + 1.0f) * (10.0f + 1.0f)) / 10.0f;
    return det > 1e-9f && (trace * trace) < THRESHOLD_RATIO * det;
}

// pipeline.rs:697-770 (f32, operation by operation)
__device__ __forceinline__ void refine3d(const float *ps, const float *cs, const float *ns, int w, int x, int y,
                                         float &odx, float &ody, float &ods, float &oi) {
    const float *c_prev = cs + (size_t)(y - 1) * w, *c_curr = cs + (size_t)y * w, *c_next = cs + (size_t)(y + 1) * w;
    const float *p_prev = ps + (size_t)(y - 1) * w, *p_curr = ps + (size_t)y * w, *p_next = ps + (size_t)(y + 1) * w;
    const float *n_prev = ns + (size_t)(y - 1) * w, *n_curr = ns + (size_t)y * w, *n_next = ns + (size_t)(y + 1) * w;
    float center = c_curr[x];
    float gx = 0.5f * (c_curr[x + 1] - c_curr[x - 1]);
    float gy = 0.5f * (c_next[x] - c_prev[x]);
    float gs = 0.5f * (n_curr[x] - p_curr[x]);
    float dxx = c_curr[x + 1] - 2.0f * center + c_curr[x - 1];
    float dyy = c_next[x] - 2.0f * center + c_prev[x];
    float dss = n_curr[x] - 2.0f * center + p_curr[x];
    float dxy = 0.25f * ((c_next[x + 1] - c_next[x - 1]) - (c_prev[x + 1] - c_prev[x - 1]));
    float dxs = 0.25f * ((n_curr[x + 1] - n_curr[x - 1]) - (p_curr[x + 1] - p_curr[x - 1]));
    float dys = 0.25f * ((n_next[x] - n_prev[x]) - (p_next[x] - p_prev[x]));
    float c00 = dyy * dss - dys * dys;
    float c01 = dys * dxs - dxy * dss;
    float c02 = dxy * dys - dyy * dxs;
    float c11 = dxx * dss - dxs * dxs;
    float c12 = dxy * dxs - dxx * dys;
    float c22 = dxx * dyy - dxy * dxy;
    float det3 = dxx * c00 + dxy * c01 + dxs * c02;
    if (fabsf(det3) > 1e-9f) {
        float inv_det = -1.0f / det3;
        float ox = (c00 * gx + c01 * gy + c02 * gs) * inv_det;
        float oy = (c01 * gx + c11 * gy + c12 * gs) * inv_det;
        float os = (c02 * gx + c12 * gy + c22 * gs) * inv_det;
        if (fabsf(ox) <= 0.5f && fabsf(oy) <= 0.5f && fabsf(os) <= 0.5f) {
            odx = ox; ody = oy; ods = os;
            oi = center + 0.5f * (gx * ox + gy * oy + gs * os);
            return;
        }
    }
    if (fabsf(c22) > 1e-9f) {
        float inv_det2 = -1.0f / c22;
        float ox = (dyy * gx - dxy * gy) * inv_det2;
        float oy = (-dxy * gx + dxx * gy) * inv_det2;
        ox = ox < -0.5f ? -0.5f : (ox > 0.5f ? 0.5f : ox);
        oy = oy < -0.5f ? -0.5f : (oy > 0.5f ? 0.5f : oy);
        float os = 0.0f;
        if (fabsf(dss) > 1e-6f) {
            os = -gs / dss;
            os = os < -0.5f ? -0.5f : (os > 0.5f ? 0.5f : os);
        }
        odx = ox; ody = oy; ods = os;
        oi = center + 0.5f * (gx * ox + gy * oy + gs * os);
        return;
    }
    odx = 0.0f; ody = 0.0f; ods = 0.0f; oi = center;
}

// 64 x ROWS interior pixels per tile (ROWS = 32 for large octaves, 16 for small ones, whose
// launches are latency bound), 256 threads per workgroup: the S searchable DoG planes of the
// tile plus a one-pixel halo are staged in LDS (every load of a thread in flight together).
// The DoG planes are not stored in HBM on the extraction path: DoG s = G(s+1) - G(s) is formed
// here from the Gaussian planes (the same single f32 subtraction the blur kernel performs,
// pipeline.rs:611), so the blur writes one plane per level instead of two.
// Phase 1: each thread tests its ROWS/4 pixels x S slices against the contrast threshold and the 8
// in-plane neighbours (the cheap tests that reject almost everything) and queues survivors
// in LDS. Phase 2: the queue is processed by all lanes in parallel -- the 18 cross-scale
// neighbours (the outer DoG planes read from global memory, only for survivors), the Hessian
// edge test and the Taylor refinement -- and accepted candidates reserve output slots with
// one atomic per workgroup.
constexpr int kExW = 66;
constexpr long kExBigLevel = 1L << 20;   // levels of at least this many pixels use 32-row tiles
// Radius of the blur producing G(S+2) from G(S+1) (the level's sigma is fixed by S: base sigma
// 1.6, k = 2^(1/S)); launch_extrema checks it against the plan's radius.
constexpr int top_radius(int S) {
    return S == 3 ? 10 : S == 4 ? 8 : S == 5 ? 7 : S == 6 ? 6 : S <= 9 ? 5 : 4;
}
// One tile whose first interior pixel is (x0, y0): the body shared by the per-octave launch and
// the grouped persistent launch.
template <int PL, int kExRows>
__device__ __forceinline__ void extrema_tile(const float *gauss, int h, int w, float thr, float mult,
                                             const EffScale &eff, double log2k, int octave, Cand *out,
                                             uint32_t *count, uint32_t cap, uint32_t *err, const Taps &top,
                                             int x0, int y0, int tid) {
    constexpr int S = PL - 2;
    constexpr int kExH = kExRows + 2;
    constexpr int per_plane = kExH * kExW;
    constexpr int NL = (per_plane + 255) / 256;
    // only the S searchable DoG planes are staged; the outer planes 0 and S+1 are read from
    // global memory for the rare survivors of the in-plane test
    __shared__ float s_dog[S * per_plane];
    // phase-1 survivors ((slice-1) << 11 | (row-1) << 6 | (column-1)): strict in-plane extrema
    // of one kind are never 8-adjacent, so at most one maximum and one minimum per 2x2 block:
    // half of the tile's pixels per slice
    static_assert(S <= 16 && kExRows <= 32 && kExW - 2 <= 64, "queue entry fields");
    __shared__ uint16_t s_q[64 * kExRows * S / 2];
    __shared__ uint32_t s_qn;
    __shared__ uint32_t s_wcnt[4], s_base, s_tcnt[4];
    // G(S+2) (see phase 2): the round's slice-S entries, then per batch of TB of them the
    // horizontal sums (TB x HR rows x 3 columns) and the 3 x 3 vertical sums
    constexpr int TR = top_radius(S), TNT = 2 * TR + 1, HR = TNT + 2, TB = 256 / HR;
    __shared__ uint16_t s_top[256];
    __shared__ float s_hrow[TB * HR * 3];
    __shared__ float s_up[TB * 9];
    const size_t P = (size_t)h * w;
    if (tid == 0) s_qn = 0;
    int nan_seen = 0;
    {
        // Gaussian planes 1..S+1 give DoG planes 1..S
        float v[S + 1][NL];
#pragma unroll
        for (int j = 0; j < NL; j++) {
            int i = tid + j * 256;
            int r = i / kExW, c = i - r * kExW;
            size_t o = (size_t)min(y0 - 1 + r, h - 1) * w + min(x0 - 1 + c, w - 1);
#pragma unroll
            for (int pl = 0; pl <= S; pl++) v[pl][j] = i < per_plane ? gauss[(size_t)(pl + 1) * P + o] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < NL; j++) {
            int i = tid + j * 256;
#pragma unroll
            for (int pl = 0; pl < S; pl++)
                if (i < per_plane) {
                    const float d = v[pl + 1][j] - v[pl][j];
                    s_dog[pl * per_plane + i] = d;
                    nan_seen |= d != d;
                }
        }
    }
    // a NaN in the staged planes (only from NaN or infinite pixels of an f32 input) switches
    // the tile to the reference's literal comparisons, where a NaN neighbour fails them all
    const bool tile_nan = __syncthreads_or(nan_seen);
    // phase 1 (pipeline.rs:800-850, the threshold and in-plane part): each thread owns RT
    // consecutive rows of one column; per slice it reads the (RT+2) x 3 window around them
    // once and decides each pixel without branches, collecting survivors in a bit mask (bit
    // (s - s0) * RT + j for up to 64 / RT slices from s0); the wave then queues them with one
    // LDS atomic. "val above all 8 neighbours" is evaluated as val > (maximum of the 8): the
    // maximum of finite floats is one of its operands, so the decision is the same as the
    // reference's 8 strict comparisons, for ties as well. The row maxima (and minima) of the
    // 3-wide window are shared by the three output rows that use them.
    {
        constexpr int RT = kExRows / 4;
        constexpr int SC = 64 / RT;   // slices per mask
        const int tx = tid & 63, lx = tx + 1;
        const int ly0 = 1 + (tid >> 6) * RT;
        // rows j of this thread that lie in the level's interior (all of them off the last tile
        // row), repeated for each slice position of a mask; empty outside the interior columns
        uint64_t rows_ok = 0;
        if (x0 + tx < w - 1) {
            uint64_t rm = 0;
#pragma unroll
            for (int j = 0; j < RT; j++)
                if (y0 + ly0 + j - 1 < h - 1) rm |= 1ull << j;
#pragma unroll
            for (int k = 0; k < SC; k++) rows_ok |= rm << (k * RT);
        }
        auto flush = [&](uint64_t found, int s0) {
            const uint32_t nf = (uint32_t)__popcll(found);
            uint32_t incl = nf;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = __shfl_up(incl, d);
                if ((tid & 63) >= d) incl += o;
            }
            const uint32_t wtot = __shfl(incl, 63);
            uint32_t pos = 0;
            if ((tid & 63) == 63 && wtot) pos = atomicAdd(&s_qn, wtot);
            pos = __shfl(pos, 63) + incl - nf;
            while (found) {
                const int b = __ffsll((unsigned long long)found) - 1;
                found &= found - 1;
                const int sl = s0 + b / RT, j = b % RT;
                s_q[pos++] = (uint16_t)(((sl - 1) << 11) | ((ly0 + j - 1) << 6) | (lx - 1));
            }
        };
        auto scan = [&](auto literal) {
            uint64_t found = 0;
#pragma unroll
            for (int s = 1; s <= S; s++) {
                const float *cs = s_dog + (s - 1) * per_plane + (ly0 - 1) * kExW + lx - 1;
                float r[RT + 2][3];
#pragma unroll
                for (int i = 0; i < RT + 2; i++)
#pragma unroll
                    for (int c = 0; c < 3; c++) r[i][c] = cs[i * kExW + c];
                if constexpr (decltype(literal)::value) {
#pragma unroll
                    for (int j = 0; j < RT; j++) {
                        const float val = r[j + 1][1];
                        bool gt = true, lt = true;
#pragma unroll
                        for (int dy = 0; dy < 3; dy++)
#pragma unroll
                            for (int dx = 0; dx < 3; dx++)
                                if (dy != 1 || dx != 1) {
                                    gt &= val > r[j + dy][dx];
                                    lt &= val < r[j + dy][dx];
                                }
                        const bool ok = fabsf(val) > thr && (val > 0.0f ? gt : lt);
                        found |= (uint64_t)ok << (((s - 1) % SC) * RT + j);
                    }
                } else {
                    float hx[RT + 2], hn[RT + 2];
#pragma unroll
                    for (int i = 0; i < RT + 2; i++) {
                        hx[i] = fmaxf(fmaxf(r[i][0], r[i][1]), r[i][2]);
                        hn[i] = fminf(fminf(r[i][0], r[i][1]), r[i][2]);
                    }
#pragma unroll
                    for (int j = 0; j < RT; j++) {
                        const float val = r[j + 1][1];
                        const float m8 = fmaxf(fmaxf(hx[j], hx[j + 2]), fmaxf(r[j + 1][0], r[j + 1][2]));
                        const float n8 = fminf(fminf(hn[j], hn[j + 2]), fminf(r[j + 1][0], r[j + 1][2]));
                        // |val| > thr and val > m8 (val > 0), or |val| > thr and -val > -n8
                        const float u = val > 0.0f ? m8 : -n8;
                        const bool ok = fabsf(val) > fmaxf(thr, u);
                        found |= (uint64_t)ok << (((s - 1) % SC) * RT + j);
                    }
                }
                if (s % SC == 0 || s == S) {
                    flush(found & rows_ok, s - (s - 1) % SC);
                    found = 0;
                }
            }
        };
        if (tile_nan) scan(std::true_type{});
        else scan(std::false_type{});
    }
    __syncthreads();
    // phase 2 (pipeline.rs:852-893); accepted candidates of a round reserve their output
    // slots with one global atomic per workgroup (ballot prefix inside the waves)
    const uint32_t nq = s_qn;
    const int lane = tid & 63, wv = tid >> 6;
    for (uint32_t q0 = 0; q0 < nq; q0 += 256) {
        // queue entry q0 + 64*wave + lane: a short queue (a few entries per tile is typical)
        // occupies as few waves as possible, and a wave without entries skips the per-entry
        // tests and refinement entirely instead of executing them for one or two live lanes
        const uint32_t qi = q0 + 64 * (uint32_t)wv + (uint32_t)lane;
        const uint32_t e = qi < nq ? s_q[qi] : 0u;
        const bool is_top = qi < nq && (int)(e >> 11) + 1 == S;
        // The upper neighbour plane of slice-S entries is DoG S+1 = G(S+2) - G(S+1). The
        // extraction path does not store G(S+2) (no other stage reads it): its 3 x 3 values
        // around each such entry are computed here from G(S+1) with that level's taps, in the
        // blur kernels' order (horizontal sums over clamped columns, then the vertical sum over
        // clamped rows, taps in index order from 0.0f), so they are the values the level's
        // blur would have written. The round's slice-S entries are compacted; per batch of TB
        // entries one thread forms the three horizontal sums of one window row and one thread
        // one vertical sum.
        float up[9];
        {
            const uint64_t tm = __ballot(is_top);
            if (lane == 0) s_tcnt[wv] = (uint32_t)__popcll(tm);
            __syncthreads();
            uint32_t my_top = (uint32_t)__popcll(tm & ((1ull << lane) - 1ull)), n_top = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (k < wv) my_top += s_tcnt[k];
                n_top += s_tcnt[k];
            }
            if (is_top) s_top[my_top] = (uint16_t)e;
            __syncthreads();
            for (uint32_t t0 = 0; t0 < n_top; t0 += TB) {
                const int nb_t = (int)min((uint32_t)TB, n_top - t0);
                if (tid < nb_t * HR) {
                    const int k = tid / HR, rr = tid - k * HR;
                    const uint32_t te = s_top[t0 + k];
                    const int ty = y0 + (int)((te >> 6) & 31u), tx = x0 + (int)(te & 63u);
                    const float *src = gauss + (size_t)(S + 1) * P + (size_t)min(max(ty - 1 - TR + rr, 0), h - 1) * w;
                    float v[TNT + 2];
#pragma unroll
                    for (int q = 0; q < TNT + 2; q++) v[q] = src[min(max(tx - 1 - TR + q, 0), w - 1)];
#pragma unroll
                    for (int cc = 0; cc < 3; cc++) {
                        float acc = 0.0f;
#pragma unroll
                        for (int q = 0; q < TNT; q++) acc += v[cc + q] * top.w[q];
                        s_hrow[(k * HR + rr) * 3 + cc] = acc;
                    }
                }
                __syncthreads();
                if (tid < nb_t * 9) {
                    const int k = tid / 9, o = tid - k * 9;
                    const float *hr = s_hrow + (k * HR + o / 3) * 3 + o % 3;
                    float acc = 0.0f;
#pragma unroll
                    for (int q = 0; q < TNT; q++) acc += hr[q * 3] * top.w[q];
                    s_up[tid] = acc;
                }
                __syncthreads();
                if (is_top && my_top - t0 < (uint32_t)nb_t) {
#pragma unroll
                    for (int o = 0; o < 9; o++) up[o] = s_up[(my_top - t0) * 9 + o];
                }
                __syncthreads();
            }
        }
        bool accept = false;
        Cand c;
        if (qi < nq) {
            const int s = (int)(e >> 11) + 1, ly = (int)((e >> 6) & 31u) + 1, lx = (int)(e & 63u) + 1;
            const int x = x0 + lx - 1, y = y0 + ly - 1;
            // 3 x 3 x 3 neighbourhood (slices s-1, s, s+1) from LDS, or for the outer planes
            // formed from the Gaussian planes in global memory; interior pixels need no clamping
            float nb[3][9];
#pragma unroll
            for (int p = 0; p < 3; p++) {
                const int pl = s - 1 + p;
                if (pl >= 1 && pl <= S) {
                    const float *q = s_dog + (pl - 1) * per_plane + (ly - 1) * kExW + lx - 1;
#pragma unroll
                    for (int dy = 0; dy < 3; dy++)
#pragma unroll
                        for (int dx = 0; dx < 3; dx++) nb[p][dy * 3 + dx] = q[dy * kExW + dx];
                } else {
                    const float *g0 = gauss + (size_t)pl * P + (size_t)(y - 1) * w + (x - 1);
                    const float *g1 = g0 + P;
                    float a[9], b[9];
#pragma unroll
                    for (int dy = 0; dy < 3; dy++)
#pragma unroll
                        for (int dx = 0; dx < 3; dx++) {
                            a[dy * 3 + dx] = pl == S + 1 ? up[dy * 3 + dx] : g1[(size_t)dy * w + dx];
                            b[dy * 3 + dx] = g0[(size_t)dy * w + dx];
                        }
#pragma unroll
                    for (int k = 0; k < 9; k++) nb[p][k] = a[k] - b[k];
                }
            }
            const float *ps = nb[0], *cs = nb[1], *ns = nb[2];
            const float val = cs[4];
            bool ext = true;
            if (val > 0.0f) {
#pragma unroll
                for (int o = 0; o < 9; o++) ext = ext && (val > ps[o]) && (val > ns[o]);
            } else {
#pragma unroll
                for (int o = 0; o < 9; o++) ext = ext && (val < ps[o]) && (val < ns[o]);
            }
            if (ext && hessian_edge_ok(cs, 3, 1, 1, val)) {
                float dx, dy, ds, interp;
                refine3d(ps, cs, ns, 3, 1, 1, dx, dy, ds, interp);
                if (fabsf(interp) > thr) {
                    // k.powf(ds): evaluated in f64 and rounded once (glibc powf agrees except in
                    // ~0.06% of arguments, where the two differ by one ulp)
                    float kpow = (float)exp2((double)ds * log2k);
                    float refined_scale = eff.v[s] * kpow;
                    float lf = roundf((float)s + ds);   // pipeline.rs:882-885
                    lf = lf < 1.0f ? 1.0f : (lf > (float)S ? (float)S : lf);
                    c.x = ((float)x + dx) * mult;
                    c.y = ((float)y + dy) * mult;
                    c.scale = fmaxf(refined_scale, 0.25f * 1.6f);
                    c.response = interp;
                    c.octave = (uint32_t)octave;
                    c.layer = (uint32_t)lf;
                    c.key = ((uint64_t)octave << 56) | ((uint64_t)s << 48) | ((uint64_t)y << 24) | (uint64_t)x;
                    accept = true;
                }
            }
        }
        const uint64_t m = __ballot(accept);
        if (lane == 0) s_wcnt[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        if (tid == 0) {
            const uint32_t tot = s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
            s_base = tot ? atomicAdd(count, tot) : 0u;
        }
        __syncthreads();
        if (accept) {
            uint32_t i = s_base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            for (int k = 0; k < wv; k++) i += s_wcnt[k];
            if (i < cap) out[i] = c;
            else atomicOr(err, kErrCandOverflow);
        }
        __syncthreads();
    }
}

template <int PL, int kExRows>
__global__ __launch_bounds__(256) void extrema_kernel(const float *gauss, int h, int w, float thr, float mult,
                                                      EffScale eff, double log2k, int octave, Cand *out,
                                                      uint32_t *count, uint32_t cap, uint32_t *err, Taps top) {
    // XCD-contiguous tile order (as the blur kernels): horizontally neighbouring tiles share
    // the cache lines of their halo columns, and land on the same XCD's L2
    const uint32_t tile = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    extrema_tile<PL, kExRows>(gauss, h, w, thr, mult, eff, log2k, octave, out, count, cap, err, top,
                              1 + (int)(tile % gridDim.x) * 64, 1 + (int)(tile / gridDim.x) * kExRows, threadIdx.x);
}

hipError_t launch_extrema(const float *gauss_octave, int h, int w, int S, float thr, float mult, const float *eff_scale,
                          double log2k, int octave, Cand *out, uint32_t *count, uint32_t cap, uint32_t *err,
                          const Taps &top_taps, int top_r, hipStream_t st) {
    if (h < 3 || w < 3 || S < 1) return hipSuccess;
    if (S < 3 || S > 10 || top_r != top_radius(S)) return hipErrorInvalidValue;
    EffScale e;
    for (int i = 0; i < kMaxLevels; i++) e.v[i] = eff_scale[i];
    const bool big = (long)h * w >= kExBigLevel;
    const int rows = big ? 32 : 16;
    dim3 grid((w - 2 + 63) / 64, (h - 2 + rows - 1) / rows);
#define PS_EXTREMA_CASE(SS)                                                                                   \
    case SS:                                                                                                  \
        if (big)                                                                                              \
            extrema_kernel<SS + 2, 32><<<grid, 256, 0, st>>>(gauss_octave, h, w, thr, mult, e, log2k, octave,  \
                                                             out, count, cap, err, top_taps);                 \
        else                                                                                                  \
            extrema_kernel<SS + 2, 16><<<grid, 256, 0, st>>>(gauss_octave, h, w, thr, mult, e, log2k, octave,  \
                                                             out, count, cap, err, top_taps);                 \
        break;
    switch (S) {
        PS_EXTREMA_CASE(3) PS_EXTREMA_CASE(4) PS_EXTREMA_CASE(5) PS_EXTREMA_CASE(6)
        PS_EXTREMA_CASE(7) PS_EXTREMA_CASE(8) PS_EXTREMA_CASE(9) PS_EXTREMA_CASE(10)
    default: return hipErrorInvalidValue;
    }
#undef PS_EXTREMA_CASE
    return hipGetLastError();
}

// ------------------------------------------------------------------ ordering + cap
// Output order (pipeline.rs:1643-1651): scan order when n <= limit (no sort); otherwise the
// stable |response|-descending sort truncated to `limit`, i.e. the order
// (|response| desc, scan key asc). For n > limit an exact radix select (four 8-bit passes
// over the |response| bit patterns, which order like the positive floats) finds T, the
// limit-th largest |response|; only candidates with |response| >= T (limit plus ties) are
// ranked. rank[i] = #{j in subset : j precedes i}; a fixed grid of workgroups walks
// (i-tile, j-chunk) pairs of the subset and adds integer partial counts, so the result is
// exact and independent of scheduling.
constexpr int kRankTile = 256;
constexpr int kRankGrid = 24;

// scratch words: [0] subset size, [1] threshold bits T, [2] by-response flag
constexpr uint32_t kSelLds = 12288;   // candidates whose |response| bits fit in LDS for the 4 passes
__device__ __forceinline__ void rank_select_body(const Cand *c, const uint32_t *count, uint32_t cap_in,
                                                 uint32_t limit, uint32_t *words) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_prefix, s_need;
    __shared__ uint32_t s_bits[kSelLds];
    const uint32_t n = min(*count, cap_in);
    if (threadIdx.x == 0) words[0] = 0;   // subset counter for rank_gather_kernel
    if (n <= limit) {
        if (threadIdx.x == 0) { words[1] = 0; words[2] = 0; }
        return;
    }
    const bool in_lds = n <= kSelLds;
    if (in_lds) {
        for (uint32_t i0 = threadIdx.x; i0 < n; i0 += 4 * 1024) {
            uint32_t v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                uint32_t i = i0 + u * 1024;
                v[u] = i < n ? __float_as_uint(fabsf(c[i].response)) : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (i0 + u * 1024 < n) s_bits[i0 + u * 1024] = v[u];
        }
    }
    if (threadIdx.x == 0) { s_prefix = 0; s_need = limit; }
    uint32_t mask = 0;
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 24 - 8 * pass;
        if (threadIdx.x < 256) hist[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t prefix = s_prefix;
        for (uint32_t i0 = threadIdx.x; i0 < n; i0 += 4 * 1024) {
            uint32_t v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                uint32_t i = i0 + u * 1024;
                v[u] = i < n ? (in_lds ? s_bits[i] : __float_as_uint(fabsf(c[i].response))) : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                // first pass (the exponent byte takes only a few values): wave-aggregated
                // increments, one LDS atomic per distinct bin in the wave; later passes spread
                // over many bins and use plain atomics
                bool act = i0 + u * 1024 < n && (v[u] & mask) == prefix;
                const uint32_t bin = (v[u] >> shift) & 255u;
                if (pass > 0) {
                    if (act) atomicAdd(&hist[bin], 1u);
                    continue;
                }
                uint64_t pending = __ballot(act);
                while (pending) {
                    const int leader = __ffsll((unsigned long long)pending) - 1;
                    const uint32_t lb = __shfl(bin, leader);
                    const uint64_t same = __ballot(act && bin == lb);
                    if ((int)(threadIdx.x & 63) == leader) atomicAdd(&hist[lb], (uint32_t)__popcll(same));
                    pending &= ~same;
                    act = act && bin != lb;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            // lane l owns buckets 255-4l .. 252-4l (descending); suffix counts by an inclusive scan
            const int lane = threadIdx.x;
            uint32_t b4[4], loc = 0;
#pragma unroll
            for (int u = 0; u < 4; u++) { b4[u] = hist[255 - 4 * lane - u]; loc += b4[u]; }
            uint32_t inc = loc;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                uint32_t t = __shfl_up(inc, o);
                if (lane >= o) inc += t;
            }
            const uint32_t need = s_need, above = inc - loc;
            if (above < need && inc >= need) {
                uint32_t acc = above;
                for (int u = 0; u < 4; u++) {
                    if (acc + b4[u] >= need) {
                        s_prefix = prefix | ((uint32_t)(255 - 4 * lane - u) << shift);
                        s_need = need - acc;
                        break;
                    }
                    acc += b4[u];
                }
            }
        }
        mask |= 255u << shift;
        __syncthreads();
    }
    if (threadIdx.x == 0) { words[1] = s_prefix; words[2] = 1; }
}

// Direct ordering of up to kFastN candidates whose scan keys pack into 31 bits (octave < 8,
// slice < 16, octave-local row and column < 4096): each candidate's output position is the
// number of candidates that precede it, counted over all pairs on one 64-bit key per candidate
// -- (~|response| bits, packed scan key) when n > limit, the packed scan key alone otherwise --
// so no threshold select or subset gather is needed. Keys are distinct (one candidate per
// octave, slice and pixel), so the positions are a permutation. A (256-candidate tile, j-chunk)
// grid writes one partial count per candidate and chunk; the scatter sums them. Larger or
// unpackable inputs take the select / gather / count path below, whose kernels then run and
// these return at once (and the other way round).
constexpr uint32_t kFastN = 32768;
constexpr uint32_t kFastChunks = 32;
constexpr int kFastPerThread = 1;   // candidates per thread (more per thread measured slower)

// k: octave << 56 | slice << 48 | row << 24 | column
__device__ __forceinline__ uint64_t fast_key(uint64_t k, float response, bool by_resp) {
    const uint32_t k31 = (uint32_t)(k >> 56) << 28 | (uint32_t)((k >> 48) & 15u) << 24 |
                         (uint32_t)((k >> 24) & 4095u) << 12 | (uint32_t)(k & 4095u);
    return by_resp ? (uint64_t)~__float_as_uint(fabsf(response)) << 32 | k31 : (uint64_t)k31;
}

__device__ __forceinline__ bool fast_order(const uint32_t *count, uint32_t cap_in, bool fast) {
    return fast && min(*count, cap_in) <= kFastN;
}

__global__ __launch_bounds__(256) void rank_fast_count_kernel(const Cand *c, const uint32_t *count, uint32_t cap_in,
                                                              uint32_t limit, uint32_t *part) {
    __shared__ uint64_t s_k[kFastN / kFastChunks];
    const uint32_t n = min(*count, cap_in);
    const uint32_t per = (n + kFastChunks - 1) / kFastChunks;
    const uint32_t j0 = blockIdx.y * per, j1 = min(j0 + per, n);
    constexpr uint32_t kTile = 256 * kFastPerThread;
    if (n > kFastN || blockIdx.x * kTile >= n || j0 >= j1) return;
    const bool by_resp = n > limit;
    // every load of the thread (its chunk entries and its own candidates) in flight together
    constexpr int kPer = kFastN / kFastChunks / 256;
    uint64_t kj[kPer], kc[kFastPerThread];
    float rj[kPer], rc[kFastPerThread];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const uint32_t j = min(j0 + threadIdx.x + 256 * u, j1 - 1);
        kj[u] = c[j].key;
        rj[u] = c[j].response;
    }
#pragma unroll
    for (int v = 0; v < kFastPerThread; v++) {
        const uint32_t i = min(blockIdx.x * kTile + threadIdx.x + 256 * v, n - 1);
        kc[v] = c[i].key;
        rc[v] = c[i].response;
    }
#pragma unroll
    for (int u = 0; u < kPer; u++)
        if (j0 + threadIdx.x + 256 * u < j1) s_k[threadIdx.x + 256 * u] = fast_key(kj[u], rj[u], by_resp);
    uint64_t ki[kFastPerThread];
    uint32_t acc[kFastPerThread];
#pragma unroll
    for (int v = 0; v < kFastPerThread; v++) {
        ki[v] = fast_key(kc[v], rc[v], by_resp);
        acc[v] = 0;
    }
    __syncthreads();
    const uint32_t m = j1 - j0;
#pragma unroll 4
    for (uint32_t t = 0; t < m; t++) {
        const uint64_t kt = s_k[t];
#pragma unroll
        for (int v = 0; v < kFastPerThread; v++) acc[v] += kt < ki[v] ? 1u : 0u;
    }
#pragma unroll
    for (int v = 0; v < kFastPerThread; v++) {
        const uint32_t i = blockIdx.x * kTile + threadIdx.x + 256 * v;
        if (i < n) part[(size_t)blockIdx.y * kFastN + i] = acc[v];
    }
}

__global__ __launch_bounds__(1024) void rank_select_kernel(const Cand *c, const uint32_t *count, uint32_t cap_in,
                                                           uint32_t limit, uint32_t *words, bool fast) {
    if (fast_order(count, cap_in, fast)) return;
    rank_select_body(c, count, cap_in, limit, words);
}

// compact record of a subset member for the rank count
struct RankRec {
    float r;
    uint32_t pad;
    uint64_t key;
};

// one workgroup reservation per 256 candidates (ballot prefix), subset in an arbitrary but
// complete order; the rank count does not depend on it
// candidates [i0, i0 + 256) of n
__device__ __forceinline__ void rank_gather_block(const Cand *c, uint32_t n, uint32_t limit, uint32_t *subset,
                                                  RankRec *rec, uint32_t *rank, uint32_t *words, uint32_t i0) {
    __shared__ uint32_t s_w[4], s_base;
    const uint32_t i = i0 + threadIdx.x;
    bool take = false;
    Cand ci;
    if (i < n) {
        ci = c[i];
        take = !(n > limit && __float_as_uint(fabsf(ci.response)) < words[1]);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t m = __ballot(take);
    if (lane == 0) s_w[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        s_base = tot ? atomicAdd(&words[0], tot) : 0u;
    }
    __syncthreads();
    if (!take) return;
    uint32_t pos = s_base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    for (int k = 0; k < wv; k++) pos += s_w[k];
    subset[pos] = i;
    rank[pos] = 0;
    RankRec r;
    r.r = fabsf(ci.response);
    r.pad = 0;
    r.key = ci.key;
    rec[pos] = r;
}

__global__ __launch_bounds__(256) void rank_gather_kernel(const Cand *c, const uint32_t *count, uint32_t cap_in,
                                                          uint32_t limit, uint32_t *subset, RankRec *rec,
                                                          uint32_t *rank, uint32_t *words, bool fast) {
    const uint32_t n = min(*count, cap_in);
    if (blockIdx.x * blockDim.x >= n || fast_order(count, cap_in, fast)) return;
    rank_gather_block(c, n, limit, subset, rec, rank, words, blockIdx.x * blockDim.x);
}

// i-tiles of 256 (one per thread) against j-chunks of kRankJ staged in LDS; a fixed grid
// walks the (i-tile, j-chunk) pairs
constexpr int kRankJ = 64;
// this workgroup's share (bx of gx i-tiles, by of gy j-chunks) of one image's rank counts
__device__ __forceinline__ void rank_count_body(const RankRec *rec, const uint32_t *words, uint32_t *rank,
                                                uint32_t bx, uint32_t gx, uint32_t by, uint32_t gy) {
    __shared__ float s_r[kRankJ];
    __shared__ uint64_t s_k[kRankJ];
    const uint32_t m = words[0];
    const bool by_resp = words[2] != 0;
    const uint32_t ntile = (m + kRankTile - 1) / kRankTile, nj = (m + kRankJ - 1) / kRankJ;
    for (uint32_t ti = bx; ti < ntile; ti += gx) {
        const uint32_t i = ti * kRankTile + threadIdx.x;
        float ri = 0.0f;
        uint64_t ki = 0;
        if (i < m) { RankRec a = rec[i]; ri = a.r; ki = a.key; }
        uint32_t acc = 0;
        for (uint32_t tj = by; tj < nj; tj += gy) {
            __syncthreads();
            if (threadIdx.x < kRankJ) {
                const uint32_t j = tj * kRankJ + threadIdx.x;
                RankRec b = j < m ? rec[j] : RankRec{0.0f, 0u, ~0ull};
                s_r[threadIdx.x] = b.r;
                s_k[threadIdx.x] = b.key;
            }
            __syncthreads();
            const uint32_t mm = min((uint32_t)kRankJ, m - tj * kRankJ);
            if (by_resp) {
#pragma unroll 16
                for (uint32_t t = 0; t < mm; t++) {
                    float r = s_r[t];
                    acc += (r > ri || (r == ri && s_k[t] < ki)) ? 1u : 0u;
                }
            } else {
#pragma unroll 16
                for (uint32_t t = 0; t < mm; t++) acc += s_k[t] < ki ? 1u : 0u;
            }
        }
        if (i < m && acc) atomicAdd(&rank[i], acc);
    }
}

__global__ __launch_bounds__(kRankTile) void rank_count_kernel(const RankRec *rec, const uint32_t *words,
                                                               uint32_t *rank, const uint32_t *count, uint32_t cap_in,
                                                               bool fast) {
    if (fast_order(count, cap_in, fast)) return;
    rank_count_body(rec, words, rank, blockIdx.x, gridDim.x, blockIdx.y, gridDim.y);
}

__device__ __forceinline__ void rank_scatter_one(const Cand *c, const uint32_t *subset, uint32_t m, uint32_t limit,
                                                 const uint32_t *rank, Cand *out, uint32_t *out_count, uint32_t i) {
    if (i == 0) *out_count = min(m, limit);
    if (i < m && rank[i] < limit) out[rank[i]] = c[subset[i]];
}

__global__ void rank_scatter_kernel(const Cand *c, const uint32_t *subset, const uint32_t *words, uint32_t limit,
                                    const uint32_t *rank, Cand *out, uint32_t *out_count, const uint32_t *count,
                                    uint32_t cap_in, bool fast, const uint32_t *part) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (fast_order(count, cap_in, fast)) {
        const uint32_t n = min(*count, cap_in);
        if (i == 0) *out_count = min(n, limit);
        if (i >= n) return;
        const uint32_t per = (n + kFastChunks - 1) / kFastChunks;
        uint32_t p[kFastChunks];
#pragma unroll
        for (uint32_t ch = 0; ch < kFastChunks; ch++) p[ch] = ch * per < n ? part[(size_t)ch * kFastN + i] : 0u;
        uint32_t r = 0;
#pragma unroll
        for (uint32_t ch = 0; ch < kFastChunks; ch++) r += p[ch];
        if (r < limit) out[r] = c[i];
        return;
    }
    rank_scatter_one(c, subset, words[0], limit, rank, out, out_count, i);
}

// scratch of one image's ordering: rank[cap_in], subset[cap_in], 64 words, RankRec[cap_in]
__device__ __forceinline__ uint32_t *rank_words(void *scratch, uint32_t cap_in) {
    return static_cast<uint32_t *>(scratch) + 2 * (size_t)cap_in;
}

size_t rank_scratch_bytes(uint32_t cap_in) {
    return ((size_t)2 * cap_in + 64) * sizeof(uint32_t) + (size_t)cap_in * sizeof(RankRec);
}

hipError_t launch_rank(const Cand *cand, const uint32_t *count, uint32_t cap_in, uint32_t limit, Cand *out,
                       uint32_t *out_count, void *scratch, bool packed_keys, hipStream_t st) {
    uint32_t *rank = static_cast<uint32_t *>(scratch);
    uint32_t *subset = rank + cap_in;
    uint32_t *words = subset + cap_in;
    RankRec *rec = reinterpret_cast<RankRec *>(words + 64);
    // the direct path's partial counts reuse the record area (kFastChunks x kFastN words)
    const bool fast = packed_keys && (size_t)cap_in * sizeof(RankRec) >= (size_t)kFastChunks * kFastN * 4;
    uint32_t *part = reinterpret_cast<uint32_t *>(rec);
    const unsigned blocks = (cap_in + 255) / 256;
    if (fast)
        rank_fast_count_kernel<<<dim3(kFastN / (256 * kFastPerThread), kFastChunks), 256, 0, st>>>(cand, count, cap_in,
                                                                                                 limit, part);
    rank_select_kernel<<<1, 1024, 0, st>>>(cand, count, cap_in, limit, words, fast);
    rank_gather_kernel<<<blocks, 256, 0, st>>>(cand, count, cap_in, limit, subset, rec, rank, words, fast);
    rank_count_kernel<<<dim3(kRankGrid, 4 * kRankGrid), kRankTile, 0, st>>>(rec, words, rank, count, cap_in, fast);
    rank_scatter_kernel<<<blocks, 256, 0, st>>>(cand, subset, words, limit, rank, out, out_count, count, cap_in, fast,
                                                part);
    return hipGetLastError();
}

// ------------------------------------------------------------------ grouped persistent launches
// A grouped launch runs one stage for many image levels at once (a level of every image of a
// batch). Its units -- output tiles, sliding-window strips, extrema tiles -- are numbered job by
// job (job j owns units [unit0_j, unit0_{j+1})), and a grid of at most as many workgroups as the
// device holds at once, a multiple of the 8 XCDs, walks them: XCD k owns one contiguous range of
// units and its workgroups stride through it, so units whose halos overlap run at about the same
// time under the same L2. The per-unit bodies are the single-level kernels' own, so every image
// gets exactly the values of its own launches.
__device__ __forceinline__ void unit_walk(uint32_t total, uint32_t &u, uint32_t &end, uint32_t &step) {
    const uint32_t k = blockIdx.x % kXcds, q = total / kXcds, r = total % kXcds;
    const uint32_t lo = k * q + min(k, r);
    end = lo + q + (k < r ? 1u : 0u);
    u = lo + blockIdx.x / kXcds;
    step = gridDim.x / kXcds;
}

// the job owning unit u: the last one with unit0 <= u (jobs sorted by unit0, job 0 at 0)
template <typename J>
__device__ __forceinline__ uint32_t find_job(const J *jobs, uint32_t n, uint32_t u) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].unit0 <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}

// a workgroup's units only increase, so after the first search the job only moves forward
template <typename J>
__device__ __forceinline__ uint32_t next_job(const J *jobs, uint32_t n, uint32_t jb, uint32_t u) {
    while (jb + 1 < n && jobs[jb + 1].unit0 <= u) jb++;
    return jb;
}

template <int R, int SRC, int kTH>
__global__ __launch_bounds__(kBlurThreads) void blur_group_kernel(const BlurJob *__restrict__ jobs, uint32_t n_jobs,
                                                                  uint32_t total, Taps taps) {
    uint32_t u, end, step;
    unit_walk(total, u, end, step);
    if (u >= end) return;
    uint32_t jb = find_job(jobs, n_jobs, u);
    for (; u < end; u += step) {
        jb = next_job(jobs, n_jobs, jb, u);
        const BlurArgs a = global_args(jobs[jb].a);
        const uint32_t t = u - jobs[jb].unit0, nx = jobs[jb].units_x;
        blur_tile<R, SRC, kTH>(a, taps, (int)(t % nx) * kTW, (int)(t / nx) * kTH, opaque_tid());
        __syncthreads();   // the next unit re-stages the LDS tiles
    }
}

template <int R, int SRC, bool CENTER>
__global__ __launch_bounds__(kBlurThreads) void blur_strip_group_kernel(const BlurJob *__restrict__ jobs, uint32_t n_jobs,
                                                                        uint32_t total, Taps taps, int chunk) {
    uint32_t u, end, step;
    unit_walk(total, u, end, step);
    if (u >= end) return;
    uint32_t jb = find_job(jobs, n_jobs, u);
    for (; u < end; u += step) {
        jb = next_job(jobs, n_jobs, jb, u);
        const BlurArgs a = global_args(jobs[jb].a);
        const uint32_t t = u - jobs[jb].unit0, nx = jobs[jb].units_x;
        blur_strip<R, SRC, CENTER>(a, taps, (int)(t % nx) * kTW, (int)(t / nx) * chunk, chunk, opaque_tid());
        __syncthreads();
    }
}

// Workgroups a grouped launch of `fn` uses for n_units units: all that fit on the device at
// once (occupancy x CUs, queried once per kernel), at most one per unit, rounded to the XCDs.
static uint32_t group_grid(const void *fn, int threads, uint32_t n_units) {
    static std::mutex mu;
    static std::map<const void *, uint32_t> resident;
    uint32_t cap = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = resident.find(fn);
        if (it != resident.end()) {
            cap = it->second;
        } else {
            int dev = 0, cus = 0, per = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, 0) != hipSuccess)
                return 0;
            cap = (uint32_t)std::max(per, 1) * (uint32_t)std::max(cus, 1);
            resident[fn] = cap;
        }
    }
    uint32_t g = std::min(cap, (n_units + kXcds - 1) / kXcds * kXcds);
    return std::max(g / kXcds * kXcds, kXcds);
}

template <int R, int SRC>
static hipError_t blur_group_src(const BlurJob *jobs, uint32_t n_jobs, uint32_t n_units, const Taps &t, int shape,
                                 bool center, hipStream_t st) {
    if (shape & kBlurSlide) {
        const int chunk = shape & (kBlurSlide - 1);
        if (chunk <= 0 || chunk % kSlideTH) return hipErrorInvalidValue;
        const void *fn = center ? (const void *)blur_strip_group_kernel<R, SRC, true>
                                : (const void *)blur_strip_group_kernel<R, SRC, false>;
        const uint32_t g = group_grid(fn, kBlurThreads, n_units);
        if (!g) return hipErrorInvalidValue;
        if (center) blur_strip_group_kernel<R, SRC, true><<<g, kBlurThreads, 0, st>>>(jobs, n_jobs, n_units, t, chunk);
        else blur_strip_group_kernel<R, SRC, false><<<g, kBlurThreads, 0, st>>>(jobs, n_jobs, n_units, t, chunk);
        return hipGetLastError();
    }
#define PS_GROUP_TH(TH)                                                                                     \
    case TH: {                                                                                              \
        const uint32_t g = group_grid((const void *)blur_group_kernel<R, SRC, TH>, kBlurThreads, n_units);  \
        if (!g) return hipErrorInvalidValue;                                                                \
        blur_group_kernel<R, SRC, TH><<<g, kBlurThreads, 0, st>>>(jobs, n_jobs, n_units, t);                \
        return hipGetLastError();                                                                           \
    }
    switch (shape) {
        PS_GROUP_TH(16) PS_GROUP_TH(32) PS_GROUP_TH(64)
    default: return hipErrorInvalidValue;
    }
#undef PS_GROUP_TH
}

template <int R>
static hipError_t blur_group_dispatch(int src_mode, const BlurJob *jobs, uint32_t n_jobs, uint32_t n_units,
                                      const Taps &t, int shape, bool center, hipStream_t st) {
    switch (src_mode) {
    case SRC_F32: return blur_group_src<R, SRC_F32>(jobs, n_jobs, n_units, t, shape, center, st);
    case SRC_DOWN2: return blur_group_src<R, SRC_DOWN2>(jobs, n_jobs, n_units, t, shape, center, st);
    case SRC_U8:
    case SRC_U8_UP:
        if constexpr (R == 4 || R == 5) {
            if (src_mode == SRC_U8) return blur_group_src<R, SRC_U8>(jobs, n_jobs, n_units, t, shape, center, st);
            return blur_group_src<R, SRC_U8_UP>(jobs, n_jobs, n_units, t, shape, center, st);
        } else {
            return hipErrorInvalidValue;
        }
    default: return hipErrorInvalidValue;
    }
}

uint32_t group_units(int h, int w, int shape) {
    if (h <= 0 || w <= 0) return 0;
    const int rows = (shape & kBlurSlide) ? (shape & (kBlurSlide - 1)) : shape;
    if (rows <= 0) return 0;
    return (uint32_t)((w + kTW - 1) / kTW) * (uint32_t)((h + rows - 1) / rows);
}

hipError_t launch_blur_group(int radius, int src_mode, const BlurJob *jobs, uint32_t n_jobs, uint32_t n_units,
                             const Taps &t, int shape, bool center, hipStream_t st) {
    if (n_jobs == 0 || n_units == 0) return hipSuccess;
    switch (radius) {
    case 1: return blur_group_dispatch<1>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 2: return blur_group_dispatch<2>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 3: return blur_group_dispatch<3>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 4: return blur_group_dispatch<4>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 5: return blur_group_dispatch<5>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 6: return blur_group_dispatch<6>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 7: return blur_group_dispatch<7>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 8: return blur_group_dispatch<8>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 9: return blur_group_dispatch<9>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 10: return blur_group_dispatch<10>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 11: return blur_group_dispatch<11>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    case 12: return blur_group_dispatch<12>(src_mode, jobs, n_jobs, n_units, t, shape, center, st);
    default: return hipErrorInvalidValue;
    }
}

template <int PL, int kExRows>
__global__ __launch_bounds__(256) void extrema_group_kernel(const ExtremaJob *__restrict__ jobs, uint32_t n_jobs, uint32_t total,
                                                            float thr, float mult, EffScale eff, double log2k,
                                                            int octave, uint32_t cap, Taps top) {
    uint32_t u, end, step;
    unit_walk(total, u, end, step);
    if (u >= end) return;
    uint32_t jb = find_job(jobs, n_jobs, u);
    for (; u < end; u += step) {
        jb = next_job(jobs, n_jobs, jb, u);
        const ExtremaJob J = jobs[jb];
        const uint32_t t = u - J.unit0;
        extrema_tile<PL, kExRows>(global_ptr(J.gauss), J.h, J.w, thr, mult, eff, log2k, octave, global_ptr(J.out),
                                  global_ptr(J.count), cap, global_ptr(J.err), top, 1 + (int)(t % J.units_x) * 64,
                                  1 + (int)(t / J.units_x) * kExRows, opaque_tid());
        __syncthreads();
    }
}

uint32_t extrema_group_units(int h, int w, int rows) {
    if (h < 3 || w < 3 || rows <= 0) return 0;
    return (uint32_t)((w - 2 + 63) / 64) * (uint32_t)((h - 2 + rows - 1) / rows);
}

hipError_t launch_extrema_group(const ExtremaJob *jobs, uint32_t n_jobs, uint32_t n_units, int S, int rows, float thr,
                                float mult, const float *eff_scale, double log2k, int octave, uint32_t cap,
                                const Taps &top_taps, int top_r, hipStream_t st) {
    if (n_jobs == 0 || n_units == 0) return hipSuccess;
    if (S < 3 || S > 10 || top_r != top_radius(S) || (rows != 16 && rows != 32)) return hipErrorInvalidValue;
    EffScale e;
    for (int i = 0; i < kMaxLevels; i++) e.v[i] = eff_scale[i];
#define PS_EXG(SS, RR)                                                                                          \
    {                                                                                                           \
        const uint32_t g = group_grid((const void *)extrema_group_kernel<SS + 2, RR>, 256, n_units);           \
        if (!g) return hipErrorInvalidValue;                                                                    \
        extrema_group_kernel<SS + 2, RR><<<g, 256, 0, st>>>(jobs, n_jobs, n_units, thr, mult, e, log2k, octave, \
                                                            cap, top_taps);                                     \
        return hipGetLastError();                                                                               \
    }
#define PS_EXG_CASE(SS)          \
    case SS:                     \
        if (rows == 32) PS_EXG(SS, 32) \
        else PS_EXG(SS, 16)
    switch (S) {
        PS_EXG_CASE(3) PS_EXG_CASE(4) PS_EXG_CASE(5) PS_EXG_CASE(6)
        PS_EXG_CASE(7) PS_EXG_CASE(8) PS_EXG_CASE(9) PS_EXG_CASE(10)
    default: return hipErrorInvalidValue;
    }
#undef PS_EXG_CASE
#undef PS_EXG
}

// ordering of many images: one select workgroup per image, then fixed grids per image
constexpr uint32_t kGatherGroup = 64, kScatterGroup = 16, kCountGroup = 8;

__global__ __launch_bounds__(1024) void rank_select_group_kernel(const RankJob *__restrict__ jobs, uint32_t cap_in,
                                                                 uint32_t limit) {
    const RankJob J = jobs[blockIdx.x];
    rank_select_body(global_ptr(J.cand), global_ptr(J.count), cap_in, limit,
                     rank_words(global_ptr(J.scratch), cap_in));
}

__global__ __launch_bounds__(256) void rank_gather_group_kernel(const RankJob *__restrict__ jobs, uint32_t cap_in, uint32_t limit) {
    const RankJob J = jobs[blockIdx.y];
    const uint32_t n = min(*global_ptr(J.count), cap_in);
    uint32_t *rank = static_cast<uint32_t *>(global_ptr(J.scratch)), *subset = rank + cap_in, *words = subset + cap_in;
    RankRec *rec = reinterpret_cast<RankRec *>(words + 64);
    for (uint32_t i0 = blockIdx.x * 256; i0 < n; i0 += gridDim.x * 256) {
        rank_gather_block(global_ptr(J.cand), n, limit, subset, rec, rank, words, i0);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kRankTile) void rank_count_group_kernel(const RankJob *__restrict__ jobs, uint32_t cap_in) {
    const RankJob J = jobs[blockIdx.z];
    uint32_t *rank = static_cast<uint32_t *>(global_ptr(J.scratch)), *words = rank + 2 * cap_in;
    rank_count_body(reinterpret_cast<const RankRec *>(words + 64), words, rank, blockIdx.x, gridDim.x, blockIdx.y,
                    gridDim.y);
}

__global__ void rank_scatter_group_kernel(const RankJob *__restrict__ jobs, uint32_t cap_in, uint32_t limit) {
    const RankJob J = jobs[blockIdx.y];
    const uint32_t *rank = static_cast<const uint32_t *>(global_ptr(J.scratch)), *subset = rank + cap_in,
                   *words = subset + cap_in;
    const uint32_t m = words[0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < max(m, 1u); i += gridDim.x * blockDim.x)
        rank_scatter_one(global_ptr(J.cand), subset, m, limit, rank, global_ptr(J.out), global_ptr(J.out_count), i);
}

hipError_t launch_rank_group(const RankJob *jobs, uint32_t n_jobs, uint32_t cap_in, uint32_t limit, hipStream_t st) {
    if (n_jobs == 0) return hipSuccess;
    rank_select_group_kernel<<<n_jobs, 1024, 0, st>>>(jobs, cap_in, limit);
    rank_gather_group_kernel<<<dim3(kGatherGroup, n_jobs), 256, 0, st>>>(jobs, cap_in, limit);
    rank_count_group_kernel<<<dim3(kCountGroup, kCountGroup, n_jobs), kRankTile, 0, st>>>(jobs, cap_in);
    rank_scatter_group_kernel<<<dim3(kScatterGroup, n_jobs), 256, 0, st>>>(jobs, cap_in, limit);
    return hipGetLastError();
}

}  // namespace ps
