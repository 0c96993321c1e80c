// Orientation assignment and 128-D descriptors, one workgroup per keypoint /
// feature (pure-sift 0.2.0 src/pipeline.rs:1206-1577).
//
// Gradients are evaluated on the fly from the octave's Gaussian level instead of
// materialising magnitude/angle planes (pipeline.rs:1168-1203): same expressions,
// and the windows never touch the zeroed border of the reference's planes. The hardware
// exponential (v_exp_f32, within 2 ulp) and square root (v_sqrt_f32, within 1 ulp) per pixel, sinf/cosf, a polynomial atan2, fused multiply-adds
// and the parallel histogram accumulation make these results tolerance-level (not bit-level)
// equal to the reference.
#include <algorithm>

#include "ps_internal.h"

#pragma clang fp contract(fast)

namespace ps {

constexpr float kTwoPi = 2.0f * 3.14159265358979323846f;

// f32::rem_euclid. fmodf(x, m) returns x exactly when |x| < |m|, which covers every call
// below (angles in (-2pi, 2pi]); the library fmodf only runs for the remaining inputs.
__device__ __forceinline__ float rem_euclid(float x, float m) {
    float r = fabsf(x) < fabsf(m) ? x : fmodf(x, m);
    return r < 0.0f ? r + fabsf(m) : r;
}

// atan2 for the gradient angle: octant reduction, a = min/max via the hardware reciprocal,
// atan(a) = a * P(a^2) with P a degree-8 polynomial fitted here (max error 2.9e-7 rad over all
// inputs, about 3 ulp -- the same order as glibc's atan2f, which the device cannot reproduce
// bit for bit anyway). Callers guarantee (x, y) != (0, 0).
__device__ __forceinline__ float grad_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mn * __builtin_amdgcn_rcpf(mx);
    const float s = a * a;
    float p = 0.00284997f;
    p = p * s + -0.01606893f;
    p = p * s + 0.042692f;
    p = p * s + -0.07504334f;
    p = p * s + 0.10640952f;
    p = p * s + -0.14203648f;
    p = p * s + 0.1999262f;
    p = p * s + -0.33333072f;
    p = p * s + 1.0f;
    p = p * a;
    if (ay > ax) p = 1.57079632679489662f - p;
    if (x < 0.0f) p = 3.14159265358979324f - p;
    return y < 0.0f ? -p : p;
}

// (p / w, p % w) for 0 <= p < 2^20 and 1 <= w < 2^10 via a float reciprocal: p/w has
// fractional part in [0, 1 - 1/w], so floor((p + 0.5) * (1/w)) is exact.
__device__ __forceinline__ void divmod_small(int p, int w, float inv_w, int &q, int &r) {
    q = (int)(((float)p + 0.5f) * inv_w);
    r = p - q * w;
}

struct LocalKp {
    const float *g;   // Gaussian level of the keypoint (octave-local pixels)
    int W, H;
    float x, y, scale;
};

// resolve_octave_and_layer + octave-local keypoint (pipeline.rs:1206-1227, :1519-1528)
__device__ __forceinline__ LocalKp local_kp(const PyrView &pv, const Cand &kp) {
    LocalKp l;
    int oct = min((int)kp.octave, pv.n_oct - 1);
    int layer = min((int)kp.layer, pv.S + 2);
    l.W = pv.w[oct];
    l.H = pv.h[oct];
    l.g = global_ptr(pv.base) + pv.off[oct] + (size_t)layer * l.W * l.H;
    float inv = 1.0f / pv.scale_factor[oct];
    l.x = kp.x * inv;
    l.y = kp.y * inv;
    l.scale = fmaxf(kp.scale * inv, 0.5f * 1.6f);
    return l;
}

// magnitude / angle of pipeline.rs:1186-1189 at an interior pixel
__device__ __forceinline__ void grad(const LocalKp &l, int x, int y, float &m, float &a) {
    const float *row = l.g + (size_t)y * l.W;
    float dx = row[x + 1] - row[x - 1];
    float dy = row[x + l.W] - row[x - l.W];
    m = sqrtf(dx * dx + dy * dy);
    a = rem_euclid(grad_atan2(dy, dx), kTwoPi);
}

// pipeline.rs:1233-1347. One workgroup of kOriThreads per keypoint. Histogram updates go to one
// of kCopies private LDS copies (wave, lane % 8), so lanes that share a copy are 8 pixels apart and
// rarely hit the same bin; the copies are summed in a fixed order (deterministic results).
// Each thread gathers the gradient samples of 4 pixels before any histogram update, so
// their loads overlap.
constexpr int kOriThreads = 128;
constexpr int kCopies = kOriThreads / 64 * 8;   // waves x 8 lane groups
constexpr int kBatch = 4;
__device__ __forceinline__ void orient_body(const PyrView &pv, const Cand *kps, uint32_t k, float *ori,
                                            uint32_t *nori, int tid) {
    __shared__ float s_whist[kCopies][37];   // +1 pad: copies of a bin sit in different LDS banks
    __shared__ float s_hist[36];
    const int lane = tid;
    const Cand kp = kps[k];
    const LocalKp l = local_kp(pv, kp);

    const float sigma_ori = fmaxf(1.5f * l.scale, 1.0f);
    const float two_sigma_sq = 2.0f * sigma_ori * sigma_ori;
    const float neg_inv_2s2 = -1.0f / two_sigma_sq;
    const int radius = max((int)ceilf(3.0f * sigma_ori), 2);
    const int cx = (int)roundf(l.x), cy = (int)roundf(l.y);
    const float bin_scale = 36.0f / kTwoPi;
    const int y_min = max(cy - radius, 1), y_max = max(min(cy + radius, l.H - 2), 1);
    const int x_min = max(cx - radius, 1), x_max = max(min(cx + radius, l.W - 2), 1);
    const float r_sq_max = (float)(radius * radius);
    for (int i = lane; i < kCopies * 37; i += kOriThreads) (&s_whist[0][0])[i] = 0.0f;
    __syncthreads();
    float *wh = s_whist[(lane >> 6) * 8 + (lane & 7)];
    const int ww = x_max - x_min + 1, hh = y_max - y_min + 1;
    const int total = (ww > 0 && hh > 0) ? ww * hh : 0;
    const float inv_ww = 1.0f / (float)max(ww, 1);
    for (int p0 = lane; p0 < total; p0 += kBatch * kOriThreads) {
        float gxa[kBatch], gxb[kBatch], gya[kBatch], gyb[kBatch], rsq[kBatch];
        bool ok[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            int p = p0 + u * kOriThreads;
            int qy, qx;
            divmod_small(p, ww, inv_ww, qy, qx);
            int yy = y_min + qy, xx = x_min + qx;
            float dx = (float)(xx - cx), dy = (float)(yy - cy);
            rsq[u] = dx * dx + dy * dy;
            ok[u] = p < total && !(rsq[u] > r_sq_max);
            const float *row = l.g + (size_t)(ok[u] ? yy : y_min) * l.W + (ok[u] ? xx : x_min);
            gxa[u] = row[1]; gxb[u] = row[-1]; gyb[u] = row[l.W]; gya[u] = row[-l.W];
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            if (!ok[u]) continue;
            float dx = gxa[u] - gxb[u], dy = gyb[u] - gya[u];    // pipeline.rs:1186-1189
            float m = __builtin_amdgcn_sqrtf(dx * dx + dy * dy);
            if (m <= 1e-7f) continue;
            float a = rem_euclid(grad_atan2(dy, dx), kTwoPi);
            float weight = __expf(rsq[u] * neg_inv_2s2);
            int bin = ((int)floorf(a * bin_scale)) % 36;
            atomicAdd(&wh[bin], m * weight);
        }
    }
    __syncthreads();
    if (lane < 36) {
        float acc = 0.0f;
        for (int c = 0; c < kCopies; c++) acc += s_whist[c][lane];
        s_hist[lane] = acc;
    }
    __syncthreads();
    if (lane >= 64) return;
    // wave 0: lane b < 36 owns smoothed bin b (pipeline.rs:1280-1287)
    const bool own = lane < 36;
    float sm = -INFINITY;
    if (own) {
        float prev = s_hist[(lane + 35) % 36], curr = s_hist[lane], next = s_hist[(lane + 1) % 36];
        sm = (prev + curr + next) * (float)(1.0 / 3.0);
    }
    // global maximum, first bin on ties (pipeline.rs:1289-1297)
    float best_val = sm;
    int best_bin = own ? lane : 64;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        float ov = __shfl_xor(best_val, o);
        int ob = __shfl_xor(best_bin, o);
        if (ov > best_val || (ov == best_val && ob < best_bin)) { best_val = ov; best_bin = ob; }
    }
    float *o = ori + (size_t)k * 4;
    if (best_val <= 1e-7f) {
        if (lane == 0) { o[0] = 0.0f; nori[k] = 1; }
        return;
    }
    // peaks >= 80% of the maximum, strict on the left, >= on the right (pipeline.rs:1303-1317)
    const float lft = __shfl(sm, (lane + 35) % 36), rgt = __shfl(sm, (lane + 1) % 36);
    const bool peak = own && !(sm < 0.80f * best_val) && sm > lft && sm >= rgt;
    const uint64_t pm = __ballot(peak);
    const int npeaks = __popcll(pm);
    // order by (value desc, bin asc), keep 4 (pipeline.rs:1323-1328)
    int rank = 0;
    for (int j = 0; j < 36; j++) {
        float vj = __shfl(sm, j);
        if (((pm >> j) & 1ull) && (vj > sm || (vj == sm && j < lane))) rank++;
    }
    if (npeaks == 0) {
        if (lane == best_bin) {
            float denom = lft - 2.0f * sm + rgt;
            float off = fabsf(denom) > 1e-6f ? 0.5f * (lft - rgt) / denom : 0.0f;
            off = off < -0.5f ? -0.5f : (off > 0.5f ? 0.5f : off);
            o[0] = rem_euclid(((float)lane + 0.5f + off) * (kTwoPi / 36.0f), kTwoPi);
            nori[k] = 1;
        }
        return;
    }
    if (peak && rank < 4) {   // parabolic refinement (pipeline.rs:1333-1345)
        float denom = lft - 2.0f * sm + rgt;
        float off = fabsf(denom) > 1e-6f ? 0.5f * (lft - rgt) / denom : 0.0f;
        off = off < -0.5f ? -0.5f : (off > 0.5f ? 0.5f : off);
        o[rank] = rem_euclid(((float)lane + 0.5f + off) * (kTwoPi / 36.0f), kTwoPi);
    }
    if (lane == 0) nori[k] = (uint32_t)min(npeaks, 4);
}

__global__ __launch_bounds__(kOriThreads) void orient_kernel(PyrView pv, const Cand *kps, const uint32_t *nkp,
                                                              uint32_t max_kp, float *ori, uint32_t *nori) {
    const uint32_t k = blockIdx.x;
    if (k >= min(*nkp, max_kp)) return;
    orient_body(pv, kps, k, ori, nori, threadIdx.x);
}

hipError_t launch_orientations(const PyrView &pv, const Cand *kps, const uint32_t *nkp, uint32_t max_kp, float *ori,
                               uint32_t *nori, hipStream_t st) {
    if (max_kp == 0) return hipSuccess;
    orient_kernel<<<max_kp, kOriThreads, 0, st>>>(pv, kps, nkp, max_kp, ori, nori);
    return hipGetLastError();
}

// Flattens (keypoint, orientation) pairs in keypoint order, each keypoint's orientations in
// peak order (flat_map_iter, pipeline.rs:1566-1571), truncated to cap (pipeline.rs:1654-1656).
constexpr int kScanThreads = 1024;
__device__ __forceinline__ void feature_scan_body(const uint32_t *nkp, const uint32_t *nori, uint32_t max_kp,
                                                  uint32_t cap, uint32_t *feat_kp, uint32_t *feat_j, uint32_t *nfeat) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    const uint32_t n = min(*nkp, max_kp);
    const uint32_t per = (max_kp + kScanThreads - 1) / kScanThreads;
    const uint32_t b = threadIdx.x * per, e = min(b + per, n);
    uint32_t local = 0;
    for (uint32_t i = b; i < e; i++) local += nori[i];
    // exclusive offset of this thread's keypoints: a scan inside the wave, then the wave totals
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t incl = local;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d);
        if (lane >= (uint32_t)d) incl += t;
    }
    if (lane == 63) s_w[wv] = incl;
    __syncthreads();
    uint32_t off = incl - local, total = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanThreads / 64; k++) {
        const uint32_t v = s_w[k];
        if (k < wv) off += v;
        total += v;
    }
    for (uint32_t i = b; i < e; i++) {
        uint32_t c = nori[i];
        for (uint32_t j = 0; j < c; j++) {
            if (off + j < cap) { feat_kp[off + j] = i; feat_j[off + j] = j; }
        }
        off += c;
    }
    if (threadIdx.x == kScanThreads - 1) *nfeat = min(total, cap);
}

__global__ __launch_bounds__(kScanThreads) void feature_scan_kernel(const uint32_t *nkp, const uint32_t *nori,
                                                                    uint32_t max_kp, uint32_t cap, uint32_t *feat_kp,
                                                                    uint32_t *feat_j, uint32_t *nfeat) {
    feature_scan_body(nkp, nori, max_kp, cap, feat_kp, feat_j, nfeat);
}

hipError_t launch_feature_scan(const uint32_t *nkp, const uint32_t *nori, uint32_t max_kp, uint32_t cap,
                               uint32_t *feat_kp, uint32_t *feat_j, uint32_t *nfeat, hipStream_t st) {
    feature_scan_kernel<<<1, kScanThreads, 0, st>>>(nkp, nori, max_kp, cap, feat_kp, feat_j, nfeat);
    return hipGetLastError();
}

// pipeline.rs:1352-1480 (+ RootSIFT, :1012-1019). One 128-thread workgroup per feature and no
// shared accumulators: a pixel that passes the window test has r_bin, c_bin in (-1, 4), so it
// belongs to exactly one of the 25 cells (floor(r_bin), floor(c_bin)) in {-1..3}^2 and only
// feeds the 4 spatial bins around that cell. Five lanes own each cell and accumulate 4 corners
// x 8 orientations into lane-private LDS regions; a fixed-order sum over the owners of the 4
// neighbouring cells then forms each of the 128 bins.
//
// The cell's pixels are found row by row, not by testing every pixel of its bounding box: along
// an image row, r_bin and c_bin (the reference's expressions, rounded operation by operation)
// are monotone in x, so the row's pixels with floor(r_bin) = r0 and floor(c_bin) = c0 inside
// the window form one interval, bounded by where four monotone tests change value. Each
// transition is located from its closed-form estimate and confirmed by evaluating the test on
// both sides (a binary search when the estimate misses, e.g. for a nearly axis-parallel
// boundary), so the intervals hold exactly the pixels the per-pixel test would keep. The five
// lanes then take every fifth pixel of the cell's row-major sequence.
constexpr int kFeatThreads = 128;   // descriptor workgroup: 5 lanes per cell, lanes 125..127 idle
constexpr int kLanesPerCell = 5;
constexpr int kPrivStride = 33;     // 32 accumulators + 1 pad
constexpr int kCellRows = 32;       // bounding-box rows of a cell kept as intervals in LDS (hist width <= ~20 px)
constexpr int kFeatBatch = 4;       // pixels whose gradient loads are issued together

// (r_rot, c_rot, r_bin, c_bin) of offset (dx, dy) from the keypoint (pipeline.rs:1380-1387), every
// operation rounded on its own (no contraction): each evaluation for one pixel gives the same
// bits, and along a row (fixed dy) every result is monotone in dx
__device__ __forceinline__ void bin_coords(float dx, float dy, float cos_t, float sin_t, float inv_hw, float &r_rot,
                                           float &c_rot, float &r_bin, float &c_bin) {
#pragma clang fp contract(off)
    r_rot = (dx * cos_t + dy * sin_t) * inv_hw;
    c_rot = (-dx * sin_t + dy * cos_t) * inv_hw;
    r_bin = r_rot + 1.5f;
    c_bin = c_rot + 1.5f;
}

// first q in [0, n] with test(q) true (test(n) taken as true), for a test that is false then
// true along q; `guess` is tried first
template <class F>
__device__ __forceinline__ int first_true(F test, int n, int guess) {
    guess = min(max(guess, 0), n);
    if ((guess == n || test(guess)) && (guess == 0 || !test(guess - 1))) return guess;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (test(mid)) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ void descriptor_body(const PyrView &pv, const Cand *kps, const float *ori,
                                                const uint32_t *feat_kp, const uint32_t *feat_j, uint32_t f, int kind,
                                                ps_feature *out_f, float *out_d, int tid) {
    __shared__ float s_priv[kFeatThreads * kPrivStride];
    __shared__ uint16_t s_rows[25 * kCellRows];   // per cell and box row: first column << 8 | length
    __shared__ float s_desc[128];
    const int lane = tid;
    const uint32_t k = feat_kp[f];
    const Cand kp = kps[k];
    const float theta = ori[(size_t)k * 4 + feat_j[f]];
    const LocalKp l = local_kp(pv, kp);
    float *priv = s_priv + lane * kPrivStride;
#pragma unroll
    for (int i = 0; i < 32; i++) priv[i] = 0.0f;
    const int cell = lane / kLanesPerCell, sub = lane - cell * kLanesPerCell;
    const int r0 = cell / 5 - 1, c0 = cell % 5 - 1;   // cell in {-1..3}^2 (lanes >= 125: none)
    const bool active = l.W >= 3 && l.H >= 3 && cell < 25;
    const float cos_t = cosf(theta), sin_t = sinf(theta);
    const float hist_width = fmaxf(3.0f * l.scale, 1.5f);
    const float inv_hw = 1.0f / hist_width;
    const float exp_scale = -1.0f / (2.0f * 2.0f * 2.0f);
    const int max_radius = max((int)ceilf(hist_width * 1.41421356237309504880f * 5.0f * 0.5f), 4);
    const int cx = (int)roundf(l.x), cy = (int)roundf(l.y);
    const int y_min = max(cy - max_radius, 1), y_max = max(min(cy + max_radius, l.H - 2), 1);
    const int x_min = max(cx - max_radius, 1), x_max = max(min(cx + max_radius, l.W - 2), 1);
    const float bin_per_rad = 8.0f / kTwoPi;
    // bounding box of the cell: r_rot in [r0 - 1.5, r0 - 0.5), c_rot likewise; image offset
    // (dx, dy) = hw (r cos - c sin, r sin + c cos); padded by 1 px against rounding
    float bx0 = 1e30f, bx1 = -1e30f, by0 = 1e30f, by1 = -1e30f;
#pragma unroll
    for (int cr = 0; cr < 2; cr++)
#pragma unroll
        for (int cc = 0; cc < 2; cc++) {
            float rr = ((float)r0 - 1.5f + (float)cr) * hist_width, c2 = ((float)c0 - 1.5f + (float)cc) * hist_width;
            float ddx = rr * cos_t - c2 * sin_t, ddy = rr * sin_t + c2 * cos_t;
            bx0 = fminf(bx0, ddx); bx1 = fmaxf(bx1, ddx);
            by0 = fminf(by0, ddy); by1 = fmaxf(by1, ddy);
        }
    const int xa = max(x_min, (int)floorf(l.x + bx0) - 1), xb = min(x_max, (int)ceilf(l.x + bx1) + 1);
    const int ya = max(y_min, (int)floorf(l.y + by0) - 1), yb = min(y_max, (int)ceilf(l.y + by1) + 1);
    const int ww = xb - xa + 1, hh = yb - ya + 1;
    const bool any = active && ww > 0 && hh > 0;
    const bool by_rows = any && ww < 256 && hh <= kCellRows;
    // the window and cell test of one pixel (pipeline.rs:1389-1396 with floor(r_bin) = r0,
    // floor(c_bin) = c0), as four tests that are monotone along a row
    auto r_lo = [&](float r_bin) { return r0 >= 0 ? r_bin >= (float)r0 : r_bin > -1.0f; };
    auto c_lo = [&](float c_bin) { return c0 >= 0 ? c_bin >= (float)c0 : c_bin > -1.0f; };
    auto bins_at = [&](int qy, int qx, float &r_rot, float &c_rot, float &r_bin, float &c_bin) {
        bin_coords((float)(xa + qx) - l.x, (float)(ya + qy) - l.y, cos_t, sin_t, inv_hw, r_rot, c_rot, r_bin, c_bin);
    };
    if (by_rows) {
        for (int qy = sub; qy < hh; qy += kLanesPerCell) {
            const float dy = (float)(ya + qy) - l.y;
            auto rb = [&](int qx) {
                float a, b, r, c;
                bin_coords((float)(xa + qx) - l.x, dy, cos_t, sin_t, inv_hw, a, b, r, c);
                return r;
            };
            auto cbv = [&](int qx) {
                float a, b, r, c;
                bin_coords((float)(xa + qx) - l.x, dy, cos_t, sin_t, inv_hw, a, b, r, c);
                return c;
            };
            // column q where the bin coordinate crosses value t (closed form), rounded up
            auto cross_r = [&](float t) {
                const float q = ((t - 1.5f) * hist_width - dy * sin_t) / cos_t + l.x - (float)xa;
                return (int)fminf(fmaxf(ceilf(q), -1.0f), (float)(ww + 1));
            };
            auto cross_c = [&](float t) {
                const float q = (dy * cos_t - (t - 1.5f) * hist_width) / sin_t + l.x - (float)xa;
                return (int)fminf(fmaxf(ceilf(q), -1.0f), (float)(ww + 1));
            };
            int lo = 0, hi = ww;   // the row's pixels: [lo, hi)
            const float tr = r0 >= 0 ? (float)r0 : -1.0f, tc = c0 >= 0 ? (float)c0 : -1.0f;
            if (cos_t > 0.0f) {        // r_bin non-decreasing along the row
                lo = max(lo, first_true([&](int q) { return r_lo(rb(q)); }, ww, cross_r(tr)));
                hi = min(hi, first_true([&](int q) { return rb(q) >= (float)(r0 + 1); }, ww, cross_r((float)(r0 + 1))));
            } else if (cos_t < 0.0f) { // non-increasing
                hi = min(hi, first_true([&](int q) { return !r_lo(rb(q)); }, ww, cross_r(tr)));
                lo = max(lo, first_true([&](int q) { return rb(q) < (float)(r0 + 1); }, ww, cross_r((float)(r0 + 1))));
            } else {
                const float r = rb(0);
                if (!(r_lo(r) && r < (float)(r0 + 1))) hi = 0;
            }
            if (sin_t < 0.0f) {        // c_bin non-decreasing along the row
                lo = max(lo, first_true([&](int q) { return c_lo(cbv(q)); }, ww, cross_c(tc)));
                hi = min(hi, first_true([&](int q) { return cbv(q) >= (float)(c0 + 1); }, ww, cross_c((float)(c0 + 1))));
            } else if (sin_t > 0.0f) { // non-increasing
                hi = min(hi, first_true([&](int q) { return !c_lo(cbv(q)); }, ww, cross_c(tc)));
                lo = max(lo, first_true([&](int q) { return cbv(q) < (float)(c0 + 1); }, ww, cross_c((float)(c0 + 1))));
            } else {
                const float c = cbv(0);
                if (!(c_lo(c) && c < (float)(c0 + 1))) hi = 0;
            }
            s_rows[cell * kCellRows + qy] = (uint16_t)(hi > lo ? (lo << 8) | (hi - lo) : 0);
        }
    }
    __syncthreads();
    if (any) {
        // gradient sample of box pixel (qy, qx): weight, orientation and trilinear accumulation
        // into the lane's 4 corners x 8 orientations (pipeline.rs:1186-1189, :1394-1440)
        auto add_pixel = [&](float gxa, float gxb, float gya, float gyb, int qy, int qx) {
            const float gdx = gxa - gxb, gdy = gyb - gya;
            const float m = __builtin_amdgcn_sqrtf(gdx * gdx + gdy * gdy);
            if (m <= 1e-7f) return;
            const float a = rem_euclid(grad_atan2(gdy, gdx), kTwoPi);
            float r_rot, c_rot, r_bin, c_bin;
            bins_at(qy, qx, r_rot, c_rot, r_bin, c_bin);
            const float wm = m * __expf((r_rot * r_rot + c_rot * c_rot) * exp_scale);
            const float o_bin = rem_euclid(a - theta, kTwoPi) * bin_per_rad;
            const int o0 = (int)floorf(o_bin);
            const float dr = r_bin - (float)r0, dc = c_bin - (float)c0, dof = o_bin - (float)o0;
            const int ob0 = o0 & 7, ob1 = (o0 + 1) & 7;
#pragma unroll
            for (int ir = 0; ir <= 1; ir++) {
                const float wr = ir == 0 ? 1.0f - dr : dr;
#pragma unroll
                for (int ic = 0; ic <= 1; ic++) {
                    const float wc = ic == 0 ? 1.0f - dc : dc;
                    const float w0 = wm * wr * wc;
                    float *acc = priv + (ir * 2 + ic) * 8;
                    acc[ob0] += w0 * (1.0f - dof);
                    acc[ob1] += w0 * dof;
                }
            }
        };
        if (by_rows) {
            // pixels sub, sub + 5, ... of the cell's rows in order, kFeatBatch at a time with
            // their gradient loads issued together
            const uint16_t *rows = s_rows + cell * kCellRows;
            int r = 0, base = 0, len = rows[0] & 255, kq = sub;
            bool more = true;
            while (more) {
                int pq[kFeatBatch];
                bool ok[kFeatBatch];
#pragma unroll
                for (int u = 0; u < kFeatBatch; u++) {
                    while (r < hh && kq >= base + len) {
                        base += len;
                        r++;
                        len = r < hh ? rows[r] & 255 : 0;
                    }
                    ok[u] = r < hh;
                    pq[u] = ok[u] ? (r << 8) | ((rows[r] >> 8) + kq - base) : (u > 0 ? pq[u - 1] : -1);
                    kq += kLanesPerCell;
                }
                more = ok[kFeatBatch - 1];
                if (!ok[0]) break;
                float gxa[kFeatBatch], gxb[kFeatBatch], gya[kFeatBatch], gyb[kFeatBatch];
#pragma unroll
                for (int u = 0; u < kFeatBatch; u++) {
                    const float *row = l.g + (size_t)(ya + (pq[u] >> 8)) * l.W + (xa + (pq[u] & 255));
                    gxa[u] = row[1]; gxb[u] = row[-1]; gyb[u] = row[l.W]; gya[u] = row[-l.W];
                }
#pragma unroll
                for (int u = 0; u < kFeatBatch; u++)
                    if (ok[u]) add_pixel(gxa[u], gxb[u], gya[u], gyb[u], pq[u] >> 8, pq[u] & 255);
            }
        } else {
            // boxes too large for the row intervals (keypoint scales far beyond what detection
            // produces): every box pixel tested and accumulated one at a time
            const int total = ww * hh;
            for (int p = sub; p < total; p += kLanesPerCell) {
                const int qy = p / ww, qx = p - qy * ww;
                float r_rot, c_rot, r_bin, c_bin;
                bins_at(qy, qx, r_rot, c_rot, r_bin, c_bin);
                if (!(r_lo(r_bin) && r_bin < (float)(r0 + 1) && c_lo(c_bin) && c_bin < (float)(c0 + 1))) continue;
                const float *row = l.g + (size_t)(ya + qy) * l.W + (xa + qx);
                add_pixel(row[1], row[-1], row[-l.W], row[l.W], qy, qx);
            }
        }
    }
    __syncthreads();
    // bin (rb, cb, ob): corner (rb - r0, cb - c0) of the cells r0 in {rb-1, rb}, c0 in {cb-1, cb}
    if (lane < 128) {
        const int rb = lane >> 5, cb = (lane >> 3) & 3, ob = lane & 7;
        float acc = 0.0f;
#pragma unroll
        for (int ir = 0; ir <= 1; ir++)
#pragma unroll
            for (int ic = 0; ic <= 1; ic++) {
                const int cr = rb - ir, ccl = cb - ic;   // owning cell, in {-1..3}
                const int owner = (cr + 1) * 5 + (ccl + 1);
                const int corner = ir * 2 + ic;
                for (int sl = 0; sl < kLanesPerCell; sl++)
                    acc += s_priv[(owner * kLanesPerCell + sl) * kPrivStride + corner * 8 + ob];
            }
        s_desc[lane] = acc;
    }
    __syncthreads();
    // Normalisation (pipeline.rs:1461-1477) by wave 0: lane l holds elements l and l + 64; the
    // sums use a fixed butterfly order (deterministic; within tolerance of the reference's
    // sequential sums, like the rest of the descriptor).
    if (lane < 64) {
        float v0 = s_desc[lane], v1 = s_desc[lane + 64];
        float n1 = v0 * v0 + v1 * v1;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) n1 += __shfl_xor(n1, o);
        n1 = sqrtf(n1);
        if (n1 > 1e-7f) {
            const float inv1 = 1.0f / n1;
            v0 = fminf(v0 * inv1, 0.2f);
            v1 = fminf(v1 * inv1, 0.2f);
            float n2 = v0 * v0 + v1 * v1;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) n2 += __shfl_xor(n2, o);
            n2 = sqrtf(n2);
            if (n2 > 1e-7f) {
                const float inv2 = 1.0f / n2;
                v0 *= inv2;
                v1 *= inv2;
            }
        }
        if (kind == PS_DESCRIPTOR_ROOTSIFT) {   // pipeline.rs:1012-1019
            float l1 = fmaxf(v0, 0.0f) + fmaxf(v1, 0.0f);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) l1 += __shfl_xor(l1, o);
            if (l1 > 1e-7f) {
                const float inv = 1.0f / l1;
                v0 = sqrtf(fmaxf(v0, 0.0f) * inv);
                v1 = sqrtf(fmaxf(v1, 0.0f) * inv);
            }
        }
        float *od = out_d + (size_t)f * 128;
        od[lane] = v0;
        od[lane + 64] = v1;
    }
    if (lane == 0) {
        ps_feature o;
        o.x = kp.x; o.y = kp.y; o.scale = kp.scale; o.response = kp.response; o.orientation = theta;
        out_f[f] = o;
    }
}

__global__ __launch_bounds__(kFeatThreads) void descriptor_kernel(PyrView pv, const Cand *kps, const float *ori,
                                                                  const uint32_t *feat_kp, const uint32_t *feat_j,
                                                                  const uint32_t *nfeat, uint32_t cap, int kind,
                                                                  ps_feature *out_f, float *out_d) {
    const uint32_t f = blockIdx.x;
    if (f >= min(*nfeat, cap)) return;
    descriptor_body(pv, kps, ori, feat_kp, feat_j, f, kind, out_f, out_d, threadIdx.x);
}

hipError_t launch_descriptors(const PyrView &pv, const Cand *kps, const float *ori, const uint32_t *feat_kp,
                              const uint32_t *feat_j, const uint32_t *nfeat, uint32_t cap, int kind, ps_feature *out_f,
                              float *out_d, hipStream_t st) {
    if (cap == 0) return hipSuccess;
    descriptor_kernel<<<cap, kFeatThreads, 0, st>>>(pv, kps, ori, feat_kp, feat_j, nfeat, cap, kind, out_f, out_d);
    return hipGetLastError();
}

__global__ void cand_to_kp_kernel(const Cand *c, const uint32_t *n, uint32_t cap, ps_keypoint *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= min(*n, cap)) return;
    Cand v = c[i];
    ps_keypoint k;
    k.x = v.x; k.y = v.y; k.scale = v.scale; k.response = v.response; k.octave = v.octave; k.layer = v.layer;
    out[i] = k;
}

hipError_t launch_cand_to_keypoints(const Cand *c, const uint32_t *n, uint32_t cap, ps_keypoint *out, hipStream_t st) {
    if (cap == 0) return hipSuccess;
    cand_to_kp_kernel<<<(cap + 255) / 256, 256, 0, st>>>(c, n, cap, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ grouped persistent launches
// The orientation and descriptor stages of many images in one launch each: the keypoints (or
// features) of all images are numbered image by image through device-side offsets, and a grid of
// resident workgroups walks them -- one keypoint or feature at a time per workgroup, as in the
// single-image kernels, whose bodies these launches run.

// off[i] = sum over j < i of ceil(min(counts[j * stride], cap) / div); off[n] = the total. One
// workgroup.
__global__ __launch_bounds__(kScanThreads) void group_offsets_kernel(const uint32_t *counts, uint32_t stride,
                                                                     uint32_t cap, uint32_t div, uint32_t n,
                                                                     uint32_t *off) {
    __shared__ uint32_t s[kScanThreads];
    const uint32_t per = (n + kScanThreads - 1) / kScanThreads;
    const uint32_t b = threadIdx.x * per, e = min(b + per, n);
    auto item = [&](uint32_t i) { return (min(counts[(size_t)i * stride], cap) + div - 1) / div; };
    uint32_t local = 0;
    for (uint32_t i = b; i < e; i++) local += item(i);
    s[threadIdx.x] = local;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0u;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t o = s[threadIdx.x] - local;
    for (uint32_t i = b; i < e; i++) {
        off[i] = o;
        o += item(i);
    }
    if (threadIdx.x == kScanThreads - 1) off[n] = s[kScanThreads - 1];
}

hipError_t launch_group_offsets(const uint32_t *counts, uint32_t stride, uint32_t cap, uint32_t div, uint32_t n,
                                uint32_t *off, hipStream_t st) {
    if (n == 0 || div == 0) return n ? hipErrorInvalidValue : hipSuccess;
    group_offsets_kernel<<<1, kScanThreads, 0, st>>>(counts, stride, cap, div, n, off);
    return hipGetLastError();
}

__global__ __launch_bounds__(kOriThreads) void orient_group_kernel(const DescJob *__restrict__ jobs, uint32_t n_jobs,
                                                                    const uint32_t *__restrict__ kp_off) {
    const uint32_t total = kp_off[n_jobs];
    uint32_t jb = blockIdx.x < total ? find_off(kp_off, n_jobs, blockIdx.x) : 0;
#pragma clang loop unroll(disable)
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        while (jb + 1 < n_jobs && kp_off[jb + 1] <= g) jb++;
        const DescJob &J = jobs[jb];
        orient_body(J.pv, global_ptr(J.kps), g - kp_off[jb], global_ptr(J.ori), global_ptr(J.nori), opaque_tid());
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanThreads) void feature_scan_group_kernel(const DescJob *__restrict__ jobs, uint32_t max_kp,
                                                                          uint32_t cap) {
    const DescJob &J = jobs[blockIdx.x];
    feature_scan_body(global_ptr(J.nkp), global_ptr(J.nori), max_kp, cap, global_ptr(J.feat_kp), global_ptr(J.feat_j),
                      global_ptr(J.nfeat));
}

__global__ __launch_bounds__(kFeatThreads) void descriptor_group_kernel(const DescJob *__restrict__ jobs, uint32_t n_jobs,
                                                                        const uint32_t *__restrict__ feat_off, int kind) {
    const uint32_t total = feat_off[n_jobs];
    uint32_t jb = blockIdx.x < total ? find_off(feat_off, n_jobs, blockIdx.x) : 0;
#pragma clang loop unroll(disable)
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        while (jb + 1 < n_jobs && feat_off[jb + 1] <= g) jb++;
        const DescJob &J = jobs[jb];
        descriptor_body(J.pv, global_ptr(J.kps), global_ptr(J.ori), global_ptr(J.feat_kp), global_ptr(J.feat_j),
                        g - feat_off[jb], kind, global_ptr(J.out_f), global_ptr(J.out_d), opaque_tid());
        __syncthreads();
    }
}

static uint32_t resident_grid(const void *fn, int threads) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, 0) != hipSuccess)
        return 0;
    return (uint32_t)std::max(per, 1) * (uint32_t)std::max(cus, 1);
}

hipError_t launch_describe_group(const DescJob *jobs, uint32_t n_jobs, uint32_t max_kp, uint32_t cap, int kind,
                                 uint32_t *kp_off, uint32_t *feat_off, const uint32_t *kp_counts,
                                 uint32_t kp_stride, const uint32_t *feat_counts, hipStream_t st) {
    if (n_jobs == 0) return hipSuccess;
    static const uint32_t g_ori = resident_grid((const void *)orient_group_kernel, kOriThreads);
    static const uint32_t g_desc = resident_grid((const void *)descriptor_group_kernel, kFeatThreads);
    if (!g_ori || !g_desc) return hipErrorInvalidValue;
    hipError_t e;
    if ((e = launch_group_offsets(kp_counts, kp_stride, max_kp, 1, n_jobs, kp_off, st)) != hipSuccess) return e;
    orient_group_kernel<<<g_ori, kOriThreads, 0, st>>>(jobs, n_jobs, kp_off);
    feature_scan_group_kernel<<<n_jobs, kScanThreads, 0, st>>>(jobs, max_kp, cap);
    if ((e = launch_group_offsets(feat_counts, 1, cap, 1, n_jobs, feat_off, st)) != hipSuccess) return e;
    descriptor_group_kernel<<<g_desc, kFeatThreads, 0, st>>>(jobs, n_jobs, feat_off, kind);
    return hipGetLastError();
}

}  // namespace ps
