// Orientation assignment and 128-D descriptors, one wavefront per keypoint /
// feature (pure-sift 0.2.0 src/pipeline.rs:1206-1577).
//
// Gradients are evaluated on the fly from the octave's Gaussian level instead of
// materialising magnitude/angle planes (pipeline.rs:1168-1203): same expressions,
// and the windows never touch the zeroed border of the reference's planes. Device
// expf/atan2f/sinf/cosf and the parallel histogram accumulation make these results
// tolerance-level (not bit-level) equal to the reference.
#include "ps_internal.h"

#pragma clang fp contract(off)

namespace ps {

constexpr float kTwoPi = 2.0f * 3.14159265358979323846f;

__device__ __forceinline__ float rem_euclid(float x, float m) {
    float r = fmodf(x, m);
    return r < 0.0f ? r + fabsf(m) : r;
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
    l.g = pv.base + pv.off[oct] + (size_t)layer * l.W * l.H;
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
    a = rem_euclid(atan2f(dy, dx), kTwoPi);
}

// pipeline.rs:1233-1347. One 64-lane workgroup per keypoint.
__global__ __launch_bounds__(64) void orient_kernel(PyrView pv, const Cand *kps, const uint32_t *nkp, uint32_t max_kp,
                                                    float *ori, uint32_t *nori) {
    __shared__ float s_hist[36];
    __shared__ float s_sm[36];
    const uint32_t k = blockIdx.x;
    const int lane = threadIdx.x;
    if (k >= min(*nkp, max_kp)) return;
    const Cand kp = kps[k];
    const LocalKp l = local_kp(pv, kp);

    const float sigma_ori = fmaxf(1.5f * l.scale, 1.0f);
    const float two_sigma_sq = 2.0f * sigma_ori * sigma_ori;
    const int radius = max((int)ceilf(3.0f * sigma_ori), 2);
    const int cx = (int)roundf(l.x), cy = (int)roundf(l.y);
    const float bin_scale = 36.0f / kTwoPi;
    const int y_min = max(cy - radius, 1), y_max = max(min(cy + radius, l.H - 2), 1);
    const int x_min = max(cx - radius, 1), x_max = max(min(cx + radius, l.W - 2), 1);
    const float r_sq_max = (float)(radius * radius);
    if (lane < 36) s_hist[lane] = 0.0f;
    __syncthreads();
    const int ww = x_max - x_min + 1, hh = y_max - y_min + 1;
    const int total = (ww > 0 && hh > 0) ? ww * hh : 0;
    for (int p = lane; p < total; p += 64) {
        int yy = y_min + p / ww, xx = x_min + p % ww;
        float dx = (float)(xx - cx), dy = (float)(yy - cy);
        float r_sq = dx * dx + dy * dy;
        if (r_sq > r_sq_max) continue;
        float m, a;
        grad(l, xx, yy, m, a);
        if (m <= 1e-7f) continue;
        float weight = expf(-r_sq / two_sigma_sq);
        int bin = ((int)floorf(a * bin_scale)) % 36;
        atomicAdd(&s_hist[bin], m * weight);
    }
    __syncthreads();
    if (lane < 36) {
        float prev = s_hist[(lane + 35) % 36], curr = s_hist[lane], next = s_hist[(lane + 1) % 36];
        s_sm[lane] = (prev + curr + next) * (float)(1.0 / 3.0);
    }
    __syncthreads();
    if (lane != 0) return;
    int best_bin = 0;
    float best_val = s_sm[0];
    for (int b = 1; b < 36; b++)
        if (s_sm[b] > best_val) { best_val = s_sm[b]; best_bin = b; }
    float *o = ori + (size_t)k * 4;
    if (best_val <= 1e-7f) { o[0] = 0.0f; nori[k] = 1; return; }
    const float split = 0.80f * best_val;
    int pb[4];
    float pv4[4];
    int np = 0;
    // collect peaks in bin order and keep the 4 best by (value desc, bin asc) (pipeline.rs:1306-1329)
    for (int b = 0; b < 36; b++) {
        float v = s_sm[b];
        if (v < split) continue;
        float lft = s_sm[(b + 35) % 36], rgt = s_sm[(b + 1) % 36];
        if (!(v > lft && v >= rgt)) continue;
        int pos = np < 4 ? np : 4;
        while (pos > 0 && pv4[pos - 1] < v) pos--;
        if (pos < 4) {
            for (int q = min(np, 3); q > pos; q--) { pb[q] = pb[q - 1]; pv4[q] = pv4[q - 1]; }
            pb[pos] = b;
            pv4[pos] = v;
            if (np < 4) np++;
        }
    }
    if (np == 0) { pb[0] = best_bin; pv4[0] = best_val; np = 1; }
    for (int i = 0; i < np; i++) {
        int bin = pb[i];
        float val = pv4[i];
        float lft = s_sm[(bin + 35) % 36], rgt = s_sm[(bin + 1) % 36];
        float denom = lft - 2.0f * val + rgt;
        float off = fabsf(denom) > 1e-6f ? 0.5f * (lft - rgt) / denom : 0.0f;
        off = off < -0.5f ? -0.5f : (off > 0.5f ? 0.5f : off);
        o[i] = rem_euclid(((float)bin + 0.5f + off) * (kTwoPi / 36.0f), kTwoPi);
    }
    nori[k] = (uint32_t)np;
}

hipError_t launch_orientations(const PyrView &pv, const Cand *kps, const uint32_t *nkp, uint32_t max_kp, float *ori,
                               uint32_t *nori, hipStream_t st) {
    if (max_kp == 0) return hipSuccess;
    orient_kernel<<<max_kp, 64, 0, st>>>(pv, kps, nkp, max_kp, ori, nori);
    return hipGetLastError();
}

// Flattens (keypoint, orientation) pairs in keypoint order, each keypoint's orientations in
// peak order (flat_map_iter, pipeline.rs:1566-1571), truncated to cap (pipeline.rs:1654-1656).
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void feature_scan_kernel(const uint32_t *nkp, const uint32_t *nori,
                                                                    uint32_t max_kp, uint32_t cap, uint32_t *feat_kp,
                                                                    uint32_t *feat_j, uint32_t *nfeat) {
    __shared__ uint32_t s[kScanThreads];
    const uint32_t n = min(*nkp, max_kp);
    const uint32_t per = (max_kp + kScanThreads - 1) / kScanThreads;
    const uint32_t b = threadIdx.x * per, e = min(b + per, n);
    uint32_t local = 0;
    for (uint32_t i = b; i < e; i++) local += nori[i];
    s[threadIdx.x] = local;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {   // Hillis-Steele inclusive scan
        uint32_t v = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0u;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t off = s[threadIdx.x] - local;
    for (uint32_t i = b; i < e; i++) {
        uint32_t c = nori[i];
        for (uint32_t j = 0; j < c; j++) {
            if (off + j < cap) { feat_kp[off + j] = i; feat_j[off + j] = j; }
        }
        off += c;
    }
    if (threadIdx.x == kScanThreads - 1) *nfeat = min(s[kScanThreads - 1], cap);
}

hipError_t launch_feature_scan(const uint32_t *nkp, const uint32_t *nori, uint32_t max_kp, uint32_t cap,
                               uint32_t *feat_kp, uint32_t *feat_j, uint32_t *nfeat, hipStream_t st) {
    feature_scan_kernel<<<1, kScanThreads, 0, st>>>(nkp, nori, max_kp, cap, feat_kp, feat_j, nfeat);
    return hipGetLastError();
}

// pipeline.rs:1352-1480 (+ RootSIFT, :1012-1019). One 64-lane workgroup per feature.
__global__ __launch_bounds__(64) void descriptor_kernel(PyrView pv, const Cand *kps, const float *ori,
                                                        const uint32_t *feat_kp, const uint32_t *feat_j,
                                                        const uint32_t *nfeat, uint32_t cap, int kind,
                                                        ps_feature *out_f, float *out_d) {
    __shared__ float s_desc[128];
    __shared__ float s_scale[2];
    const uint32_t f = blockIdx.x;
    const int lane = threadIdx.x;
    if (f >= min(*nfeat, cap)) return;
    const uint32_t k = feat_kp[f];
    const Cand kp = kps[k];
    const float theta = ori[(size_t)k * 4 + feat_j[f]];
    const LocalKp l = local_kp(pv, kp);
    s_desc[lane] = 0.0f;
    s_desc[lane + 64] = 0.0f;
    __syncthreads();
    if (l.W >= 3 && l.H >= 3) {
        const float cos_t = cosf(theta), sin_t = sinf(theta);
        const float hist_width = fmaxf(3.0f * l.scale, 1.5f);
        const float inv_hw = 1.0f / hist_width;
        const float exp_scale = -1.0f / (2.0f * 2.0f * 2.0f);
        const int max_radius = max((int)ceilf(hist_width * 1.41421356237309504880f * 5.0f * 0.5f), 4);
        const int cx = (int)roundf(l.x), cy = (int)roundf(l.y);
        const int y_min = max(cy - max_radius, 1), y_max = max(min(cy + max_radius, l.H - 2), 1);
        const int x_min = max(cx - max_radius, 1), x_max = max(min(cx + max_radius, l.W - 2), 1);
        const float bin_per_rad = 8.0f / kTwoPi;
        const int ww = x_max - x_min + 1, hh = y_max - y_min + 1;
        const int total = (ww > 0 && hh > 0) ? ww * hh : 0;
        for (int p = lane; p < total; p += 64) {
            int yy = y_min + p / ww, xx = x_min + p % ww;
            float dy = (float)yy - l.y, dx = (float)xx - l.x;
            float r_rot = (dx * cos_t + dy * sin_t) * inv_hw;
            float c_rot = (-dx * sin_t + dy * cos_t) * inv_hw;
            float r_bin = r_rot + 4.0f * 0.5f - 0.5f;
            float c_bin = c_rot + 4.0f * 0.5f - 0.5f;
            if (r_bin <= -1.0f || r_bin >= 4.0f || c_bin <= -1.0f || c_bin >= 4.0f) continue;
            float m, a;
            grad(l, xx, yy, m, a);
            if (m <= 1e-7f) continue;
            float wm = m * expf((r_rot * r_rot + c_rot * c_rot) * exp_scale);
            float o_bin = rem_euclid(a - theta, kTwoPi) * bin_per_rad;
            int r0 = (int)floorf(r_bin), c0 = (int)floorf(c_bin), o0 = (int)floorf(o_bin);
            float dr = r_bin - (float)r0, dc = c_bin - (float)c0, dof = o_bin - (float)o0;
#pragma unroll
            for (int ir = 0; ir <= 1; ir++) {
                int rb = r0 + ir;
                if (rb < 0 || rb >= 4) continue;
                float wr = ir == 0 ? 1.0f - dr : dr;
#pragma unroll
                for (int ic = 0; ic <= 1; ic++) {
                    int cb = c0 + ic;
                    if (cb < 0 || cb >= 4) continue;
                    float wc = ic == 0 ? 1.0f - dc : dc;
#pragma unroll
                    for (int io = 0; io <= 1; io++) {
                        int ob = (o0 + io) & 7;
                        float wo = io == 0 ? 1.0f - dof : dof;
                        atomicAdd(&s_desc[(rb * 4 + cb) * 8 + ob], wm * wr * wc * wo);
                    }
                }
            }
        }
    }
    __syncthreads();
    // normalisation sums run sequentially in element order, as the reference's iterator sums
    if (lane == 0) {
        float n1 = 0.0f;
        for (int i = 0; i < 128; i++) n1 += s_desc[i] * s_desc[i];
        n1 = sqrtf(n1);
        s_scale[0] = n1;
        s_scale[1] = 1.0f / n1;
    }
    __syncthreads();
    const bool norm_ok = s_scale[0] > 1e-7f;
    if (norm_ok) {
        float inv1 = s_scale[1];
        s_desc[lane] = fminf(s_desc[lane] * inv1, 0.2f);
        s_desc[lane + 64] = fminf(s_desc[lane + 64] * inv1, 0.2f);
    }
    __syncthreads();
    if (norm_ok && lane == 0) {
        float n2 = 0.0f;
        for (int i = 0; i < 128; i++) n2 += s_desc[i] * s_desc[i];
        n2 = sqrtf(n2);
        s_scale[0] = n2;
        s_scale[1] = 1.0f / n2;
    }
    __syncthreads();
    if (norm_ok && s_scale[0] > 1e-7f) {
        float inv2 = s_scale[1];
        s_desc[lane] *= inv2;
        s_desc[lane + 64] *= inv2;
    }
    __syncthreads();
    if (kind == PS_DESCRIPTOR_ROOTSIFT) {
        if (lane == 0) {
            float l1 = 0.0f;
            for (int i = 0; i < 128; i++) l1 += fmaxf(s_desc[i], 0.0f);
            s_scale[0] = l1;
            s_scale[1] = 1.0f / l1;
        }
        __syncthreads();
        if (s_scale[0] > 1e-7f) {
            float inv = s_scale[1];
            s_desc[lane] = sqrtf(fmaxf(s_desc[lane], 0.0f) * inv);
            s_desc[lane + 64] = sqrtf(fmaxf(s_desc[lane + 64], 0.0f) * inv);
        }
        __syncthreads();
    }
    float *od = out_d + (size_t)f * 128;
    od[lane] = s_desc[lane];
    od[lane + 64] = s_desc[lane + 64];
    if (lane == 0) {
        ps_feature o;
        o.x = kp.x; o.y = kp.y; o.scale = kp.scale; o.response = kp.response; o.orientation = theta;
        out_f[f] = o;
    }
}

hipError_t launch_descriptors(const PyrView &pv, const Cand *kps, const float *ori, const uint32_t *feat_kp,
                              const uint32_t *feat_j, const uint32_t *nfeat, uint32_t cap, int kind, ps_feature *out_f,
                              float *out_d, hipStream_t st) {
    if (cap == 0) return hipSuccess;
    descriptor_kernel<<<cap, 64, 0, st>>>(pv, kps, ori, feat_kp, feat_j, nfeat, cap, kind, out_f, out_d);
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

}  // namespace ps
