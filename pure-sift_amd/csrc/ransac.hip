// Homography LO-RANSAC (pipeline.rs:2663-2755) seeded by similarity RANSAC
// (pipeline.rs:2055-2191), one workgroup per problem, for a batch of problems.
//
// Exactness: every f32 / f64 expression is evaluated in the reference's order with
// separate roundings (no contraction), sums run sequentially in index order, and the
// SplitMix64 stream (pipeline.rs:1862-1868) is reproduced draw for draw, so inliers and
// models equal the sequential algorithm's. The adaptive iteration bound uses the device
// f64 log (glibc's may differ in the last ulp, which moves ceil() only when the quotient
// lies within ~1e-15 of an integer).
//
// Parallel structure. The sample stream never depends on the evaluation (draw k of the
// stream is SplitMix64 at position k; which draws form an iteration depends only on the
// draws), so hypotheses are evaluated in chunks of 256, one per thread (sample, degeneracy
// test, 4-point normalised DLT with the cyclic Jacobi eigen-solver, sequential scoring over
// all matches). The chunk is then replayed in order by the whole workgroup: the
// accept rule, local optimisation on improvement (cooperative: ordered inlier compaction,
// per-entry sequential AtA sums, sequential error sums) and the adaptive bound, exactly as
// the sequential loop would, stopping at the bound.
#include <algorithm>

#include "ps_internal.h"

#pragma clang fp contract(off)

namespace ps {

namespace {

constexpr int kRT = 256;            // threads per problem
constexpr uint32_t kStage = 512;   // points per staging chunk
constexpr uint32_t kDraws = 1024;  // sampler draws per buffer
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// k-th output (k = 0, 1, ...) of the generator started at `seed`
__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t k) { return mix64(seed + (k + 1) * kGamma); }
__device__ __forceinline__ double unit_draw(uint64_t z) { return (double)(z >> 11) * (1.0 / 9007199254740992.0); }

// partition_point(c < u), clamped to n - 1 (pipeline.rs:2031-2036)
__device__ uint32_t cdf_pick(const double *cdf, uint32_t n, double u) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    return lo < n - 1 ? lo : n - 1;
}
__device__ __forceinline__ uint32_t uniform_pick(uint32_t n, double u) {
    uint64_t i = (uint64_t)(u * (double)n);
    return i < n ? (uint32_t)i : n - 1;
}

// pipeline.rs:1979-2002
__device__ uint32_t adaptive_iters(double inl, double total, int s, double conf, uint32_t min_it, uint32_t cap) {
    if (total == 0.0 || inl == 0.0) return cap;
    double w = inl / total;
    if (w < 1e-3) w = 1e-3;
    if (w > 1.0 - 1e-6) w = 1.0 - 1e-6;
    double p = 1.0, a = w;
    int b = s;
    for (;;) {
        if (b & 1) p *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    if (p < 1e-12) p = 1e-12;
    if (p > 1.0 - 1e-9) p = 1.0 - 1e-9;
    double num = log(1.0 - conf), den = log(1.0 - p);
    if (fabs(den) < 1e-12) return min_it;
    double est = ceil(num / den);
    if (!isfinite(est)) return cap;
    double e = est <= 0.0 ? 0.0 : est;
    if (e < (double)min_it) e = (double)min_it;
    if (e > (double)cap) e = (double)cap;
    return (uint32_t)e;
}

struct Sim {
    float a, b, tx, ty;
};

// pipeline.rs:1873-1897 (sequential f32 error sum)
__device__ uint32_t eval_sim(const ps_feature_match *m, uint32_t n, float a, float b, float tx, float ty, float tsq,
                             float *err) {
    uint32_t count = 0;
    float es = 0.0f;
#pragma unroll 8
    for (uint32_t i = 0; i < n; i++) {
        float px = a * m[i].qx - b * m[i].qy + tx;
        float py = b * m[i].qx + a * m[i].qy + ty;
        float dx = px - m[i].tx, dy = py - m[i].ty;
        float e = dx * dx + dy * dy;
        if (e <= tsq) {
            count++;
            es += e;
        }
    }
    *err = es;
    return count;
}

__device__ __forceinline__ bool sim_inlier(const ps_feature_match &p, const Sim &s, float tsq) {
    float dx = (s.a * p.qx - s.b * p.qy + s.tx) - p.tx, dy = (s.b * p.qx + s.a * p.qy + s.ty) - p.ty;
    return dx * dx + dy * dy <= tsq;
}

// pipeline.rs:2276-2353: cyclic Jacobi on the symmetric 9x9 AtA; eigenvector of the
// smallest eigenvalue (first index on ties). Every loop has a constant trip count and is
// unrolled, so all indices are compile-time constants and the matrix (diagonal + upper
// triangle: the reference writes a[p][r] and a[r][p] with the same value) and the
// eigenvectors stay in registers instead of per-lane scratch memory. Same operations in
// the same order as the reference.
__device__ constexpr int tri9(int p, int q) { return p * 9 - p * (p + 1) / 2 + (q - p - 1); }   // p < q

__device__ bool smallest_eigvec9(double (&a)[9][9], double (&h)[9]) {
    double d[9], u[36], v[9][9];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        d[i] = a[i][i];
#pragma unroll
        for (int j = 0; j < 9; j++) v[i][j] = i == j ? 1.0 : 0.0;
#pragma unroll
        for (int j = i + 1; j < 9; j++) u[tri9(i, j)] = a[i][j];
    }
    for (int sweep = 0; sweep < 20; sweep++) {
        double mo = 0.0;
#pragma unroll
        for (int k = 0; k < 36; k++) {
            const double av = fabs(u[k]);
            if (av > mo) mo = av;
        }
        if (mo < 1e-13) break;
#pragma unroll
        for (int p = 0; p < 8; p++)
#pragma unroll
            for (int q = p + 1; q < 9; q++) {
                const double apq = u[tri9(p, q)];
                if (fabs(apq) < 1e-15) continue;
                const double app = d[p], aqq = d[q];
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                            : -1.0 / (-tau + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                d[p] = app - t * apq;
                d[q] = aqq + t * apq;
                u[tri9(p, q)] = 0.0;
#pragma unroll
                for (int r = 0; r < 9; r++) {
                    if (r != p && r != q) {
                        const int ip = r < p ? tri9(r, p) : tri9(p, r), iq = r < q ? tri9(r, q) : tri9(q, r);
                        const double arp = u[ip], arq = u[iq];
                        u[ip] = c * arp - s * arq;
                        u[iq] = s * arp + c * arq;
                    }
                    const double vrp = v[r][p], vrq = v[r][q];
                    v[r][p] = c * vrp - s * vrq;
                    v[r][q] = s * vrp + c * vrq;
                }
            }
    }
    int mi = 0;
    double me = d[0];
#pragma unroll
    for (int i = 1; i < 9; i++)
        if (d[i] < me) {
            me = d[i];
            mi = i;
        }
    bool ok = true;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        double x = v[r][0];
#pragma unroll
        for (int i = 1; i < 9; i++) x = mi == i ? v[r][i] : x;   // select, no dynamic index
        h[r] = x;
        ok = ok && isfinite(x);
    }
    return ok;
}

// The same cyclic Jacobi, run by one wave for a matrix held in LDS (row-major 9x9, symmetric):
// lane r < 9 keeps row r of the matrix and row r of the eigenvector matrix. Each rotation's
// angle is computed identically by every lane from the broadcast d[p], d[q], a[p][q]; lane r
// updates its own pair (a[r][p], a[r][q]) and (v[r][p], v[r][q]) with the reference's
// expressions, and lanes p and q then take their rows from the updated columns (symmetry), so
// every element goes through exactly the operations of smallest_eigvec9. Writes the
// eigenvector to h (LDS) from lanes 0..8 and returns the wave-uniform success flag.
// lane `src` (uniform) of x, through the scalar registers
__device__ __forceinline__ double readlane_f64(double x, int src) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ bool smallest_eigvec9_wave(const double *A, double *h) {
    const int lane = threadIdx.x & 63;
    const int r = lane < 9 ? lane : 8;
    const bool mine = lane < 9;
    double row[9], vr[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        row[k] = A[r * 9 + k];
        vr[k] = k == r ? 1.0 : 0.0;
    }
    for (int sweep = 0; sweep < 20; sweep++) {
        double mo = 0.0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const double av = fabs(row[k]);
            if (mine && k > r && av > mo) mo = av;
        }
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            const double om = __shfl_xor(mo, o);
            if (om > mo) mo = om;
        }
        mo = readlane_f64(mo, 0);
        if (mo < 1e-13) break;
#pragma unroll
        for (int p = 0; p < 8; p++)
#pragma unroll
            for (int q = p + 1; q < 9; q++) {
                const double apq = readlane_f64(row[p], q);   // a[q][p] = a[p][q], kept current by lane q
                if (fabs(apq) < 1e-15) continue;
                const double app = readlane_f64(row[p], p), aqq = readlane_f64(row[q], q);
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                            : -1.0 / (-tau + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
                const double dp = app - t * apq, dq = aqq + t * apq;
                if (r != p && r != q) {
                    const double arp = row[p], arq = row[q];
                    row[p] = c * arp - sn * arq;
                    row[q] = sn * arp + c * arq;
                }
                {
                    const double vrp = vr[p], vrq = vr[q];
                    vr[p] = c * vrp - sn * vrq;
                    vr[q] = sn * vrp + c * vrq;
                }
                // rows p and q from the updated columns p and q
#pragma unroll
                for (int k = 0; k < 9; k++) {
                    if (k == p || k == q) continue;
                    const double kp = readlane_f64(row[p], k), kq = readlane_f64(row[q], k);
                    if (r == p) row[k] = kp;
                    if (r == q) row[k] = kq;
                }
                if (r == p) {
                    row[p] = dp;
                    row[q] = 0.0;
                }
                if (r == q) {
                    row[q] = dq;
                    row[p] = 0.0;
                }
            }
    }
    // smallest eigenvalue, first index on ties
    int mi = 0;
    double me = readlane_f64(row[0], 0);
#pragma unroll
    for (int i = 1; i < 9; i++) {
        const double di = readlane_f64(row[i], i);
        if (di < me) {
            me = di;
            mi = i;
        }
    }
    double x = vr[0];
#pragma unroll
    for (int i = 1; i < 9; i++) x = mi == i ? vr[i] : x;
    const bool fin = !mine || isfinite(x);
    if (mine) h[r] = x;
    return __all(fin);
}

// pipeline.rs:2357-2404
__device__ bool degenerate_tri(const double (&px)[4], const double (&py)[4]) {
    const int T[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    for (int t = 0; t < 4; t++) {
        int i = T[t][0], j = T[t][1], k = T[t][2];
        if (!isfinite(px[i]) || !isfinite(py[i]) || !isfinite(px[j]) || !isfinite(py[j]) || !isfinite(px[k]) ||
            !isfinite(py[k]))
            return true;
        double dx1 = px[j] - px[i], dy1 = py[j] - py[i], dx2 = px[k] - px[i], dy2 = py[k] - py[i];
        double dx3 = px[k] - px[j], dy3 = py[k] - py[j];
        double cross = fabs(dx1 * dy2 - dy1 * dx2);
        double e1 = dx1 * dx1 + dy1 * dy1, e2 = dx2 * dx2 + dy2 * dy2, e3 = dx3 * dx3 + dy3 * dy3;
        double me = fmax(fmax(e1, e2), e3);
        if (cross < 4.0 || cross < 1e-3 * me) return true;
    }
    return false;
}

// Normalisation + AtA + eigenvector + denormalisation (pipeline.rs:2408-2528), for the
// 4 points of one hypothesis (per thread).
__device__ bool fit_dlt4(const ps_feature_match *m, const uint32_t (&idx)[4], double (&H)[3][3]) {
    const double inv_n = 1.0 / 4.0;
    double cxq = 0, cyq = 0, cxt = 0, cyt = 0;
    for (int i = 0; i < 4; i++) {
        const ps_feature_match &p = m[idx[i]];
        double qx = p.qx, qy = p.qy, tx = p.tx, ty = p.ty;
        if (!isfinite(qx) || !isfinite(qy) || !isfinite(tx) || !isfinite(ty)) return false;
        cxq += qx;
        cyq += qy;
        cxt += tx;
        cyt += ty;
    }
    cxq *= inv_n;
    cyq *= inv_n;
    cxt *= inv_n;
    cyt *= inv_n;
    double mdq = 0, mdt = 0;
    for (int i = 0; i < 4; i++) {
        const ps_feature_match &p = m[idx[i]];
        double dxq = (double)p.qx - cxq, dyq = (double)p.qy - cyq, dxt = (double)p.tx - cxt, dyt = (double)p.ty - cyt;
        mdq += sqrt(dxq * dxq + dyq * dyq);
        mdt += sqrt(dxt * dxt + dyt * dyt);
    }
    mdq *= inv_n;
    mdt *= inv_n;
    if (!isfinite(mdq) || !isfinite(mdt) || mdq < 1e-6 || mdt < 1e-6) return false;
    const double SQ2 = 1.41421356237309504880;
    double sq = SQ2 / mdq, stt = SQ2 / mdt;
    double ata[9][9];
    for (int a = 0; a < 9; a++)
        for (int b = 0; b < 9; b++) ata[a][b] = 0.0;
    for (int i = 0; i < 4; i++) {
        const ps_feature_match &p = m[idx[i]];
        double x = ((double)p.qx - cxq) * sq, y = ((double)p.qy - cyq) * sq;
        double xp = ((double)p.tx - cxt) * stt, yp = ((double)p.ty - cyt) * stt;
        double r1[9] = {-x, -y, -1.0, 0.0, 0.0, 0.0, x * xp, y * xp, xp};
        double r2[9] = {0.0, 0.0, 0.0, -x, -y, -1.0, x * yp, y * yp, yp};
        for (int a = 0; a < 9; a++)
            for (int b = a; b < 9; b++) ata[a][b] += r1[a] * r1[b] + r2[a] * r2[b];
    }
    for (int a = 0; a < 9; a++)
        for (int b = 0; b < a; b++) ata[a][b] = ata[b][a];
    double hv[9];
    if (!smallest_eigvec9(ata, hv)) return false;
    double hn[3][3] = {{hv[0], hv[1], hv[2]}, {hv[3], hv[4], hv[5]}, {hv[6], hv[7], hv[8]}};
    double tq[3][3] = {{sq, 0.0, -sq * cxq}, {0.0, sq, -sq * cyq}, {0.0, 0.0, 1.0}};
    double ist = 1.0 / stt;
    double tti[3][3] = {{ist, 0.0, cxt}, {0.0, ist, cyt}, {0.0, 0.0, 1.0}};
    double tmp[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) tmp[r][c] = hn[r][0] * tq[0][c] + hn[r][1] * tq[1][c] + hn[r][2] * tq[2][c];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) H[r][c] = tti[r][0] * tmp[0][c] + tti[r][1] * tmp[1][c] + tti[r][2] * tmp[2][c];
    if (fabs(H[2][2]) > 1e-12) {
        double inv = 1.0 / H[2][2];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) H[r][c] *= inv;
    }
    return true;
}

// squared reprojection error of one match, or -1 when |w| < 1e-9 (pipeline.rs:2550-2576)
__device__ __forceinline__ double reproj_err(const ps_feature_match &p, const double *H) {
    double qx = p.qx, qy = p.qy;
    double w = H[6] * qx + H[7] * qy + H[8];
    if (fabs(w) < 1e-9) return -1.0;
    double iw = 1.0 / w;
    double px = (H[0] * qx + H[1] * qy + H[2]) * iw;
    double py = (H[3] * qx + H[4] * qy + H[5]) * iw;
    double dx = px - (double)p.tx, dy = py - (double)p.ty;
    return dx * dx + dy * dy;
}

// sequential count / error / sampling-mass of one model (one thread)
__device__ uint32_t eval_h_seq(const ps_feature_match *m, uint32_t n, const double *H, double tsq, double *err,
                               const float *wts, double *mass) {
    uint32_t c = 0;
    double es = 0.0, ms = 0.0;
#pragma unroll 8
    for (uint32_t i = 0; i < n; i++) {
        double e = reproj_err(m[i], H);
        if (e >= 0.0 && e <= tsq) {
            c++;
            es += e;
            if (wts) ms += (double)wts[i];
        }
    }
    *err = es;
    if (mass) *mass = ms;
    return c;
}

// The sample that starts at draws[0]: distinct match indices in draw order until K of them
// (valid) or A attempts (invalid); returns the number of draws it consumes. The sampler's
// sequential grouping is the chain of such samples, each starting where the previous ended.
template <int K, int A>
__device__ __forceinline__ uint32_t scan_sample(const uint32_t *draws, uint32_t (&q)[K], bool &valid) {
    uint32_t ch = 0, at = 0;
#pragma unroll
    for (int j = 0; j < K; j++) q[j] = 0;
    while (true) {
        const uint32_t c = draws[at++];
        bool dup = false;
#pragma unroll
        for (int j = 0; j < K; j++) dup |= (uint32_t)j < ch && q[j] == c;
        if (!dup) {
#pragma unroll
            for (int j = 0; j < K; j++)
                if ((uint32_t)j == ch) q[j] = c;
            ch++;
        }
        if (ch == (uint32_t)K || at == (uint32_t)A) break;
    }
    valid = ch == (uint32_t)K;
    return at;
}

struct Hyp {
    double H[9];
    double err, mass;
    uint32_t count, valid;
};

struct Shared {
    Hyp hyp[kRT];
    uint32_t quad[kRT][4];
    uint32_t qvalid[kRT];
    uint32_t draws[kDraws + 16];   // a buffer of sampler draws plus one sample's worth of lookahead
    uint8_t len[kDraws];           // draws consumed by the sample starting at each position
    uint32_t start[kRT];           // buffer position of each sample of the batch
    uint32_t scan_c[kRT];          // replay filter: prefix-best count (+ validity bit 31)
    uint32_t cand[kRT];            // replay filter: hypotheses that may improve the best
    uint32_t n_cand;
    uint32_t wave_cnt[kRT / 64];
    uint32_t n_quads, n_draws, draw_pos, total;
    double red[kRT];
    // chunk staging for the order-dependent sums: all threads load / compute per-point
    // terms into it, thread 0 adds them in point order (no global loads on its serial path)
    union {
        double d[4][kStage];
        float f[4][2 * kStage];
    } stage;
};

// Ordered compaction of the indices i in [0, n) whose predicate holds, into out[];
// every thread calls it; returns the count.
template <typename Pred>
__device__ uint32_t ordered_compact(uint32_t n, uint32_t *out, Shared &S, Pred pred) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t base = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += kRT) {
        const uint32_t i = c0 + tid;
        const bool p = i < n && pred(i);
        const uint64_t mask = __ballot(p);
        if (lane == 0) S.wave_cnt[w] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t off = base;
        for (int v = 0; v < w; v++) off += S.wave_cnt[v];
        off += (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (p) out[off] = i;
        uint32_t tot = 0;
        for (int v = 0; v < kRT / 64; v++) tot += S.wave_cnt[v];
        base += tot;
        __syncthreads();
    }
    return base;
}

// pipeline.rs:1900-1974 over m[idx[0..n)]: the sequential f32 sums done by thread 0 from the
// staging buffer (the points, then the per-point products computed in parallel, in chunks).
// Every thread calls it; the model is broadcast.
__device__ bool fit_procrustes_block(const ps_feature_match *m, const uint32_t *idx, uint32_t n, Sim *out,
                                     Shared &S) {
    __shared__ float s_r[4];
    if (n == 0) return false;
    if (n == 1) {
        *out = Sim{1.0f, 0.0f, m[idx[0]].dx, m[idx[0]].dy};
        return true;
    }
    const int tid = threadIdx.x;
    float(*F)[2 * kStage] = S.stage.f;
    const float inv_n = 1.0f / (float)n;
    float mqx = 0, mqy = 0, mtx = 0, mty = 0;   // thread 0's running sums
    for (uint32_t c0 = 0; c0 < n; c0 += kStage) {
        const uint32_t cn = min(kStage, n - c0);
        for (uint32_t k = tid; k < cn; k += kRT) {
            const ps_feature_match &p = m[idx[c0 + k]];
            F[0][k] = p.qx;
            F[1][k] = p.qy;
            F[2][k] = p.tx;
            F[3][k] = p.ty;
        }
        __syncthreads();
        if (tid == 0) {
#pragma unroll 8
            for (uint32_t k = 0; k < cn; k++) {
                mqx += F[0][k];
                mqy += F[1][k];
                mtx += F[2][k];
                mty += F[3][k];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        s_r[0] = mqx * inv_n;
        s_r[1] = mqy * inv_n;
        s_r[2] = mtx * inv_n;
        s_r[3] = mty * inv_n;
    }
    __syncthreads();
    const float cqx = s_r[0], cqy = s_r[1], ctx = s_r[2], cty = s_r[3];
    float sqq = 0, na = 0, nb = 0;   // thread 0
    for (uint32_t c0 = 0; c0 < n; c0 += kStage) {
        const uint32_t cn = min(kStage, n - c0);
        for (uint32_t k = tid; k < cn; k += kRT) {
            const ps_feature_match &p = m[idx[c0 + k]];
            const float qx = p.qx - cqx, qy = p.qy - cqy, tx = p.tx - ctx, ty = p.ty - cty;
            F[0][k] = qx * qx + qy * qy;
            F[1][k] = qx * tx + qy * ty;
            F[2][k] = qx * ty - qy * tx;
        }
        __syncthreads();
        if (tid == 0) {
#pragma unroll 8
            for (uint32_t k = 0; k < cn; k++) {
                sqq += F[0][k];
                na += F[1][k];
                nb += F[2][k];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        float a = 1.0f, b = 0.0f;
        if (sqq > 1e-5f) {
            const float af = na / sqq, bf = nb / sqq;
            const float s2 = af * af + bf * bf;
            if (s2 >= 0.04f && s2 <= 25.0f) {
                a = af;
                b = bf;
            }
        }
        s_r[0] = a;
        s_r[1] = b;
        s_r[2] = ctx - (a * cqx - b * cqy);
        s_r[3] = cty - (b * cqx + a * cqy);
    }
    __syncthreads();
    *out = Sim{s_r[0], s_r[1], s_r[2], s_r[3]};
    __syncthreads();
    return true;
}

// Entry j of the two DLT rows of a normalised correspondence (x, y) -> (xp, yp):
// r1 = (-x, -y, -1, 0, 0, 0, x xp, y xp, xp), r2 = (0, 0, 0, -x, -y, -1, x yp, y yp, yp),
// selected without a dynamically indexed array (which would live in scratch memory).
__device__ __forceinline__ double dlt_row1(int j, double x, double y, double xp) {
    const double v[9] = {-x, -y, -1.0, 0.0, 0.0, 0.0, x * xp, y * xp, xp};
    double r = v[0];
#pragma unroll
    for (int i = 1; i < 9; i++) r = j == i ? v[i] : r;
    return r;
}
__device__ __forceinline__ double dlt_row2(int j, double x, double y, double yp) {
    const double v[9] = {0.0, 0.0, 0.0, -x, -y, -1.0, x * yp, y * yp, yp};
    double r = v[0];
#pragma unroll
    for (int i = 1; i < 9; i++) r = j == i ? v[i] : r;
    return r;
}

// Cooperative DLT over m[idx[0..n)] (n >= 4), same arithmetic as the sequential fit. The points
// pass through the staging buffer in chunks: thread 0 adds the centroid and mean-distance
// terms in point order (the distances are computed in parallel), each of the 45 AtA entries is
// summed by one thread in point order from the staged normalised coordinates, and thread 0
// takes the eigenvector. Result broadcast through Hout (shared).
__device__ bool fit_dlt_block(const ps_feature_match *m, const uint32_t *idx, uint32_t n, double *Hout, Shared &S) {
    __shared__ double s_norm[6];
    __shared__ double s_ata[45];
    __shared__ int s_ok, s_bad;
    const int tid = threadIdx.x;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    double cxq = 0, cyq = 0, cxt = 0, cyt = 0;   // thread 0's running sums
    for (uint32_t c0 = 0; c0 < n; c0 += kStage) {
        const uint32_t cn = min(kStage, n - c0);
        for (uint32_t k = tid; k < cn; k += kRT) {
            const ps_feature_match &p = m[idx[c0 + k]];
            const double qx = p.qx, qy = p.qy, tx = p.tx, ty = p.ty;
            if (!isfinite(qx) || !isfinite(qy) || !isfinite(tx) || !isfinite(ty)) s_bad = 1;
            S.stage.d[0][k] = qx;
            S.stage.d[1][k] = qy;
            S.stage.d[2][k] = tx;
            S.stage.d[3][k] = ty;
        }
        __syncthreads();
        if (tid == 0 && !s_bad) {
#pragma unroll 8
            for (uint32_t k = 0; k < cn; k++) {
                cxq += S.stage.d[0][k];
                cyq += S.stage.d[1][k];
                cxt += S.stage.d[2][k];
                cyt += S.stage.d[3][k];
            }
        }
        __syncthreads();
    }
    // a non-finite point fails the fit (the sequential loop stops at the first one)
    if (s_bad) {
        __syncthreads();
        return false;
    }
    const double inv_n = 1.0 / (double)n;
    if (tid == 0) {
        s_norm[0] = cxq * inv_n;
        s_norm[1] = cyq * inv_n;
        s_norm[2] = cxt * inv_n;
        s_norm[3] = cyt * inv_n;
    }
    __syncthreads();
    {
        const double mxq = s_norm[0], myq = s_norm[1], mxt = s_norm[2], myt = s_norm[3];
        double mdq = 0, mdt = 0;
        for (uint32_t c0 = 0; c0 < n; c0 += kStage) {
            const uint32_t cn = min(kStage, n - c0);
            for (uint32_t k = tid; k < cn; k += kRT) {
                const ps_feature_match &p = m[idx[c0 + k]];
                const double dxq = (double)p.qx - mxq, dyq = (double)p.qy - myq, dxt = (double)p.tx - mxt,
                             dyt = (double)p.ty - myt;
                S.stage.d[0][k] = sqrt(dxq * dxq + dyq * dyq);
                S.stage.d[1][k] = sqrt(dxt * dxt + dyt * dyt);
            }
            __syncthreads();
            if (tid == 0) {
#pragma unroll 8
                for (uint32_t k = 0; k < cn; k++) {
                    mdq += S.stage.d[0][k];
                    mdt += S.stage.d[1][k];
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            mdq *= inv_n;
            mdt *= inv_n;
            const bool ok = !(!isfinite(mdq) || !isfinite(mdt) || mdq < 1e-6 || mdt < 1e-6);
            const double SQ2 = 1.41421356237309504880;
            s_norm[4] = ok ? SQ2 / mdq : 0.0;
            s_norm[5] = ok ? SQ2 / mdt : 0.0;
            s_ok = ok;
        }
        __syncthreads();
    }
    if (!s_ok) {
        __syncthreads();
        return false;
    }
    {
        // entry (a, b), a <= b, in row-major upper-triangle order
        int a = 0, r = tid < 45 ? tid : 0;
        while (r >= 9 - a) {
            r -= 9 - a;
            a++;
        }
        const int b = a + r;
        const double cxq = s_norm[0], cyq = s_norm[1], cxt = s_norm[2], cyt = s_norm[3], sq = s_norm[4],
                     stt = s_norm[5];
        double acc = 0.0;
        for (uint32_t c0 = 0; c0 < n; c0 += kStage) {
            const uint32_t cn = min(kStage, n - c0);
            for (uint32_t k = tid; k < cn; k += kRT) {
                const ps_feature_match &p = m[idx[c0 + k]];
                S.stage.d[0][k] = ((double)p.qx - cxq) * sq;
                S.stage.d[1][k] = ((double)p.qy - cyq) * sq;
                S.stage.d[2][k] = ((double)p.tx - cxt) * stt;
                S.stage.d[3][k] = ((double)p.ty - cyt) * stt;
            }
            __syncthreads();
            if (tid < 45) {
#pragma unroll 8
                for (uint32_t k = 0; k < cn; k++) {
                    const double x = S.stage.d[0][k], y = S.stage.d[1][k], xp = S.stage.d[2][k], yp = S.stage.d[3][k];
                    acc += dlt_row1(a, x, y, xp) * dlt_row1(b, x, y, xp) +
                           dlt_row2(a, x, y, yp) * dlt_row2(b, x, y, yp);
                }
            }
            __syncthreads();
        }
        if (tid < 45) s_ata[tid] = acc;
    }
    __syncthreads();
    // the symmetric AtA, then its smallest eigenvector by wave 0
    __shared__ double s_full[81], s_hv[9];
    __shared__ int s_eig_ok;
    if (tid < 81) {
        const int a = tid / 9, b = tid % 9, lo = a < b ? a : b, hi = a < b ? b : a;
        s_full[tid] = s_ata[lo * 9 - lo * (lo - 1) / 2 + (hi - lo)];
    }
    __syncthreads();
    if (tid < 64) {
        const bool eok = smallest_eigvec9_wave(s_full, s_hv);
        if (tid == 0) s_eig_ok = eok;
    }
    __syncthreads();
    if (tid == 0) {
        double hv[9];
        for (int i = 0; i < 9; i++) hv[i] = s_hv[i];
        int ok = s_eig_ok;
        if (ok) {
            const double cxq = s_norm[0], cyq = s_norm[1], cxt = s_norm[2], cyt = s_norm[3], sq = s_norm[4],
                         stt = s_norm[5];
            double hn[3][3] = {{hv[0], hv[1], hv[2]}, {hv[3], hv[4], hv[5]}, {hv[6], hv[7], hv[8]}};
            double tq[3][3] = {{sq, 0.0, -sq * cxq}, {0.0, sq, -sq * cyq}, {0.0, 0.0, 1.0}};
            double ist = 1.0 / stt;
            double tti[3][3] = {{ist, 0.0, cxt}, {0.0, ist, cyt}, {0.0, 0.0, 1.0}};
            double tmp[3][3], H[3][3];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++)
                    tmp[r][c] = hn[r][0] * tq[0][c] + hn[r][1] * tq[1][c] + hn[r][2] * tq[2][c];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++)
                    H[r][c] = tti[r][0] * tmp[0][c] + tti[r][1] * tmp[1][c] + tti[r][2] * tmp[2][c];
            if (fabs(H[2][2]) > 1e-12) {
                double inv = 1.0 / H[2][2];
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) H[r][c] *= inv;
            }
            for (int i = 0; i < 9; i++) Hout[i] = H[i / 3][i % 3];
        }
        s_ok = ok;
    }
    __syncthreads();
    const bool ok = s_ok;
    __syncthreads();
    return ok;
}

// Cooperative sequential-order scoring: per-match errors (and weights) computed in parallel
// into the staging buffer, chunk by chunk, and added by thread 0 in index order. Returns the
// count; *err / *mass broadcast through shared memory.
__device__ uint32_t eval_h_block(const ps_feature_match *m, uint32_t n, const double *H, double tsq,
                                 const float *wts, double *err, double *mass, Shared &S) {
    __shared__ uint32_t s_c;
    __shared__ double s_e, s_m;
    const int tid = threadIdx.x;
    uint32_t c = 0;
    double es = 0.0, ms = 0.0;
    for (uint32_t c0 = 0; c0 < n; c0 += kStage) {
        const uint32_t cn = min(kStage, n - c0);
        for (uint32_t k = tid; k < cn; k += kRT) {
            const double e = reproj_err(m[c0 + k], H);
            S.stage.d[0][k] = (e >= 0.0 && e <= tsq) ? e : -1.0;
            if (wts) S.stage.d[1][k] = (double)wts[c0 + k];
        }
        __syncthreads();
        if (tid == 0) {
#pragma unroll 8
            for (uint32_t k = 0; k < cn; k++) {
                const double e = S.stage.d[0][k];
                if (e >= 0.0) {
                    c++;
                    es += e;
                    if (wts) ms += S.stage.d[1][k];
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        s_c = c;
        s_e = es;
        s_m = ms;
    }
    __syncthreads();
    const uint32_t cc = s_c;
    *err = s_e;
    if (mass) *mass = s_m;
    __syncthreads();
    return cc;
}

// pipeline.rs:2605-2630 (two refits on the current inliers while they improve)
__device__ void local_opt_block(const ps_feature_match *m, uint32_t n, double *H, uint32_t *bc, double *be,
                                double tsq, uint32_t *ibuf, Shared &S) {
    __shared__ double s_R[9];
    for (int it = 0; it < 2; it++) {
        const uint32_t ni = ordered_compact(n, ibuf, S, [&](uint32_t i) {
            double e = reproj_err(m[i], H);
            return e >= 0.0 && e <= tsq;
        });
        if (ni < 4) break;
        if (!fit_dlt_block(m, ibuf, ni, s_R, S)) break;
        double re;
        const uint32_t rc = eval_h_block(m, n, s_R, tsq, nullptr, &re, nullptr, S);
        if (rc > *bc || (rc == *bc && re < *be)) {
            __syncthreads();
            if (threadIdx.x < 9) H[threadIdx.x] = s_R[threadIdx.x];
            *bc = rc;
            *be = re;
            __syncthreads();
        } else {
            break;
        }
    }
}

}  // namespace

struct RansacParamsDev {
    float reproj;
    int sampler;          // 0: reference 1/ratio weights, 1: uniform, 2: caller weights
    double confidence;
    uint32_t min_iters, max_iters;
    int local_opt, sim_seed_model, adaptive;
    int sim_only;         // stop after the similarity stage (its inliers stay in the workspace)
};

namespace {

__global__ __launch_bounds__(kRT) void ransac_kernel(const ps_feature_match *matches, const float *weights,
                                                     const ps_ransac_problem *probs, RansacParamsDev prm,
                                                     double *work_d, uint32_t *work_u, uint32_t *inliers,
                                                     ps_ransac_result *outs) {
    __shared__ Shared S;
    __shared__ double s_H[9];
    __shared__ uint32_t s_flag;
    const int tid = threadIdx.x;
    const ps_ransac_problem P = probs[blockIdx.x];
    const uint32_t n = P.n;
    const ps_feature_match *m = matches + P.offset;
    const float *wts = prm.sampler == 2 ? weights + P.offset : nullptr;
    double *cdf = work_d + (size_t)P.offset * 2;           // n doubles (then n unused)
    uint32_t *ibuf = work_u + (size_t)P.offset * 2;         // n u32
    uint32_t *sim_inl = ibuf + n;                           // n u32
    uint32_t *out_inl = inliers + P.offset;
    ps_ransac_result &O = outs[blockIdx.x];

    if (tid == 0) {
        O.has_model = 0;
        O.n_inliers = 0;
        O.hyps = 0;
        O.has_sim = 0;
        O.n_sim = 0;
    }
    if (n == 0) return;

    // ---------------- similarity RANSAC (pipeline.rs:2055-2191)
    const float fth = fmaxf(prm.reproj, 0.5f), ftsq = fth * fth;
    Sim best_sim{1.0f, 0.0f, 0.0f, 0.0f};
    bool have_sim = false;
    uint32_t n_sim = 0;
    // reference weights 1 / clamp(ratio) (pipeline.rs:2006-2028), prefix sums in f64 in order
    // (weights computed in parallel into the staging buffer, the running sum by thread 0 in
    // order, the normalisation in parallel)
    auto build_cdf = [&](bool ref_weights) {
        double acc = 0.0;   // thread 0
        for (uint32_t c0 = 0; c0 < n; c0 += kStage) {
            const uint32_t cn = min(kStage, n - c0);
            for (uint32_t k = tid; k < cn; k += kRT) {
                const uint32_t i = c0 + k;
                double wv;
                if (!ref_weights) {
                    wv = (double)wts[i];
                } else {
                    float rr = m[i].ratio;
                    double r = isfinite(rr) ? (double)rr : 1.0;
                    if (isfinite(rr)) {
                        if (r < 0.05) r = 0.05;
                        if (r > 1.0) r = 1.0;
                    }
                    wv = 1.0 / r;
                }
                S.stage.d[0][k] = wv;
            }
            __syncthreads();
            if (tid == 0) {
#pragma unroll 8
                for (uint32_t k = 0; k < cn; k++) {
                    acc += S.stage.d[0][k];
                    cdf[c0 + k] = acc;
                }
                S.red[0] = acc;
            }
            __syncthreads();
        }
        const double total = S.red[0];
        if (total > 0.0) {
            const double inv = 1.0 / total;
            for (uint32_t i = tid; i < n; i += kRT) cdf[i] *= inv;
        }
        __syncthreads();
        if (tid == 0) cdf[n - 1] = 1.0;
        __syncthreads();
    };

    if (prm.sim_seed_model) {
        if (n == 1) {
            best_sim = Sim{1.0f, 0.0f, m[0].dx, m[0].dy};
            have_sim = true;
            n_sim = 1;
            if (tid == 0) sim_inl[0] = 0;
        } else {
            // the 32 smallest (ratio, index): ratios order as their sign-adjusted bit patterns
            __shared__ uint64_t s_sorted[32];
            __shared__ uint64_t s_red[kRT];
            const uint32_t ns = n < 32 ? n : 32;
            uint64_t prev = 0;
            for (uint32_t r = 0; r < ns; r++) {
                uint64_t mine = ~0ull;
                for (uint32_t i = tid; i < n; i += kRT) {
                    uint32_t b = __float_as_uint(m[i].ratio);
                    b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
                    uint64_t key = ((uint64_t)b << 32) | i;
                    if ((r == 0 || key > prev) && key < mine) mine = key;
                }
                s_red[tid] = mine;
                __syncthreads();
                for (int o = kRT / 2; o > 0; o >>= 1) {
                    if (tid < o && s_red[tid + o] < s_red[tid]) s_red[tid] = s_red[tid + o];
                    __syncthreads();
                }
                prev = s_red[0];
                if (tid == 0) s_sorted[r] = prev;
                __syncthreads();
            }
            // initial translation candidates (pipeline.rs:2080-2101), one per thread
            __shared__ uint32_t s_cc[32];
            __shared__ float s_ce[32];
            if (tid < (int)ns) {
                const uint32_t si = (uint32_t)s_sorted[tid];
                float e;
                s_cc[tid] = eval_sim(m, n, 1.0f, 0.0f, m[si].dx, m[si].dy, ftsq, &e);
                s_ce[tid] = e;
            }
            __syncthreads();
            {
                const uint32_t s0 = (uint32_t)s_sorted[0];
                best_sim = Sim{1.0f, 0.0f, m[s0].dx, m[s0].dy};
            }
            uint32_t bc = s_cc[0];
            float berr = s_ce[0];
            for (uint32_t t = 0; t < ns; t++) {
                const uint32_t si = (uint32_t)s_sorted[t];
                if (s_cc[t] > bc || (s_cc[t] == bc && s_ce[t] < berr)) {
                    bc = s_cc[t];
                    berr = s_ce[t];
                    best_sim = Sim{1.0f, 0.0f, m[si].dx, m[si].dy};
                }
            }
            build_cdf(true);
            uint32_t max_it = adaptive_iters((double)bc, (double)n, 2, 0.999, 64, 1000), it = 0;
            __shared__ uint32_t s_sc[kRT], s_sv[kRT];
            __shared__ float s_se[kRT], s_sm[kRT][4];
            while (it < max_it) {
                // hypotheses it+1 .. it+kRT, one per thread
                const uint32_t t = it + tid;   // 0-based iteration index
                s_sv[tid] = 0;
                if (t < max_it) {   // past the (shrinking) bound a hypothesis is never replayed
                    uint32_t i1 = cdf_pick(cdf, n, unit_draw(splitmix_at(P.sim_seed, 2ull * t)));
                    uint32_t i2 = cdf_pick(cdf, n, unit_draw(splitmix_at(P.sim_seed, 2ull * t + 1)));
                    if (i2 == i1) i2 = (i1 + 1) % n;
                    const ps_feature_match &m1 = m[i1], &m2 = m[i2];
                    float dxq = m2.qx - m1.qx, dyq = m2.qy - m1.qy, dq = dxq * dxq + dyq * dyq;
                    uint32_t valid = 0, c = 0;
                    float e = 0.0f, a = 0.0f, b = 0.0f, tx = 0.0f, ty = 0.0f;
                    if (!(dq < 16.0f)) {
                        float dxt = m2.tx - m1.tx, dyt = m2.ty - m1.ty;
                        a = (dxq * dxt + dyq * dyt) / dq;
                        b = (dxq * dyt - dyq * dxt) / dq;
                        float s2 = a * a + b * b;
                        if (s2 >= 0.04f && s2 <= 25.0f) {
                            tx = m1.tx - (a * m1.qx - b * m1.qy);
                            ty = m1.ty - (b * m1.qx + a * m1.qy);
                            c = eval_sim(m, n, a, b, tx, ty, ftsq, &e);
                            valid = 1;
                        }
                    }
                    s_sv[tid] = valid;
                    s_sc[tid] = c;
                    s_se[tid] = e;
                    s_sm[tid][0] = a;
                    s_sm[tid][1] = b;
                    s_sm[tid][2] = tx;
                    s_sm[tid][3] = ty;
                }
                __syncthreads();
                // replay in order (every thread, identically)
                for (int k = 0; k < kRT && it < max_it; k++) {
                    it++;
                    if (!s_sv[k]) continue;
                    if (s_sc[k] > bc || (s_sc[k] == bc && s_se[k] < berr)) {
                        bc = s_sc[k];
                        berr = s_se[k];
                        best_sim = Sim{s_sm[k][0], s_sm[k][1], s_sm[k][2], s_sm[k][3]};
                        max_it = adaptive_iters((double)bc, (double)n, 2, 0.999, 64, 1000);
                    }
                }
                __syncthreads();
            }
            // final inliers, refit, refine (pipeline.rs:2150-2191)
            uint32_t ni = ordered_compact(n, sim_inl, S, [&](uint32_t i) { return sim_inlier(m[i], best_sim, ftsq); });
            __shared__ float s_ref[4];
            __shared__ int s_has_ref;
            {
                Sim ref;
                const bool hr = fit_procrustes_block(m, sim_inl, ni, &ref, S);
                if (tid == 0) {
                    s_has_ref = hr;
                    s_ref[0] = ref.a;
                    s_ref[1] = ref.b;
                    s_ref[2] = ref.tx;
                    s_ref[3] = ref.ty;
                }
            }
            __syncthreads();
            const bool has_ref = s_has_ref;
            const Sim ref{s_ref[0], s_ref[1], s_ref[2], s_ref[3]};
            if (has_ref) {
                const uint32_t nr = ordered_compact(n, ibuf, S, [&](uint32_t i) { return sim_inlier(m[i], ref, ftsq); });
                if (nr >= ni) {
                    for (uint32_t i = tid; i < nr; i += kRT) sim_inl[i] = ibuf[i];
                    ni = nr;
                }
                __syncthreads();
            }
            {
                Sim fin;
                const bool hf = fit_procrustes_block(m, sim_inl, ni, &fin, S);
                if (tid == 0) {
                    if (hf) {
                        s_ref[0] = fin.a;
                        s_ref[1] = fin.b;
                        s_ref[2] = fin.tx;
                        s_ref[3] = fin.ty;
                    }
                    s_has_ref = hf || has_ref;
                }
            }
            __syncthreads();
            have_sim = s_has_ref;
            best_sim = Sim{s_ref[0], s_ref[1], s_ref[2], s_ref[3]};
            n_sim = ni;
            __syncthreads();
        }
        if (tid == 0) {
            O.has_sim = have_sim;
            O.n_sim = n_sim;
            O.sim[0] = best_sim.a;
            O.sim[1] = best_sim.b;
            O.sim[2] = best_sim.tx;
            O.sim[3] = best_sim.ty;
        }
    }
    if (prm.sim_only) {   // the similarity result alone: its inliers go to the output list
        for (uint32_t i = tid; i < n_sim; i += kRT) out_inl[i] = sim_inl[i];
        if (tid == 0) O.n_inliers = n_sim;
        return;
    }

    // ---------------- homography LO-RANSAC (pipeline.rs:2663-2755)
    const double thresh = (double)fth, tsq = thresh * thresh;
    bool have = false;
    uint32_t bc = 0;
    double be = INFINITY;
    if (prm.sim_seed_model && have_sim) {
        if (tid < 9) {
            const double Hs[9] = {(double)best_sim.a, -(double)best_sim.b, (double)best_sim.tx,
                                  (double)best_sim.b, (double)best_sim.a, (double)best_sim.ty, 0.0, 0.0, 1.0};
            s_H[tid] = Hs[tid];
        }
        __syncthreads();
        have = true;
        bc = eval_h_block(m, n, s_H, tsq, nullptr, &be, nullptr, S);
    }
    if (n < 4) {
        if (prm.sim_seed_model) {
            for (uint32_t i = tid; i < n_sim; i += kRT) out_inl[i] = sim_inl[i];
            if (tid == 0) {
                O.n_inliers = n_sim;
                if (have) {
                    for (int i = 0; i < 9; i++) O.h[i] = (float)s_H[i];
                    O.has_model = 1;
                }
            }
        }
        return;
    }
    int sampler = prm.sampler;
    double wsum = 0.0;
    if (sampler == 2) {
        // the caller's weights need usable mass: every weight finite and non-negative and a
        // finite positive total; otherwise the problem falls back to uniform sampling (the
        // reference's 1 / ratio sampler always has positive mass, pipeline.rs:2006-2028)
        double acc = 0.0;   // thread 0, in order
        if (tid == 0) S.total = 0;
        __syncthreads();
        for (uint32_t c0 = 0; c0 < n; c0 += kStage) {
            const uint32_t cn = min(kStage, n - c0);
            for (uint32_t k = tid; k < cn; k += kRT) {
                const double wv = (double)wts[c0 + k];
                if (!(wv >= 0.0) || !isfinite(wv)) S.total = 1u;
                S.stage.d[0][k] = wv;
            }
            __syncthreads();
            if (tid == 0) {
#pragma unroll 8
                for (uint32_t k = 0; k < cn; k++) acc += S.stage.d[0][k];
                S.red[0] = acc;
            }
            __syncthreads();
        }
        wsum = S.red[0];
        const bool usable = S.total == 0u && isfinite(wsum) && wsum > 0.0;
        __syncthreads();
        if (!usable) {
            sampler = 1;
            wsum = 0.0;
        }
    }
    if (sampler == 0) {
        if (!prm.sim_seed_model || n == 1) build_cdf(true);   // otherwise the similarity stage built it
    } else if (sampler == 2) {
        build_cdf(false);
    }
    const uint32_t min_it = prm.min_iters, cap = prm.max_iters;
    uint32_t max_it;
    auto bound_from = [&](uint32_t cnt, double bmass) -> uint32_t {
        if (!prm.adaptive) return cap;
        if (sampler == 2)
            return wsum > 0 ? adaptive_iters((double)(uint64_t)(bmass / wsum * 1e9), 1e9, 4, prm.confidence, min_it, cap)
                            : cap;
        return adaptive_iters((double)cnt, (double)n, 4, prm.confidence, min_it, cap);
    };
    {
        double bmass = 0.0;
        if (prm.adaptive && sampler == 2 && have) {
            double e2;
            eval_h_block(m, n, s_H, tsq, wts, &e2, &bmass, S);
        }
        if (!prm.adaptive) max_it = cap;
        else if (sampler == 2) max_it = have && wsum > 0 ? bound_from(bc, bmass) : cap;
        else max_it = bound_from(bc, 0.0);
    }
    uint32_t it = 0;
    // the homography SplitMix64 stream: the buffer holds draws [base, base + kDraws + 16) and the
    // next sample starts at stream position pos (both uniform)
    uint64_t base = 0, pos = 0;
    bool filled = false;
    while (it < max_it) {
        // 1. the next kRT samples (4 distinct matches within 16 draws each), the sequential
        // grouping done in parallel: every buffer position gets the length of a sample starting
        // there, thread 0 follows the chain of sample starts from pos, and each thread rebuilds
        // one sample; a chain that reaches the end of the buffer refills it from there
        uint32_t nq = 0;
        while (nq < (uint32_t)kRT) {
            if (!filled || pos >= base + kDraws) {
                base = pos;
                filled = true;
                for (uint32_t k = tid; k < kDraws + 16; k += kRT) {
                    const double u = unit_draw(splitmix_at(P.hom_seed, base + k));
                    S.draws[k] = sampler == 1 ? uniform_pick(n, u) : cdf_pick(cdf, n, u);
                }
                __syncthreads();
                for (uint32_t k = tid; k < kDraws; k += kRT) {
                    uint32_t qq[4];
                    bool v;
                    S.len[k] = (uint8_t)scan_sample<4, 16>(S.draws + k, qq, v);
                }
                __syncthreads();
            }
            if (tid == 0) {
                uint64_t p = pos;
                uint32_t k = nq;
                while (p < base + kDraws && k < (uint32_t)kRT) {
                    S.start[k++] = (uint32_t)(p - base);
                    p += S.len[p - base];
                }
                S.n_quads = k;
                S.total = (uint32_t)(p - base);
            }
            __syncthreads();
            const uint32_t nq2 = S.n_quads;
            if ((uint32_t)tid >= nq && (uint32_t)tid < nq2) {
                uint32_t qq[4];
                bool v;
                scan_sample<4, 16>(S.draws + S.start[tid], qq, v);
                S.qvalid[tid] = v;
#pragma unroll
                for (int j = 0; j < 4; j++) S.quad[tid][j] = qq[j];
            }
            pos = base + S.total;
            nq = nq2;
            __syncthreads();
        }
        // 2. evaluate the kRT hypotheses, one per thread
        {
            Hyp &h = S.hyp[tid];
            h.valid = 0;
            h.count = 0;
            h.err = 0.0;
            h.mass = 0.0;
            // the bound only shrinks as the best improves (not so for the weighted sampler's
            // mass-based bound), so a hypothesis past it is never replayed: skip its fit
            if (S.qvalid[tid] && (sampler == 2 || it + (uint32_t)tid < max_it)) {
                const uint32_t idx[4] = {S.quad[tid][0], S.quad[tid][1], S.quad[tid][2], S.quad[tid][3]};
                double qx[4], qy[4], tx[4], ty[4];
                for (int i = 0; i < 4; i++) {
                    qx[i] = m[idx[i]].qx;
                    qy[i] = m[idx[i]].qy;
                    tx[i] = m[idx[i]].tx;
                    ty[i] = m[idx[i]].ty;
                }
                if (!degenerate_tri(qx, qy) && !degenerate_tri(tx, ty)) {
                    double Hc[3][3];
                    if (fit_dlt4(m, idx, Hc)) {
                        for (int i = 0; i < 9; i++) h.H[i] = Hc[i / 3][i % 3];
                        double e, ms = 0.0;
                        h.count = eval_h_seq(m, n, h.H, tsq, &e, wts, &ms);
                        h.err = e;
                        h.mass = ms;
                        h.valid = 1;
                    }
                }
            }
        }
        __syncthreads();
        // 3. replay in order. Only a hypothesis better than every earlier valid one of the batch
        // and than the best at the batch start can improve the best (the best never gets worse),
        // so a parallel prefix-best scan picks those; the others only advance the count, which
        // is settled at the end exactly as the sequential loop would leave it.
        {
            const Hyp &mine = S.hyp[tid];
            const bool mv = mine.valid;
            uint32_t pc = mv ? (mine.count | 0x80000000u) : 0u;   // prefix best, inclusive
            double pe = mv ? mine.err : 0.0;
            S.scan_c[tid] = pc;
            S.red[tid] = pe;
            __syncthreads();
            for (int o = 1; o < kRT; o <<= 1) {
                uint32_t oc = 0;
                double oe = 0.0;
                if (tid >= o) {
                    oc = S.scan_c[tid - o];
                    oe = S.red[tid - o];
                }
                __syncthreads();
                // the better of (oc, oe) and (pc, pe); an invalid one (bit 31 clear) never wins
                if ((oc >> 31) && (!(pc >> 31) || (oc & 0x7fffffffu) > (pc & 0x7fffffffu) ||
                                   ((oc & 0x7fffffffu) == (pc & 0x7fffffffu) && oe < pe))) {
                    pc = oc;
                    pe = oe;
                }
                S.scan_c[tid] = pc;
                S.red[tid] = pe;
                __syncthreads();
            }
            bool cnd = mv && (mine.count > bc || (mine.count == bc && mine.err < be));
            if (cnd && tid > 0) {
                const uint32_t qc = S.scan_c[tid - 1];
                const double qe = S.red[tid - 1];
                if (qc >> 31) cnd = mine.count > (qc & 0x7fffffffu) || (mine.count == (qc & 0x7fffffffu) && mine.err < qe);
            }
            if (tid == 0) S.n_cand = 0;
            __syncthreads();
            const uint32_t nc = ordered_compact((uint32_t)kRT, S.cand, S, [&](uint32_t) { return cnd; });
            const uint32_t it0 = it;   // the sequential loop visits k = 0, 1, ... while it0 + k < max_it
            int k_last = -1;
            for (uint32_t ci = 0; ci < nc; ci++) {
                const uint32_t k = S.cand[ci];
                if (it0 + k >= max_it) break;
                k_last = (int)k;
                const Hyp &h = S.hyp[k];
                uint32_t c = h.count;
                double e = h.err;
                if (c > bc || (c == bc && e < be)) {
                    __syncthreads();
                    if (tid < 9) s_H[tid] = h.H[tid];
                    __syncthreads();
                    if (prm.local_opt) local_opt_block(m, n, s_H, &c, &e, tsq, ibuf, S);
                    bc = c;
                    be = e;
                    have = true;
                    if (prm.adaptive) {
                        if (sampler == 2) {
                            double e2, bmass = 0.0;
                            eval_h_block(m, n, s_H, tsq, wts, &e2, &bmass, S);
                            max_it = wsum > 0 ? bound_from(bc, bmass) : cap;
                        } else {
                            max_it = bound_from(bc, 0.0);
                        }
                    }
                    __syncthreads();
                }
            }
            // the iterations it counts: up to the batch end or the (final) bound, at least past
            // the last visited hypothesis
            int64_t jend = (int64_t)max_it - (int64_t)it0;
            if (jend < (int64_t)(k_last + 1)) jend = k_last + 1;
            if (jend > kRT) jend = kRT;
            it = it0 + (uint32_t)jend;
        }
        __syncthreads();
    }
    uint32_t ni;
    if (!have) {
        ni = n_sim;
        for (uint32_t i = tid; i < ni; i += kRT) out_inl[i] = sim_inl[i];
    } else {
        if (prm.local_opt) local_opt_block(m, n, s_H, &bc, &be, tsq, ibuf, S);
        ni = ordered_compact(n, out_inl, S, [&](uint32_t i) {
            double e = reproj_err(m[i], s_H);
            return e >= 0.0 && e <= tsq;
        });
    }
    if (tid == 0) {
        O.hyps = it;
        O.n_inliers = ni;
        if (have) {
            for (int i = 0; i < 9; i++) O.h[i] = (float)s_H[i];
            O.has_model = 1;
        }
    }
    (void)s_flag;
}

}  // namespace

size_t ransac_work_bytes(size_t total_matches) {
    return total_matches * 2 * sizeof(double) + total_matches * 2 * sizeof(uint32_t) + 512;
}


hipError_t launch_ransac(const ps_feature_match *matches, const float *weights, const ps_ransac_problem *probs,
                         uint32_t n_probs, const ps_ransac_params &p, size_t total_matches, void *work,
                         uint32_t *inliers, ps_ransac_result *outs, hipStream_t st, bool sim_only) {
    if (n_probs == 0) return hipSuccess;
    RansacParamsDev prm;
    prm.reproj = p.reproj_threshold;
    prm.sampler = p.sampler;
    prm.confidence = p.confidence;
    prm.min_iters = p.min_iters;
    prm.max_iters = p.max_iters;
    prm.local_opt = p.local_opt;
    prm.sim_seed_model = p.sim_seed_model;
    prm.adaptive = p.adaptive;
    prm.sim_only = sim_only ? 1 : 0;
    double *wd = static_cast<double *>(work);
    uint32_t *wu = reinterpret_cast<uint32_t *>(wd + total_matches * 2);
    ransac_kernel<<<n_probs, kRT, 0, st>>>(matches, weights, probs, prm, wd, wu, inliers, outs);
    return hipGetLastError();
}

}  // namespace ps
