/*
 * ps_oracle_catalog.c -- CPU oracle (test infrastructure) for the feature catalog.
 *
 * A plain-C restatement of pure-sift 0.2.0 src/catalog.rs and the parts of
 * src/pipeline.rs it calls: the PSIFTCAT binary format (catalog.rs:200-540), the
 * u8 descriptor quantisation pair (pipeline.rs:1024-1044), the coarse vote pass
 * with its inverted-index and exhaustive paths (catalog.rs:595-633, :707-728),
 * summarize_tracking (pipeline.rs:2194-2272) and search_features
 * (catalog.rs:576-702). Never linked into the product library.
 *
 * How each part is pinned (tests/test_oracle_catalog.py):
 *  - format: the byte layout by closed form (33-byte header, 20 + id bytes per
 *    entry, 164 bytes per feature), field-by-field decoding of a file assembled
 *    in Python with struct.pack, a write/parse round trip, every error branch.
 *  - u8 quantisation: numpy closed forms (round half away from zero, clamp,
 *    sequential f32 norm); dequantised rows have unit norm.
 *  - votes: the inverted-index path equals a brute-force "some catalog code
 *    within 24 bits that also shares one byte in the same position" predicate and
 *    the exhaustive path the "within 24 bits" predicate (numpy), with planted
 *    codes at 24 / 25 bits and at 16 bits with no shared byte (voted only by the
 *    exhaustive path).
 *  - tracking: closed forms on pure translations, rotations and scalings, the
 *    even/odd median rule, the empty-match defaults.
 *  - search: the reference's behavioural tests (tests/pipeline_tests.rs:839-877,
 *    1384-1436, 1438-1483; tests/coverage_gaps.rs:29-58) on synthetic scenes, the
 *    shortlist order (votes desc, id asc) and hit order rules on crafted catalogs.
 */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ps_oracle.h"

#define COARSE_HAMMING_MAX_BITS 24u       /* catalog.rs:32 */
#define MIH_MIN_ENTRIES 32u               /* catalog.rs:596: inverted index above 32 entries */
#define DEFAULT_RANSAC_REPROJ 3.0f        /* pipeline.rs:1050 */
#define DEFAULT_SIM_SEED 0xD1B54A32D192ED03ull   /* pipeline.rs:173 */
#define F32_DEG_PER_RAD 57.2957795130823208767981548141051703f   /* Rust f32::to_degrees */

/* ------------------------------------------------------------------ u8 descriptors */

/* pipeline.rs:1024-1030: (v * 512).round().clamp(0, 255) as u8 (NaN -> 0) */
void psr_quantize_u8(const float *desc, size_t n, uint8_t *out) {
    for (size_t i = 0; i < n * PSR_DESC_LEN; i++) {
        float r = roundf(desc[i] * 512.0f);
        if (r < 0.0f) r = 0.0f;
        if (r > 255.0f) r = 255.0f;
        out[i] = (r == r) ? (uint8_t)r : 0;
    }
}

/* pipeline.rs:1034-1044: b / 512, renormalised when the sequential f32 norm > 1e-7 */
void psr_dequantize_u8(const uint8_t *q, size_t n, float *out) {
    for (size_t r = 0; r < n; r++) {
        const uint8_t *s = q + r * PSR_DESC_LEN;
        float *o = out + r * PSR_DESC_LEN;
        float acc = 0.0f;
        for (int i = 0; i < PSR_DESC_LEN; i++) {
            o[i] = (float)s[i] * (1.0f / 512.0f);
            acc += o[i] * o[i];
        }
        float norm = sqrtf(acc);
        if (norm > 1e-7f) {
            float inv = 1.0f / norm;
            for (int i = 0; i < PSR_DESC_LEN; i++) o[i] *= inv;
        }
    }
}

/* ------------------------------------------------------------------ PSIFTCAT */

typedef struct {
    const uint8_t *p;
    size_t left;
} rd;

static int take(rd *r, void *dst, size_t n) {
    if (r->left < n) return -1;
    memcpy(dst, r->p, n);
    r->p += n;
    r->left -= n;
    return 0;
}

/* std::str::from_utf8 acceptance (RFC 3629: no overlongs, no surrogates, <= U+10FFFF) */
static int utf8_valid(const uint8_t *s, size_t n) {
    size_t i = 0;
    while (i < n) {
        uint8_t c = s[i];
        if (c < 0x80) { i++; continue; }
        size_t need;
        uint8_t lo = 0x80, hi = 0xBF;
        if (c >= 0xC2 && c <= 0xDF) need = 1;
        else if (c == 0xE0) { need = 2; lo = 0xA0; }
        else if (c >= 0xE1 && c <= 0xEC) need = 2;
        else if (c == 0xED) { need = 2; hi = 0x9F; }
        else if (c >= 0xEE && c <= 0xEF) need = 2;
        else if (c == 0xF0) { need = 3; lo = 0x90; }
        else if (c >= 0xF1 && c <= 0xF3) need = 3;
        else if (c == 0xF4) { need = 3; hi = 0x8F; }
        else return 0;
        if (i + need >= n) return 0;
        if (s[i + 1] < lo || s[i + 1] > hi) return 0;
        for (size_t k = 2; k <= need; k++)
            if (s[i + k] < 0x80 || s[i + k] > 0xBF) return 0;
        i += need + 1;
    }
    return 1;
}

static int fail(char *err, size_t cap, const char *msg) {
    if (err && cap) snprintf(err, cap, "%s", msg);
    return -1;
}

void psr_catalog_free(psr_catalog *c) {
    if (!c) return;
    free((void *)c->id_blob);
    free((void *)c->id_off);
    free((void *)c->width);
    free((void *)c->height);
    free((void *)c->off);
    free((void *)c->feat);
    free((void *)c->sig);
    free((void *)c->desc);
    memset(c, 0, sizeof *c);
}

/* catalog.rs:447-540 load_binary over an in-memory file. Descriptors are dequantised
 * (pipeline.rs:1034-1044), signatures kept as stored. Trailing bytes are ignored, as there. */
int psr_catalog_parse(const uint8_t *buf, size_t len, psr_catalog *out, char *err, size_t cap) {
    memset(out, 0, sizeof *out);
    rd r = {buf, len};
    char magic[8];
    if (take(&r, magic, 8)) return fail(err, cap, "Unexpected EOF while reading binary catalog");
    if (memcmp(magic, "PSIFTCAT", 8) != 0) return fail(err, cap, "Invalid binary catalog magic header");
    uint32_t version, octaves, scales, kind32;
    float thr;
    uint8_t kind;
    uint64_t n_entries;
    if (take(&r, &version, 4)) return fail(err, cap, "Unexpected EOF while reading binary catalog");
    if (version != 1) {
        char m[96];
        snprintf(m, sizeof m, "Unsupported binary catalog version %u (expected 1)", version);
        return fail(err, cap, m);
    }
    if (take(&r, &octaves, 4) || take(&r, &scales, 4) || take(&r, &thr, 4) || take(&r, &kind, 1))
        return fail(err, cap, "Unexpected EOF while reading binary catalog");
    if (kind > 1) {
        char m[96];
        snprintf(m, sizeof m, "Unknown descriptor kind %u in binary catalog", (unsigned)kind);
        return fail(err, cap, m);
    }
    kind32 = kind;
    if (take(&r, &n_entries, 8)) return fail(err, cap, "Unexpected EOF while reading binary catalog");

    /* first pass: sizes (every entry needs >= 20 bytes, which bounds n_entries by len) */
    rd s = r;
    size_t nf = 0, ids = 0;
    for (uint64_t e = 0; e < n_entries; e++) {
        uint32_t idl, w, h;
        uint64_t fc;
        if (take(&s, &idl, 4)) return fail(err, cap, "Unexpected EOF while reading binary catalog");
        if (s.left < idl) return fail(err, cap, "Truncated entry id in binary catalog");
        if (!utf8_valid(s.p, idl)) return fail(err, cap, "Invalid UTF-8 entry id in binary catalog");
        s.p += idl;
        s.left -= idl;
        ids += idl;
        if (take(&s, &w, 4) || take(&s, &h, 4) || take(&s, &fc, 8))
            return fail(err, cap, "Unexpected EOF while reading binary catalog");
        if (fc > s.left / 164) return fail(err, cap, "Unexpected EOF while reading binary catalog");
        s.p += fc * 164;
        s.left -= fc * 164;
        nf += fc;
    }
    char *blob = (char *)malloc(ids ? ids : 1);
    uint64_t *id_off = (uint64_t *)malloc(sizeof(uint64_t) * (n_entries + 1));
    uint32_t *wid = (uint32_t *)malloc(sizeof(uint32_t) * (n_entries ? n_entries : 1));
    uint32_t *hei = (uint32_t *)malloc(sizeof(uint32_t) * (n_entries ? n_entries : 1));
    uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * (n_entries + 1));
    psr_feature *feat = (psr_feature *)malloc(sizeof(psr_feature) * (nf ? nf : 1));
    uint64_t *sig = (uint64_t *)malloc(16 * (nf ? nf : 1));
    float *desc = (float *)malloc(sizeof(float) * PSR_DESC_LEN * (nf ? nf : 1));
    size_t f = 0, b = 0;
    id_off[0] = 0;
    off[0] = 0;
    for (uint64_t e = 0; e < n_entries; e++) {   /* sizes checked by the first pass */
        uint32_t idl = 0;
        uint64_t fc = 0;
        take(&r, &idl, 4);
        take(&r, blob + b, idl);
        b += idl;
        id_off[e + 1] = b;
        take(&r, &wid[e], 4);
        take(&r, &hei[e], 4);
        take(&r, &fc, 8);
        for (uint64_t k = 0; k < fc; k++, f++) {
            float v[5];
            uint8_t q[PSR_DESC_LEN];
            take(&r, v, 20);
            take(&r, &sig[2 * f], 16);
            take(&r, q, PSR_DESC_LEN);
            feat[f].x = v[0];
            feat[f].y = v[1];
            feat[f].scale = v[2];
            feat[f].response = v[3];
            feat[f].orientation = v[4];
            psr_dequantize_u8(q, 1, desc + f * PSR_DESC_LEN);
        }
        off[e + 1] = f;
    }
    out->octaves = octaves;
    out->scales = scales;
    out->threshold = thr;
    out->kind = kind32;
    out->n_entries = n_entries;
    out->n_features = nf;
    out->id_blob = blob;
    out->id_off = id_off;
    out->width = wid;
    out->height = hei;
    out->off = off;
    out->feat = feat;
    out->sig = sig;
    out->desc = desc;
    return 0;
}

/* catalog.rs:400-445 save_binary into buf (cap bytes); returns the size, or the needed
 * size when cap is smaller (nothing written then). Descriptors are quantised to u8
 * (pipeline.rs:1024-1030) and the stored signatures written as they are. */
size_t psr_catalog_write(const psr_catalog *c, uint8_t *buf, size_t cap) {
    size_t need = 8 + 4 + 4 + 4 + 4 + 1 + 8;
    for (size_t e = 0; e < c->n_entries; e++)
        need += 4 + (c->id_off[e + 1] - c->id_off[e]) + 4 + 4 + 8 + 164 * (c->off[e + 1] - c->off[e]);
    if (cap < need) return need;
    uint8_t *p = buf;
#define PUT(src, n) do { memcpy(p, (src), (n)); p += (n); } while (0)
    const uint32_t version = 1;
    const uint8_t kind = (uint8_t)c->kind;
    const uint64_t ne = c->n_entries;
    PUT("PSIFTCAT", 8);
    PUT(&version, 4);
    PUT(&c->octaves, 4);
    PUT(&c->scales, 4);
    PUT(&c->threshold, 4);
    PUT(&kind, 1);
    PUT(&ne, 8);
    for (size_t e = 0; e < c->n_entries; e++) {
        const uint32_t idl = (uint32_t)(c->id_off[e + 1] - c->id_off[e]);
        const uint64_t fc = c->off[e + 1] - c->off[e];
        PUT(&idl, 4);
        PUT(c->id_blob + c->id_off[e], idl);
        PUT(&c->width[e], 4);
        PUT(&c->height[e], 4);
        PUT(&fc, 8);
        for (uint64_t k = c->off[e]; k < c->off[e + 1]; k++) {
            uint8_t q[PSR_DESC_LEN];
            PUT(&c->feat[k], 20);
            PUT(&c->sig[2 * k], 16);
            psr_quantize_u8(c->desc + k * PSR_DESC_LEN, 1, q);
            PUT(q, PSR_DESC_LEN);
        }
    }
#undef PUT
    return need;
}

/* ------------------------------------------------------------------ coarse votes */

static inline uint32_t ham128(const uint64_t *a, const uint64_t *b) {
    return (uint32_t)(__builtin_popcountll(a[0] ^ b[0]) + __builtin_popcountll(a[1] ^ b[1]));
}

/* catalog.rs:280-297 ensure_mih_allocated / index_entry_signatures: 16 inverted tables, one
 * per signature byte, each with 256 buckets of (entry, signature) pairs in indexing order
 * (entry by entry, signatures in order), stored as one CSR array per table. */
struct psr_mih {
    size_t stride;              /* pairs per table (n_features) */
    size_t start[16 * 257];     /* bucket k of table t: [start[t*257+k], start[t*257+k+1]) */
    uint32_t *pair_e, *pair_s;  /* 16 * stride each */
};

void *psr_catalog_index(const psr_catalog *c) {
    const size_t nf = c->n_features, stride = nf ? nf : 1;
    struct psr_mih *m = (struct psr_mih *)calloc(1, sizeof *m);
    m->stride = stride;
    m->pair_e = (uint32_t *)malloc(sizeof(uint32_t) * 16 * stride);
    m->pair_s = (uint32_t *)malloc(sizeof(uint32_t) * 16 * stride);
    for (size_t f = 0; f < nf; f++)
        for (int band = 0; band < 16; band++) {
            unsigned key = (unsigned)((c->sig[2 * f + band / 8] >> (8 * (band % 8))) & 0xFF);
            m->start[band * 257 + key + 1]++;
        }
    for (int band = 0; band < 16; band++)
        for (int k = 0; k < 256; k++) m->start[band * 257 + k + 1] += m->start[band * 257 + k];
    size_t *fill = (size_t *)malloc(sizeof(size_t) * 16 * 256);
    for (int band = 0; band < 16; band++)
        for (int k = 0; k < 256; k++) fill[band * 256 + k] = m->start[band * 257 + k];
    for (size_t e = 0; e < c->n_entries; e++)
        for (uint64_t f = c->off[e]; f < c->off[e + 1]; f++)
            for (int band = 0; band < 16; band++) {
                unsigned key = (unsigned)((c->sig[2 * f + band / 8] >> (8 * (band % 8))) & 0xFF);
                size_t slot = band * stride + fill[band * 256 + key]++;
                m->pair_e[slot] = (uint32_t)e;
                m->pair_s[slot] = (uint32_t)(f - c->off[e]);
            }
    free(fill);
    return m;
}

void psr_catalog_index_free(void *p) {
    struct psr_mih *m = (struct psr_mih *)p;
    if (!m) return;
    free(m->pair_e);
    free(m->pair_s);
    free(m);
}

/* catalog.rs:597-620: a query visits the bucket of each of its bytes and gives an entry at
 * most one vote (the seen stamp). Uses c->mih when set, else indexes on the spot. */
static void votes_mih(const psr_catalog *c, const uint64_t *q, size_t nq, uint32_t *votes) {
    const struct psr_mih *m = (const struct psr_mih *)c->mih;
    void *own = NULL;
    if (!m) m = (const struct psr_mih *)(own = psr_catalog_index(c));
    size_t *seen = (size_t *)calloc(c->n_entries ? c->n_entries : 1, sizeof(size_t));
    memset(votes, 0, sizeof(uint32_t) * c->n_entries);
    for (size_t qi = 0; qi < nq; qi++) {
        const size_t stamp = qi + 1;
        const uint64_t *qs = q + 2 * qi;
        for (int band = 0; band < 16; band++) {
            unsigned key = (unsigned)((qs[band / 8] >> (8 * (band % 8))) & 0xFF);
            for (size_t k = m->start[band * 257 + key]; k < m->start[band * 257 + key + 1]; k++) {
                size_t slot = band * m->stride + k;
                uint32_t e = m->pair_e[slot];
                if (seen[e] == stamp) continue;
                const uint64_t *es = c->sig + 2 * (c->off[e] + m->pair_s[slot]);
                if (ham128(qs, es) <= COARSE_HAMMING_MAX_BITS) {
                    seen[e] = stamp;
                    votes[e]++;
                }
            }
        }
    }
    free(seen);
    psr_catalog_index_free(own);
}

/* catalog.rs:595-633: mih != 0 -> the inverted-index path (taken by the reference for
 * more than 32 entries), else count_coarse_hamming_votes per entry (:707-728). */
void psr_catalog_votes(const psr_catalog *c, const uint64_t *q, size_t nq, int mih, uint32_t *votes) {
    if (mih) {
        votes_mih(c, q, nq, votes);
        return;
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t e = 0; e < c->n_entries; e++)
        votes[e] = (uint32_t)psr_count_coarse_hamming_votes(q, nq, c->sig + 2 * c->off[e],
                                                            (size_t)(c->off[e + 1] - c->off[e]));
}

static int id_cmp(const psr_catalog *c, uint32_t a, uint32_t b) {
    size_t la = c->id_off[a + 1] - c->id_off[a], lb = c->id_off[b + 1] - c->id_off[b];
    int r = memcmp(c->id_blob + c->id_off[a], c->id_blob + c->id_off[b], la < lb ? la : lb);
    if (r) return r;
    return la < lb ? -1 : (la > lb ? 1 : 0);
}

/* stable merge sort of n items of `size` bytes */
typedef int (*cmp_fn)(const void *a, const void *b, const void *ctx);
static void stable_sort(void *base, size_t n, size_t size, cmp_fn cmp, const void *ctx) {
    if (n < 2) return;
    uint8_t *a = (uint8_t *)base, *tmp = (uint8_t *)malloc(n * size);
    for (size_t w = 1; w < n; w *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * w) {
            size_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            size_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                if (cmp(a + j * size, a + i * size, ctx) < 0) memcpy(tmp + size * k++, a + size * j++, size);
                else memcpy(tmp + size * k++, a + size * i++, size);
            }
            while (i < mid) memcpy(tmp + size * k++, a + size * i++, size);
            while (j < hi) memcpy(tmp + size * k++, a + size * j++, size);
        }
        memcpy(a, tmp, n * size);
    }
    free(tmp);
}

typedef struct { uint32_t entry, votes; } cand_t;

static int cand_cmp(const void *pa, const void *pb, const void *ctx) {
    const cand_t *a = (const cand_t *)pa, *b = (const cand_t *)pb;
    if (a->votes != b->votes) return a->votes > b->votes ? -1 : 1;
    return id_cmp((const psr_catalog *)ctx, a->entry, b->entry);
}

/* catalog.rs:594-638: the candidate list from the votes (entries with 0 votes dropped on
 * the inverted-index path), sorted by (votes desc, id asc) and cut to max(shortlist_k, 1)
 * when longer. Returns the count; cand_entry / cand_votes have capacity n_entries. */
size_t psr_catalog_shortlist(const psr_catalog *c, const uint32_t *votes, int mih, size_t shortlist_k,
                             uint32_t *cand_entry, uint32_t *cand_votes) {
    cand_t *cs = (cand_t *)malloc(sizeof(cand_t) * (c->n_entries ? c->n_entries : 1));
    size_t n = 0;
    for (size_t e = 0; e < c->n_entries; e++)
        if (!mih || votes[e] > 0) cs[n++] = (cand_t){(uint32_t)e, votes[e]};
    size_t keep = shortlist_k > 1 ? shortlist_k : 1;
    if (n > keep) {
        stable_sort(cs, n, sizeof(cand_t), cand_cmp, c);
        n = keep;
    }
    for (size_t i = 0; i < n; i++) {
        cand_entry[i] = cs[i].entry;
        cand_votes[i] = cs[i].votes;
    }
    free(cs);
    return n;
}

/* ------------------------------------------------------------------ tracking summary */

static int f32_cmp(const void *pa, const void *pb, const void *ctx) {
    float a = *(const float *)pa, b = *(const float *)pb;
    return a < b ? -1 : (a > b ? 1 : 0);   /* partial_cmp, unordered -> Equal */
}

/* pipeline.rs:2194-2206 (stable sort, as slice::sort_by) */
static float median_f32(float *v, size_t n) {
    if (n == 0) return 0.0f;
    stable_sort(v, n, sizeof(float), f32_cmp, NULL);
    size_t mid = n / 2;
    return (n % 2 == 1) ? v[mid] : 0.5f * (v[mid - 1] + v[mid]);
}

/* pipeline.rs:2210-2272 */
void psr_summarize_tracking(const psr_feature *q, const psr_feature *t, const psr_match *m, size_t n,
                            psr_tracking *out) {
    memset(out, 0, sizeof *out);
    out->mean_scale_ratio = 1.0f;
    if (n == 0) return;
    uint32_t *inl = (uint32_t *)malloc(sizeof(uint32_t) * n);
    psr_similarity model;
    int has = 0;
    size_t ni = psr_filter_matches_ransac(m, n, DEFAULT_RANSAC_REPROJ, DEFAULT_SIM_SEED, inl, &model, &has);
    size_t na = ni ? ni : n;
    float inv_n = 1.0f / (float)na;
    float sdx = 0.0f, sdy = 0.0f, sscale = 0.0f, ssin = 0.0f, scos = 0.0f;
    float *dxs = (float *)malloc(sizeof(float) * na), *dys = (float *)malloc(sizeof(float) * na);
    for (size_t k = 0; k < na; k++) {
        const psr_match *mm = &m[ni ? inl[k] : k];
        sdx += mm->dx;
        sdy += mm->dy;
        dxs[k] = mm->dx;
        dys[k] = mm->dy;
        const psr_feature *qq = &q[mm->query_index], *tt = &t[mm->target_index];
        if (qq->scale > 1e-6f) sscale += tt->scale / qq->scale;
        float d = tt->orientation - qq->orientation;
        ssin += sinf(d);
        scos += cosf(d);
    }
    out->raw_match_count = n;
    out->match_count = na;
    out->inlier_ratio = (float)na / (float)n;
    out->mean_dx = sdx * inv_n;
    out->mean_dy = sdy * inv_n;
    out->median_dx = median_f32(dxs, na);
    out->median_dy = median_f32(dys, na);
    out->mean_scale_ratio = sscale * inv_n;
    out->mean_rotation_deg = atan2f(ssin, scos) * F32_DEG_PER_RAD;
    free(dys);
    free(dxs);
    free(inl);
}

/* ------------------------------------------------------------------ search */

typedef struct {
    const psr_catalog *c;
} hit_ctx;

static int hit_cmp(const void *pa, const void *pb, const void *ctx) {
    const psr_catalog_hit *a = (const psr_catalog_hit *)pa, *b = (const psr_catalog_hit *)pb;
    if (a->match_count != b->match_count) return a->match_count > b->match_count ? -1 : 1;
    if (a->inlier_ratio > b->inlier_ratio) return -1;   /* partial_cmp desc, unordered -> Equal */
    if (a->inlier_ratio < b->inlier_ratio) return 1;
    if (a->coarse_votes != b->coarse_votes) return a->coarse_votes > b->coarse_votes ? -1 : 1;
    return id_cmp(((const hit_ctx *)ctx)->c, a->entry, b->entry);
}

/* catalog.rs:576-702 search_features. out needs capacity min(top_k, n_entries). */
size_t psr_catalog_search(const psr_catalog *c, const psr_feature *qf, const float *qd, size_t nq,
                          size_t top_k, size_t shortlist_k, float ratio, float ransac_thresh,
                          psr_catalog_hit *out) {
    if (c->n_entries == 0 || nq == 0 || top_k == 0) return 0;
    uint64_t *qs = (uint64_t *)malloc(16 * nq);
    psr_quantize_u128(qd, nq, qs);
    const int mih = c->n_entries > MIH_MIN_ENTRIES;
    uint32_t *votes = (uint32_t *)malloc(sizeof(uint32_t) * c->n_entries);
    psr_catalog_votes(c, qs, nq, mih, votes);
    uint32_t *ce = (uint32_t *)malloc(sizeof(uint32_t) * c->n_entries);
    uint32_t *cv = (uint32_t *)malloc(sizeof(uint32_t) * c->n_entries);
    size_t nc = psr_catalog_shortlist(c, votes, mih, shortlist_k, ce, cv);
    psr_catalog_hit *hits = (psr_catalog_hit *)calloc(nc ? nc : 1, sizeof(psr_catalog_hit));
    int *keep = (int *)calloc(nc ? nc : 1, sizeof(int));
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t i = 0; i < nc; i++) {
        const uint32_t e = ce[i];
        const size_t nt = (size_t)(c->off[e + 1] - c->off[e]);
        if (nt < 2) continue;
        const psr_feature *tf = c->feat + c->off[e];
        const float *td = c->desc + c->off[e] * PSR_DESC_LEN;
        psr_match *m = (psr_match *)malloc(sizeof(psr_match) * nq);
        size_t nm = psr_match_descriptors(qf, qd, nq, tf, td, nt, ratio, 0, m);
        if (nm == 0) { free(m); continue; }
        psr_tracking tr;
        psr_summarize_tracking(qf, tf, m, nm, &tr);
        uint32_t *inl = (uint32_t *)malloc(sizeof(uint32_t) * nm);
        psr_similarity sim;
        int has_sim = 0, has_h = 0;
        size_t n_sim = psr_filter_matches_ransac(m, nm, ransac_thresh, DEFAULT_SIM_SEED, inl, &sim, &has_sim);
        float h[9];
        size_t n_h = psr_homography_ransac_ref(m, nm, ransac_thresh, 0, 1, inl, h, &has_h);
        size_t best = n_h > n_sim ? n_h : n_sim;
        psr_catalog_hit *o = &hits[i];
        o->entry = e;
        o->coarse_votes = cv[i];
        o->raw_match_count = tr.raw_match_count;
        o->match_count = best;
        o->inlier_ratio = tr.raw_match_count ? (float)best / (float)tr.raw_match_count : tr.inlier_ratio;
        o->mean_dx = tr.mean_dx;
        o->mean_dy = tr.mean_dy;
        o->median_dx = tr.median_dx;
        o->median_dy = tr.median_dy;
        o->mean_scale_ratio = tr.mean_scale_ratio;
        o->mean_rotation_deg = tr.mean_rotation_deg;
        o->has_transform = (uint32_t)has_sim;
        if (has_sim) {
            o->transform[0] = sim.a;
            o->transform[1] = sim.b;
            o->transform[2] = sim.tx;
            o->transform[3] = sim.ty;
            o->transform[4] = sim.scale;
            o->transform[5] = sim.rotation_deg;
        }
        o->has_homography = (uint32_t)has_h;
        if (has_h) memcpy(o->homography, h, sizeof h);
        keep[i] = 1;
        free(inl);
        free(m);
    }
    size_t nh = 0;
    for (size_t i = 0; i < nc; i++)
        if (keep[i]) hits[nh++] = hits[i];
    hit_ctx ctx = {c};
    stable_sort(hits, nh, sizeof(psr_catalog_hit), hit_cmp, &ctx);
    if (nh > top_k) nh = top_k;
    memcpy(out, hits, nh * sizeof(psr_catalog_hit));
    free(keep);
    free(hits);
    free(cv);
    free(ce);
    free(votes);
    free(qs);
    return nh;
}
