// C ABI of the GPU-resident feature catalog (catalog.rs): PSIFTCAT load / save, entry
// upsert / remove, the coarse vote pass and search_features.
//
// Device layout: the features of all entries back to back in entry order — ps_feature
// records (20 B), signatures (two u64 words, 16 B) and f32 descriptors (512 B) — plus
// per-entry u32 feature offsets and counts. A search runs on one engine's stream:
// query signatures -> vote kernel over every catalog signature -> (host) shortlist ->
// per shortlisted entry the exact matcher -> (host) problem list -> one batched RANSAC
// launch (similarity + homography; a similarity-only launch at 3 px when the threshold
// differs) -> one tracking-summary launch -> (host) hit ordering.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_common.h"

using namespace pshost;

struct ps_catalog {
    uint32_t octaves = 3, scales = 3;
    float threshold = 0.015f;
    uint8_t kind = 0;
    std::vector<std::string> ids;
    std::vector<uint32_t> width, height;
    std::vector<uint32_t> off{0};                     // n_entries + 1 feature offsets
    std::unordered_map<std::string, size_t> index;    // id -> entry (last wins, catalog.rs:284-297)
    DevBuf feat, desc, sig, off_d;
    size_t n_features() const { return off.back(); }
};

namespace {

constexpr uint32_t kMihMinEntries = 32;                     // catalog.rs:596
constexpr float kDefaultReproj = 3.0f;                      // pipeline.rs:1050
constexpr uint64_t kSimSeed = 0xD1B54A32D192ED03ull;        // pipeline.rs:173
constexpr uint64_t kHomSeed = 0x517CC1B727220A95ull;        // pipeline.rs:175
constexpr size_t kMaxQueries = 262144;                      // matcher limit per call
constexpr size_t kMaxVoteQueries = (size_t)65535 * 256;     // vote grid: 256-query blocks in y
constexpr float kDegPerRad = 57.2957795130823208767981548141051703f;   // f32::to_degrees

void free_buf(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}

void free_catalog(ps_catalog *c) {
    free_buf(c->feat);
    free_buf(c->desc);
    free_buf(c->sig);
    free_buf(c->off_d);
}

// std::str::from_utf8 acceptance: RFC 3629, no overlongs, no surrogates, <= U+10FFFF
bool utf8_valid(const uint8_t *s, size_t n) {
    size_t i = 0;
    while (i < n) {
        const uint8_t c = s[i];
        if (c < 0x80) {
            i++;
            continue;
        }
        size_t need;
        uint8_t lo = 0x80, hi = 0xBF;
        if (c >= 0xC2 && c <= 0xDF) need = 1;
        else if (c == 0xE0) need = 2, lo = 0xA0;
        else if (c >= 0xE1 && c <= 0xEC) need = 2;
        else if (c == 0xED) need = 2, hi = 0x9F;
        else if (c >= 0xEE && c <= 0xEF) need = 2;
        else if (c == 0xF0) need = 3, lo = 0x90;
        else if (c >= 0xF1 && c <= 0xF3) need = 3;
        else if (c == 0xF4) need = 3, hi = 0x8F;
        else return false;
        if (i + need >= n) return false;
        if (s[i + 1] < lo || s[i + 1] > hi) return false;
        for (size_t k = 2; k <= need; k++)
            if (s[i + k] < 0x80 || s[i + k] > 0xBF) return false;
        i += need + 1;
    }
    return true;
}

struct Reader {
    const uint8_t *p;
    size_t left;
    bool take(void *dst, size_t n) {
        if (left < n) return false;
        std::memcpy(dst, p, n);
        p += n;
        left -= n;
        return true;
    }
};

ps_status eof() { return set_err(PS_E_FORMAT, "Unexpected EOF while reading binary catalog"); }

// One entry of a catalog being rebuilt: an existing entry (old >= 0) or caller data.
struct EntrySrc {
    int64_t old = -1;
    const ps_feature *hf = nullptr;
    const float *hd = nullptr;
    uint32_t n = 0;
};

ps_status upload_offsets(ps_catalog *c, hipStream_t st) {
    const size_t ne = c->ids.size();
    PS_TRY(ensure(c->off_d, (ne + 1) * sizeof(uint32_t)));
    PS_HIP(hipMemcpyAsync(c->off_d.p, c->off.data(), (ne + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    return PS_OK;
}

// Lays the device arrays out again for the entry list `src` (new entry order), copying
// runs of consecutive existing entries with one copy per array.
ps_status rebuild(ps_catalog *c, ps_engine *e, const std::vector<EntrySrc> &src) {
    std::vector<uint32_t> off(src.size() + 1, 0);
    for (size_t k = 0; k < src.size(); k++) {
        const uint64_t n = src[k].old >= 0 ? c->off[src[k].old + 1] - c->off[src[k].old] : src[k].n;
        const uint64_t t = (uint64_t)off[k] + n;
        if (t >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "catalog larger than 2^32 - 1 features");
        off[k + 1] = (uint32_t)t;
    }
    const size_t nf = off.back();
    DevBuf f, d, s;
    ps_status st = ensure(f, std::max<size_t>(nf, 1) * sizeof(ps_feature));
    if (st == PS_OK) st = ensure(d, std::max<size_t>(nf, 1) * 128 * sizeof(float));
    if (st == PS_OK) st = ensure(s, std::max<size_t>(nf, 1) * 16);
    auto fail = [&](ps_status r) {
        (void)hipStreamSynchronize(e->stream);   // nothing queued may still use the new arrays
        free_buf(f);
        free_buf(d);
        free_buf(s);
        return r;
    };
    if (st != PS_OK) return fail(st);
    hipStream_t hs = e->stream;
    auto copy_run = [&](size_t dst, size_t srcf, size_t n) -> hipError_t {
        hipError_t r = hipMemcpyAsync(as<ps_feature>(f) + dst, as<ps_feature>(c->feat) + srcf, n * sizeof(ps_feature),
                                      hipMemcpyDeviceToDevice, hs);
        if (r == hipSuccess)
            r = hipMemcpyAsync(as<float>(d) + dst * 128, as<float>(c->desc) + srcf * 128, n * 512,
                               hipMemcpyDeviceToDevice, hs);
        if (r == hipSuccess)
            r = hipMemcpyAsync(as<uint64_t>(s) + dst * 2, as<uint64_t>(c->sig) + srcf * 2, n * 16,
                               hipMemcpyDeviceToDevice, hs);
        return r;
    };
    size_t k = 0;
    while (k < src.size()) {
        if (src[k].old >= 0) {
            size_t j = k + 1;
            while (j < src.size() && src[j].old == src[j - 1].old + 1) j++;
            const size_t n = off[j] - off[k];
            if (n) {
                hipError_t r = copy_run(off[k], c->off[src[k].old], n);
                if (r != hipSuccess) return fail(set_err(PS_E_DEVICE, std::string("catalog copy: ") + hipGetErrorString(r)));
            }
            k = j;
        } else {
            const uint32_t n = src[k].n;
            if (n) {
                hipError_t r = hipMemcpyAsync(as<ps_feature>(f) + off[k], src[k].hf, n * sizeof(ps_feature),
                                              hipMemcpyHostToDevice, hs);
                if (r == hipSuccess)
                    r = hipMemcpyAsync(as<float>(d) + (size_t)off[k] * 128, src[k].hd, (size_t)n * 512,
                                       hipMemcpyHostToDevice, hs);
                if (r == hipSuccess)
                    r = ps::launch_quantize_u128(as<float>(d) + (size_t)off[k] * 128, n,
                                                 as<uint64_t>(s) + (size_t)off[k] * 2, hs);
                if (r != hipSuccess) return fail(set_err(PS_E_DEVICE, std::string("catalog upload: ") + hipGetErrorString(r)));
            }
            k++;
        }
    }
    hipError_t r = hipStreamSynchronize(hs);
    if (r != hipSuccess) return fail(set_err(PS_E_DEVICE, std::string("catalog rebuild: ") + hipGetErrorString(r)));
    free_buf(c->feat);
    free_buf(c->desc);
    free_buf(c->sig);
    c->feat = f;
    c->desc = d;
    c->sig = s;
    c->off = off;
    return PS_OK;
}

void reindex(ps_catalog *c) {
    c->index.clear();
    for (size_t i = 0; i < c->ids.size(); i++) c->index[c->ids[i]] = i;
}

// catalog.rs:447-540 on the host: header and entries into c, the feature records split
// into the three arrays that go to the device (u8 descriptors still quantised).
struct Parsed {
    std::vector<ps_feature> feat;
    std::vector<uint64_t> sig;
    std::vector<uint8_t> q;
};

ps_status parse_host(ps_catalog *c, Parsed &out, const uint8_t *bytes, size_t len) {
    Reader r{bytes, len};
    char magic[8];
    if (!r.take(magic, 8)) return eof();
    if (std::memcmp(magic, "PSIFTCAT", 8) != 0) return set_err(PS_E_FORMAT, "Invalid binary catalog magic header");
    uint32_t version;
    if (!r.take(&version, 4)) return eof();
    if (version != 1)
        return set_err(PS_E_FORMAT, "Unsupported binary catalog version " + std::to_string(version) + " (expected 1)");
    uint8_t kind;
    uint64_t ne;
    if (!r.take(&c->octaves, 4) || !r.take(&c->scales, 4) || !r.take(&c->threshold, 4) || !r.take(&kind, 1))
        return eof();
    if (kind > 1)
        return set_err(PS_E_FORMAT, "Unknown descriptor kind " + std::to_string(kind) + " in binary catalog");
    c->kind = kind;
    if (!r.take(&ne, 8)) return eof();
    // pass 1: validate and size (every entry takes >= 20 bytes, so ne is bounded by len)
    Reader s = r;
    uint64_t nf = 0;
    for (uint64_t k = 0; k < ne; k++) {
        uint32_t idl, w, h;
        uint64_t fc;
        if (!s.take(&idl, 4)) return eof();
        if (s.left < idl) return set_err(PS_E_FORMAT, "Truncated entry id in binary catalog");
        if (!utf8_valid(s.p, idl)) return set_err(PS_E_FORMAT, "Invalid UTF-8 entry id in binary catalog");
        s.p += idl;
        s.left -= idl;
        if (!s.take(&w, 4) || !s.take(&h, 4) || !s.take(&fc, 8)) return eof();
        if (fc > s.left / 164) return eof();
        s.p += fc * 164;
        s.left -= fc * 164;
        nf += fc;
    }
    if (nf >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "catalog larger than 2^32 - 1 features");
    if (ne >= 0x7fffffffull) return set_err(PS_E_INVALID_ARG, "catalog larger than 2^31 - 1 entries");
    // pass 2 (sizes already checked)
    out.feat.resize(nf);
    out.sig.resize(2 * nf);
    out.q.resize((size_t)nf * 128);
    c->ids.resize(ne);
    c->width.resize(ne);
    c->height.resize(ne);
    c->off.assign(ne + 1, 0);
    size_t f = 0;
    for (uint64_t k = 0; k < ne; k++) {
        uint32_t idl = 0;
        uint64_t fc = 0;
        r.take(&idl, 4);
        c->ids[k].assign(reinterpret_cast<const char *>(r.p), idl);
        r.p += idl;
        r.left -= idl;
        r.take(&c->width[k], 4);
        r.take(&c->height[k], 4);
        r.take(&fc, 8);
        for (uint64_t j = 0; j < fc; j++, f++) {
            std::memcpy(&out.feat[f], r.p, 20);
            std::memcpy(&out.sig[2 * f], r.p + 20, 16);
            std::memcpy(&out.q[f * 128], r.p + 36, 128);
            r.p += 164;
            r.left -= 164;
        }
        c->off[k + 1] = (uint32_t)f;
    }
    reindex(c);
    return PS_OK;
}

// the parsed arrays to the device; descriptors dequantised there (pipeline.rs:1034-1044)
ps_status upload_parsed(ps_catalog *c, ps_engine *e, const Parsed &in) {
    hipStream_t st = e->stream;
    const size_t nf = in.feat.size(), nfa = std::max<size_t>(nf, 1);
    PS_TRY(ensure(c->feat, nfa * sizeof(ps_feature)));
    PS_TRY(ensure(c->desc, nfa * 512));
    PS_TRY(ensure(c->sig, nfa * 16));
    DevBuf qd;
    ps_status rs = ensure(qd, nfa * 128);
    if (rs != PS_OK) return rs;
    hipError_t he = hipSuccess;
    if (nf) {
        he = hipMemcpyAsync(c->feat.p, in.feat.data(), nf * sizeof(ps_feature), hipMemcpyHostToDevice, st);
        if (he == hipSuccess) he = hipMemcpyAsync(c->sig.p, in.sig.data(), nf * 16, hipMemcpyHostToDevice, st);
        if (he == hipSuccess) he = hipMemcpyAsync(qd.p, in.q.data(), nf * 128, hipMemcpyHostToDevice, st);
        if (he == hipSuccess) he = ps::launch_dequant_u8(as<uint8_t>(qd), (uint32_t)nf, as<float>(c->desc), st);
    }
    if (he == hipSuccess) {
        rs = upload_offsets(c, st);
        he = hipStreamSynchronize(st);
    }
    free_buf(qd);
    if (rs != PS_OK) return rs;
    if (he != hipSuccess) return set_err(PS_E_DEVICE, std::string("catalog load: ") + hipGetErrorString(he));
    return PS_OK;
}

bool mkdirs(const std::string &dir) {
    if (dir.empty()) return true;
    struct stat sb;
    if (stat(dir.c_str(), &sb) == 0) return S_ISDIR(sb.st_mode);
    const size_t slash = dir.find_last_of('/');
    if (slash != std::string::npos && slash > 0 && !mkdirs(dir.substr(0, slash))) return false;
    return mkdir(dir.c_str(), 0777) == 0 || errno == EEXIST;
}

// ---------------------------------------------------------------- search

struct Cand {
    uint32_t entry, votes;
};

// catalog.rs:597-620 / 626-633 on the device; votes copied to the host
ps_status coarse_votes(ps_engine *e, const ps_catalog *c, const float *d_qd, size_t nq, std::vector<uint32_t> &votes) {
    const size_t ne = c->ids.size();
    votes.assign(ne, 0);
    if (ne == 0) return PS_OK;
    hipStream_t st = e->stream;
    PS_TRY(ensure(e->aux[AUX_CS_0], std::max<size_t>(nq, 1) * 16));
    PS_TRY(ensure(e->aux[AUX_CS_1], ne * sizeof(uint32_t)));
    uint64_t *codes = as<uint64_t>(e->aux[AUX_CS_0]);
    uint32_t *dv = as<uint32_t>(e->aux[AUX_CS_1]);
    if (nq) PS_HIP(ps::launch_quantize_u128(d_qd, (uint32_t)nq, codes, st));
    PS_HIP(ps::launch_catalog_votes(as<uint64_t>(const_cast<DevBuf &>(c->sig)), as<uint32_t>(const_cast<DevBuf &>(c->off_d)),
                                    (uint32_t)ne, codes, (uint32_t)nq, ne > kMihMinEntries, dv, st));
    PS_HIP(hipMemcpyAsync(votes.data(), dv, ne * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PS_HIP(hipStreamSynchronize(st));
    return PS_OK;
}

ps_status search_impl(ps_engine *e, const ps_catalog *c, const ps_feature *d_qf, const float *d_qd, size_t nq,
                      size_t top_k, size_t shortlist_k, float ratio, float thr, ps_catalog_hit *out, size_t cap,
                      size_t *out_count) {
    *out_count = 0;
    const size_t ne = c->ids.size();
    if (ne == 0 || nq == 0 || top_k == 0) return PS_OK;   // catalog.rs:584-586
    if (nq > kMaxQueries) return set_err(PS_E_INVALID_ARG, "at most 262144 query features per catalog search");
    hipStream_t st = e->stream;
    DevBuf &cat_feat = const_cast<DevBuf &>(c->feat), &cat_desc = const_cast<DevBuf &>(c->desc);

    // 1. coarse votes and the shortlist (catalog.rs:594-638)
    std::vector<uint32_t> votes;
    PS_TRY(coarse_votes(e, c, d_qd, nq, votes));
    const bool mih = ne > kMihMinEntries;
    std::vector<Cand> cands;
    cands.reserve(ne);
    for (size_t k = 0; k < ne; k++)
        if (!mih || votes[k] > 0) cands.push_back({(uint32_t)k, votes[k]});
    const size_t keep = std::max<size_t>(shortlist_k, 1);
    if (cands.size() > keep) {
        std::stable_sort(cands.begin(), cands.end(), [&](const Cand &a, const Cand &b) {
            if (a.votes != b.votes) return a.votes > b.votes;
            return c->ids[a.entry] < c->ids[b.entry];
        });
        cands.resize(keep);
    }
    // entries with fewer than 2 features give no hit (catalog.rs:651-653)
    std::vector<Cand> slots;
    for (const Cand &k : cands)
        if (c->off[k.entry + 1] - c->off[k.entry] >= 2) slots.push_back(k);
    if (slots.empty()) return PS_OK;
    const size_t ns = slots.size();
    if ((uint64_t)ns * nq >= 0x7fffffffull) return set_err(PS_E_INVALID_ARG, "shortlist x queries too large");
    const size_t total = ns * nq;

    // 2. exact ratio match of the queries against every shortlisted entry
    PS_TRY(ensure(e->aux[AUX_CS_2], total * sizeof(ps_feature_match)));
    PS_TRY(ensure(e->aux[AUX_CS_3], (ns + 1) * sizeof(uint32_t)));
    ps_feature_match *dm = as<ps_feature_match>(e->aux[AUX_CS_2]);
    uint32_t *dcnt = as<uint32_t>(e->aux[AUX_CS_3]);   // [0, ns): match counts
    // one grouped match: the query set against each shortlisted entry's features
    std::vector<ps_feature_set> tsets(ns + 1);
    std::vector<ps_match_problem> mprob(ns);
    uint32_t set_cap = (uint32_t)nq;
    for (size_t k = 0; k < ns; k++) {
        const uint32_t ent = slots[k].entry, o = c->off[ent], nt = c->off[ent + 1] - o;
        tsets[k] = ps_feature_set{o, nt};
        mprob[k] = ps_match_problem{0u, (uint32_t)k, (uint32_t)(k * nq), 0u};
        set_cap = std::max(set_cap, nt);
    }
    tsets[ns] = ps_feature_set{0u, (uint32_t)nq};   // the query set
    PS_TRY(ensure(e->aux[AUX_CS_12], (ns + 1) * sizeof(ps_feature_set) + ns * sizeof(ps_match_problem)));
    ps_feature_set *dsets = as<ps_feature_set>(e->aux[AUX_CS_12]);
    ps_match_problem *dmp = reinterpret_cast<ps_match_problem *>(dsets + ns + 1);
    for (size_t k = 0; k < ns; k++) mprob[k].query_set = (uint32_t)ns;
    PS_HIP(hipMemcpyAsync(dsets, tsets.data(), (ns + 1) * sizeof(ps_feature_set), hipMemcpyHostToDevice, st));
    PS_HIP(hipMemcpyAsync(dmp, mprob.data(), ns * sizeof(ps_match_problem), hipMemcpyHostToDevice, st));
    ps::MatchSetsArgs ma{};
    ma.qf = d_qf;
    ma.qd = d_qd;
    ma.q_rows = nq;
    ma.qsets = dsets;
    ma.n_qsets = (uint32_t)ns + 1;
    ma.tf = as<ps_feature>(cat_feat);
    ma.td = as<float>(cat_desc);
    ma.t_rows = c->off[c->ids.size()];
    ma.tsets = dsets;
    ma.n_tsets = (uint32_t)ns;
    ma.cap = set_cap;
    ma.probs = dmp;
    ma.n_probs = (uint32_t)ns;
    ma.ratio = ratio;
    ma.shared = false;
    PS_TRY(enqueue_match_sets(e, ma, e->aux[AUX_CS_13], dm, dcnt));
    // 3. RANSAC problems and tracking jobs of every slot, built on the device from the match
    // counts (catalog.rs:656-671; a slot without matches runs an empty problem and is dropped
    // from the hits below)
    const size_t np = ns;
    const bool same_thr = thr == kDefaultReproj;
    const size_t wbytes = ps::ransac_work_bytes(total);
    PS_TRY(ensure(e->aux[AUX_CS_4], np * sizeof(ps_ransac_problem)));
    PS_TRY(ensure(e->aux[AUX_CS_5], np * sizeof(ps::TrackJob)));
    PS_TRY(ensure(e->aux[AUX_CS_6], wbytes));
    PS_TRY(ensure(e->aux[AUX_CS_7], total * sizeof(uint32_t)));
    PS_TRY(ensure(e->aux[AUX_CS_8], 2 * np * sizeof(ps_ransac_result)));
    PS_TRY(ensure(e->aux[AUX_CS_9], np * sizeof(ps::TrackOut)));
    if (!same_thr) PS_TRY(ensure(e->aux[AUX_CS_10], wbytes));
    ps_ransac_problem *dprob = as<ps_ransac_problem>(e->aux[AUX_CS_4]);
    ps::TrackJob *djob = as<ps::TrackJob>(e->aux[AUX_CS_5]);
    ps_ransac_result *dres = as<ps_ransac_result>(e->aux[AUX_CS_8]);
    ps::TrackOut *dtrk = as<ps::TrackOut>(e->aux[AUX_CS_9]);
    PS_HIP(ps::launch_catalog_problems(dcnt, dsets, (uint32_t)ns, (uint32_t)nq, kSimSeed, kHomSeed, dprob, djob, st));
    ps_ransac_params prm;
    std::memset(&prm, 0, sizeof prm);
    prm.reproj_threshold = thr;
    prm.sampler = PS_SAMPLER_REFERENCE;
    prm.local_opt = 1;
    prm.sim_seed_model = 1;
    prm.adaptive = 1;
    prm.confidence = 0.999;   // pipeline.rs:2689
    prm.min_iters = 64;
    prm.max_iters = 2000;
    // filter_matches_ransac + estimate_homography_ransac at the search threshold: the
    // homography's similarity seed stage is the same similarity RANSAC (pipeline.rs:2676-2678)
    PS_HIP(ps::launch_ransac(dm, nullptr, dprob, (uint32_t)np, prm, total, e->aux[AUX_CS_6].p,
                             as<uint32_t>(e->aux[AUX_CS_7]), dres, st));
    const ps_ransac_result *sim_res = dres;
    const void *sim_work = e->aux[AUX_CS_6].p;
    if (!same_thr) {
        // summarize_tracking's own similarity RANSAC at the default 3 px (pipeline.rs:2227)
        ps_ransac_params p3 = prm;
        p3.reproj_threshold = kDefaultReproj;
        PS_HIP(ps::launch_ransac(dm, nullptr, dprob, (uint32_t)np, p3, total, e->aux[AUX_CS_10].p,
                                 as<uint32_t>(e->aux[AUX_CS_7]), dres + np, st, true));
        sim_res = dres + np;
        sim_work = e->aux[AUX_CS_10].p;
    }
    PS_HIP(ps::launch_tracking(dm, djob, (uint32_t)np, sim_res, sim_work, total, d_qf, as<ps_feature>(cat_feat),
                               dtrk, st));
    std::vector<ps_ransac_result> res(np);
    std::vector<ps::TrackOut> trk(np);
    std::vector<uint32_t> mcnt(np);
    PS_HIP(hipMemcpyAsync(res.data(), dres, np * sizeof(ps_ransac_result), hipMemcpyDeviceToHost, st));
    PS_HIP(hipMemcpyAsync(trk.data(), dtrk, np * sizeof(ps::TrackOut), hipMemcpyDeviceToHost, st));
    PS_HIP(hipMemcpyAsync(mcnt.data(), dcnt, np * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PS_HIP(hipStreamSynchronize(st));

    // 4. hits of the entries with matches (catalog.rs:664-700)
    std::vector<ps_catalog_hit> hits;
    hits.reserve(np);
    for (size_t p = 0; p < np; p++) {
        if (mcnt[p] == 0) continue;
        const Cand &k = slots[p];
        const ps_ransac_result &R = res[p];
        const ps::TrackOut &T = trk[p];
        hits.emplace_back();
        ps_catalog_hit &h = hits.back();
        std::memset(&h, 0, sizeof h);
        const uint32_t raw = mcnt[p], best = std::max(R.n_inliers, R.n_sim);
        h.entry = k.entry;
        h.coarse_votes = k.votes;
        h.raw_match_count = raw;
        h.match_count = best;
        h.inlier_ratio = (float)best / (float)raw;
        h.mean_dx = T.mean_dx;
        h.mean_dy = T.mean_dy;
        h.median_dx = T.median_dx;
        h.median_dy = T.median_dy;
        h.mean_scale_ratio = T.mean_scale_ratio;
        h.mean_rotation_deg = T.mean_rotation_deg;
        h.has_transform = R.has_sim ? 1 : 0;
        if (R.has_sim) {
            // pipeline.rs:1962-1963
            const float a = R.sim[0], b = R.sim[1];
            h.transform[0] = a;
            h.transform[1] = b;
            h.transform[2] = R.sim[2];
            h.transform[3] = R.sim[3];
            // f32 libm calls, as the reference's f32::sqrt / f32::atan2 (no promotion to double)
            const float rot = ::atan2f(b, a);
            h.transform[4] = ::sqrtf(a * a + b * b);
            h.transform[5] = rot * kDegPerRad;
        }
        h.has_homography = R.has_model ? 1 : 0;
        if (R.has_model) std::memcpy(h.homography, R.h, sizeof h.homography);
    }
    std::stable_sort(hits.begin(), hits.end(), [&](const ps_catalog_hit &a, const ps_catalog_hit &b) {
        if (a.match_count != b.match_count) return a.match_count > b.match_count;
        if (a.inlier_ratio != b.inlier_ratio) return a.inlier_ratio > b.inlier_ratio;
        if (a.coarse_votes != b.coarse_votes) return a.coarse_votes > b.coarse_votes;
        return c->ids[a.entry] < c->ids[b.entry];
    });
    const size_t nh = std::min(hits.size(), top_k);
    *out_count = nh;
    if (nh > cap) return set_err(PS_E_CAPACITY, "output capacity smaller than the hit count");
    if (nh) std::memcpy(out, hits.data(), nh * sizeof(ps_catalog_hit));
    return PS_OK;
}

ps_status check_search_args(const ps_catalog *c, const void *q, const void *qd, size_t nq, const void *out,
                            size_t cap, const size_t *out_count) {
    if (!c || !out_count || (nq && (!q || !qd)) || (cap && !out)) return set_err(PS_E_INVALID_ARG, "null argument");
    return PS_OK;
}

}  // namespace

extern "C" {

ps_status ps_catalog_create(uint32_t octaves, uint32_t scales, float contrast_threshold, uint8_t descriptor_kind,
                            ps_catalog **out) {
    PS_TRY(check_fail_injection());
    if (!out) return set_err(PS_E_INVALID_ARG, "null argument");
    *out = nullptr;
    if (descriptor_kind > 1) return set_err(PS_E_INVALID_ARG, "descriptor_kind must be 0 (Sift) or 1 (RootSift)");
    Borrow b;
    PS_TRY(b.st);
    PS_HIP(hipSetDevice(b.e->device));
    ps_catalog *c = new ps_catalog;
    c->octaves = octaves;
    c->scales = scales;
    c->threshold = contrast_threshold;
    c->kind = descriptor_kind;
    ps_status s = upload_offsets(c, b.e->stream);
    if (s == PS_OK && hipStreamSynchronize(b.e->stream) != hipSuccess) s = set_err(PS_E_DEVICE, "catalog create");
    if (s != PS_OK) {
        free_catalog(c);
        delete c;
        return s;
    }
    *out = c;
    return PS_OK;
}

ps_status ps_catalog_load_binary(const uint8_t *bytes, size_t len, ps_catalog **out) {
    PS_TRY(check_fail_injection());
    if (!out || (len && !bytes)) return set_err(PS_E_INVALID_ARG, "null argument");
    *out = nullptr;
    ps_catalog *c = new ps_catalog;
    Parsed parsed;
    ps_status s = parse_host(c, parsed, bytes, len);   // format errors need no device
    if (s == PS_OK) {
        Borrow b;
        s = b.st;
        if (s == PS_OK && hipSetDevice(b.e->device) != hipSuccess) s = set_err(PS_E_DEVICE, "hipSetDevice");
        if (s == PS_OK) s = upload_parsed(c, b.e, parsed);
        if (s != PS_OK && b.e) (void)hipStreamSynchronize(b.e->stream);
    }
    if (s != PS_OK) {
        free_catalog(c);
        delete c;
        return s;
    }
    *out = c;
    return PS_OK;
}

ps_status ps_catalog_load_file(const char *path, ps_catalog **out) {
    PS_TRY(check_fail_injection());
    if (!path || !out) return set_err(PS_E_INVALID_ARG, "null argument");
    *out = nullptr;
    FILE *fp = std::fopen(path, "rb");
    if (!fp)
        return set_err(PS_E_IO, std::string("Failed to read binary catalog `") + path + "`: " + std::strerror(errno));
    std::vector<uint8_t> buf;
    uint8_t chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, fp)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    const bool bad = std::ferror(fp) != 0;
    std::fclose(fp);
    if (bad) return set_err(PS_E_IO, std::string("Failed to read binary catalog `") + path + "`");
    return ps_catalog_load_binary(buf.data(), buf.size(), out);
}

ps_status ps_catalog_save_binary(const ps_catalog *c, const char *path) {
    PS_TRY(check_fail_injection());
    if (!c || !path) return set_err(PS_E_INVALID_ARG, "null argument");
    const size_t nf = c->n_features(), ne = c->ids.size();
    std::vector<ps_feature> feat(nf);
    std::vector<float> desc(nf * 128);
    std::vector<uint64_t> sig(2 * nf);
    if (nf) {
        Borrow b;
        PS_TRY(b.st);
        PS_HIP(hipSetDevice(b.e->device));
        hipStream_t st = b.e->stream;
        PS_HIP(hipMemcpyAsync(feat.data(), c->feat.p, nf * sizeof(ps_feature), hipMemcpyDeviceToHost, st));
        PS_HIP(hipMemcpyAsync(desc.data(), c->desc.p, nf * 512, hipMemcpyDeviceToHost, st));
        PS_HIP(hipMemcpyAsync(sig.data(), c->sig.p, nf * 16, hipMemcpyDeviceToHost, st));
        PS_HIP(hipStreamSynchronize(st));
    }
    std::vector<uint8_t> buf;
    auto put = [&](const void *p, size_t n) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        buf.insert(buf.end(), b, b + n);
    };
    const uint32_t version = 1;
    const uint64_t ne64 = ne;
    put("PSIFTCAT", 8);
    put(&version, 4);
    put(&c->octaves, 4);
    put(&c->scales, 4);
    put(&c->threshold, 4);
    put(&c->kind, 1);
    put(&ne64, 8);
    for (size_t k = 0; k < ne; k++) {
        const uint32_t idl = (uint32_t)c->ids[k].size();
        const uint64_t fc = c->off[k + 1] - c->off[k];
        put(&idl, 4);
        put(c->ids[k].data(), idl);
        put(&c->width[k], 4);
        put(&c->height[k], 4);
        put(&fc, 8);
        for (uint32_t f = c->off[k]; f < c->off[k + 1]; f++) {
            put(&feat[f], 20);
            put(&sig[2 * f], 16);
            uint8_t q[128];
            for (int i = 0; i < 128; i++) {   // pipeline.rs:1024-1030
                float r = std::round(desc[(size_t)f * 128 + i] * 512.0f);
                if (r < 0.0f) r = 0.0f;
                if (r > 255.0f) r = 255.0f;
                q[i] = (r == r) ? (uint8_t)r : 0;
            }
            put(q, 128);
        }
    }
    // catalog.rs:204-244: create the parent directory, write a temporary file, rename
    const std::string p(path);
    const size_t slash = p.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "" : p.substr(0, slash);
    const std::string name = slash == std::string::npos ? p : p.substr(slash + 1);
    if (!mkdirs(dir))
        return set_err(PS_E_IO, "Failed to create catalog directory `" + dir + "`: " + std::strerror(errno));
    const std::string tmp = (dir.empty() ? "" : dir + "/") + (name.empty() ? "catalog" : name) + ".tmp." +
                            std::to_string((long)getpid());
    FILE *fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return set_err(PS_E_IO, "Failed to create temporary catalog `" + tmp + "`: " + std::strerror(errno));
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
    const bool closed = std::fclose(fp) == 0;
    if (!ok || !closed) {
        std::remove(tmp.c_str());
        return set_err(PS_E_IO, "Failed to write catalog `" + tmp + "`");
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        const std::string why = std::strerror(errno);
        std::remove(tmp.c_str());
        return set_err(PS_E_IO, "Failed to atomically replace catalog `" + p + "`: " + why);
    }
    return PS_OK;
}

ps_status ps_catalog_destroy(ps_catalog *c) {
    if (!c) return PS_OK;
    free_catalog(c);
    delete c;
    return PS_OK;
}

ps_status ps_catalog_get_info(const ps_catalog *c, ps_catalog_info *out) {
    if (!c || !out) return set_err(PS_E_INVALID_ARG, "null argument");
    std::memset(out, 0, sizeof *out);
    out->n_entries = c->ids.size();
    out->n_features = c->n_features();
    out->octaves = c->octaves;
    out->scales = c->scales;
    out->contrast_threshold = c->threshold;
    out->descriptor_kind = c->kind;
    return PS_OK;
}

ps_status ps_catalog_entry(const ps_catalog *c, size_t idx, ps_catalog_entry_info *out) {
    if (!c || !out) return set_err(PS_E_INVALID_ARG, "null argument");
    if (idx >= c->ids.size()) return set_err(PS_E_INVALID_ARG, "entry index out of range");
    out->id = c->ids[idx].data();
    out->id_len = c->ids[idx].size();
    out->width = c->width[idx];
    out->height = c->height[idx];
    out->feature_count = c->off[idx + 1] - c->off[idx];
    return PS_OK;
}

ps_status ps_catalog_entry_features(const ps_catalog *c, size_t idx, ps_feature *out_feat, float *out_desc,
                                    uint64_t *out_sig, size_t cap) {
    PS_TRY(check_fail_injection());
    if (!c) return set_err(PS_E_INVALID_ARG, "null argument");
    if (idx >= c->ids.size()) return set_err(PS_E_INVALID_ARG, "entry index out of range");
    const size_t o = c->off[idx], n = c->off[idx + 1] - o;
    if (n > cap) return set_err(PS_E_CAPACITY, "output capacity smaller than the entry's feature count");
    if (n == 0) return PS_OK;
    Borrow b;
    PS_TRY(b.st);
    PS_HIP(hipSetDevice(b.e->device));
    hipStream_t st = b.e->stream;
    if (out_feat)
        PS_HIP(hipMemcpyAsync(out_feat, as<ps_feature>(const_cast<DevBuf &>(c->feat)) + o, n * sizeof(ps_feature),
                              hipMemcpyDeviceToHost, st));
    if (out_desc)
        PS_HIP(hipMemcpyAsync(out_desc, as<float>(const_cast<DevBuf &>(c->desc)) + o * 128, n * 512,
                              hipMemcpyDeviceToHost, st));
    if (out_sig)
        PS_HIP(hipMemcpyAsync(out_sig, as<uint64_t>(const_cast<DevBuf &>(c->sig)) + o * 2, n * 16,
                              hipMemcpyDeviceToHost, st));
    PS_HIP(hipStreamSynchronize(st));
    return PS_OK;
}

ps_status ps_catalog_upsert(ps_catalog *c, const char *id, size_t id_len, uint32_t width, uint32_t height,
                            const ps_feature *feat, const float *desc, size_t n) {
    PS_TRY(check_fail_injection());
    if (!c || (id_len && !id) || (n && (!feat || !desc))) return set_err(PS_E_INVALID_ARG, "null argument");
    if (!utf8_valid(reinterpret_cast<const uint8_t *>(id), id_len))
        return set_err(PS_E_INVALID_ARG, "entry id is not valid UTF-8");
    if (n >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "too many features");
    const std::string key(id ? id : "", id_len);
    Borrow b;
    PS_TRY(b.st);
    PS_HIP(hipSetDevice(b.e->device));
    const auto it = c->index.find(key);
    const size_t ne = c->ids.size();
    std::vector<EntrySrc> src(ne);
    for (size_t k = 0; k < ne; k++) src[k].old = (int64_t)k;
    EntrySrc fresh;
    fresh.hf = feat;
    fresh.hd = desc;
    fresh.n = (uint32_t)n;
    const size_t at = it != c->index.end() ? it->second : ne;
    if (at < ne) src[at] = fresh;
    else src.push_back(fresh);
    PS_TRY(rebuild(c, b.e, src));
    if (at < ne) {
        c->width[at] = width;
        c->height[at] = height;
    } else {
        c->ids.push_back(key);
        c->width.push_back(width);
        c->height.push_back(height);
        c->index[key] = at;
    }
    PS_TRY(upload_offsets(c, b.e->stream));
    PS_HIP(hipStreamSynchronize(b.e->stream));
    return PS_OK;
}

ps_status ps_catalog_remove(ps_catalog *c, const char *id, size_t id_len, int32_t *removed) {
    PS_TRY(check_fail_injection());
    if (!c || !removed || (id_len && !id)) return set_err(PS_E_INVALID_ARG, "null argument");
    *removed = 0;
    const auto it = c->index.find(std::string(id ? id : "", id_len));
    if (it == c->index.end()) return PS_OK;
    const size_t at = it->second, last = c->ids.size() - 1;
    Borrow b;
    PS_TRY(b.st);
    PS_HIP(hipSetDevice(b.e->device));
    // Vec::swap_remove: the last entry takes the removed one's place
    std::vector<EntrySrc> src(last);
    for (size_t k = 0; k < last; k++) src[k].old = (int64_t)(k == at ? last : k);
    PS_TRY(rebuild(c, b.e, src));
    if (at != last) {
        c->ids[at] = c->ids[last];
        c->width[at] = c->width[last];
        c->height[at] = c->height[last];
    }
    c->ids.pop_back();
    c->width.pop_back();
    c->height.pop_back();
    reindex(c);
    PS_TRY(upload_offsets(c, b.e->stream));
    PS_HIP(hipStreamSynchronize(b.e->stream));
    *removed = 1;
    return PS_OK;
}

ps_status ps_catalog_coarse_votes(const ps_catalog *c, const float *q_desc, size_t nq, uint32_t *out_votes) {
    PS_TRY(check_fail_injection());
    if (!c || (nq && !q_desc) || (!c->ids.empty() && !out_votes)) return set_err(PS_E_INVALID_ARG, "null argument");
    if (nq > kMaxVoteQueries) return set_err(PS_E_INVALID_ARG, "at most 65535 * 256 queries per vote call");
    if (c->ids.empty()) return PS_OK;
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->aux[AUX_CS_11], std::max<size_t>(nq, 1) * 512));
    if (nq)
        PS_HIP(hipMemcpyAsync(e->aux[AUX_CS_11].p, q_desc, nq * 512, hipMemcpyHostToDevice, e->stream));
    std::vector<uint32_t> votes;
    PS_TRY(coarse_votes(e, c, as<float>(e->aux[AUX_CS_11]), nq, votes));
    std::memcpy(out_votes, votes.data(), votes.size() * sizeof(uint32_t));
    return PS_OK;
}

ps_status ps_catalog_search(const ps_catalog *c, const ps_feature *q, const float *q_desc, size_t nq, size_t top_k,
                            size_t shortlist_k, float ratio, float ransac_threshold, ps_catalog_hit *out, size_t cap,
                            size_t *out_count) {
    TraceRange tr("ps.catalog_search");
    PS_TRY(check_fail_injection());
    PS_TRY(check_search_args(c, q, q_desc, nq, out, cap, out_count));
    *out_count = 0;
    if (c->ids.empty() || nq == 0 || top_k == 0) return PS_OK;
    if (nq > kMaxQueries) return set_err(PS_E_INVALID_ARG, "at most 262144 query features per catalog search");
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->qf, nq * sizeof(ps_feature)));
    PS_TRY(ensure(e->qd, nq * 512));
    PS_HIP(hipMemcpyAsync(e->qf.p, q, nq * sizeof(ps_feature), hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->qd.p, q_desc, nq * 512, hipMemcpyHostToDevice, e->stream));
    return search_impl(e, c, as<ps_feature>(e->qf), as<float>(e->qd), nq, top_k, shortlist_k, ratio,
                       ransac_threshold, out, cap, out_count);
}

ps_status ps_engine_catalog_search(ps_engine *e, const ps_catalog *c, const ps_feature *d_q, const float *d_q_desc,
                                   size_t nq, size_t top_k, size_t shortlist_k, float ratio, float ransac_threshold,
                                   ps_catalog_hit *out, size_t cap, size_t *out_count) {
    PS_TRY(check_fail_injection());
    if (!e) return set_err(PS_E_INVALID_ARG, "null engine");
    PS_TRY(check_search_args(c, d_q, d_q_desc, nq, out, cap, out_count));
    PS_HIP(hipSetDevice(e->device));
    return search_impl(e, c, d_q, d_q_desc, nq, top_k, shortlist_k, ratio, ransac_threshold, out, cap, out_count);
}

}  // extern "C"
