// Checkpoint / resume of the match-scoring MLP (the learned part of guided RANSAC): the f32
// master parameters and, for resuming training, the Adam moments and step, in a versioned
// little-endian file written atomically (temporary file + rename, as the catalog's PSIFTCAT
// files are, catalog.rs:203-245).
//
// Layout: "PSMLPCKP" | u32 version (1) | u32 4 | u32 dims[4] = 8, 64, 64, 1 | u32 n_params
// (4801) | u32 adam_step | u32 flags (bit 0: moments follow) | u32 0 | f32 params[n] |
// [f32 m[n] | f32 v[n]] | u64 FNV-1a of every preceding byte.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_common.h"

using namespace pshost;

namespace {

constexpr char kMagic[8] = {'P', 'S', 'M', 'L', 'P', 'C', 'K', 'P'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kDims[4] = {PS_MLP_FEATURES, 64, 64, 1};
constexpr size_t kHeader = 8 + 4 * 10;

uint64_t fnv1a(const unsigned char *p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

void put_u32(std::vector<unsigned char> &b, uint32_t v) {
    for (int i = 0; i < 4; i++) b.push_back((unsigned char)(v >> (8 * i)));
}

uint32_t get_u32(const unsigned char *p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

void put_f32s(std::vector<unsigned char> &b, const float *v, size_t n) {
    const size_t at = b.size();
    b.resize(at + n * 4);
    for (size_t i = 0; i < n; i++) {
        uint32_t u;
        std::memcpy(&u, &v[i], 4);
        for (int k = 0; k < 4; k++) b[at + 4 * i + k] = (unsigned char)(u >> (8 * k));
    }
}

void get_f32s(const unsigned char *p, float *v, size_t n) {
    for (size_t i = 0; i < n; i++) {
        const uint32_t u = get_u32(p + 4 * i);
        std::memcpy(&v[i], &u, 4);
    }
}

}  // namespace

extern "C" {

ps_status ps_mlp_checkpoint_save(const char *path, const float *params, const float *m, const float *v,
                                 uint32_t adam_step) {
    PS_TRY(check_fail_injection());
    if (!path || !params || (!m != !v)) return set_err(PS_E_INVALID_ARG, "null argument");
    const size_t n = PS_MLP_PARAMS;
    const uint32_t flags = m ? 1u : 0u;
    std::vector<unsigned char> b(kMagic, kMagic + 8);
    put_u32(b, kVersion);
    put_u32(b, 4);
    for (uint32_t d : kDims) put_u32(b, d);
    put_u32(b, (uint32_t)n);
    put_u32(b, adam_step);
    put_u32(b, flags);
    put_u32(b, 0);
    put_f32s(b, params, n);
    if (flags) {
        put_f32s(b, m, n);
        put_f32s(b, v, n);
    }
    const uint64_t h = fnv1a(b.data(), b.size());
    for (int i = 0; i < 8; i++) b.push_back((unsigned char)(h >> (8 * i)));
    const std::string tmp = std::string(path) + ".tmp";
    FILE *fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return set_err(PS_E_IO, "Failed to create temporary checkpoint `" + tmp + "`: " + std::strerror(errno));
    const bool ok = std::fwrite(b.data(), 1, b.size(), fp) == b.size();
    if (std::fclose(fp) != 0 || !ok) {
        std::remove(tmp.c_str());
        return set_err(PS_E_IO, "Failed to write checkpoint `" + tmp + "`");
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        const std::string why = std::strerror(errno);
        std::remove(tmp.c_str());
        return set_err(PS_E_IO, "Failed to atomically replace checkpoint `" + std::string(path) + "`: " + why);
    }
    return PS_OK;
}

ps_status ps_mlp_checkpoint_load(const char *path, float *params, float *m, float *v, uint32_t *adam_step,
                                 uint32_t *flags) {
    PS_TRY(check_fail_injection());
    if (!path || !params || (!m != !v)) return set_err(PS_E_INVALID_ARG, "null argument");
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return set_err(PS_E_IO, "Failed to read checkpoint `" + std::string(path) + "`: " + std::strerror(errno));
    std::vector<unsigned char> b;
    unsigned char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, fp)) > 0) b.insert(b.end(), chunk, chunk + got);
    const bool bad = std::ferror(fp) != 0;
    std::fclose(fp);
    if (bad) return set_err(PS_E_IO, "Failed to read checkpoint `" + std::string(path) + "`");
    if (b.size() < kHeader + 8) return set_err(PS_E_FORMAT, "Truncated MLP checkpoint");
    if (std::memcmp(b.data(), kMagic, 8) != 0) return set_err(PS_E_FORMAT, "Invalid MLP checkpoint magic header");
    const uint32_t version = get_u32(&b[8]);
    if (version != kVersion)
        return set_err(PS_E_FORMAT, "Unsupported MLP checkpoint version " + std::to_string(version) + " (expected 1)");
    if (get_u32(&b[12]) != 4) return set_err(PS_E_FORMAT, "MLP checkpoint layer count differs from 8-64-64-1");
    for (int i = 0; i < 4; i++)
        if (get_u32(&b[16 + 4 * i]) != kDims[i])
            return set_err(PS_E_FORMAT, "MLP checkpoint layer sizes differ from 8-64-64-1");
    const uint32_t n = get_u32(&b[32]), step = get_u32(&b[36]), fl = get_u32(&b[40]);
    if (n != PS_MLP_PARAMS) return set_err(PS_E_FORMAT, "MLP checkpoint holds " + std::to_string(n) + " parameters (expected 4801)");
    if (fl & ~1u) return set_err(PS_E_FORMAT, "Unknown MLP checkpoint flags " + std::to_string(fl));
    const size_t want = kHeader + (size_t)n * 4 * ((fl & 1u) ? 3 : 1) + 8;
    if (b.size() != want) return set_err(PS_E_FORMAT, "MLP checkpoint size differs from its header");
    uint64_t h = 0;
    for (int i = 0; i < 8; i++) h |= (uint64_t)b[want - 8 + i] << (8 * i);
    if (h != fnv1a(b.data(), want - 8)) return set_err(PS_E_FORMAT, "MLP checkpoint checksum mismatch");
    get_f32s(&b[kHeader], params, n);
    if ((fl & 1u) && m) {
        get_f32s(&b[kHeader + (size_t)n * 4], m, n);
        get_f32s(&b[kHeader + (size_t)n * 8], v, n);
    }
    if (adam_step) *adam_step = step;
    if (flags) *flags = fl;
    return PS_OK;
}

}  // extern "C"
