// C ABI of the 128-bit code functions: quantisation, coarse votes, top-k shortlist, rerank.
#include <algorithm>
#include <cstring>

#include "host_common.h"

using namespace pshost;

namespace {

ps_status check_topk_args(size_t n, size_t nq, uint32_t k) {
    if (k == 0 || k > ps::kHamMaxK) return set_err(PS_E_INVALID_ARG, "k must be in 1..=1024");
    if (n >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "catalog larger than 2^32 - 1 codes");
    if (nq >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "too many queries");
    return PS_OK;
}

// host fill of the (UINT32_MAX, 255) rows when the catalog is empty
void fill_empty(uint32_t *idx, uint8_t *dist, size_t cnt) {
    for (size_t i = 0; i < cnt; i++) idx[i] = 0xffffffffu;
    std::memset(dist, 0xff, cnt);
}

ps_status enqueue_topk(ps_engine *e, const uint64_t *cat, size_t n, const uint64_t *qs, size_t nq, uint32_t k,
                       uint32_t *idx, uint8_t *dist) {
    const uint32_t chunk = (uint32_t)std::min<size_t>(nq, ps::kHamQueryChunk);
    PS_TRY(ensure(e->aux[AUX_HAM_KEYS], ps::hamming_topk_scratch_bytes(chunk)));
    PS_HIP(ps::launch_hamming_topk(cat, (uint32_t)n, qs, (uint32_t)nq, k, e->aux[AUX_HAM_KEYS].p, idx, dist,
                                   e->stream));
    return PS_OK;
}

}  // namespace

extern "C" {

ps_status ps_quantize_u128(const float *desc, size_t n, uint64_t *out_codes) {
    PS_TRY(check_fail_injection());
    if (n && (!desc || !out_codes)) return set_err(PS_E_INVALID_ARG, "null argument");
    if (n >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "too many descriptors");
    if (n == 0) return PS_OK;
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->aux[AUX_RR_QD], n * 128 * sizeof(float)));
    PS_TRY(ensure(e->aux[AUX_HAM_Q], n * 16));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RR_QD].p, desc, n * 128 * sizeof(float), hipMemcpyHostToDevice, e->stream));
    PS_HIP(ps::launch_quantize_u128(as<float>(e->aux[AUX_RR_QD]), (uint32_t)n, as<uint64_t>(e->aux[AUX_HAM_Q]),
                                    e->stream));
    PS_HIP(hipMemcpyAsync(out_codes, e->aux[AUX_HAM_Q].p, n * 16, hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

ps_status ps_count_coarse_hamming_votes(const uint64_t *q_codes, size_t nq, const uint64_t *e_codes, size_t ne,
                                        size_t *out_votes) {
    PS_TRY(check_fail_injection());
    if (!out_votes || (nq && !q_codes) || (ne && !e_codes)) return set_err(PS_E_INVALID_ARG, "null argument");
    *out_votes = 0;
    if (nq >= 0xffffffffull || ne >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "too many codes");
    if (ne == 0 || nq == 0) return PS_OK;   // catalog.rs:708-710
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->aux[AUX_HAM_Q], nq * 16));
    PS_TRY(ensure(e->aux[AUX_HAM_CAT], ne * 16));
    PS_TRY(ensure(e->aux[AUX_HAM_CNT], (nq + 1) * sizeof(uint32_t)));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_HAM_Q].p, q_codes, nq * 16, hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_HAM_CAT].p, e_codes, ne * 16, hipMemcpyHostToDevice, e->stream));
    uint32_t *sc = as<uint32_t>(e->aux[AUX_HAM_CNT]);
    PS_HIP(ps::launch_hamming_votes(as<uint64_t>(e->aux[AUX_HAM_Q]), (uint32_t)nq, as<uint64_t>(e->aux[AUX_HAM_CAT]),
                                    (uint32_t)ne, sc, e->stream));
    PS_HIP(hipMemcpyAsync(&e->host_counts[0], sc + nq, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    *out_votes = e->host_counts[0];
    return PS_OK;
}

ps_status ps_hamming_topk(const uint64_t *catalog, size_t n, const uint64_t *q_codes, size_t nq, uint32_t k,
                          uint32_t *out_idx, uint8_t *out_dist) {
    PS_TRY(check_fail_injection());
    PS_TRY(check_topk_args(n, nq, k));
    if (nq && (!q_codes || !out_idx || !out_dist)) return set_err(PS_E_INVALID_ARG, "null argument");
    if (n && !catalog) return set_err(PS_E_INVALID_ARG, "null argument");
    if (nq == 0) return PS_OK;
    if (n == 0) {
        fill_empty(out_idx, out_dist, nq * k);
        return PS_OK;
    }
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->aux[AUX_HAM_CAT], n * 16));
    PS_TRY(ensure(e->aux[AUX_HAM_Q], nq * 16));
    PS_TRY(ensure(e->aux[AUX_HAM_IDX], nq * k * sizeof(uint32_t)));
    PS_TRY(ensure(e->aux[AUX_HAM_DIST], nq * k));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_HAM_CAT].p, catalog, n * 16, hipMemcpyHostToDevice, e->stream));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_HAM_Q].p, q_codes, nq * 16, hipMemcpyHostToDevice, e->stream));
    PS_TRY(enqueue_topk(e, as<uint64_t>(e->aux[AUX_HAM_CAT]), n, as<uint64_t>(e->aux[AUX_HAM_Q]), nq, k,
                        as<uint32_t>(e->aux[AUX_HAM_IDX]), as<uint8_t>(e->aux[AUX_HAM_DIST])));
    PS_HIP(hipMemcpyAsync(out_idx, e->aux[AUX_HAM_IDX].p, nq * k * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          e->stream));
    PS_HIP(hipMemcpyAsync(out_dist, e->aux[AUX_HAM_DIST].p, nq * k, hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

ps_status ps_rerank_l2(const float *q_desc, size_t nq, const float *cat_desc, size_t n, const uint32_t *shortlist,
                       uint32_t k, uint32_t *out_idx, float *out_d2) {
    PS_TRY(check_fail_injection());
    PS_TRY(check_topk_args(n, nq, k));
    if (nq && (!q_desc || !shortlist || !out_idx || !out_d2)) return set_err(PS_E_INVALID_ARG, "null argument");
    if (n && !cat_desc) return set_err(PS_E_INVALID_ARG, "null argument");
    if (nq == 0) return PS_OK;
    Borrow b;
    PS_TRY(b.st);
    ps_engine *e = b.e;
    PS_HIP(hipSetDevice(e->device));
    PS_TRY(ensure(e->aux[AUX_RR_QD], nq * 128 * sizeof(float)));
    PS_TRY(ensure(e->aux[AUX_RR_CD], std::max<size_t>(n, 1) * 128 * sizeof(float)));
    PS_TRY(ensure(e->aux[AUX_HAM_IDX], nq * k * sizeof(uint32_t)));
    PS_TRY(ensure(e->aux[AUX_RR_IDX], nq * k * sizeof(uint32_t)));
    PS_TRY(ensure(e->aux[AUX_RR_DIST], nq * k * sizeof(float)));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_RR_QD].p, q_desc, nq * 128 * sizeof(float), hipMemcpyHostToDevice, e->stream));
    if (n)
        PS_HIP(hipMemcpyAsync(e->aux[AUX_RR_CD].p, cat_desc, n * 128 * sizeof(float), hipMemcpyHostToDevice,
                              e->stream));
    PS_HIP(hipMemcpyAsync(e->aux[AUX_HAM_IDX].p, shortlist, nq * k * sizeof(uint32_t), hipMemcpyHostToDevice,
                          e->stream));
    PS_HIP(ps::launch_rerank_l2(as<float>(e->aux[AUX_RR_QD]), (uint32_t)nq, as<float>(e->aux[AUX_RR_CD]), (uint32_t)n,
                                as<uint32_t>(e->aux[AUX_HAM_IDX]), k, as<uint32_t>(e->aux[AUX_RR_IDX]),
                                as<float>(e->aux[AUX_RR_DIST]), e->stream));
    PS_HIP(hipMemcpyAsync(out_idx, e->aux[AUX_RR_IDX].p, nq * k * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          e->stream));
    PS_HIP(hipMemcpyAsync(out_d2, e->aux[AUX_RR_DIST].p, nq * k * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    PS_HIP(hipStreamSynchronize(e->stream));
    return PS_OK;
}

ps_status ps_engine_quantize_u128(ps_engine *e, const float *d_desc, size_t n, uint64_t *d_codes) {
    PS_TRY(check_fail_injection());
    if (!e || (n && (!d_desc || !d_codes))) return set_err(PS_E_INVALID_ARG, "null argument");
    if (n >= 0xffffffffull) return set_err(PS_E_INVALID_ARG, "too many descriptors");
    PS_HIP(hipSetDevice(e->device));
    PS_HIP(ps::launch_quantize_u128(d_desc, (uint32_t)n, d_codes, e->stream));
    return PS_OK;
}

ps_status ps_engine_hamming_topk(ps_engine *e, const uint64_t *d_catalog, size_t n, const uint64_t *d_q_codes,
                                 size_t nq, uint32_t k, uint32_t *d_idx, uint8_t *d_dist) {
    PS_TRY(check_fail_injection());
    if (!e) return set_err(PS_E_INVALID_ARG, "null engine");
    PS_TRY(check_topk_args(n, nq, k));
    if (nq && (!d_q_codes || !d_idx || !d_dist)) return set_err(PS_E_INVALID_ARG, "null argument");
    if (n && !d_catalog) return set_err(PS_E_INVALID_ARG, "null argument");
    if (nq == 0) return PS_OK;
    PS_HIP(hipSetDevice(e->device));
    if (n == 0) {
        PS_HIP(ps::launch_fill_u32(d_idx, 0xffffffffu, nq * k, e->stream));
        PS_HIP(hipMemsetAsync(d_dist, 0xff, nq * k, e->stream));
        return PS_OK;
    }
    return enqueue_topk(e, d_catalog, n, d_q_codes, nq, k, d_idx, d_dist);
}

ps_status ps_engine_rerank_l2(ps_engine *e, const float *d_q_desc, size_t nq, const float *d_cat_desc, size_t n,
                              const uint32_t *d_shortlist, uint32_t k, uint32_t *d_idx, float *d_d2) {
    PS_TRY(check_fail_injection());
    if (!e) return set_err(PS_E_INVALID_ARG, "null engine");
    PS_TRY(check_topk_args(n, nq, k));
    if (nq && (!d_q_desc || !d_shortlist || !d_idx || !d_d2)) return set_err(PS_E_INVALID_ARG, "null argument");
    if (nq == 0) return PS_OK;
    PS_HIP(hipSetDevice(e->device));
    PS_HIP(ps::launch_rerank_l2(d_q_desc, (uint32_t)nq, d_cat_desc, (uint32_t)n, d_shortlist, k, d_idx, d_d2,
                                e->stream));
    return PS_OK;
}

}  // extern "C"
