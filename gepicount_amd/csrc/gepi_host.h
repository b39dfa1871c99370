/*
 * gepi_host.h -- host-side state shared by the library's HIP translation
 * units: the resident panel, device tables and small helpers.
 */
#ifndef GEPI_HOST_H
#define GEPI_HOST_H

#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/gepicount.h"
#include "gepi_device.h"
#include "gepi_internal.h"
#include "gepi_score.h"
#include "gepi_stage.h"

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return gepi_fail(e_ == hipErrorOutOfMemory ? GEPI_ERR_NOMEM : GEPI_ERR_DEVICE, \
                             "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),        \
                             __FILE__, __LINE__);                                         \
    } while (0)

namespace gepi_host {

using gepi::Cand;
using gepi::CompressWord;
using gepi::Pend;
using gepi::ScanState;

struct DeviceTables {
    int32_t tmax = -1;
    double *lhi = nullptr, *llo = nullptr, *lfact = nullptr;
};

} // namespace gepi_host

struct gepi_panel {
    using DeviceTables = gepi_host::DeviceTables;
    int device = 0;
    int64_t n_var = 0;
    int32_t n_samples = 0;
    int32_t w64 = 0;
    uint64_t *d_nat = nullptr;
    hipStream_t stream = nullptr;
    /* per-scan workspace, grown on demand and reused */
    uint32_t *d_cls = nullptr;
    size_t cls_bytes = 0;
    uint64_t *d_clsk = nullptr;
    size_t clsk_bytes = 0;
    int32_t *d_perm = nullptr;
    size_t perm_bytes = 0;
    gepi::CompressWord *d_cw = nullptr; /* fp4 path: class compression words */
    size_t cw_bytes = 0;
    gepi::PairRec *d_park = nullptr; /* fp4 wave tiles, Fisher: parked pairs */
    size_t park_bytes = 0;
    gepi::Cand *d_cand[2] = {nullptr, nullptr};
    gepi::Pend *d_pend = nullptr;
    unsigned long long buf_cap = 0;  /* records in each candidate buffer */
    unsigned long long pend_cap = 0; /* records in the pending list       */
    gepi::ScanState *d_state = nullptr;
    uint64_t *d_small = nullptr;
    DeviceTables tabs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};


namespace gepi_host {

/* Makes `device` current; fails unless it is a gfx950. */
int set_device(int device);

/* Compute units of a device seen by set_device (256 if unknown). */
int32_t device_cus(int device);

/* Workgroups per CU of the persistent batched fp4 count launch (0: one
 * tile per workgroup); dflt unless GEPI_F4_WG_PER_CU overrides it. */
int f4_wg_per_cu(int dflt);

/* Grows the t*log2(t) and ln(t!) tables to cover t <= tmax. */
int ensure_tables(DeviceTables &tabs, int32_t tmax, hipStream_t stream);

template <typename T>
int grow(T **ptr, size_t *have, size_t need) {
    if (need <= *have && *ptr) return GEPI_OK;
    hipFree(*ptr);
    *ptr = nullptr;
    *have = 0;
    HIP_TRY(hipMalloc(ptr, need));
    *have = need;
    return GEPI_OK;
}


/* Sorts candidate records by the canonical key and writes the first
 * min(k, size) as results; returns how many were written.                   */
inline int64_t candidates_to_results(std::vector<gepi::Cand> &host, int64_t k, gepi_result *tops) {
    std::sort(host.begin(), host.end(), [](const gepi::Cand &a, const gepi::Cand &b) {
        const uint64_t ha = gepi::order_bits(a.score), hb = gepi::order_bits(b.score);
        return ha < hb || (ha == hb && a.idx < b.idx);
    });
    const int64_t n = std::min<int64_t>((int64_t)host.size(), k);
    for (int64_t i = 0; i < n; i++) {
        const gepi::Cand &c = host[(size_t)i];
        gepi_result &r = tops[i];
        r.sA = (int64_t)(c.idx >> 32);
        r.sB = (int64_t)((c.idx >> 3) & 0x1FFFFFFFull);
        r.fun = (int32_t)(c.idx & 7);
        r.w = (int32_t)c.w;
        r.x = (int32_t)c.x;
        r.y = (int32_t)c.y;
        r.z = (int32_t)c.z;
        r.obj = c.score;
        r.reserved = 0;
    }
    return n;
}

/* Class compression words of a scan: [2][2 * w64], class 0 then class 1;
 * mask bits at or beyond n_samples are dropped.  Returns (n0, n1).          */
inline void build_compress_words(const uint64_t *mask_ctrl, const uint64_t *mask_case,
                                 int32_t n_samples, int32_t w64,
                                 std::vector<gepi::CompressWord> &cw, int32_t &n0, int32_t &n1) {
    const int32_t nw32 = 2 * w64;
    cw.resize((size_t)2 * std::max(nw32, 1));
    int32_t cnt[2] = {0, 0};
    for (int c = 0; c < 2; c++) {
        const uint64_t *mk = c == 0 ? mask_ctrl : mask_case;
        for (int32_t j = 0; j < nw32; j++) {
            uint32_t m = 0;
            if (mk) m = (uint32_t)(mk[j >> 1] >> (32 * (j & 1)));
            const int32_t lo = 32 * j;
            if (lo >= n_samples)
                m = 0;
            else if (n_samples - lo < 32)
                m &= (1u << (n_samples - lo)) - 1u;
            cw[(size_t)c * nw32 + j] = gepi::make_compress_word(m);
            cnt[c] += (int32_t)cw[(size_t)c * nw32 + j].cnt;
        }
    }
    n0 = cnt[0];
    n1 = cnt[1];
}

} // namespace gepi_host

#endif
