/*
 * gepi_device.h -- device-side records, scan state and the fused per-pair
 * filter shared by the count kernels.
 */
#ifndef GEPI_DEVICE_H
#define GEPI_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gepicount.h"
#include "gepi_score.h"

namespace gepi {

constexpr int TV = 32;              /* popcount kernel: variants per tile side       */
constexpr int TM = 64;              /* MFMA kernel: variants per tile side           */
constexpr int SEL_SMALL = 1024;     /* radix select finishes in LDS below this size  */
constexpr int MFMA_LOW_MAX = 16383; /* class count limit of the 0x01-encoded class   */

/* A candidate for the top-k: exact score plus counts.  32 bytes. */
struct Cand {
    double score;
    uint64_t idx; /* sA << 32 | sB << 3 | fun */
    uint32_t w, x, y, z;
};

/* An evaluation that passed the single-precision filter.  16 bytes. */
struct Pend {
    uint16_t w, x, y, z;
    uint64_t idx;
};

struct ScanState {
    double thr_score;          /* admission key: (thr_score, thr_idx), inclusive   */
    uint64_t thr_idx;
    float filt_k;              /* entropy filter coefficient, 0 = admit everything */
    uint32_t overflow;         /* candidate buffer overflowed in this band          */
    unsigned long long count;  /* records in the active candidate buffer           */
    unsigned long long count2; /* records written by a compaction                  */
    unsigned long long pend_count;
    uint32_t pend_overflow;
    uint32_t pad1;
    /* radix select */
    uint64_t key_hi, key_lo;
    uint32_t bits_done, rank, matched, done, gather_n, pad0;
    uint32_t hist[256];
};

struct CountArgs {
    const uint32_t *planes;   /* popcount layout [n_var_pad][3][wc] dwords          */
    const uint64_t *planes_k; /* MFMA layout [n_ks][n_var_pad][3] 64-sample words   */
    int64_t tile_begin;       /* first linear tile of the band                      */
    int32_t n_tiles;          /* tiles in the band (= grid size)                    */
    int32_t n_var;
    int32_t n_var_pad;
    int32_t wc;               /* dwords per plane row (multiple of 32)              */
    int32_t w0;               /* class-0 dwords (multiple of 4)                     */
    int32_t wused;            /* dwords holding data (multiple of 4)                */
    int32_t n_ks;             /* 64-sample K steps (wused / 2)                      */
    int32_t ks0;              /* K steps of class 0                                 */
    int32_t cls0_high;        /* MFMA: class 0 encoded as 0x80 (else class 1 is)    */
    int32_t pad;
    ScanState *st;
    Pend *pend;
    unsigned long long pcap;
    const double *lfact;      /* ln t! (Fisher filter)                              */
};

__host__ __device__ inline void tile_of(int64_t t, int32_t &ta, int32_t &tb) {
    /* t = ta*(ta+1)/2 + tb, 0 <= tb <= ta */
    double r = sqrt(8.0 * (double)t + 1.0);
    int64_t a = (int64_t)((r - 1.0) * 0.5);
    while (a * (a + 1) / 2 > t) a--;
    while ((a + 1) * (a + 2) / 2 <= t) a++;
    ta = (int32_t)a;
    tb = (int32_t)(t - a * (a + 1) / 2);
}

/* Workgroup -> tile of the band.  Consecutive workgroups are dispatched to
 * the 8 XCDs round robin; each XCD gets a contiguous run of tiles so that
 * tiles sharing a row panel share that XCD's L2.                            */
__device__ __forceinline__ int64_t band_tile(int64_t tile_begin, int32_t n_tiles) {
    const int32_t gidx = blockIdx.x;
    const int32_t xcd = gidx & 7, slot = gidx >> 3;
    const int32_t per = n_tiles >> 3, rem = n_tiles & 7;
    return tile_begin + (int64_t)xcd * per + (xcd < rem ? xcd : rem) + slot;
}

__device__ __forceinline__ void push_pending(const CountArgs &g, int32_t w, int32_t x, int32_t y,
                                             int32_t z, uint64_t idx) {
    unsigned long long pos = atomicAdd(&g.st->pend_count, 1ull);
    if (pos < g.pcap) {
        Pend p;
        p.w = (uint16_t)w;
        p.x = (uint16_t)x;
        p.y = (uint16_t)y;
        p.z = (uint16_t)z;
        p.idx = idx;
        g.pend[pos] = p;
    } else {
        atomicOr(&g.st->pend_overflow, 1u);
    }
}

/* Per-scan filter constants, read once per workgroup. */
struct FilterConsts {
    double thr_s;
    float kf;
};

/* Filters the 20 (or 12 on the diagonal) evaluations of variant pair (A, B)
 * from the nine plane products of each class, c[cls][p][q] =
 * popc(planeA[p] & planeB[q] & class cls) with p, q in {allele1, allele2,
 * valid}.  Survivors are appended to the pending list.
 *
 * Entropy: -I <= thr needs I >= -thr =: It, and I_bits <= log2(1 + chi2/m)
 * (Jensen on KL), so a survivor must satisfy
 *     (wz - xy)^2 >= (2^It - 1) * y * z * xx * (m - xx),
 * evaluated in float with a 2^-16 relative margin (float error < 5e-7);
 * wz - xy == 0 is an independent table with I == 0.
 * Fisher: p >= P(observed table), so log10 P(obs) <= thr is required.      */
template <int FFUN, bool ROLLED>
__device__ __forceinline__ void filter_pair(const CountArgs &g, const FilterConsts &fc, int32_t A,
                                            int32_t B, const int32_t (&c)[2][3][3]) {
    const int32_t y = c[0][2][2], z = c[1][2][2];
    const int32_t m = y + z;
    const float kyz = fc.kf * (float)(y * z) * (1.0f - 1.0f / 65536.0f);
    /* row pairs (2A+ia, 2B+ib); on the diagonal (2A, 2A+1) is not a pair.  */
    auto row_pair = [&](int ia, int ib, int32_t n0, int32_t n1, int32_t a0, int32_t a1, int32_t b0,
                        int32_t b1) {
        const int32_t ws[5] = {n0, b0 - n0, a0 - n0, y - a0 - b0 + n0, a0 + b0 - 2 * n0};
        const int32_t xs[5] = {n1, b1 - n1, a1 - n1, z - a1 - b1 + n1, a1 + b1 - 2 * n1};
        const uint64_t base = pack_idx(2 * (int64_t)A + ia, 2 * (int64_t)B + ib, 0);
#pragma unroll
        for (int f = 0; f < 5; f++) {
            const int32_t w = ws[f], x = xs[f];
            bool pass;
            if (FFUN == GEPI_FFUNC_ENTROPY) {
                const int32_t xx = w + x;
                const int32_t q = xx * (m - xx);
                const int32_t d = w * z - x * y;
                const float fd = (float)d;
                pass = (fc.kf == 0.0f) || (d != 0 && fd * fd >= kyz * (float)q);
            } else {
                pass = true;
                if (fc.thr_s < 0.0) {
                    const double *lf = g.lfact;
                    double lp = ((lf[y] + lf[z]) + (lf[w + x] + lf[m - w - x])) -
                                (((lf[w] + lf[y - w]) + (lf[x] + lf[z - x])) + lf[m]);
                    pass = lp * 0.4342944819032518 - 1e-6 <= fc.thr_s;
                }
            }
            if (pass) push_pending(g, w, x, y, z, base | (uint64_t)f);
        }
    };
    if (ROLLED) {
        /* A short runtime loop keeps the code of callers with many pairs per
         * thread small.  The per-row-pair operands advance through rotating
         * registers (no indexed array, which the compiler would place in
         * scratch memory): row pair rp = 2*ia + ib.                        */
        int32_t n0a = c[0][0][0], n0b = c[0][0][1], n0c = c[0][1][0], n0d = c[0][1][1];
        int32_t n1a = c[1][0][0], n1b = c[1][0][1], n1c = c[1][1][0], n1d = c[1][1][1];
        int32_t a0a = c[0][0][2], a0b = c[0][1][2], a1a = c[1][0][2], a1b = c[1][1][2];
        int32_t b0a = c[0][2][0], b0b = c[0][2][1], b1a = c[1][2][0], b1b = c[1][2][1];
#pragma unroll 1
        for (int rp = 0; rp < 4; rp++) {
            const int ia = rp >> 1, ib = rp & 1;
            if (!(A == B && ia < ib)) row_pair(ia, ib, n0a, n1a, a0a, a1a, b0a, b1a);
            n0a = n0b; n0b = n0c; n0c = n0d;
            n1a = n1b; n1b = n1c; n1c = n1d;
            if (ib) { a0a = a0b; a1a = a1b; }
            int32_t t = b0a; b0a = b0b; b0b = t;
            t = b1a; b1a = b1b; b1b = t;
        }
    } else {
#pragma unroll
        for (int rp = 0; rp < 4; rp++) {
            const int ia = rp >> 1, ib = rp & 1;
            if (A == B && ia < ib) continue;
            row_pair(ia, ib, c[0][ia][ib], c[1][ia][ib], c[0][ia][2], c[1][ia][2], c[0][2][ib],
                     c[1][2][ib]);
        }
    }
}

} // namespace gepi

#endif
