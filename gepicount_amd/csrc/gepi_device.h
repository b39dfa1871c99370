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
constexpr int64_t TILE_EVALS_MAX = (int64_t)TM * TM * 20; /* evaluations of the largest tile */

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
    const uint64_t *planes_k; /* int8: [n_ks][n_var_pad][3] 64-sample words;
                               * fp4: [n_u][n_var_pad*3][4] uint4 super-step rows
                               * (K steps of 128 samples: 4 per super-step for
                               * single scans, 3 in the batch engine; see
                               * k_class_f4)                                      */
    const int32_t *pc3;       /* fp4 single scans: [2][n_var_pad*3] set bits of
                               * each plane row in its class's fourth K steps
                               * (k_class_f4); unused by the batch engine         */
    int64_t tile_begin;       /* first linear tile of the band                      */
    int32_t n_tiles;          /* tiles in the band (= grid size)                    */
    int32_t n_var;
    int32_t n_var_pad;
    int32_t wc;               /* dwords per plane row (multiple of 32)              */
    int32_t w0;               /* class-0 dwords (multiple of 4)                     */
    int32_t wused;            /* dwords holding data (multiple of 4)                */
    int32_t n_ks;             /* K steps: 64 samples (int8), 128 samples (fp4)      */
    int32_t ks0;              /* K steps of class 0 (fp4: class 1 starts at
                               * super-step f4_super(ks0))                         */
    int32_t cls0_high;        /* MFMA: class 0 encoded as 0x80 (else class 1 is)    */
    int32_t f32;              /* fp4: class sizes for which the filter may work on
                               * the fp32 counts (f32_counts)                    */
    int32_t n0, n1;           /* class sizes (bounds of y and z: Fisher filter)   */
    ScanState *st;
    Pend *pend;
    unsigned long long pcap;
    const double *lfact;      /* ln t! (Fisher filter)                              */
    struct PairRec *park;     /* fp4 wave tiles, Fisher: [park_slots][park_cap]
                               * regions for parked variant pairs (null: none)   */
    uint32_t *park_busy;      /* [park_slots] claim flags, then 8 claim counters  */
    uint32_t park_cap, park_slots;
};

/* The XCD (0-7) this wave runs on, from the hardware register. */
__device__ __forceinline__ uint32_t xcc_id() {
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7u; /* HW_REG_XCC_ID */
}

/* A parking region of CountArgs::park for this workgroup (thread 0): the
 * first free slot, from a rotating start, among the slots of the XCD it
 * runs on.  The XCDs' L2 caches are not coherent with one another, so a
 * region written and read through one XCD's L2 must not pass through
 * another's within a launch (a region reused across XCDs read stale
 * records when another XCD's dirty lines were written back over it).
 * Every claim is released before its workgroup ends and the host keeps
 * more slots per XCD than workgroups can be resident on one, so a free
 * slot is found; the probes are bounded all the same (PARK_NONE: no
 * region, the pairs are finished in the lane).                           */
constexpr uint32_t PARK_NONE = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t park_claim(uint32_t *busy, uint32_t slots) {
    const uint32_t x = xcc_id(), per = slots >> 3, base = x * per;
    uint32_t s = atomicAdd(&busy[slots + x], 1u) % per;
    for (uint32_t t = 0; t < 4 * per; t++) {
        if (atomicCAS(&busy[base + s], 0u, 1u) == 0u) return base + s;
        s = s + 1 == per ? 0u : s + 1;
    }
    return PARK_NONE;
}

__host__ __device__ inline void tile_of(int64_t t, int32_t &ta, int32_t &tb) {
    /* t = ta*(ta+1)/2 + tb, 0 <= tb <= ta */
    double r = sqrt(8.0 * (double)t + 1.0);
    int64_t a = (int64_t)((r - 1.0) * 0.5);
    while (a * (a + 1) / 2 > t) a--;
    while ((a + 1) * (a + 2) / 2 <= t) a++;
    ta = (int32_t)a;
    tb = (int32_t)(t - a * (a + 1) / 2);
}

/* Workgroup tile t of a scan with n_a A blocks: A variants [64ta, 64ta+64)
 * against B variants [32tb, 32tb+32), 0 <= tb <= 2ta+1 (row ta holds 2ta+2
 * tiles, so n_a(n_a+1) in all).  Tiles are numbered in panels of
 * F4_PANEL rows; panel rows [a0, a0+P) take indices [a0(a0+1),
 * (a0+P)(a0+P+1)), first the rectangle tb < 2a0+2 that all of them share,
 * column by column, then the rest of each row (row a0+i has 2i more tiles)
 * in row order.  Workgroups running together on one XCD take consecutive
 * tiles, so they share 8 A blocks and 8 B blocks in that XCD's L2; in row
 * order they shared one A block and read 64 B blocks that the next row
 * needs again only after the whole row has gone through.                 */
constexpr int F4_PANEL = 8;
__host__ __device__ inline void tile_of_f4(int64_t t, int32_t n_a, int32_t &ta, int32_t &tb) {
    double r = sqrt(4.0 * (double)t + 1.0);
    int64_t a = (int64_t)((r - 1.0) * 0.5);
    while (a * (a + 1) > t) a--;
    while ((a + 1) * (a + 2) <= t) a++;
    const int64_t a0 = a - a % F4_PANEL;
    const int64_t P = n_a - a0 < F4_PANEL ? n_a - a0 : F4_PANEL;
    const int64_t W = 2 * a0 + 2;
    const int64_t l = t - a0 * (a0 + 1);
    int64_t ra, rb;
    if (l < P * W) {
        rb = P == F4_PANEL ? l / F4_PANEL : l / P; /* a shift for a full panel */
        ra = a0 + (l - rb * P);
    } else {
        const int64_t e = l - P * W; /* i(i-1) tiles of the rest precede row a0+i */
        int64_t i = 1;
        while ((i + 1) * i <= e) i++;
        ra = a0 + i;
        rb = W + e - i * (i - 1);
    }
    ta = (int32_t)ra;
    tb = (int32_t)rb;
}

/* Linear workgroup tile of variant pair (a >= b).  Square tiles of side tv
 * (popcount, int8 MFMA): ta*(ta+1)/2 + tb.  fp4 tiles of 64 x 32 variants
 * (n_a4 > 0, the scan's A blocks): the inverse of tile_of_f4 at ta = a/64,
 * tb = b/32.                                                              */
__host__ __device__ inline int64_t tile_index(int64_t a, int64_t b, int32_t tv, int32_t n_a4) {
    if (n_a4 > 0) {
        const int64_t ta = a / 64, tb = b / 32;
        const int64_t a0 = ta - ta % F4_PANEL;
        const int64_t P = n_a4 - a0 < F4_PANEL ? n_a4 - a0 : F4_PANEL;
        const int64_t W = 2 * a0 + 2, i = ta - a0;
        return a0 * (a0 + 1) + (tb < W ? tb * P + i : P * W + i * (i - 1) + (tb - W));
    }
    const int64_t ta = a / tv, tb = b / tv;
    return ta * (ta + 1) / 2 + tb;
}

/* Workgroup -> tile of the band.  Consecutive workgroups are dispatched to
 * the 8 XCDs round robin; each XCD gets a contiguous run of tiles, so that
 * the tiles running together on it are neighbours in the tile order and
 * share panel rows in that XCD's L2.                                      */
__device__ __forceinline__ int64_t band_tile(int64_t tile_begin, int32_t n_tiles) {
    const int32_t gidx = blockIdx.x;
    const int32_t xcd = gidx & 7, slot = gidx >> 3;
    const int32_t per = n_tiles >> 3, rem = n_tiles & 7;
    return tile_begin + (int64_t)xcd * per + (xcd < rem ? xcd : rem) + slot;
}

/* Survivors of one workgroup are staged in LDS (LDS atomics) and written to
 * the pending list with one global reservation per workgroup; past
 * PEND_LDS they go to the pending list directly.                          */
constexpr int PEND_LDS = 256;
struct PendStage {
    Pend rec[PEND_LDS];
    uint32_t n;
    uint32_t pad;
    unsigned long long base;
};

/* Atomics on the scan state through global-address-space pointers: the
 * state pointer may come from a descriptor in memory, where the compiler
 * cannot tell its address space, and global atomics on a uniform address
 * let it combine a wave's atomics into one (generic ones it does not). */
typedef __attribute__((address_space(1))) unsigned long long gu64_t;
typedef __attribute__((address_space(1))) uint32_t gu32_t;
__device__ __forceinline__ unsigned long long st_add(unsigned long long *p, unsigned long long v) {
    return __hip_atomic_fetch_add((gu64_t *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_or(uint32_t *p, uint32_t v) {
    __hip_atomic_fetch_or((gu32_t *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void push_global(const CountArgs &g, const Pend &p) {
    unsigned long long pos = st_add(&g.st->pend_count, 1ull);
    if (pos < g.pcap)
        g.pend[pos] = p;
    else
        st_or(&g.st->pend_overflow, 1u);
}

__device__ __forceinline__ void push_pending(const CountArgs &g, PendStage &ps, int32_t w,
                                             int32_t x, int32_t y, int32_t z, uint64_t idx) {
    Pend p;
    p.w = (uint16_t)w;
    p.x = (uint16_t)x;
    p.y = (uint16_t)y;
    p.z = (uint16_t)z;
    p.idx = idx;
    const uint32_t pos = atomicAdd(&ps.n, 1u);
    if (pos < (uint32_t)PEND_LDS)
        ps.rec[pos] = p;
    else
        push_global(g, p);
}

/* Called by every thread of the workgroup at kernel start (before the first
 * barrier) and after the last push, respectively.                          */
__device__ __forceinline__ void stage_init(PendStage &ps) {
    if (threadIdx.x == 0) ps.n = 0;
}

/* The flush proper: every push is complete (the caller has passed a
 * barrier after the last push).                                           */
__device__ __forceinline__ void stage_flush_after_barrier(const CountArgs &g, PendStage &ps) {
    if (threadIdx.x == 0) {
        const uint32_t n = ps.n < (uint32_t)PEND_LDS ? ps.n : (uint32_t)PEND_LDS;
        ps.n = n;
        ps.base = n ? st_add(&g.st->pend_count, (unsigned long long)n) : 0ull;
    }
    __syncthreads();
    const uint32_t n = ps.n;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const unsigned long long pos = ps.base + i;
        if (pos < g.pcap)
            g.pend[pos] = ps.rec[i];
        else
            st_or(&g.st->pend_overflow, 1u);
    }
}

__device__ __forceinline__ void stage_flush(const CountArgs &g, PendStage &ps) {
    __syncthreads();
    stage_flush_after_barrier(g, ps);
}

/* Per-scan filter constants, read once per workgroup (filter_consts). */
struct FilterConsts {
    double thr_s; /* admission score threshold                                 */
    float kf;     /* entropy: chi-square coefficient, 0 = admit everything      */
    float a_s;    /* Fisher: -(thr + 1e-6) ln 10 - ln(4 n0 n1) / 2 + 0.9189,
                   * less a float error margin                                 */
};

__device__ __forceinline__ FilterConsts filter_consts(const ScanState *st, int32_t n0, int32_t n1) {
    FilterConsts fc;
    fc.thr_s = st->thr_score;
    fc.kf = st->filt_k;
    const double a = -(fc.thr_s + 1e-6) * 2.302585092994046;
    const double A = a - 0.5 * log(4.0 * (double)(n0 > 1 ? n0 : 1) * (double)(n1 > 1 ? n1 : 1)) +
                     0.9189;
    /* the float roundings of A and of the per-pair arithmetic stay far
     * under this margin, which keeps the derived bound conservative        */
    fc.a_s = fc.thr_s < 0.0 ? (float)(A - (fabs(A) + fabs(a) + 10.0) * 1e-6) : 0.0f;
    return fc;
}

/* Counts of the four row pairs (2A+ia, 2B+ib) of a variant pair, rp = 2ia+ib,
 * for function f: with n = N11, a = NA, b = NB and v = V of one class,
 *   f = 0 AND n, 1 NAND b-n, 2 ANDN a-n, 3 NANDN v-a-b+n, 4 XOR a+b-2n.      */
__device__ __forceinline__ int32_t fun_count(int f, int32_t n, int32_t a, int32_t b, int32_t v) {
    return f == 0 ? n : f == 1 ? b - n : f == 2 ? a - n : f == 3 ? v - a - b + n : a + b - 2 * n;
}

/* Fisher's exact check of one evaluation: log10 P(obs) <= thr (p >=
 * P(observed table)) from the log-factorial table, with the admission's
 * 1e-6 margin.                                                             */
__device__ __forceinline__ bool fisher_admits(const CountArgs &g, const FilterConsts &fc, int32_t w,
                                              int32_t x, int32_t y, int32_t z) {
    const int32_t m = y + z;
    const double *lf = g.lfact;
    const double lp = ((lf[y] + lf[z]) + (lf[w + x] + lf[m - w - x])) -
                      (((lf[w] + lf[y - w]) + (lf[x] + lf[z - x])) + lf[m]);
    return lp * 0.4342944819032518 - 1e-6 <= fc.thr_s;
}

/* Appends the evaluations whose bit 5*rp+f is set in mask to the pending
 * list.  Runs only for lanes with a survivor.  The row pair's operands
 * advance through rotating registers: selecting among the counts with a
 * runtime index would place them in scratch memory.  Fisher: the survivors
 * of the chi-square bound are checked against log10 P(obs) <= thr (p >=
 * P(observed table)) with the log-factorial table first.                 */
template <int FFUN>
__device__ __forceinline__ void push_mask(const CountArgs &g, PendStage &ps, const FilterConsts &fc,
                                          int32_t A, int32_t B, const int32_t (&c)[2][3][3],
                                          uint32_t mask) {
    const int32_t y = c[0][2][2], z = c[1][2][2];
    int32_t n0a = c[0][0][0], n0b = c[0][0][1], n0c = c[0][1][0], n0d = c[0][1][1];
    int32_t n1a = c[1][0][0], n1b = c[1][0][1], n1c = c[1][1][0], n1d = c[1][1][1];
    int32_t a0 = c[0][0][2], a1 = c[1][0][2];
    int32_t b0 = c[0][2][0], b1 = c[1][2][0], b0o = c[0][2][1], b1o = c[1][2][1];
#pragma unroll 1
    for (int rp = 0; rp < 4; rp++) {
        uint32_t mb = (mask >> (5 * rp)) & 31u;
        const uint64_t base = pack_idx(2 * (int64_t)A + (rp >> 1), 2 * (int64_t)B + (rp & 1), 0);
        while (mb) {
            const int f = __builtin_ctz(mb);
            mb &= mb - 1;
            const int32_t w = fun_count(f, n0a, a0, b0, y), x = fun_count(f, n1a, a1, b1, z);
            if (FFUN == GEPI_FFUNC_FISHERPV && fc.thr_s < 0.0 && !fisher_admits(g, fc, w, x, y, z))
                continue;
            push_pending(g, ps, w, x, y, z, base | (uint64_t)f);
        }
        n0a = n0b; n0b = n0c; n0c = n0d;
        n1a = n1b; n1b = n1c; n1c = n1d;
        if (rp == 1) { a0 = c[0][1][2]; a1 = c[1][1][2]; }
        int32_t t = b0; b0 = b0o; b0o = t;
        t = b1; b1 = b1o; b1o = t;
    }
}

/* Filters the 20 (or 12 on the diagonal) evaluations of variant pair (A, B)
 * from the nine plane products of each class, c[cls][p][q] =
 * popc(planeA[p] & planeB[q] & class cls) with p, q in {allele1, allele2,
 * valid}.  Survivors are staged for the pending list.
 *
 * Both objectives reduce to a chi-square test.  With I the mutual
 * information of the class x f table (nats) and chi2/m = D^2/(y z q),
 * D = wz - xy, q = xx(m - xx), Jensen on the KL divergence gives
 * I <= ln(1 + chi2/m), so I >= T requires D^2 >= expm1(T) * y * z * q.
 * Entropy: -I_bits <= thr needs T = -thr * ln 2 (kf = 2^-thr - 1, made on
 * the host side of the threshold update, rounded down).
 * Fisher: p >= P(obs) = C(y,w) C(z,x) / C(m,xx); e^{nH(k/n)} / sqrt(2n) <=
 * C(n,k) for the numerator and C(m,xx) <= e^{mH} sqrt(m / (2 pi xx (m-xx)))
 * (1 <= xx <= m-1; H in nats) give P(obs) >= e^{-mI} sqrt(2 pi (m-1) / m) /
 * (2 sqrt(yz)), so log10 P(obs) - 1e-6 <= thr needs
 *     T = (-(thr + 1e-6) ln 10 - ln(4yz) / 2 + 0.9189 - 1/m) / m
 * (y, z clamped to >= 1; ln(2 pi (m-1)/m) / 2 >= 0.9189 - 1/m);
 * its survivors are then checked exactly against the log-factorial table.
 * The test is s = fd^2 - RN(K*y*z*(1 - 2^-16) * q) > bias with fd = float(D)
 * squared exactly inside one fma.  fd is D exactly when |D| <= 2^24 (always
 * so on the float path) and within 2^-24 relative otherwise; the right side
 * is within 2^-21 relative; the 2^-16 shrink of K covers both, so every
 * table meeting the exact bound passes.  D == 0 (an independent table,
 * I == 0) fails whenever K > 0, and K == 0 (bias -1) admits everything.
 * D and xx are linear in the three per-row-pair quantities (N11, NA, NB):
 * up to sign, D_f = {Dn, Db-Dn, Da-Dn, (Da-Dn)+Db, (Da-Dn)+(Db-Dn)} with
 * Dq = q0*z - q1*y (only D^2 enters the test), and xx_f likewise from
 * S = q0 + q1.
 * A row pair is flagged when the largest of its five s exceeds the bias;
 * the bits of the functions are rebuilt, from the same expressions, only
 * for the lanes with a flagged row pair.
 *
 * ROLLED: a runtime loop over the four row pairs keeps the code of callers
 * with many pairs per thread small; its operands advance through rotating
 * registers (an indexed array would be placed in scratch memory).
 * DEFER: a variant pair with a flagged row pair is parked in pl (its counts)
 * for pair_drain instead of being finished in the lane.                    */
template <typename DT>
struct RowTest {
    float kyz, mf, bias;
    float kyzm; /* float counts: kyz * mf, exact (see filter_pair) */
    /* s of the five functions of one row pair.  Every D below is one
     * function's w z - x y up to sign, so |D| <= y z: exact for int32 and,
     * when y z <= 2^24, for float.                                         */
    __device__ __forceinline__ void s5(DT Dn, float Sn, DT Da, float Sa, DT Db, float Sb,
                                       float (&s)[5]) const {
        const DT Dnb = Db - Dn, Dna = Da - Dn;
        const DT D[5] = {Dn, Dnb, Dna, Dna + Db, Dna + Dnb};
        const float Sab = Sa + Sb;
        const float X[5] = {Sn, Sb - Sn, Sa - Sn, (mf - Sab) + Sn, Sab - 2.0f * Sn};
#pragma unroll
        for (int f = 0; f < 5; f++) {
            float r;
            if constexpr (sizeof(DT) == sizeof(float) && DT(0.5f) != DT(0)) {
                /* kyz (m - X) exactly in one fma, then one rounding */
                r = X[f] * __builtin_fmaf(-kyz, X[f], kyzm);
            } else {
                r = kyz * (X[f] * (mf - X[f]));
            }
            const float fd = (float)D[f];
            s[f] = __builtin_fmaf(fd, fd, -r);
        }
    }
    __device__ __forceinline__ bool any(DT Dn, float Sn, DT Da, float Sa, DT Db, float Sb) const {
        float s[5];
        s5(Dn, Sn, Da, Sa, Db, Sb, s);
        return fmaxf(fmaxf(fmaxf(s[0], s[1]), fmaxf(s[2], s[3])), s[4]) > bias;
    }
    __device__ __forceinline__ uint32_t bits(DT Dn, float Sn, DT Da, float Sa, DT Db, float Sb) const {
        float s[5];
        s5(Dn, Sn, Da, Sa, Db, Sb, s);
        uint32_t b = 0;
#pragma unroll
        for (int f = 0; f < 5; f++) b |= s[f] > bias ? (1u << f) : 0u;
        return b;
    }
};

/* Class sizes for which the fp4 filter works on its fp32 accumulators:
 * every product of a class-0 and a class-1 count exact in float (n0 n1 <=
 * 2^24) and m <= 16384 (the RowTest's exact kyz (m - X)).                  */
__host__ __device__ inline bool f32_counts(int64_t n0, int64_t n1) {
    return n0 * n1 <= ((int64_t)1 << 24) && n0 + n1 <= 16384;
}

/* Count arithmetic of the filter.  int32 counts (< 2^16): D = q0 z - q1 y
 * with 24-bit products.  float counts (the fp4 path's accumulators, exact
 * integers, f32_counts): every product q0 z, q1 y is at most n0 n1 <= 2^24
 * (a count of class c is at most n_c) and so exact, and fma gives D exactly
 * with no conversion.                                                     */
__device__ __forceinline__ int32_t dq_of(int32_t q0, int32_t q1, int32_t y, int32_t z) {
    return __mul24(q0, z) - __mul24(q1, y);
}
__device__ __forceinline__ float dq_of(float q0, float q1, float y, float z) {
    return __builtin_fmaf(q0, z, -(q1 * y));
}
__device__ __forceinline__ float sq_of(int32_t q0, int32_t q1) { return (float)(q0 + q1); }
__device__ __forceinline__ float sq_of(float q0, float q1) { return q0 + q1; }
__device__ __forceinline__ float yz_of(int32_t y, int32_t z) { return (float)__mul24(y, z); }
__device__ __forceinline__ float yz_of(float y, float z) { return y * z; }

struct PairList;
template <int FFUN, bool ROLLED, bool DEFER = false, typename CT, typename PL = PairList>
__device__ __forceinline__ void filter_pair(const CountArgs &g, PendStage &ps, const FilterConsts &fc,
                                            int32_t A, int32_t B, const CT (&c)[2][3][3],
                                            PL *pl = nullptr, const CT (*tot)[3] = nullptr,
                                            uint32_t rows_known = 0);

/* Variant pairs with a flagged row pair, parked in LDS: the lanes that
 * flag one are few and scattered over the waves, and finishing a pair (the
 * function bits, Fisher's exact check, the pending list) in the lane costs
 * each wave its full length for them.  After the tile the workgroup
 * finishes the parked pairs one per thread (pair_drain); past PAIR_LDS the
 * lane finishes at once.  Counts are below 2^16 (class sizes <= 46,340). */
constexpr int PAIR_LDS = 512;
struct PairRec {
    uint16_t c[2][3][3];
    int32_t A, B;
    uint32_t pad;
};
struct PairList {
    PairRec rec[PAIR_LDS];
    uint32_t n;
    uint32_t pad[3];
    __device__ __forceinline__ uint32_t cap() const { return PAIR_LDS; }
    __device__ __forceinline__ PairRec &at(uint32_t i) { return rec[i]; }
};

/* The same list with its records in global memory (a workgroup's own
 * region, L2-resident while the workgroup runs) and its counter in LDS. */
struct PairListG {
    uint32_t *n; /* LDS */
    PairRec *rec;
    uint32_t cap_;
    __device__ __forceinline__ uint32_t cap() const { return cap_; }
    __device__ __forceinline__ PairRec &at(uint32_t i) { return rec[i]; }
};
__device__ __forceinline__ uint32_t park_reserve(PairList *pl) { return atomicAdd(&pl->n, 1u); }
__device__ __forceinline__ uint32_t park_reserve(PairListG *pl) { return atomicAdd(pl->n, 1u); }

template <int FFUN, bool ROLLED, bool DEFER, typename CT, typename PL>
__device__ __forceinline__ void filter_pair(const CountArgs &g, PendStage &ps, const FilterConsts &fc,
                                            int32_t A, int32_t B, const CT (&c)[2][3][3],
                                            PL *pl, const CT (*tot)[3], uint32_t rows_known) {
    /* tot: the products' totals over both classes, c[0] + c[1], when the
     * caller holds them (the wave tiles' accumulators), else summed here */
    const CT y = c[0][2][2], z = c[1][2][2];
    const CT m = tot ? tot[2][2] : y + z;
    const float mf = (float)m;
    float K = fc.kf;
    if (FFUN == GEPI_FFUNC_FISHERPV) {
        K = 0.0f;
        if (fc.thr_s < 0.0) {
            /* T = (a - ln(4yz)/2 + 0.9189 - 1/m) / m at ln(4yz) <= ln(4 n0 n1)
             * (a_s; y <= n0, z <= n1), and K = T (1 + T/2) <= expm1(T): both
             * only lower K; 0.9189 - 1/m <= ln(2 pi (m-1)/m) / 2 comes from
             * C(m, xx) <= e^{m H} sqrt(m / (2 pi xx (m - xx))), 1 <= xx <= m-1
             * (the other tables have P(obs) = 1 and D = 0) */
            const float r = __builtin_amdgcn_rcpf(m > 1 ? mf : 1.0f);
            const float T = (fc.a_s - r) * r;
            if (T > 0.0f) K = T * (1.0f + 0.5f * T) * (1.0f - 2e-5f);
        }
    }
    RowTest<CT> rt;
    rt.kyz = K * yz_of(y, z) * (1.0f - 1.0f / 65536.0f);
    rt.mf = mf;
    rt.kyzm = 0.0f;
    if constexpr (sizeof(CT) == sizeof(float) && CT(0.5f) != CT(0)) {
        /* float counts come with m <= 16384 (14 bits): kyz cut down to 10
         * significant bits makes kyz m and kyz X exact, so that fma(-kyz, X,
         * kyz m) is kyz (m - X) exactly; cutting only lowers the bound */
        rt.kyz = __uint_as_float(__float_as_uint(rt.kyz) & 0xFFFFC000u);
        rt.kyzm = rt.kyz * mf;
    }
    rt.bias = K == 0.0f ? -1.0f : 0.0f;
    auto dq = [&](CT q0, CT q1) { return dq_of(q0, q1, y, z); };
    /* S of product (p, q): both classes' count */
    auto sq = [&](int p, int q) { return tot ? (float)tot[p][q] : sq_of(c[0][p][q], c[1][p][q]); };
    const CT Da0 = dq(c[0][0][2], c[1][0][2]), Da1 = dq(c[0][1][2], c[1][1][2]);
    const CT Db0 = dq(c[0][2][0], c[1][2][0]), Db1 = dq(c[0][2][1], c[1][2][1]);
    const float Sa0 = sq(0, 2), Sa1 = sq(1, 2);
    const float Sb0 = sq(2, 0), Sb1 = sq(2, 1);
    /* row pair rp = 2ia + ib flagged in bit rp; rows_known: flagged before
     * the pair was parked (that test is as conservative as this one) */
    uint32_t rows = 0;
    if (rows_known) {
        rows = rows_known;
    } else if (ROLLED) {
        CT n0a = c[0][0][0], n0b = c[0][0][1], n0c = c[0][1][0], n0d = c[0][1][1];
        CT n1a = c[1][0][0], n1b = c[1][0][1], n1c = c[1][1][0], n1d = c[1][1][1];
        CT Da = Da0, Db = Db0, Dbo = Db1;
        float Sa = Sa0, Sb = Sb0, Sbo = Sb1;
#pragma unroll 1
        for (int rp = 0; rp < 4; rp++) {
            rows |= rt.any(dq(n0a, n1a), sq_of(n0a, n1a), Da, Sa, Db, Sb) ? 1u << rp : 0u;
            n0a = n0b; n0b = n0c; n0c = n0d;
            n1a = n1b; n1b = n1c; n1c = n1d;
            if (rp == 1) { Da = Da1; Sa = Sa1; }
            CT t = Db; Db = Dbo; Dbo = t;
            float u = Sb; Sb = Sbo; Sbo = u;
        }
    } else {
        rows = (rt.any(dq(c[0][0][0], c[1][0][0]), sq(0, 0), Da0, Sa0, Db0, Sb0) ? 1u : 0u) |
               (rt.any(dq(c[0][0][1], c[1][0][1]), sq(0, 1), Da0, Sa0, Db1, Sb1) ? 2u : 0u) |
               (rt.any(dq(c[0][1][0], c[1][1][0]), sq(1, 0), Da1, Sa1, Db0, Sb0) ? 4u : 0u) |
               (rt.any(dq(c[0][1][1], c[1][1][1]), sq(1, 1), Da1, Sa1, Db1, Sb1) ? 8u : 0u);
    }
    if (A == B) rows &= ~2u; /* (2A, 2A+1) is not a row pair */
    if (DEFER && rows) {
        const uint32_t pos = park_reserve(pl);
        if (pos < pl->cap()) {
            PairRec &r = pl->at(pos);
#pragma unroll
            for (int k = 0; k < 2; k++)
#pragma unroll
                for (int p = 0; p < 3; p++)
#pragma unroll
                    for (int q = 0; q < 3; q++) r.c[k][p][q] = (uint16_t)c[k][p][q];
            r.A = A;
            r.B = B;
            r.pad = rows; /* the drain starts from these rows */
            return;
        }
    }
    if (rows) {
        /* rare: the function bits of the flagged row pairs */
        uint32_t mask = 0;
        if (rows & 1u)
            mask |= rt.bits(dq(c[0][0][0], c[1][0][0]), sq(0, 0), Da0, Sa0, Db0, Sb0);
        if (rows & 2u)
            mask |= rt.bits(dq(c[0][0][1], c[1][0][1]), sq(0, 1), Da0, Sa0, Db1, Sb1) << 5;
        if (rows & 4u)
            mask |= rt.bits(dq(c[0][1][0], c[1][1][0]), sq(1, 0), Da1, Sa1, Db0, Sb0) << 10;
        if (rows & 8u)
            mask |= rt.bits(dq(c[0][1][1], c[1][1][1]), sq(1, 1), Da1, Sa1, Db1, Sb1) << 15;
        int32_t ci[2][3][3];
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++) ci[k][p][q] = (int32_t)c[k][p][q];
        push_mask<FFUN>(g, ps, fc, A, B, ci, mask);
    }
}

/* Finishes the parked pairs, one per thread, with the int32 filter (as
 * conservative as the test that parked them, so it keeps every evaluation
 * that the exact top-k needs).  Every thread of the workgroup, after a
 * barrier that follows the last park; the caller clears pl.n after a later
 * barrier.                                                                 */
template <int FFUN>
__device__ __forceinline__ void pair_drain(const CountArgs &g, PendStage &ps, const FilterConsts &fc,
                                           const PairRec *rec, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const PairRec &r = rec[i];
        int32_t c[2][3][3];
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++) c[k][p][q] = r.c[k][p][q];
        filter_pair<FFUN, true>(g, ps, fc, r.A, r.B, c, (PairList *)nullptr,
                                (const int32_t(*)[3]) nullptr, r.pad);
    }
}

/* Waits until every global store of this wave has been acknowledged by L2
 * (s_waitcnt vmcnt(0); expcnt and lgkmcnt at their maxima: no wait). */
__device__ __forceinline__ void f4_wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

} // namespace gepi

#endif
