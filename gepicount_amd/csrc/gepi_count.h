/*
 * gepi_count.h -- the two all-pairs count kernels (replace k1_count /
 * k1_count_sm, src/kernels/kernel1.cu:60-284).  Both compute, for every
 * variant pair (A >= B) of a tile and each class, the nine plane products
 * c[cls][p][q] = popc(planeA[p] & planeB[q]) over the class's samples, then
 * run the fused filter (gepi_device.h) on the 20 evaluations they determine.
 *
 * k_count_popc  VALU path: 32x32-variant tiles staged in LDS in 32-dword K
 *               chunks; each thread owns 2x2 variant pairs and accumulates
 *               and+popcount over 32-bit words (classes are word ranges).
 * k_count_mfma  matrix-core path: the products are a Gram matrix of 0/1
 *               plane rows, computed with v_mfma_i32_16x16x64_i8 on bytes
 *               expanded from the packed bits in LDS.  One class is encoded
 *               as 0x01 and the other as 0x80 (-128), so one int32
 *               accumulator holds c_low + 16384 * c_high exactly (c_low <=
 *               16383): both classes in one pass over K.  64x64-variant
 *               workgroup tiles, 4 waves of 32x32, 64-sample K steps double
 *               buffered in LDS.
 * k_count_f4    matrix-core path on fp4 (e2m1) operands: the layout keeps 24
 *               samples per 32-bit word, at nibble bits 0-2, so that one AND
 *               per dword makes a fragment (a set bit at nibble bit p is
 *               0.5 * 2^p); v_mfma_scale_f32_16x16x128_f8f6f4 with block
 *               scale 2^(1-p) makes every product exactly 1.0, and the fp32
 *               accumulators hold exact counts (< 2^24).  Classes use
 *               separate accumulators (no class size limit).  64x32-variant
 *               workgroup tiles, 4 waves of 32x16, 384-sample super-steps
 *               (three K steps from the same words) double buffered in LDS.
 */
#ifndef GEPI_COUNT_H
#define GEPI_COUNT_H

#include "gepi_device.h"

namespace gepi {

/* ---------------------------------------------------------------------- */
/* popcount path                                                           */

constexpr int KC = 32;              /* dwords per K chunk staged in LDS          */
constexpr int VSTRIDE = 3 * KC + 4; /* LDS dwords per variant (bank spread)      */
constexpr int POPC_THREADS = 256;   /* 16 x 16                                   */

template <int CLS>
__device__ __forceinline__ void popc_step(const uint32_t *__restrict__ sA,
                                          const uint32_t *__restrict__ sB, int ty, int tx, int kk,
                                          uint32_t (&acc)[2][2][2][3][3]) {
    uint4 a[2][3], b[2][3];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int p = 0; p < 3; p++) {
            a[i][p] = *reinterpret_cast<const uint4 *>(sA + (ty + 16 * i) * VSTRIDE + p * KC + kk);
            b[i][p] = *reinterpret_cast<const uint4 *>(sB + (tx + 16 * i) * VSTRIDE + p * KC + kk);
        }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    uint32_t s = acc[CLS][i][j][p][q];
                    s += __popc(a[i][p].x & b[j][q].x);
                    s += __popc(a[i][p].y & b[j][q].y);
                    s += __popc(a[i][p].z & b[j][q].z);
                    s += __popc(a[i][p].w & b[j][q].w);
                    acc[CLS][i][j][p][q] = s;
                }
}

template <int FFUN>
__global__ void __launch_bounds__(POPC_THREADS) k_count_popc(CountArgs g) {
    __shared__ __attribute__((aligned(16))) uint32_t sA[TV * VSTRIDE];
    __shared__ __attribute__((aligned(16))) uint32_t sB[TV * VSTRIDE];
    __shared__ PendStage stage;
    stage_init(stage);
    int32_t ta, tb;
    tile_of(band_tile(g.tile_begin, g.n_tiles), ta, tb);

    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    uint32_t acc[2][2][2][3][3];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int p = 0; p < 3; p++)
#pragma unroll
                    for (int q = 0; q < 3; q++) acc[c][i][j][p][q] = 0;

    const uint32_t *baseA = g.planes + (int64_t)ta * TV * 3 * g.wc;
    const uint32_t *baseB = g.planes + (int64_t)tb * TV * 3 * g.wc;
    const int32_t nchunk = (g.wused + KC - 1) / KC;
    for (int32_t c = 0; c < nchunk; c++) {
        const int32_t c0 = c * KC;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 3; r++) {
            int idx = tid + r * POPC_THREADS; /* 0..767: (variant, plane, quad) */
            int row = idx >> 3, quad = idx & 7;
            int v = row / 3, p = row - 3 * v;
            int64_t goff = ((int64_t)v * 3 + p) * g.wc + c0 + quad * 4;
            int soff = v * VSTRIDE + p * KC + quad * 4;
            *reinterpret_cast<uint4 *>(sA + soff) = *reinterpret_cast<const uint4 *>(baseA + goff);
            *reinterpret_cast<uint4 *>(sB + soff) = *reinterpret_cast<const uint4 *>(baseB + goff);
        }
        __syncthreads();
        int32_t end = g.wused - c0;
        if (end > KC) end = KC;
        int32_t split = g.w0 - c0;
        if (split < 0) split = 0;
        if (split > end) split = end;
        for (int kk = 0; kk < split; kk += 4) popc_step<0>(sA, sB, ty, tx, kk, acc);
        for (int kk = split; kk < end; kk += 4) popc_step<1>(sA, sB, ty, tx, kk, acc);
    }

    const FilterConsts fc = filter_consts(g.st, g.n0, g.n1);
#pragma unroll
    for (int i = 0; i < 2; i++) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int32_t A = ta * TV + ty + 16 * i;
            const int32_t B = tb * TV + tx + 16 * j;
            if (A >= g.n_var || B > A) continue;
            int32_t c[2][3][3];
#pragma unroll
            for (int k = 0; k < 2; k++)
#pragma unroll
                for (int p = 0; p < 3; p++)
#pragma unroll
                    for (int q = 0; q < 3; q++) c[k][p][q] = (int32_t)acc[k][i][j][p][q];
            filter_pair<FFUN, false>(g, stage, fc, A, B, c);
        }
    }
    stage_flush(g, stage);
}

/* ---------------------------------------------------------------------- */
/* MFMA path                                                               */

typedef int32_t v4i __attribute__((ext_vector_type(4)));

constexpr int MF_THREADS = 256;  /* 4 waves, 2 x 2 of 32 x 32 variants          */
constexpr int MF_ROWS = 2 * TM * 3; /* LDS rows per K step: A then B, 3 planes  */
constexpr int MF_ROW = 64;       /* bytes per LDS row: 64 samples, unpadded     */

/* 32 packed samples -> 32 bytes at an LDS half row: bit b becomes byte b,
 * 0x01 (shift 0) or 0x80 (shift 7).  Nibble n spreads with
 * (n * 0x204081) & 0x01010101.                                            */
__device__ __forceinline__ void expand_half(uint8_t *__restrict__ dst, uint32_t bits, int shift) {
    uint32_t d[8];
#pragma unroll
    for (int i = 0; i < 8; i++) d[i] = ((((bits >> (4 * i)) & 0xFu) * 0x204081u) & 0x01010101u) << shift;
    *reinterpret_cast<uint4 *>(dst) = make_uint4(d[0], d[1], d[2], d[3]);
    *reinterpret_cast<uint4 *>(dst + 16) = make_uint4(d[4], d[5], d[6], d[7]);
}

template <int FFUN>
__global__ void __launch_bounds__(MF_THREADS, 2) k_count_mfma(CountArgs g) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[2][MF_ROWS * MF_ROW];
    __shared__ PendStage stage;
    stage_init(stage);
    int32_t ta, tb;
    tile_of(band_tile(g.tile_begin, g.n_tiles), ta, tb);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wa = wave >> 1, wb = wave & 1;
    const int l16 = lane & 15, lg = lane >> 4;

    v4i acc[2][2][3][3];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++) acc[i][j][p][q] = v4i{0, 0, 0, 0};

    /* Staging.  A K step holds 384 LDS rows of 64 samples: row e < 192 is A
     * variant e/3 plane e%3, row e >= 192 is B variant (e-192)/3.  Thread t
     * expands half rows h = t, t+256, t+512 (row h/2, samples 32*(h%2)..):
     * the 32-bit word at dword offset h of the step's A or B block.         */
    const uint32_t *srcA = reinterpret_cast<const uint32_t *>(g.planes_k) + (int64_t)ta * TM * 6;
    const uint32_t *srcB = reinterpret_cast<const uint32_t *>(g.planes_k) + (int64_t)tb * TM * 6;
    const int64_t sstride = (int64_t)g.n_var_pad * 6; /* dwords per K step */
    const uint32_t *src[3];
    int dst[3];
#pragma unroll
    for (int u = 0; u < 3; u++) {
        const int h = tid + 256 * u;                  /* 0..767 */
        src[u] = h < 384 ? srcA + h : srcB + (h - 384);
        dst[u] = (h >> 1) * MF_ROW + 32 * (h & 1);
    }
    const int last = g.n_ks > 0 ? g.n_ks - 1 : 0; /* n_ks == 0: no K loop, step 0 exists */
    uint32_t pf1[3], pf2[3];
    auto shift_of = [&](int s) { return ((s < g.ks0) == (g.cls0_high != 0)) ? 7 : 0; };

    {
        uint32_t r[3];
#pragma unroll
        for (int u = 0; u < 3; u++) r[u] = src[u][0];
        const int sh = shift_of(0);
#pragma unroll
        for (int u = 0; u < 3; u++) expand_half(&lds[0][dst[u]], r[u], sh);
        const int64_t o1 = (int64_t)(last < 1 ? last : 1) * sstride;
        const int64_t o2 = (int64_t)(last < 2 ? last : 2) * sstride;
#pragma unroll
        for (int u = 0; u < 3; u++) {
            pf1[u] = src[u][o1];
            pf2[u] = src[u][o2];
        }
    }
    __syncthreads();

    for (int32_t s = 0; s < g.n_ks; s++) {
        const int buf = s & 1;
        const uint8_t *L = lds[buf];
        v4i a[2][3], b[2][3];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int p = 0; p < 3; p++) {
                a[i][p] = *reinterpret_cast<const v4i *>(
                    L + ((32 * wa + 16 * i + l16) * 3 + p) * MF_ROW + 16 * lg);
                b[i][p] = *reinterpret_cast<const v4i *>(
                    L + (192 + (32 * wb + 16 * i + l16) * 3 + p) * MF_ROW + 16 * lg);
            }
        /* next step's expansion (harmless extra copy on the last step: the
         * other buffer is not read again) interleaved with this step's MFMAs */
        const int sh = shift_of(s + 1);
        uint8_t *W = lds[buf ^ 1];
#pragma unroll
        for (int u = 0; u < 3; u++) expand_half(&W[dst[u]], pf1[u], sh);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int p = 0; p < 3; p++)
#pragma unroll
                    for (int q = 0; q < 3; q++)
                        acc[i][j][p][q] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                            a[i][p], b[j][q], acc[i][j][p][q], 0, 0, 0);
        const int64_t on = (int64_t)(s + 3 <= last ? s + 3 : last) * sstride;
#pragma unroll
        for (int u = 0; u < 3; u++) {
            pf1[u] = pf2[u];
            pf2[u] = src[u][on];
        }
#pragma unroll
        for (int k = 0; k < 36; k++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); /* 1 MFMA      */
            __builtin_amdgcn_sched_group_barrier(0x002, 3, 0); /* 3 VALU      */
        }
        __builtin_amdgcn_sched_group_barrier(0x200, 6, 0);     /* DS writes   */
        __syncthreads();
    }

    /* epilogue: lane holds D[4*lg + r][l16] of each 16x16 block, i.e. A
     * variant 32wa + 16i + 4lg + r against B variant 32wb + 16j + l16.     */
    const FilterConsts fc = filter_consts(g.st, g.n0, g.n1);
    const bool c0_high = g.cls0_high != 0; /* class 0 held in the high part */
#pragma unroll
    for (int i = 0; i < 2; i++) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int32_t A = ta * TM + 32 * wa + 16 * i + 4 * lg + r;
                const int32_t B = tb * TM + 32 * wb + 16 * j + l16;
                if (A >= g.n_var || B > A) continue;
                int32_t c[2][3][3];
#pragma unroll
                for (int p = 0; p < 3; p++)
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        const int32_t S = acc[i][j][p][q][r];
                        const int32_t hi = S >> 14, lo = S & 16383;
                        c[0][p][q] = c0_high ? hi : lo;
                        c[1][p][q] = c0_high ? lo : hi;
                    }
                filter_pair<FFUN, true>(g, stage, fc, A, B, c);
            }
        }
    }
    stage_flush(g, stage);
}

/* ---------------------------------------------------------------------- */
/* fp4 MFMA path                                                           */

typedef int32_t v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int F4_THREADS = 256; /* 4 waves, 2 x 2 of 32 x 16 variants             */
constexpr int TB4 = 32;         /* B variants per workgroup tile (A: TM = 64)     */
constexpr int F4_RA = TM * 3;   /* A plane rows of a tile                         */
constexpr int F4_ROWS = F4_RA + TB4 * 3; /* plane rows of a tile: A, then B      */
constexpr int F4_UNITS = F4_ROWS * 4; /* uint4 of a staged super-step             */
constexpr int F4_FMT = 4;       /* operand format code of e2m1                    */

/* fp4 layout geometry: a super-step is 4 K steps of 128 samples (single
 * scans; the batch engine's layout has 3, f4_super3, see count_f4_run);
 * each class starts on a super-step, class 0 at 0 and class 1 at
 * f4_super(ks0).                                                          */
__host__ __device__ inline int32_t f4_super(int32_t ks) { return (ks + 3) / 4; }
__host__ __device__ inline int32_t f4_super3(int32_t ks) { return (ks + 2) / 3; }

/* LDS layout of a staged super-step: the 288 rows of 4 uint4 unpadded, uint4
 * c of row r at unit 4r + (c ^ ((r >> 2) & 3)).  A fragment read (16 lanes:
 * rows 3 apart, one column) and the staging stores (consecutive units) hit
 * distinct bank groups whether the LDS serves 8, 16 or 32 lanes of a
 * 16-byte access at a time.  Thread t of the workgroup stores units t + 256j
 * of the A rows and of the B rows; the uint4 it loads for them is row
 * t/4 + 64j, column (t%4) ^ ((t/16)%4), i.e. offset f4_src(t) + 256j.     */
__device__ __forceinline__ int f4_unit(int r, int c) { return 4 * r + (c ^ ((r >> 2) & 3)); }
__device__ __forceinline__ int f4_src(int t) { return 4 * (t >> 2) + ((t & 3) ^ ((t >> 4) & 3)); }

/* One lane's fp4 fragment of K step p < 3 of a super-step: the lane's four
 * layout dwords masked to nibble bit p.  A set bit is e2m1 0.5 * 2^p; with
 * block scale 2^(1-p) on both operands every product of two set bits is
 * exactly 1.0.  The order of samples within a K step does not change a dot
 * product, and A and B fragments use the same order.                      */
__device__ __forceinline__ v8i frag_f4(const uint4 &w, uint32_t m) {
    return v8i{(int32_t)(w.x & m), (int32_t)(w.y & m), (int32_t)(w.z & m), (int32_t)(w.w & m),
               0, 0, 0, 0};
}

/* K step 3 is the e2m1 sign bit: its fragment keeps nibble bit 3 and sets
 * bit 0, so every slot (padding included) is -0.5 when its sample bit is
 * set and +0.5 when it is clear.  With block scale 1 a product is
 * 0.25 (1 - 2a)(1 - 2b) = 1/4 - a/2 - b/2 + ab, so over the n3 slots of a
 * class's fourth K steps the accumulator gains n3/4 - pa/2 - pb/2 + AND (pa,
 * pb: the two rows' set bits there, from k_class_f4).  The accumulators
 * start at pa/2 + pb/2 - n3/4 (f4_corr_init) and end at the plain count;
 * every partial sum is a multiple of 1/4 far below 2^22, so exact.        */
__device__ __forceinline__ v8i frag_f4s(const uint4 &w, uint32_t ones) {
    return v8i{(int32_t)((w.x & 0x88888888u) | ones), (int32_t)((w.y & 0x88888888u) | ones),
               (int32_t)((w.z & 0x88888888u) | ones), (int32_t)((w.w & 0x88888888u) | ones), 0, 0,
               0, 0};
}

/* 0x11111111 in a vector register the optimizer cannot see through: with
 * the mask in a scalar register, (w & mask) | ones is then one
 * v_and_or_b32 (a three-operand instruction may read one scalar register
 * and no literal; with both constants it becomes an and and an or).      */
__device__ __forceinline__ uint32_t f4_ones_vgpr() {
    uint32_t ones = 0x11111111u;
    asm("" : "+v"(ones));
    return ones;
}

/* Word 3 of a gfx9 buffer resource: 32-bit data format, raw access. */
constexpr int F4_RSRC_WORD3 = 0x00020000;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 as_u4(u32x4 v) { return make_uint4(v[0], v[1], v[2], v[3]); }

/* Block scale (E8M0) of K step k: 2^(1-k) for k < 3, 1 for the sign step. */
__device__ __forceinline__ constexpr int f4_scale(int k) { return k < 3 ? 128 - k : 127; }

/* The fourth-K-step correction of a tile's 288 plane rows, per class: half
 * the row's set bits there (LDS, written before the tile's first barrier). */
struct F4Corr {
    float h[2][F4_ROWS];
};

/* Global plane row of tile row r (A rows first, then B rows). */
__device__ __forceinline__ int64_t f4_tile_row(int r, int32_t ta, int32_t tb) {
    return r < F4_RA ? (int64_t)ta * F4_RA + r : (int64_t)tb * (TB4 * 3) + (r - F4_RA);
}

/* Starting accumulators of class c (ks: its K steps): element (i, p, q, r)
 * is A row (32wa + 16i + 4lg + r)*3 + p against B row (16wb + l16)*3 + q. */
__device__ __forceinline__ void f4_corr_init(v4f (&acc)[2][3][3], const F4Corr &cr, int c,
                                             int32_t ks, int wa, int wb, int lg, int l16) {
    const float n3q = 32.0f * (float)(ks / 4); /* n3 / 4, n3 = 128 per full super-step */
    float hB[3];
#pragma unroll
    for (int q = 0; q < 3; q++) hB[q] = cr.h[c][F4_RA + (16 * wb + l16) * 3 + q] - n3q;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int p = 0; p < 3; p++) {
            float hA[4];
#pragma unroll
            for (int r = 0; r < 4; r++) hA[r] = cr.h[c][(32 * wa + 16 * i + 4 * lg + r) * 3 + p];
#pragma unroll
            for (int q = 0; q < 3; q++)
                acc[i][p][q] = v4f{hA[0] + hB[q], hA[1] + hB[q], hA[2] + hB[q], hA[3] + hB[q]};
        }
}

/* One 64 x 32-variant workgroup tile of the fp4 path (tile numbering as in
 * tile_of_f4).  Called by every thread of a 256-thread workgroup.  F32: the
 * filter works on the fp32 accumulators directly (the scan's class sizes
 * pass f32_counts), else on their int32 conversions.                       */
template <int FFUN, bool F32>
__device__ __forceinline__ void count_f4_tile(const CountArgs &g, int64_t tile) {
    __shared__ __attribute__((aligned(16))) uint4 lds[2][F4_UNITS];
    __shared__ PendStage stage;
    __shared__ F4Corr corr;
    /* Fisher flags a variant pair in about 13% of lanes: those pairs are
     * parked and finished one per thread after the tile (entropy's few are
     * finished in the lane, and its kernels hold no parking space) */
    constexpr bool PARK = FFUN == GEPI_FFUNC_FISHERPV;
    __shared__ std::conditional_t<PARK, PairList, uint32_t> park;
    PairList *pairs = PARK ? reinterpret_cast<PairList *>(&park) : nullptr;
    stage_init(stage);
    if (PARK && threadIdx.x == 0) pairs->n = 0;
    int32_t ta, tb;
    tile_of_f4(tile, g.n_var_pad / TM, ta, tb);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wa = wave >> 1, wb = wave & 1;
    const int l16 = lane & 15, lg = lane >> 4;

    /* super-steps: class 0 [0, u1), class 1 [u1, n_u) */
    const int32_t ks1 = g.n_ks - g.ks0;
    const int32_t u1 = f4_super(g.ks0), n_u = u1 + f4_super(ks1);
    const int32_t last = n_u > 0 ? n_u - 1 : 0; /* n_u == 0: no K loop, super-step 0 exists */

    /* Staging.  A super-step of a tile is 288 plane rows of 64 bytes (4
     * uint4): A rows 0..191 (variant 64ta + r/3, plane r%3) and B rows, both
     * contiguous in the layout; thread t stores LDS units t + 256j of the A
     * rows (j < 3) and of the B rows (t + 256j < 384), see f4_unit.        */
    /* the tile's rows start at workgroup-uniform addresses (scalar
     * registers); a thread adds only its own 32-bit byte offset, so the
     * loads take the scalar-base form with no 64-bit address arithmetic    */
    const char *lay = reinterpret_cast<const char *>(g.planes_k);
    const int64_t ustride_b = (int64_t)g.n_var_pad * 3 * 64; /* bytes per super-step */
    const char *tA = lay + (int64_t)ta * (F4_RA * 64);
    const char *tB = lay + (int64_t)tb * (TB4 * 3 * 64);
    const uint32_t so = (uint32_t)f4_src(tid) * 16u;
    const bool hasB1 = tid < 128;
    const int dA = tid, dB = F4_RA * 4 + tid;
    constexpr int D64 = 256; /* 64 rows further down */
    uint4 pa0, pa1, pa2, pb0, pb1 = make_uint4(0u, 0u, 0u, 0u);
    auto fetch = [&](int32_t u) {
        /* buffer loads: the super-step's A and B rows as two scalar buffer
         * resources (12,288 and 6,144 bytes), each thread adding its own
         * 32-bit offset; no vector address arithmetic per super-step      */
        const int64_t o = (int64_t)u * ustride_b;
        const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char *>(tA + o), (short)0, F4_RA * 64, F4_RSRC_WORD3);
        const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char *>(tB + o), (short)0, TB4 * 3 * 64, F4_RSRC_WORD3);
        pa0 = as_u4(__builtin_amdgcn_raw_buffer_load_b128(ra, (int)so, 0, 0));
        pa1 = as_u4(__builtin_amdgcn_raw_buffer_load_b128(ra, (int)so + 4096, 0, 0));
        pa2 = as_u4(__builtin_amdgcn_raw_buffer_load_b128(ra, (int)so + 8192, 0, 0));
        pb0 = as_u4(__builtin_amdgcn_raw_buffer_load_b128(rb, (int)so, 0, 0));
        if (hasB1) pb1 = as_u4(__builtin_amdgcn_raw_buffer_load_b128(rb, (int)so + 4096, 0, 0));
    };
    auto put = [&](uint4 *W) {
        W[dA] = pa0;
        W[dA + D64] = pa1;
        W[dA + 2 * D64] = pa2;
        W[dB] = pb0;
        if (hasB1) W[dB + D64] = pb1;
    };
    fetch(0);
    /* the fourth-K-step corrections of the tile's rows (none without samples:
     * the class layout was not built) */
    const int64_t prow = (int64_t)g.n_var_pad * 3;
    for (int r = tid; r < F4_ROWS; r += F4_THREADS) {
        const int64_t gr = f4_tile_row(r, ta, tb);
        corr.h[0][r] = g.n_ks > 0 ? 0.5f * (float)g.pc3[gr] : 0.0f;
        corr.h[1][r] = g.n_ks > 0 ? 0.5f * (float)g.pc3[prow + gr] : 0.0f;
    }
    put(lds[0]);
    fetch(last < 1 ? last : 1);
    __syncthreads();

    /* lane's fragment rows: A (i, p) is row (32wa + 16i + l16)*3 + p, B q is
     * row 192 + (16wb + l16)*3 + q; the lane takes 16-byte column lg.      */
    int oA[3], oB[3]; /* A (i, p) at oA[p] + 192 i: 48 rows keep the swizzle */
#pragma unroll
    for (int p = 0; p < 3; p++) {
        oA[p] = f4_unit((32 * wa + l16) * 3 + p, lg);
        oB[p] = f4_unit(F4_RA + (16 * wb + l16) * 3 + p, lg);
    }
    const uint32_t ones = f4_ones_vgpr();
    /* super-step u into one class's accumulators, its first np K steps
     * (the rest of a class's last super-step is padding); the staging
     * pipeline runs on across the two calls.                               */
    auto step = [&](v4f (&acc)[2][3][3], int32_t u, int32_t np) {
        const uint4 *L = lds[u & 1];
        uint4 wA[2][3], wB[3];
        /* B first: the first MFMA needs B and A[0][0] only */
#pragma unroll
        for (int q = 0; q < 3; q++) wB[q] = L[oB[q]];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int p = 0; p < 3; p++) wA[i][p] = L[oA[p] + 192 * i];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k < np) {
                const uint32_t m = 0x11111111u << k;
                v8i b[3];
#pragma unroll
                for (int q = 0; q < 3; q++) b[q] = k < 3 ? frag_f4(wB[q], m) : frag_f4s(wB[q], ones);
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int p = 0; p < 3; p++) {
                        const v8i a = k < 3 ? frag_f4(wA[i][p], m) : frag_f4s(wA[i][p], ones);
#pragma unroll
                        for (int q = 0; q < 3; q++)
                            acc[i][p][q] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                                a, b[q], acc[i][p][q], F4_FMT, F4_FMT, 0, f4_scale(k), 0,
                                f4_scale(k));
                    }
            }
        }
        /* stage super-step u+1 (a harmless extra copy on the last one: that
         * buffer is not read again), prefetch u+2.  After the MFMAs, so that
         * a wave waiting for its loads has its MFMAs already issued; the
         * other buffer's last reads were in super-step u-1, before the
         * barrier that ended it.                                           */
        put(lds[(u & 1) ^ 1]);
        fetch(u + 2 <= last ? u + 2 : last);
        __syncthreads();
    };
    v4f acc0[2][3][3], acc1[2][3][3];
    f4_corr_init(acc0, corr, 0, g.ks0, wa, wb, lg, l16);
    for (int32_t u = 0; u < u1; u++) {
        const int32_t r = g.ks0 - 4 * u;
        step(acc0, u, r < 4 ? r : 4);
    }
    f4_corr_init(acc1, corr, 1, ks1, wa, wb, lg, l16);
    for (int32_t u = u1; u < n_u; u++) {
        const int32_t r = ks1 - 4 * (u - u1);
        step(acc1, u, r < 4 ? r : 4);
    }

    /* epilogue: lane holds D[4*lg + r][l16] of each 16x16 block, i.e. A
     * variant 32wa + 16i + 4lg + r against B variant 16wb + l16.           */
    const FilterConsts fc = filter_consts(g.st, g.n0, g.n1);
#pragma unroll
    for (int i = 0; i < 2; i++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int32_t A = ta * TM + 32 * wa + 16 * i + 4 * lg + r;
            const int32_t B = tb * TB4 + 16 * wb + l16;
            if (A >= g.n_var || B > A) continue;
            if constexpr (F32) {
                float c[2][3][3];
#pragma unroll
                for (int k = 0; k < 2; k++)
#pragma unroll
                    for (int p = 0; p < 3; p++)
#pragma unroll
                        for (int q = 0; q < 3; q++)
                            c[k][p][q] = k == 0 ? acc0[i][p][q][r] : acc1[i][p][q][r];
                filter_pair<FFUN, false, PARK>(g, stage, fc, A, B, c, pairs);
            } else {
                int32_t c[2][3][3];
#pragma unroll
                for (int k = 0; k < 2; k++)
#pragma unroll
                    for (int p = 0; p < 3; p++)
#pragma unroll
                        for (int q = 0; q < 3; q++)
                            c[k][p][q] = (int32_t)(k == 0 ? acc0[i][p][q][r] : acc1[i][p][q][r]);
                filter_pair<FFUN, false, PARK>(g, stage, fc, A, B, c, pairs);
            }
        }
    }
    if constexpr (PARK) {
        __syncthreads();
        pair_drain<FFUN>(g, stage, fc, pairs->rec,
                         pairs->n < (uint32_t)PAIR_LDS ? pairs->n : (uint32_t)PAIR_LDS);
    }
    stage_flush(g, stage);
}

/* Persistent fp4 count loop (batched scans, whose tiles are short and come
 * from many scans).  Its layout has three K steps per super-step (nibble
 * bits 0-2, f4_super3): the sign-bit fourth step, which needs the
 * per-row correction, measured 11% slower on batches of Fisher scans
 * (larger code in this already large kernel) and no faster on entropy
 * batches.  A launch has a few workgroups per CU; each takes tiles
 * from the work counter of its XCD (workgroups are
 * dispatched to the 8 XCDs round robin, XCD = blockIdx.x % 8) until that
 * XCD's share of the band is done.  Per tile: the super-step pipeline into
 * the class accumulators, then the fused filter; the first super-step of
 * the next tile is fetched before the filter runs, so its latency is
 * hidden.  Survivors stay staged in LDS across tiles of the same scan
 * until the stage is half full.
 *
 * Src provides grab(n) (thread 0 only; n = the workgroup's earlier grabs:
 * its next job, ok = 0 when there is none) and argp(item), the scan's count
 * arguments in global memory.  The arguments are read through that pointer
 * (uniform, into a kernel-argument table): the compiler then keeps them in
 * scalar registers and combines a wave's atomic adds on the pending-list
 * counter into one, which matters when many evaluations survive.
 * Work counters come in two sets of 8 used by alternate launches; each
 * launch zeroes the other set for the next one (no memset, capture safe).  */
struct F4Job {
    int64_t tile;
    int32_t item;
    int32_t ok;
};

/* A job read back from LDS by every lane, made wave-uniform (scalar
 * registers) for the compiler. */
__device__ __forceinline__ F4Job uniform_job(const F4Job &j) {
    F4Job o;
    const uint64_t t = (uint64_t)j.tile;
    o.tile = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(t >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)t));
    o.item = __builtin_amdgcn_readfirstlane(j.item);
    o.ok = __builtin_amdgcn_readfirstlane(j.ok);
    return o;
}

template <int FFUN, bool HALF, class Src>
__device__ __forceinline__ void count_f4_run(const Src &src) {
    /* HALF: a job is half of a 64 x 32-variant tile, its A variants 32h ..
     * 32h + 31 (job unit 2 * tile + h), and each wave takes 16 x 16
     * variants (one A block); else a job is the whole tile and each wave
     * 32 x 16 (two A blocks).  Same tile numbering and bands either way. */
    constexpr int IB = HALF ? 1 : 2;        /* A blocks of 16 variants per wave */
    constexpr int RA = HALF ? F4_RA / 2 : F4_RA; /* A plane rows of a job        */
    constexpr int UNITS = (RA + TB4 * 3) * 4;    /* uint4 of a staged super-step  */
    __shared__ __attribute__((aligned(16))) uint4 lds[2][UNITS];
    __shared__ PendStage stage;
    __shared__ F4Job s_job[2];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wa = wave >> 1, wb = wave & 1;
    const int l16 = lane & 15, lg = lane >> 4;

    /* Staging.  A super-step of a job is RA + 96 plane rows of 64 bytes (4
     * uint4): A rows 0..RA-1 (plane r%3 of the job's A variant r/3) and B
     * rows, both contiguous in the layout; thread t stores LDS units t +
     * 256j of the A rows (t + 256j < 4 RA) and of the B rows (t + 256j <
     * 384), see f4_unit (RA and 192 are multiples of 16 rows, so the
     * swizzle of f4_src holds for every block).                           */
    const int si = f4_src(tid);
    const bool hasB1 = tid < 128;
    const bool hasA1 = !HALF || tid < 128;
    const int dA = tid, dB = RA * 4 + tid;
    constexpr int D64 = 256; /* 64 rows further down */
    /* lane's fragment rows: A (i, p) is row (16 IB wa + 16i + l16)*3 + p, B q
     * is row RA + (16wb + l16)*3 + q; the lane takes 16-byte column lg.    */
    int oA[3], oB[3]; /* A (i, p) at oA[p] + 192 i: 48 rows keep the swizzle */
#pragma unroll
    for (int p = 0; p < 3; p++) {
        oA[p] = f4_unit((16 * IB * wa + l16) * 3 + p, lg);
        oB[p] = f4_unit(RA + (16 * wb + l16) * 3 + p, lg);
    }

    /* geometry of one tile */
    struct Geo {
        const char *tA, *tB; /* the tile's A and B rows of super-step 0 (uniform) */
        int64_t ustride_b;   /* bytes per super-step */
        int32_t ta, tb, a0, u1, n_u, last, ks1; /* a0: the job's first A variant in the tile */
    };
    const uint32_t so = (uint32_t)si * 16u;
    auto geo_of = [&](const CountArgs &g, int64_t unit) {
        Geo o;
        const int64_t tile = HALF ? unit >> 1 : unit;
        tile_of_f4(tile, g.n_var_pad / TM, o.ta, o.tb);
        o.a0 = HALF ? (int32_t)(unit & 1) * 32 : 0;
        const char *lay = reinterpret_cast<const char *>(g.planes_k);
        o.ustride_b = (int64_t)g.n_var_pad * 3 * 64;
        o.tA = lay + ((int64_t)o.ta * TM + o.a0) * (3 * 64);
        o.tB = lay + (int64_t)o.tb * (TB4 * 3 * 64);
        if constexpr (FFUN != GEPI_FFUNC_ENTROPY) {
            /* Fisher: per-thread pointers (see fetch) */
            o.tA += so;
            o.tB += so;
        }
        o.ks1 = g.n_ks - g.ks0;
        o.u1 = f4_super3(g.ks0);
        o.n_u = o.u1 + f4_super3(o.ks1);
        o.last = o.n_u > 0 ? o.n_u - 1 : 0; /* n_u == 0: no K loop, super-step 0 exists */
        return o;
    };
    uint4 pa0, pa1, pa2, pb0, pb1 = make_uint4(0u, 0u, 0u, 0u);
    auto fetch = [&](const Geo &o, int32_t u) {
        const int64_t off = (int64_t)(u < o.last ? u : o.last) * o.ustride_b;
        if constexpr (FFUN != GEPI_FFUNC_ENTROPY) {
            /* Fisher: plain loads through per-thread pointers.  The buffer
             * form measured 8% slower on Fisher batches, whose kernel is at
             * the register limit (it moved spills into the tile loop).     */
            const uint4 *a = reinterpret_cast<const uint4 *>(o.tA + off);
            const uint4 *b = reinterpret_cast<const uint4 *>(o.tB + off);
            pa0 = a[0];
            if (hasA1) pa1 = a[256];
            if (!HALF) pa2 = a[512];
            pb0 = b[0];
            if (hasB1) pb1 = b[256];
            return;
        }
        /* scalar buffer resources over the super-step's A and B rows, as in
         * count_f4_tile */
        const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char *>(o.tA + off), (short)0, RA * 64, F4_RSRC_WORD3);
        const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char *>(o.tB + off), (short)0, TB4 * 3 * 64, F4_RSRC_WORD3);
        pa0 = as_u4(__builtin_amdgcn_raw_buffer_load_b128(ra, (int)so, 0, 0));
        if (hasA1) pa1 = as_u4(__builtin_amdgcn_raw_buffer_load_b128(ra, (int)so + 4096, 0, 0));
        if (!HALF) pa2 = as_u4(__builtin_amdgcn_raw_buffer_load_b128(ra, (int)so + 8192, 0, 0));
        pb0 = as_u4(__builtin_amdgcn_raw_buffer_load_b128(rb, (int)so, 0, 0));
        if (hasB1) pb1 = as_u4(__builtin_amdgcn_raw_buffer_load_b128(rb, (int)so + 4096, 0, 0));
    };
    auto put = [&](uint4 *W) {
        W[dA] = pa0;
        if (hasA1) W[dA + D64] = pa1;
        if (!HALF) W[dA + 2 * D64] = pa2;
        W[dB] = pb0;
        if (hasB1) W[dB + D64] = pb1;
    };

    if (tid == 0) {
        s_job[0] = src.grab(0);
        stage.n = 0;
    }
    __syncthreads();
    F4Job job = uniform_job(s_job[0]);
    if (!job.ok) return;
    const CountArgs *gp = src.argp(job.item);
    Geo geo = geo_of(*gp, job.tile);
    fetch(geo, 0);
    put(lds[0]);
    fetch(geo, 1);
    if (tid == 0) s_job[1] = src.grab(1);
    __syncthreads();
    if (geo.n_u == 0) __syncthreads();

    for (int32_t t = 0;; t++) {
        const CountArgs &g = *gp;
        /* super-step u into one class's accumulators, its first np K steps
         * (the rest of a class's last super-step is padding); the staging
         * pipeline runs on across the two phases.                          */
        auto step = [&](v4f (&acc)[IB][3][3], int32_t u, int32_t np) {
            const uint4 *L = lds[u & 1];
            uint4 wA[IB][3], wB[3];
            /* B first: the first MFMA needs B and A[0][0] only */
#pragma unroll
            for (int q = 0; q < 3; q++) wB[q] = L[oB[q]];
#pragma unroll
            for (int i = 0; i < IB; i++)
#pragma unroll
                for (int p = 0; p < 3; p++) wA[i][p] = L[oA[p] + 192 * i];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                if (k < np) {
                    const uint32_t m = 0x11111111u << k;
                    v8i b[3];
#pragma unroll
                    for (int q = 0; q < 3; q++) b[q] = frag_f4(wB[q], m);
#pragma unroll
                    for (int i = 0; i < IB; i++)
#pragma unroll
                        for (int p = 0; p < 3; p++) {
                            const v8i a = frag_f4(wA[i][p], m);
#pragma unroll
                            for (int q = 0; q < 3; q++)
                                acc[i][p][q] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                                    a, b[q], acc[i][p][q], F4_FMT, F4_FMT, 0, 128 - k, 0,
                                    128 - k);
                        }
                }
            }
            /* stage super-step u+1 (a harmless extra copy on the last one:
             * that buffer is not read again), prefetch u+2; after the MFMAs,
             * as in count_f4_tile */
            put(lds[(u & 1) ^ 1]);
            fetch(geo, u + 2);
            __syncthreads();
        };
        v4f acc0[IB][3][3], acc1[IB][3][3];
#pragma unroll
        for (int i = 0; i < IB; i++)
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++) acc0[i][p][q] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int32_t u = 0; u < geo.u1; u++) {
            const int32_t r = g.ks0 - 3 * u;
            step(acc0, u, r < 3 ? r : 3);
        }
#pragma unroll
        for (int i = 0; i < IB; i++)
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++) acc1[i][p][q] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int32_t u = geo.u1; u < geo.n_u; u++) {
            const int32_t r = geo.ks1 - 3 * (u - geo.u1);
            step(acc1, u, r < 3 ? r : 3);
        }

        /* the filter's constants are read before the next tile's loads are
         * issued: waits on global loads are in issue order */
        const FilterConsts fc = filter_consts(g.st, g.n0, g.n1);
        /* next job (written before this tile's first barrier) and its first
         * super-step, in flight during the filter */
        const F4Job nj = uniform_job(s_job[(t + 1) & 1]);
        const CountArgs *gpn = gp;
        Geo geon;
        if (nj.ok) {
            gpn = src.argp(nj.item);
            geon = geo_of(*gpn, nj.tile);
            fetch(geon, 0);
        }

        /* filter: lane holds D[4*lg + r][l16] of each 16x16 block, i.e. A
         * variant a0 + 16 IB wa + 16i + 4lg + r against B variant 16wb + l16. */
#pragma unroll
        for (int i = 0; i < IB; i++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int32_t A = geo.ta * TM + geo.a0 + 16 * IB * wa + 16 * i + 4 * lg + r;
                const int32_t B = geo.tb * TB4 + 16 * wb + l16;
                if (A >= g.n_var || B > A) continue;
                if (g.f32) {
                    float c[2][3][3];
#pragma unroll
                    for (int k = 0; k < 2; k++)
#pragma unroll
                        for (int p = 0; p < 3; p++)
#pragma unroll
                            for (int q = 0; q < 3; q++)
                                c[k][p][q] = k == 0 ? acc0[i][p][q][r] : acc1[i][p][q][r];
                    filter_pair<FFUN, false>(g, stage, fc, A, B, c);
                } else {
                    int32_t c[2][3][3];
#pragma unroll
                    for (int k = 0; k < 2; k++)
#pragma unroll
                        for (int p = 0; p < 3; p++)
#pragma unroll
                            for (int q = 0; q < 3; q++)
                                c[k][p][q] =
                                    (int32_t)(k == 0 ? acc0[i][p][q][r] : acc1[i][p][q][r]);
                    filter_pair<FFUN, false>(g, stage, fc, A, B, c);
                }
            }
        }
        if (!nj.ok) break;
        /* the staged survivors go out when the scan changes or the stage is
         * half full (thread 0 may clamp stage.n while others read it here;
         * both values give the same answer)                               */
        __syncthreads();
        const bool flush = nj.item != job.item || stage.n >= (uint32_t)(PEND_LDS / 2);
        if (flush) stage_flush_after_barrier(g, stage);
        put(lds[0]);
        fetch(geon, 1);
        __syncthreads(); /* super-step 0 staged; the flush's copies are done */
        if (tid == 0) {
            if (flush) stage.n = 0;
            s_job[t & 1] = src.grab(t + 2);
        }
        job = nj;
        gp = gpn;
        geo = geon;
        if (geo.n_u == 0) __syncthreads();
    }
    stage_flush(*gp, stage);
}

/* Work counters of a launch: set (parity) of two sets of 8. */
__device__ __forceinline__ void zero_other_counters(uint32_t *wq, int32_t parity) {
    if (blockIdx.x == 0 && threadIdx.x < 8) wq[8 * (parity ^ 1) + threadIdx.x] = 0u;
}

/* One tile of the band per workgroup (tiles as band_tile maps them). */
template <int FFUN, bool F32>
__global__ void __launch_bounds__(F4_THREADS, 2) k_count_f4(CountArgs g) {
    count_f4_tile<FFUN, F32>(g, band_tile(g.tile_begin, g.n_tiles));
}

/* pair_drain over records in global memory written by other waves of the
 * workgroup after they waited for their stores (f4_wait_stores) and a
 * barrier: each record read as three 16-byte buffer loads with the sc0
 * (workgroup-coherent) policy, which miss the CU's L1 (it may hold the
 * region's lines from a workgroup that used it earlier) and read the XCD's
 * L2, where the stores are.                                               */
template <int FFUN>
__device__ __forceinline__ void pair_drain_l2(const CountArgs &g, PendStage &ps,
                                              const FilterConsts &fc, const PairRec *rec,
                                              uint32_t n, uint32_t cap) {
    static_assert(sizeof(PairRec) == 48, "a parked pair is three 16-byte loads");
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<PairRec *>(rec), (short)0, (int)(cap * sizeof(PairRec)), 0x00020000);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        u32x4 w[3];
#pragma unroll
        for (int k = 0; k < 3; k++)
            w[k] = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(i * 48 + 16 * k), 0, 1);
        PairRec r;
        __builtin_memcpy(&r, w, sizeof r);
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

/* ---------------------------------------------------------------------- */
/* fp4 wave-tile path (single scans)                                      */

typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int F4W_THREADS = 256; /* 4 waves: two 64 x 32-variant tiles, 32 x 32 each */

/* Class-0 counts of a wave tile, two 16-bit counts per dword (in LDS). */
constexpr int F4W_PARK = 72;

/* Fisher's parked variant pairs per workgroup (two tiles) of the wave-tile
 * kernel, in a global region the workgroup claims. */
constexpr uint32_t F4W_PAIR_CAP = 1024;

/* The automatic choice takes the wave tiles for scans of at least this many
 * 128-sample K steps (both classes): measured against the LDS-staged kernel
 * at 20,000 variants, 0.99-1.09x its time up to 1,024 samples, 0.88-0.96x
 * from 1,536 on. */
constexpr int32_t F4W_MIN_KSTEPS = 10;
/* Entropy scans on wave tiles park flagged pairs from this many K steps of
 * 128 samples (Fisher scans always do).  Measured, entropy top-100 at 20,000
 * variants and top-1,000 at 40,000: parking took 1.13-1.38x the time at 24-40
 * K steps, 0.96x and 1.10x at 48, 0.96x and 0.96x at 64 (the large scan,
 * 65 K steps: 0.98x). */
constexpr int32_t F4W_PARK_MIN_KSTEPS = 64;
static_assert(GEPI_MAX_SAMPLES <= 65535, "a class count must fit 16 bits");

/* Workgroup -> its two tiles (2u, 2u + 1 of the band), units numbered as
 * band_tile numbers tiles: each XCD takes a contiguous run of units.      */
__device__ __forceinline__ int64_t band_unit(int32_t n_units) {
    const int32_t gidx = blockIdx.x;
    const int32_t xcd = gidx & 7, slot = gidx >> 3;
    const int32_t per = n_units >> 3, rem = n_units & 7;
    return (int64_t)xcd * per + (xcd < rem ? xcd : rem) + slot;
}

/* Two 64 x 32-variant tiles of tile_of_f4 per workgroup, each as two
 * 32 x 32-variant wave tiles, two waves per SIMD.  The products come from
 * v_mfma_scale_f32_32x32x64_f8f6f4, which issues half as many instructions
 * per product as the 16x16x128 form; the two waves of a SIMD hide each
 * other's VALU work (fragment ANDs, the filter), which within one wave runs
 * between its MFMAs rather than beside them.
 *
 * Operands come from the grouped 3-step layout (k_class_f4<3, true>): the
 * 32 variants of a group, one plane and one 16-byte column of a super-step
 * are 512 contiguous bytes, so every fragment load of a wave reads 1 KB
 * contiguous (lane i + 32h: row i, column h + 2cp, the K half of the
 * 32x32x64 operand) straight into registers: no LDS staging and no barrier
 * in the K loop.  Per super-step and column pair cp, three 64-sample K
 * chunks (nibble bits 0-2): A block p is plane p of the wave's 32 A
 * variants, B block q plane q of the 32 B variants, 9 MFMAs per chunk;
 * after a column pair's chunks its six registers are reloaded with the next
 * super-step's.  Lane l holds, in accumulator (p, q) element 4j + r, A
 * variant 8j + 4(l / 32) + r of the wave's group against B variant l % 32,
 * so every lane has all 9 products of its 16 variant pairs.
 *
 * Classes follow each other in the K order (class 0 from super-step 0,
 * class 1 from f4_super3(ks0)); the accumulators run on across the class
 * boundary, where the class-0 counts (below 2^16: GEPI_MAX_SAMPLES) go to
 * LDS, two per dword; class 1 is
 * the final accumulator less the class-0 count.  Padding K steps of a
 * class's last super-step are zero in the layout and add nothing.  The
 * filter runs per lane on its 16 pairs.  PARK: a flagged variant pair
 * (Fisher: about one in eight; entropy: one in a few hundred once the
 * threshold has settled, yet in a third of a wave's filter steps some lane
 * has one) is parked in a global region the workgroup claims (park_claim;
 * the LDS holds the class-0 counts) and finished one per thread after the
 * filter (pair_drain), so that no wave runs the function bits and
 * pending-list pushes for one lane; past g.park_cap, or without regions
 * (g.park null), in the lane.  The claim, the store wait and two barriers
 * cost every workgroup a few microseconds, repaid only by long K loops
 * (F4W_PARK_MIN_KSTEPS); without PARK every lane finishes its own pairs.
 *
 * The workgroup's unit (tiles tile_begin + 2 unit and 2 unit + 1 of the
 * band) comes from the caller: k_count_f4w numbers units as band_tile
 * numbers tiles, k_count_f4w_batch takes a (scan, unit) slot.  The
 * workgroup-shared state lives in F4WShared.                             */
struct F4WShared {
    uint32_t c0lds[4][F4W_PARK * 64];
    PendStage stage;
    uint32_t n_park, park_slot;
};

template <int FFUN, bool F32, bool PARK>
__device__ __forceinline__ void count_f4w_unit(const CountArgs &g, int64_t unit, F4WShared &sh) {
    PendStage &stage = sh.stage;
    stage_init(stage);
    if (PARK && threadIdx.x == 0) {
        sh.n_park = 0;
        sh.park_slot = g.park ? park_claim(g.park_busy, g.park_slots) : PARK_NONE;
    }
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t tile = g.tile_begin + 2 * unit + (wave >> 1);
    const bool live = tile < g.tile_begin + g.n_tiles;
    int32_t ta = 0, tb = 0;
    if (live) tile_of_f4(tile, g.n_var_pad / TM, ta, tb);
    const int32_t gA = 2 * ta + (wave & 1), gB = tb;
    const int32_t u1 = f4_super3(g.ks0), n_u = u1 + f4_super3(g.n_ks - g.ks0);
    const int32_t last = n_u > 0 ? n_u - 1 : 0; /* n_u == 0: super-step 0 exists */
    const char *lay = reinterpret_cast<const char *>(g.planes_k);
    const int64_t ustride = (int64_t)g.n_var_pad * 192; /* bytes per super-step */
    const int loff = (lane >> 5) * 512 + (lane & 31) * 16;
    uint32_t *P = sh.c0lds[wave] + lane; /* dword d of this lane at P[64 d] */

    /* c[cp][p]: A plane p, column h + 2cp; c[cp][3 + q]: B plane q */
    uint4 c[2][6];
    auto fetch = [&](int32_t u, int cp) {
        const char *base = lay + (int64_t)u * ustride;
        const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char *>(base + (int64_t)gA * 6144), (short)0, 6144, F4_RSRC_WORD3);
        const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char *>(base + (int64_t)gB * 6144), (short)0, 6144, F4_RSRC_WORD3);
#pragma unroll
        for (int p = 0; p < 3; p++) {
            c[cp][p] = as_u4(__builtin_amdgcn_raw_buffer_load_b128(ra, loff + p * 2048 + cp * 1024, 0, 0));
            c[cp][3 + p] =
                as_u4(__builtin_amdgcn_raw_buffer_load_b128(rb, loff + p * 2048 + cp * 1024, 0, 0));
        }
    };
    v16f acc[3][3];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[p][q][e] = 0.0f;
    auto chunks = [&](int cp) {
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const uint32_t m = 0x11111111u << b;
            v8i fb[3];
#pragma unroll
            for (int q = 0; q < 3; q++) fb[q] = frag_f4(c[cp][3 + q], m);
#pragma unroll
            for (int p = 0; p < 3; p++) {
                const v8i fa = frag_f4(c[cp][p], m);
#pragma unroll
                for (int q = 0; q < 3; q++)
                    acc[p][q] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                        fa, fb[q], acc[p][q], F4_FMT, F4_FMT, 0, f4_scale(b), 0, f4_scale(b));
            }
        }
    };
    auto sstep = [&](int32_t u) {
        const int32_t un = u + 1 < last ? u + 1 : last;
        chunks(0);
        fetch(un, 0);
        chunks(1);
        fetch(un, 1);
    };
    const int32_t nu_live = live ? n_u : 0;
    const int32_t u1_live = live ? u1 : 0;
    if (live) {
        fetch(0, 0);
        fetch(0, 1);
    }
    int32_t u = 0;
    for (; u < u1_live; u++) sstep(u);
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int e = 0; e < 16; e += 2)
                P[((p * 3 + q) * 8 + e / 2) * 64] =
                    (uint32_t)acc[p][q][e] | ((uint32_t)acc[p][q][e + 1] << 16);
    for (; u < nu_live; u++) sstep(u);

    const FilterConsts fc = filter_consts(g.st, g.n0, g.n1);
    __syncthreads(); /* the stage counter is cleared, the park region claimed */
    const bool parks = PARK && sh.park_slot != PARK_NONE;
    PairListG pl{&sh.n_park, parks ? g.park + (int64_t)sh.park_slot * g.park_cap : nullptr,
                 parks ? g.park_cap : 0u};
    const int32_t B = gB * 32 + (lane & 31);
    const int32_t Ab = gA * 32 + 4 * (lane >> 5);
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int32_t A = Ab + 8 * (e >> 2) + (e & 3);
        if (!live || A >= g.n_var || B > A) continue;
        if constexpr (F32) {
            float cn[2][3][3], tot[3][3];
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const uint32_t w = P[((p * 3 + q) * 8 + e / 2) * 64];
                    cn[0][p][q] = (float)((e & 1) ? (w >> 16) : (w & 0xFFFFu));
                    cn[1][p][q] = acc[p][q][e] - cn[0][p][q];
                    tot[p][q] = acc[p][q][e];
                }
            filter_pair<FFUN, false, PARK>(g, stage, fc, A, B, cn, &pl, tot);
        } else {
            int32_t cn[2][3][3];
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const uint32_t w = P[((p * 3 + q) * 8 + e / 2) * 64];
                    cn[0][p][q] = (int32_t)((e & 1) ? (w >> 16) : (w & 0xFFFFu));
                    cn[1][p][q] = (int32_t)acc[p][q][e] - cn[0][p][q];
                }
            filter_pair<FFUN, false, PARK>(g, stage, fc, A, B, cn, &pl);
        }
    }
    if constexpr (PARK) {
        /* The parks are global stores read by other waves after the
         * barrier, whose workgroup-scope release does not wait for global
         * stores (it issues s_barrier before s_waitcnt vmcnt(0)): the wait
         * comes first here, and the drain reads past the CU's L1
         * (pair_drain_l2); the region belongs to this XCD (park_claim). */
        f4_wait_stores();
        __syncthreads();
        pair_drain_l2<FFUN>(g, stage, fc, pl.rec, sh.n_park < pl.cap_ ? sh.n_park : pl.cap_,
                            pl.cap_);
        __syncthreads(); /* the region is read: free it */
        if (threadIdx.x == 0 && parks) atomicExch(&g.park_busy[sh.park_slot], 0u);
    }
    stage_flush(g, stage);
}

template <int FFUN, bool F32, bool PARK>
__global__ void __launch_bounds__(F4W_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_count_f4w(CountArgs g) {
    __shared__ F4WShared sh;
    count_f4w_unit<FFUN, F32, PARK>(g, band_unit((g.n_tiles + 1) >> 1), sh);
}

} // namespace gepi

#endif
