/*
 * gepi_scan.hip -- device side of libgepicount: panel staging, the all-pairs
 * count kernel with fused scoring, and the on-device streaming top-k.
 *
 * Replaces src/kernels/kernel1.cu (launch_process_gpu_1 :295-505, k1_count
 * :182-284, k1_count_sm :60-180), src/kernels/entropyD.cu:31-50,
 * src/kernels/fisher_d.cu:290-311, src/kernels/misc1.cu (result buffers,
 * thrust sort, transferBack, mergeResults :35-295) and src/kernels/setup1.h.
 *
 * Data layout in HBM
 *   panel   [n_var][3][W64] u64: per variant the allele-1 plane (row 2i of the
 *           reader, masked by valid), the allele-2 plane (row 2i+1) and the
 *           valid plane, in file sample order.  Resident across scans.
 *   classes [n_var_pad][3][Wc] u32: the same planes with samples reordered
 *           per scan: class-0 (control) samples first, then class-1 (case),
 *           each segment zero-padded to 128 bits.  Class masks disappear:
 *           a class is a word range.  n_var_pad is a multiple of the tile.
 *   cand    [cap] {score, idx, w, x, y, z}: candidates for the top-k.
 *
 * Counting.  For variant pair (A, B) and class c the kernel forms the nine
 * popcounts prod[p][q] = popc(planeA[p] & planeB[q]) over the class's words
 * (p, q in {allele1, allele2, valid}).  Every row pair of the two variants
 * and every Boolean function follows exactly, because allele planes are
 * subsets of their valid plane:
 *   N11 = prod[ia][ib], NA = prod[ia][V], NB = prod[V][ib], V = prod[V][V]
 *   AND = N11, NAND = NB-N11, ANDN = NA-N11, NANDN = V-NA-NB+N11,
 *   XOR = NA+NB-2*N11,    y = V(class 0), z = V(class 1).
 *
 * Scoring.  Each evaluation passes a conservative single-precision filter
 * against the current admission threshold; survivors get the exact score
 * (gepi_score.h) and are appended when their key (score, sA, sB, fun) is <=
 * the threshold key.  Entropy filter: -I <= t requires I >= -t, and
 * I_bits <= log2(1 + chi2/m) gives (wz - xy)^2 >= (2^I - 1) y z xx (m - xx).
 * Fisher filter: p >= P(observed table).
 *
 * Top-k.  A band of tiles is counted, then a radix select over the 128-bit
 * keys finds the k-th smallest key, the buffer is compacted to those k and
 * the key becomes the new threshold, so only k records ever leave the GPU.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gepicount.h"
#include "gepi_internal.h"
#include "gepi_score.h"

#define GEPI_VERSION "gepicount-mi355x 0.1.0"

/* ------------------------------------------------------------------------ */
/* errors                                                                    */
static thread_local char g_err[1024] = "";

int gepi_fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char *gepi_last_error(void) { return g_err; }
extern "C" const char *gepi_version(void) { return GEPI_VERSION; }

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return gepi_fail(e_ == hipErrorOutOfMemory ? GEPI_ERR_NOMEM : GEPI_ERR_DEVICE, \
                             "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),        \
                             __FILE__, __LINE__);                                         \
    } while (0)

/* ------------------------------------------------------------------------ */
/* constants and device structs                                              */
namespace {

constexpr int TV = 32;                 /* variants per tile side                 */
constexpr int KC = 32;                 /* dwords per K chunk staged in LDS       */
constexpr int VSTRIDE = 3 * KC + 4;    /* LDS dwords per variant (bank spread)   */
constexpr int COUNT_THREADS = 256;     /* 16 x 16                                */
constexpr int SEL_SMALL = 1024;        /* finish the select in LDS below this    */

struct Cand {
    double score;
    uint64_t idx;
    uint32_t w, x, y, z;
};
static_assert(sizeof(Cand) == 32, "candidate record is 32 bytes");

struct ScanState {
    double thr_score;            /* admission threshold key: (thr_score, thr_idx) */
    uint64_t thr_idx;
    float filt_k;                /* entropy filter coefficient, 0 = pass all      */
    uint32_t overflow;
    unsigned long long count;    /* records appended to the active buffer         */
    unsigned long long count2;   /* records written by the compaction             */
    unsigned long long exact;    /* exact evaluations performed                   */
    /* radix select */
    uint64_t key_hi, key_lo;
    uint32_t bits_done, rank, matched, done, gather_n, pad0;
    uint32_t hist[256];
};

struct CountArgs {
    const uint32_t *planes;      /* [n_var_pad][3][wc]                            */
    int64_t tile_begin;          /* first linear tile of the band                 */
    int32_t n_tiles;             /* tiles in the band (= grid size)               */
    int32_t n_var;
    int32_t wc;                  /* dwords per plane row (multiple of KC)         */
    int32_t w0;                  /* class-0 dwords                                */
    int32_t wused;               /* dwords holding data (multiple of 4)           */
    int32_t ffun;
    ScanState *st;
    Cand *cand;
    unsigned long long cap;
    const double *lhi, *llo;     /* t*log2 t, double-double                       */
    const double *lfact;         /* ln t!                                         */
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

/* ------------------------------------------------------------------------ */
/* panel staging kernels                                                     */

/* .bed payload -> [n_var][3][W64] planes (BinPlinkReader.h:171-201). */
__global__ void k_decode_bed(const uint8_t *__restrict__ bed, int64_t n_var, int32_t n_samples,
                             int32_t w64, int64_t bps, uint64_t *__restrict__ nat) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n_var * w64) return;
    int64_t v = gid / w64;
    int32_t wd = (int32_t)(gid % w64);
    const uint8_t *src = bed + v * bps;
    uint64_t a1 = 0, a2 = 0, vv = 0;
    int32_t s0 = wd * 64;
    int32_t s1 = s0 + 64 < n_samples ? s0 + 64 : n_samples;
    for (int32_t s = s0; s < s1; s += 4) {
        uint32_t byte = src[s >> 2];
        for (int32_t q = 0; q < 4 && s + q < s1; q++) {
            uint32_t code = (byte >> (2 * q)) & 3u;
            uint64_t bit = 1ull << ((s + q) & 63);
            if (code == 1u) continue;
            vv |= bit;
            if (code != 3u) a1 |= bit;
            if (code != 0u) a2 |= bit;
        }
    }
    uint64_t *dst = nat + (v * 3) * w64 + wd;
    dst[0] = a1;
    dst[w64] = a2;
    dst[2 * w64] = vv;
}

/* Class layout: out[v][p][d] bit b = panel bit perm[32*d + b] (or 0). */
__global__ void __launch_bounds__(256) k_gather_class(const uint64_t *__restrict__ nat,
                                                       int32_t n_var, int32_t w64,
                                                       const int32_t *__restrict__ perm,
                                                       int32_t wc, int32_t wused,
                                                       uint32_t *__restrict__ out) {
    extern __shared__ uint64_t srow[];
    const int32_t v = blockIdx.x;
    const bool real = v < n_var;
    for (int32_t i = threadIdx.x; i < 3 * w64; i += blockDim.x)
        srow[i] = real ? nat[(int64_t)v * 3 * w64 + i] : 0ull;
    __syncthreads();
    for (int32_t i = threadIdx.x; i < 3 * wc; i += blockDim.x) {
        int32_t p = i / wc, d = i - p * wc;
        uint32_t word = 0;
        if (real && d < wused) {
            const int32_t *pp = perm + 32 * d;
            const uint64_t *row = srow + p * w64;
#pragma unroll 8
            for (int32_t b = 0; b < 32; b++) {
                int32_t s = pp[b];
                if (s >= 0) word |= (uint32_t)((row[s >> 6] >> (s & 63)) & 1ull) << b;
            }
        }
        out[((int64_t)v * 3 + p) * wc + d] = word;
    }
}

/* ------------------------------------------------------------------------ */
/* count kernel                                                              */

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

__device__ __forceinline__ void admit(const CountArgs &g, double s, uint64_t idx, double thr_s,
                                      uint64_t thr_i, int32_t w, int32_t x, int32_t y, int32_t z) {
    if (s < thr_s || (s == thr_s && idx <= thr_i)) {
        unsigned long long pos = atomicAdd(&g.st->count, 1ull);
        if (pos < g.cap) {
            Cand c;
            c.score = s;
            c.idx = idx;
            c.w = (uint32_t)w;
            c.x = (uint32_t)x;
            c.y = (uint32_t)y;
            c.z = (uint32_t)z;
            g.cand[pos] = c;
        } else {
            atomicOr(&g.st->overflow, 1u);
        }
    }
}

template <int FFUN>
__global__ void __launch_bounds__(COUNT_THREADS) k_count_popc(CountArgs g) {
    __shared__ __attribute__((aligned(16))) uint32_t sA[TV * VSTRIDE];
    __shared__ __attribute__((aligned(16))) uint32_t sB[TV * VSTRIDE];

    /* XCD-aware band mapping: consecutive workgroups land on different XCDs
     * (round robin); give each XCD a contiguous run of tiles so that tiles
     * sharing a row panel share that XCD's L2.                              */
    const int32_t nb = g.n_tiles;
    const int32_t gidx = blockIdx.x;
    const int32_t xcd = gidx & 7, slot = gidx >> 3;
    const int32_t per = nb >> 3, rem = nb & 7;
    const int64_t t = g.tile_begin + (int64_t)xcd * per + (xcd < rem ? xcd : rem) + slot;
    int32_t ta, tb;
    tile_of(t, ta, tb);

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
            int idx = tid + r * COUNT_THREADS; /* 0..767: (variant, plane, quad) */
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

    /* epilogue: scoring of 4 variant pairs x (4 or 3 row pairs) x 5 functions */
    const double thr_s = g.st->thr_score;
    const uint64_t thr_i = g.st->thr_idx;
    const float kf = g.st->filt_k;
    const float lo_margin = 1.0f - 1.0f / 65536.0f;
    uint32_t n_exact = 0;
#pragma unroll
    for (int i = 0; i < 2; i++) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int32_t A = ta * TV + ty + 16 * i;
            const int32_t B = tb * TV + tx + 16 * j;
            if (A >= g.n_var || B > A) continue;
            const int32_t y = (int32_t)acc[0][i][j][2][2];
            const int32_t z = (int32_t)acc[1][i][j][2][2];
            const int32_t m = y + z;
            const float kyz = kf * (float)(y * z);
#pragma unroll
            for (int ia = 0; ia < 2; ia++) {
#pragma unroll
                for (int ib = 0; ib < 2; ib++) {
                    if (A == B && ia < ib) continue;
                    const int32_t n0 = (int32_t)acc[0][i][j][ia][ib];
                    const int32_t a0 = (int32_t)acc[0][i][j][ia][2];
                    const int32_t b0 = (int32_t)acc[0][i][j][2][ib];
                    const int32_t n1 = (int32_t)acc[1][i][j][ia][ib];
                    const int32_t a1 = (int32_t)acc[1][i][j][ia][2];
                    const int32_t b1 = (int32_t)acc[1][i][j][2][ib];
                    const int64_t sA_ = 2 * (int64_t)A + ia;
                    const int64_t sB_ = 2 * (int64_t)B + ib;
                    int32_t ws[5] = {n0, b0 - n0, a0 - n0, y - a0 - b0 + n0, a0 + b0 - 2 * n0};
                    int32_t xs[5] = {n1, b1 - n1, a1 - n1, z - a1 - b1 + n1, a1 + b1 - 2 * n1};
#pragma unroll
                    for (int f = 0; f < 5; f++) {
                        const int32_t w = ws[f], x = xs[f];
                        bool pass;
                        if (FFUN == GEPI_FFUNC_ENTROPY) {
                            const int32_t xx = w + x;
                            const int32_t q = xx * (m - xx);
                            const int32_t d = w * z - x * y;
                            const float fd = (float)d;
                            /* d == 0 is an independent table (I = 0): it only
                             * qualifies while the threshold admits everything */
                            pass = (kf == 0.0f) || (d != 0 && fd * fd >= kyz * (float)q * lo_margin);
                        } else {
                            pass = true;
                            if (thr_s < 0.0) {
                                const double *lf = g.lfact;
                                double lp = ((lf[y] + lf[z]) + (lf[w + x] + lf[m - w - x])) -
                                            (((lf[w] + lf[y - w]) + (lf[x] + lf[z - x])) + lf[m]);
                                pass = lp * 0.4342944819032518 - 1e-6 <= thr_s;
                            }
                        }
                        if (pass) {
                            n_exact++;
                            const double s =
                                (FFUN == GEPI_FFUNC_ENTROPY)
                                    ? gepi::entropy_exact(w, x, y, z, g.lhi, g.llo)
                                    : gepi::fisher_exact(w, x, y, z);
                            admit(g, s, gepi::pack_idx(sA_, sB_, f), thr_s, thr_i, w, x, y, z);
                        }
                    }
                }
            }
        }
    }
    if (n_exact) atomicAdd(&g.st->exact, (unsigned long long)n_exact);
}

/* ------------------------------------------------------------------------ */
/* top-k maintenance: radix select over 128-bit keys (score bits, idx)       */

struct SelPrefix {
    uint64_t key_hi, key_lo;
    uint32_t bits_done;
};

__device__ __forceinline__ SelPrefix load_prefix(const ScanState *st) {
    SelPrefix s;
    s.key_hi = st->key_hi;
    s.key_lo = st->key_lo;
    s.bits_done = st->bits_done;
    return s;
}

__device__ __forceinline__ bool key_match(uint64_t hi, uint64_t lo, const SelPrefix &s) {
    uint32_t b = s.bits_done;
    if (b == 0) return true;
    if (b < 64) return (hi >> (64 - b)) == (s.key_hi >> (64 - b));
    if (hi != s.key_hi) return false;
    if (b == 64) return true;
    return (lo >> (128 - b)) == (s.key_lo >> (128 - b));
}

__device__ __forceinline__ uint32_t key_digit(uint64_t hi, uint64_t lo, uint32_t b) {
    return b < 64 ? (uint32_t)((hi >> (56 - b)) & 0xFF) : (uint32_t)((lo >> (120 - b)) & 0xFF);
}

__global__ void __launch_bounds__(256) k_sel_hist(const Cand *__restrict__ cand, ScanState *st) {
    if (st->done) return;
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const SelPrefix s = load_prefix(st);
    const unsigned long long n = st->count;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        uint64_t hi = gepi::order_bits(cand[i].score), lo = cand[i].idx;
        if (key_match(hi, lo, s)) atomicAdd(&h[key_digit(hi, lo, s.bits_done)], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&st->hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_sel_pick(ScanState *st) {
    if (threadIdx.x != 0 || st->done) return;
    uint32_t cum = 0, d = 0;
    for (d = 0; d < 256; d++) {
        uint32_t c = st->hist[d];
        if (cum + c >= st->rank) break;
        cum += c;
    }
    if (d == 256) { /* rank beyond the matching records: cannot happen for rank <= count */
        st->done = 3;
        return;
    }
    uint32_t b = st->bits_done;
    if (b < 64)
        st->key_hi |= (uint64_t)d << (56 - b);
    else
        st->key_lo |= (uint64_t)d << (120 - b);
    st->rank -= cum;
    st->matched = st->hist[d];
    st->bits_done = b + 8;
    for (int i = 0; i < 256; i++) st->hist[i] = 0;
    if (st->bits_done == 128)
        st->done = 1;
    else if (st->matched <= SEL_SMALL)
        st->done = 2;
}

__global__ void __launch_bounds__(256) k_sel_gather(const Cand *__restrict__ cand, ScanState *st,
                                                     uint64_t *small) {
    if (st->done != 2) return;
    const SelPrefix s = load_prefix(st);
    const unsigned long long n = st->count;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
         i < n; i += (unsigned long long)gridDim.x * blockDim.x) {
        uint64_t hi = gepi::order_bits(cand[i].score), lo = cand[i].idx;
        if (key_match(hi, lo, s)) {
            uint32_t p = atomicAdd(&st->gather_n, 1u);
            if (p < SEL_SMALL) {
                small[2 * p] = hi;
                small[2 * p + 1] = lo;
            }
        }
    }
}

__global__ void __launch_bounds__(1024) k_sel_small(ScanState *st, const uint64_t *small) {
    if (st->done != 2) return;
    __shared__ uint64_t kh[SEL_SMALL], kl[SEL_SMALL];
    const uint32_t n = st->gather_n < SEL_SMALL ? st->gather_n : SEL_SMALL;
    const int t = threadIdx.x;
    kh[t] = t < (int)n ? small[2 * t] : ~0ull;
    kl[t] = t < (int)n ? small[2 * t + 1] : ~0ull;
    __syncthreads();
    for (int size = 2; size <= SEL_SMALL; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            int partner = t ^ stride;
            if (partner > t) {
                bool up = (t & size) == 0;
                bool gt = kh[t] > kh[partner] || (kh[t] == kh[partner] && kl[t] > kl[partner]);
                if (gt == up) {
                    uint64_t a = kh[t], b = kl[t];
                    kh[t] = kh[partner];
                    kl[t] = kl[partner];
                    kh[partner] = a;
                    kl[partner] = b;
                }
            }
            __syncthreads();
        }
    }
    if (t == 0) {
        uint32_t r = st->rank;
        st->key_hi = kh[r - 1];
        st->key_lo = kl[r - 1];
        st->bits_done = 128;
        st->done = 1;
    }
}

/* Keeps records with key <= selected key, optionally dropping a tile range. */
__global__ void __launch_bounds__(256) k_compact(const Cand *__restrict__ src, ScanState *st,
                                                  Cand *__restrict__ dst, int64_t drop_t0,
                                                  int64_t drop_t1) {
    const uint64_t khi = st->key_hi, klo = st->key_lo;
    const unsigned long long n = st->count;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        Cand c = src[i];
        uint64_t hi = gepi::order_bits(c.score);
        bool keep = hi < khi || (hi == khi && c.idx <= klo);
        if (keep && drop_t1 > drop_t0) {
            int64_t a = (int64_t)(c.idx >> 32) >> 1;
            int64_t b = (int64_t)((c.idx >> 3) & 0x1FFFFFFFull) >> 1;
            int64_t ta = a / TV, tb = b / TV;
            int64_t t = ta * (ta + 1) / 2 + tb;
            keep = t < drop_t0 || t >= drop_t1;
        }
        if (keep) {
            unsigned long long p = atomicAdd(&st->count2, 1ull);
            dst[p] = c;
        }
    }
}

/* After a compaction: adopt the compacted buffer and the new threshold. */
__global__ void k_finalize(ScanState *st, int32_t ffun) {
    if (threadIdx.x != 0) return;
    st->count = st->count2;
    st->count2 = 0;
    double thr = gepi::from_order_bits(st->key_hi);
    st->thr_score = thr;
    st->thr_idx = st->key_lo;
    float kf = 0.0f;
    if (ffun == GEPI_FFUNC_ENTROPY) {
        double it = -thr - 1e-12;
        if (it > 0.0) {
            double k = expm1(it * 0.6931471805599453) * (1.0 - 1e-6);
            kf = (float)k;
            if ((double)kf > k) kf = nextafterf(kf, 0.0f);
        }
    }
    st->filt_k = kf;
}

/* Resets the select state for a new select over `count` records. */
__global__ void k_sel_reset(ScanState *st, uint32_t rank) {
    if (threadIdx.x != 0) return;
    st->key_hi = 0;
    st->key_lo = 0;
    st->bits_done = 0;
    st->rank = rank;
    st->matched = 0;
    st->done = 0;
    st->gather_n = 0;
    for (int i = 0; i < 256; i++) st->hist[i] = 0;
}

/* ------------------------------------------------------------------------ */
/* host side                                                                 */

struct DeviceTables {
    int32_t tmax = -1;
    double *lhi = nullptr, *llo = nullptr, *lfact = nullptr;
};

} // namespace

struct gepi_panel {
    int device = 0;
    int64_t n_var = 0;
    int32_t n_samples = 0;
    int32_t w64 = 0;
    uint64_t *d_nat = nullptr;
    hipStream_t stream = nullptr;
    /* per-scan workspace, grown on demand */
    uint32_t *d_cls = nullptr;
    size_t cls_bytes = 0;
    int32_t *d_perm = nullptr;
    size_t perm_bytes = 0;
    Cand *d_cand[2] = {nullptr, nullptr};
    unsigned long long cand_cap = 0;
    ScanState *d_state = nullptr;
    uint64_t *d_small = nullptr;
    DeviceTables tabs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

int set_device(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return gepi_fail(GEPI_ERR_DEVICE, "no HIP device available (%s)",
                         e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    if (device < 0 || device >= n)
        return gepi_fail(GEPI_ERR_ARG, "device %d out of range (%d visible)", device, n);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return gepi_fail(GEPI_ERR_DEVICE, "device %d is %s; this build targets gfx950", device,
                         prop.gcnArchName);
    return GEPI_OK;
}

int ensure_tables(gepi_panel *p, int32_t tmax) {
    if (tmax <= p->tabs.tmax) return GEPI_OK;
    int32_t n = tmax + 1;
    std::vector<double> hi(n), lo(n), lf(n);
    gepi_build_tables(n, hi.data(), lo.data(), lf.data());
    hipFree(p->tabs.lhi);
    hipFree(p->tabs.llo);
    hipFree(p->tabs.lfact);
    p->tabs = DeviceTables();
    HIP_TRY(hipMalloc(&p->tabs.lhi, sizeof(double) * n));
    HIP_TRY(hipMalloc(&p->tabs.llo, sizeof(double) * n));
    HIP_TRY(hipMalloc(&p->tabs.lfact, sizeof(double) * n));
    HIP_TRY(hipMemcpy(p->tabs.lhi, hi.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->tabs.llo, lo.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->tabs.lfact, lf.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    p->tabs.tmax = tmax;
    return GEPI_OK;
}

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

int panel_alloc(gepi_panel *p, int device, int64_t n_var, int32_t n_samples) {
    p->device = device;
    p->n_var = n_var;
    p->n_samples = n_samples;
    p->w64 = (n_samples + 63) / 64;
    size_t bytes = (size_t)(n_var > 0 ? n_var : 1) * 3 * (size_t)(p->w64 > 0 ? p->w64 : 1) * 8;
    HIP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&p->d_nat, bytes));
    HIP_TRY(hipMemsetAsync(p->d_nat, 0, bytes, p->stream));
    HIP_TRY(hipMalloc(&p->d_state, sizeof(ScanState)));
    HIP_TRY(hipMalloc(&p->d_small, sizeof(uint64_t) * 2 * SEL_SMALL));
    HIP_TRY(hipEventCreate(&p->ev0));
    HIP_TRY(hipEventCreate(&p->ev1));
    return GEPI_OK;
}

void panel_free(gepi_panel *p) {
    if (!p) return;
    hipSetDevice(p->device);
    if (p->stream) hipStreamSynchronize(p->stream);
    hipFree(p->d_nat);
    hipFree(p->d_cls);
    hipFree(p->d_perm);
    hipFree(p->d_cand[0]);
    hipFree(p->d_cand[1]);
    hipFree(p->d_state);
    hipFree(p->d_small);
    hipFree(p->tabs.lhi);
    hipFree(p->tabs.llo);
    hipFree(p->tabs.lfact);
    if (p->ev0) hipEventDestroy(p->ev0);
    if (p->ev1) hipEventDestroy(p->ev1);
    if (p->stream) hipStreamDestroy(p->stream);
    delete p;
}

int check_pairing(const uint64_t *valid, int64_t n_rows, int32_t words) {
    for (int64_t i = 0; i + 1 < n_rows; i += 2) {
        const uint64_t *a = valid + (size_t)i * words, *b = a + words;
        if (memcmp(a, b, sizeof(uint64_t) * (size_t)words) != 0)
            return gepi_fail(GEPI_ERR_LAYOUT,
                             "rows %lld and %lld do not share valid bits; rows must come in "
                             "allele pairs of one variant",
                             (long long)i, (long long)(i + 1));
    }
    return GEPI_OK;
}

int validate_params(const gepi_params *par) {
    if (!par) return gepi_fail(GEPI_ERR_ARG, "null params");
    if (par->topmost < 0) return gepi_fail(GEPI_ERR_ARG, "topmost must be >= 0");
    if (par->ffun != GEPI_FFUNC_ENTROPY && par->ffun != GEPI_FFUNC_FISHERPV)
        return gepi_fail(GEPI_ERR_ARG, "unknown objective %d", par->ffun);
    if (par->count_path < 0 || par->count_path > 2)
        return gepi_fail(GEPI_ERR_ARG, "unknown count path %d", par->count_path);
    return GEPI_OK;
}

/* Select the rank-th smallest key (1-based) among st->count records of src. */
void launch_select(gepi_panel *p, const Cand *src, uint32_t rank, unsigned long long count) {
    k_sel_reset<<<1, 1, 0, p->stream>>>(p->d_state, rank);
    unsigned long long blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    for (int pass = 0; pass < 16; pass++) {
        k_sel_hist<<<(unsigned)blocks, 256, 0, p->stream>>>(src, p->d_state);
        k_sel_pick<<<1, 1, 0, p->stream>>>(p->d_state);
    }
    k_sel_gather<<<(unsigned)blocks, 256, 0, p->stream>>>(src, p->d_state, p->d_small);
    k_sel_small<<<1, SEL_SMALL, 0, p->stream>>>(p->d_state, p->d_small);
}

struct HostState {
    unsigned long long count, exact;
    uint32_t overflow;
};

int read_state(gepi_panel *p, HostState *hs) {
    ScanState s;
    HIP_TRY(hipMemcpyAsync(&s, p->d_state, sizeof(ScanState), hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    hs->count = s.count;
    hs->exact = s.exact;
    hs->overflow = s.overflow;
    return GEPI_OK;
}

int run_scan(gepi_panel *p, const uint64_t *mask_ctrl, const uint64_t *mask_case,
             const gepi_params *par, gepi_result *tops, int64_t *n_tops,
             gepi_scan_stats *stats) {
    auto t_start = std::chrono::steady_clock::now();
    int rc;
    if ((rc = validate_params(par)) != GEPI_OK) return rc;
    if (!n_tops) return gepi_fail(GEPI_ERR_ARG, "null n_tops");
    *n_tops = 0;
    if (par->topmost > 0 && !tops) return gepi_fail(GEPI_ERR_ARG, "null tops");
    if (par->count_path == 2)
        return gepi_fail(GEPI_ERR_ARG, "MFMA count path not available in this build");
    if ((rc = set_device(p->device)) != GEPI_OK) return rc;

    gepi_scan_stats st_local;
    memset(&st_local, 0, sizeof st_local);
    const int64_t n_var = p->n_var;
    const int64_t R = 2 * n_var;
    const int64_t total_evals = 5 * (R * (R + 1) / 2);
    st_local.evals = total_evals;
    st_local.count_path = 1;

    /* class layout (sample permutation) */
    std::vector<int32_t> cls0, cls1;
    for (int32_t s = 0; s < p->n_samples; s++) {
        uint64_t bit = 1ull << (s & 63);
        if (mask_ctrl && (mask_ctrl[s >> 6] & bit)) cls0.push_back(s);
        if (mask_case && (mask_case[s >> 6] & bit)) cls1.push_back(s);
    }
    const int32_t n0 = (int32_t)cls0.size(), n1 = (int32_t)cls1.size();
    st_local.n_class0 = n0;
    st_local.n_class1 = n1;
    if ((int64_t)n0 + n1 > GEPI_MAX_SAMPLES)
        return gepi_fail(GEPI_ERR_LIMIT, "%d class samples exceed the limit of %d", n0 + n1,
                         GEPI_MAX_SAMPLES);
    const int64_t k = par->topmost < total_evals ? par->topmost : total_evals;
    if (k == 0) {
        if (stats) *stats = st_local;
        return GEPI_OK;
    }
    if (k > (int64_t)(1u << 30)) return gepi_fail(GEPI_ERR_LIMIT, "topmost too large");
    const int32_t w0 = ((n0 + 127) / 128) * 4; /* dwords, 128-bit padded */
    const int32_t w1 = ((n1 + 127) / 128) * 4;
    const int32_t wused = w0 + w1;
    const int32_t wc = ((wused + KC - 1) / KC) * KC > 0 ? ((wused + KC - 1) / KC) * KC : KC;
    std::vector<int32_t> perm((size_t)wc * 32, -1);
    for (int32_t i = 0; i < n0; i++) perm[i] = cls0[i];
    for (int32_t i = 0; i < n1; i++) perm[(size_t)w0 * 32 + i] = cls1[i];
    const int64_t nt_side = (n_var + TV - 1) / TV;
    const int64_t n_var_pad = nt_side * TV;
    const int64_t n_tiles = nt_side * (nt_side + 1) / 2;

    if ((rc = ensure_tables(p, n0 + n1 + 1)) != GEPI_OK) return rc;
    if ((rc = grow(&p->d_cls, &p->cls_bytes, (size_t)n_var_pad * 3 * wc * 4)) != GEPI_OK) return rc;
    if ((rc = grow(&p->d_perm, &p->perm_bytes, perm.size() * 4)) != GEPI_OK) return rc;
    unsigned long long cap = (unsigned long long)std::max<int64_t>(4 * k, 1 << 23);
    if (const char *env = getenv("GEPI_CAND_CAP")) {
        /* capacity override for exercising the overflow recovery in tests */
        long long want = atoll(env);
        cap = (unsigned long long)std::max<long long>(want, 2 * k + 1024);
    }
    if (cap > p->cand_cap) {
        hipFree(p->d_cand[0]);
        hipFree(p->d_cand[1]);
        p->d_cand[0] = p->d_cand[1] = nullptr;
        p->cand_cap = 0;
        HIP_TRY(hipMalloc(&p->d_cand[0], sizeof(Cand) * cap));
        HIP_TRY(hipMalloc(&p->d_cand[1], sizeof(Cand) * cap));
        p->cand_cap = cap;
    }

    HIP_TRY(hipMemcpyAsync(p->d_perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice,
                           p->stream));
    HIP_TRY(hipEventRecord(p->ev0, p->stream));
    {
        size_t lds = (size_t)3 * (p->w64 > 0 ? p->w64 : 1) * 8;
        k_gather_class<<<(unsigned)n_var_pad, 256, lds, p->stream>>>(
            p->d_nat, (int32_t)n_var, p->w64, p->d_perm, wc, wused, p->d_cls);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(p->ev1, p->stream));

    ScanState init;
    memset(&init, 0, sizeof init);
    init.thr_score = INFINITY;
    init.thr_idx = ~0ull;
    init.filt_k = 0.0f;
    HIP_TRY(hipMemcpyAsync(p->d_state, &init, sizeof init, hipMemcpyHostToDevice, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    {
        float ms = 0;
        hipEventElapsedTime(&ms, p->ev0, p->ev1);
        st_local.ms_prepare = ms;
    }

    int cur = 0; /* active candidate buffer */
    CountArgs ca;
    ca.planes = p->d_cls;
    ca.n_var = (int32_t)n_var;
    ca.wc = wc;
    ca.w0 = w0;
    ca.wused = wused;
    ca.ffun = par->ffun;
    ca.st = p->d_state;
    ca.cap = cap;
    ca.lhi = p->tabs.lhi;
    ca.llo = p->tabs.llo;
    ca.lfact = p->tabs.lfact;

    const int64_t evals_per_tile = (int64_t)TV * TV * 20;
    int64_t max_band = 65536;
    if (par->max_grid_area > 0)
        max_band = std::max<int64_t>(1, ((int64_t)par->max_grid_area << 20) / evals_per_tile);
    int64_t band = std::max<int64_t>(1, std::min<int64_t>(max_band, (2 * k) / evals_per_tile + 1));
    double ms_count = 0.0, ms_sel = 0.0;
    int64_t t0 = 0;
    unsigned long long count = 0, exact_total = 0;
    int bands = 0;
    while (t0 < n_tiles) {
        int64_t t1 = std::min<int64_t>(n_tiles, t0 + band);
        ca.tile_begin = t0;
        ca.n_tiles = (int32_t)(t1 - t0);
        ca.cand = p->d_cand[cur];
        HIP_TRY(hipEventRecord(p->ev0, p->stream));
        if (par->ffun == GEPI_FFUNC_ENTROPY)
            k_count_popc<GEPI_FFUNC_ENTROPY><<<(unsigned)(t1 - t0), COUNT_THREADS, 0, p->stream>>>(ca);
        else
            k_count_popc<GEPI_FFUNC_FISHERPV><<<(unsigned)(t1 - t0), COUNT_THREADS, 0, p->stream>>>(ca);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(p->ev1, p->stream));
        HostState hs;
        if ((rc = read_state(p, &hs)) != GEPI_OK) return rc;
        float ms = 0;
        hipEventElapsedTime(&ms, p->ev0, p->ev1);
        ms_count += ms;
        bands++;
        unsigned long long added = hs.count - count;
        auto ts = std::chrono::steady_clock::now();
        if (hs.overflow) {
            /* Threshold from what was kept, drop this band's records, redo it. */
            unsigned long long kept = std::min<unsigned long long>(hs.count, cap);
            ScanState fix;
            HIP_TRY(hipMemcpyAsync(&fix, p->d_state, sizeof fix, hipMemcpyDeviceToHost, p->stream));
            HIP_TRY(hipStreamSynchronize(p->stream));
            fix.count = kept;
            fix.overflow = 0;
            HIP_TRY(hipMemcpyAsync(p->d_state, &fix, sizeof fix, hipMemcpyHostToDevice, p->stream));
            launch_select(p, p->d_cand[cur], (uint32_t)k, kept);
            k_compact<<<4096, 256, 0, p->stream>>>(p->d_cand[cur], p->d_state, p->d_cand[cur ^ 1],
                                                   t0, t1);
            k_finalize<<<1, 1, 0, p->stream>>>(p->d_state, par->ffun);
            cur ^= 1;
            HostState h2;
            if ((rc = read_state(p, &h2)) != GEPI_OK) return rc;
            count = h2.count;
            band = std::max<int64_t>(1, band / 4);
            ms_sel += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts)
                          .count();
            continue; /* same t0 */
        }
        if (hs.count > (unsigned long long)k) {
            launch_select(p, p->d_cand[cur], (uint32_t)k, hs.count);
            k_compact<<<4096, 256, 0, p->stream>>>(p->d_cand[cur], p->d_state, p->d_cand[cur ^ 1],
                                                   0, 0);
            k_finalize<<<1, 1, 0, p->stream>>>(p->d_state, par->ffun);
            HIP_TRY(hipGetLastError());
            cur ^= 1;
            count = (unsigned long long)k;
        } else {
            count = hs.count;
        }
        exact_total = hs.exact;
        ms_sel +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
        /* size the next band so its survivors fit comfortably */
        int64_t tiles_done = t1 - t0;
        double rate = (double)added / (double)tiles_done;
        int64_t room = (int64_t)(cap - (unsigned long long)k) / 2;
        int64_t next = rate > 0.0 ? (int64_t)((double)room / rate) : max_band;
        band = std::max<int64_t>(1, std::min<int64_t>({max_band, next, band * 4}));
        t0 = t1;
    }
    /* final extraction: exactly k records */
    auto ts = std::chrono::steady_clock::now();
    std::vector<Cand> host((size_t)count);
    if (count > 0)
        HIP_TRY(hipMemcpyAsync(host.data(), p->d_cand[cur], sizeof(Cand) * count,
                               hipMemcpyDeviceToHost, p->stream));
    HostState hs;
    if ((rc = read_state(p, &hs)) != GEPI_OK) return rc;
    exact_total = hs.exact;
    std::sort(host.begin(), host.end(), [](const Cand &a, const Cand &b) {
        uint64_t ha = gepi::order_bits(a.score), hb = gepi::order_bits(b.score);
        return ha < hb || (ha == hb && a.idx < b.idx);
    });
    int64_t n = std::min<int64_t>((int64_t)host.size(), k);
    for (int64_t i = 0; i < n; i++) {
        const Cand &c = host[(size_t)i];
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
    *n_tops = n;
    ms_sel += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
    st_local.ms_count_kernels = ms_count;
    st_local.ms_select = ms_sel;
    st_local.exact_evals = (int64_t)exact_total;
    st_local.bands = bands;
    st_local.ms_scan =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
            .count();
    if (n != k)
        return gepi_fail(GEPI_ERR_INTERNAL, "top-k extraction returned %lld of %lld records",
                         (long long)n, (long long)k);
    if (stats) *stats = st_local;
    return GEPI_OK;
}

int upload_planes(gepi_panel *p, const uint64_t *bits, const uint64_t *valid, int64_t n_rows,
                  int32_t words) {
    const int64_t n_var = n_rows / 2;
    std::vector<uint64_t> nat((size_t)(n_var > 0 ? n_var : 1) * 3 * (words > 0 ? words : 1), 0);
    for (int64_t v = 0; v < n_var; v++) {
        const uint64_t *b1 = bits + (size_t)(2 * v) * words, *b2 = b1 + words;
        const uint64_t *v1 = valid + (size_t)(2 * v) * words;
        uint64_t *o = nat.data() + (size_t)v * 3 * words;
        for (int32_t w = 0; w < words; w++) {
            o[w] = b1[w] & v1[w];
            o[words + w] = b2[w] & v1[w];
            o[2 * words + w] = v1[w];
        }
    }
    if (n_var > 0 && words > 0)
        HIP_TRY(hipMemcpy(p->d_nat, nat.data(), nat.size() * 8, hipMemcpyHostToDevice));
    return GEPI_OK;
}

} // namespace

/* ------------------------------------------------------------------------ */
/* C ABI                                                                     */

extern "C" void gepi_params_init(gepi_params *par) {
    if (!par) return;
    memset(par, 0, sizeof *par);
    par->gpu_num = 0;
    par->ffun = GEPI_FFUNC_ENTROPY;
    par->topmost = GEPI_DEFAULT_TOPMOST;
    par->count_path = 0;
    par->blk_x = par->blk_y = 16;
    par->max_grid_area = 0;
}

extern "C" int gepi_device_count(int32_t *count) {
    if (!count) return gepi_fail(GEPI_ERR_ARG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return GEPI_OK;
}

extern "C" int gepi_panel_create(const uint64_t *bit_data, const uint64_t *valid_data,
                                 int64_t n_rows, int32_t n_samples, int32_t device,
                                 gepi_panel **out) {
    if (!out) return gepi_fail(GEPI_ERR_ARG, "null out");
    *out = nullptr;
    if (n_rows < 0 || n_samples < 0) return gepi_fail(GEPI_ERR_ARG, "negative size");
    if (n_rows % 2) return gepi_fail(GEPI_ERR_LAYOUT, "odd row count %lld", (long long)n_rows);
    if (n_rows >= (1LL << 29)) return gepi_fail(GEPI_ERR_LIMIT, "too many rows");
    const int32_t words = (n_samples + 63) / 64;
    if (n_rows > 0 && words > 0 && (!bit_data || !valid_data))
        return gepi_fail(GEPI_ERR_ARG, "null plane data");
    int rc;
    if (n_rows > 0 && words > 0 && (rc = check_pairing(valid_data, n_rows, words)) != GEPI_OK)
        return rc;
    if ((rc = set_device(device)) != GEPI_OK) return rc;
    gepi_panel *p = new gepi_panel();
    if ((rc = panel_alloc(p, device, n_rows / 2, n_samples)) != GEPI_OK ||
        (rc = upload_planes(p, bit_data, valid_data, n_rows, words)) != GEPI_OK) {
        panel_free(p);
        return rc;
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess) {
        panel_free(p);
        return gepi_fail(GEPI_ERR_DEVICE, "panel upload failed");
    }
    *out = p;
    return GEPI_OK;
}

extern "C" int gepi_panel_create_bed(const uint8_t *bed_payload, int64_t n_variants,
                                     int32_t n_samples, int32_t device, gepi_panel **out) {
    if (!out) return gepi_fail(GEPI_ERR_ARG, "null out");
    *out = nullptr;
    if (n_variants < 0 || n_samples < 0) return gepi_fail(GEPI_ERR_ARG, "negative size");
    if (n_variants >= (1LL << 28)) return gepi_fail(GEPI_ERR_LIMIT, "too many variants");
    const int64_t bps = ((int64_t)n_samples + 3) / 4;
    if (n_variants > 0 && bps > 0 && !bed_payload) return gepi_fail(GEPI_ERR_ARG, "null payload");
    int rc;
    if ((rc = set_device(device)) != GEPI_OK) return rc;
    gepi_panel *p = new gepi_panel();
    if ((rc = panel_alloc(p, device, n_variants, n_samples)) != GEPI_OK) {
        panel_free(p);
        return rc;
    }
    const int64_t bytes = n_variants * bps;
    if (bytes > 0 && p->w64 > 0) {
        uint8_t *d_bed = nullptr;
        /* pad so the decoder's last 4-sample read of a row stays in bounds */
        if (hipMalloc(&d_bed, (size_t)bytes + 16) != hipSuccess) {
            panel_free(p);
            return gepi_fail(GEPI_ERR_NOMEM, "bed staging allocation failed");
        }
        hipMemsetAsync(d_bed + bytes, 0, 16, p->stream);
        hipMemcpyAsync(d_bed, bed_payload, (size_t)bytes, hipMemcpyHostToDevice, p->stream);
        int64_t threads = n_variants * p->w64;
        k_decode_bed<<<(unsigned)((threads + 255) / 256), 256, 0, p->stream>>>(
            d_bed, n_variants, n_samples, p->w64, bps, p->d_nat);
        hipError_t e = hipStreamSynchronize(p->stream);
        hipFree(d_bed);
        if (e != hipSuccess || hipGetLastError() != hipSuccess) {
            panel_free(p);
            return gepi_fail(GEPI_ERR_DEVICE, "bed decode failed: %s", hipGetErrorString(e));
        }
    }
    *out = p;
    return GEPI_OK;
}

extern "C" int gepi_panel_destroy(gepi_panel *panel) {
    if (!panel) return gepi_fail(GEPI_ERR_ARG, "null panel");
    panel_free(panel);
    return GEPI_OK;
}

extern "C" int gepi_panel_download(gepi_panel *p, uint64_t *bit_data, uint64_t *valid_data) {
    if (!p || !bit_data || !valid_data) return gepi_fail(GEPI_ERR_ARG, "null argument");
    int rc;
    if ((rc = set_device(p->device)) != GEPI_OK) return rc;
    const int32_t W = p->w64;
    std::vector<uint64_t> nat((size_t)p->n_var * 3 * W);
    if (!nat.empty())
        HIP_TRY(hipMemcpy(nat.data(), p->d_nat, nat.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t v = 0; v < p->n_var; v++)
        for (int32_t w = 0; w < W; w++) {
            const uint64_t *o = nat.data() + (size_t)v * 3 * W;
            bit_data[(size_t)(2 * v) * W + w] = o[w];
            bit_data[(size_t)(2 * v + 1) * W + w] = o[W + w];
            valid_data[(size_t)(2 * v) * W + w] = o[2 * W + w];
            valid_data[(size_t)(2 * v + 1) * W + w] = o[2 * W + w];
        }
    return GEPI_OK;
}

extern "C" int gepi_scan(gepi_panel *panel, const uint64_t *mask_ctrl, const uint64_t *mask_case,
                         const gepi_params *par, gepi_result *tops, int64_t *n_tops,
                         gepi_scan_stats *stats) {
    if (!panel) return gepi_fail(GEPI_ERR_ARG, "null panel");
    return run_scan(panel, mask_ctrl, mask_case, par, tops, n_tops, stats);
}

extern "C" int gepi_process_gpu_1(const uint64_t *bit_data, const uint64_t *valid_data,
                                  const uint64_t *mask_ctrl, const uint64_t *mask_case,
                                  int64_t n_rows, int32_t n_samples, const gepi_params *par,
                                  gepi_result *tops, int64_t *n_tops, gepi_scan_stats *stats) {
    int rc;
    if ((rc = validate_params(par)) != GEPI_OK) return rc;
    gepi_panel *p = nullptr;
    if ((rc = gepi_panel_create(bit_data, valid_data, n_rows, n_samples, par->gpu_num, &p)) !=
        GEPI_OK)
        return rc;
    rc = run_scan(p, mask_ctrl, mask_case, par, tops, n_tops, stats);
    panel_free(p);
    return rc;
}
