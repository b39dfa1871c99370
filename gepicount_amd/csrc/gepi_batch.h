/*
 * gepi_batch.h -- kernels of the batched scan engine.  Many scans (panels
 * and phenotype vectors of different sizes) advance together in rounds; a
 * round is three launches whatever the number of scans:
 *
 *   k_count_f4_batch  one band of fp4 count tiles per scan in one persistent
 *                     grid: workgroups take (scan, tile) slots from per-XCD
 *                     work counters through a descriptor table
 *   k_exact_batch     exact scores of each scan's filter survivors
 *                     (blockIdx.y = scan)
 *   k_topk_wg         one 1024-thread workgroup per scan: radix select of
 *                     the k-th key, in-place compaction to k records, new
 *                     threshold, overflow recovery, a report for the host
 *
 * A round can also run as two captured HIP graphs (the work-counter clear
 * and the count launches; the exact and top-k launches) replayed every
 * round: the grids are then fixed by the descriptor capacity, a scan's
 * band lives in the descriptor tables, and an empty descriptor slot does
 * nothing.  k_fill_u32 clears the work counters inside the graph.
 *
 * There is no counterpart in the reference, which runs one scan per process
 * (src/gEpiCount.cpp:84-87).
 */
#ifndef GEPI_BATCH_H
#define GEPI_BATCH_H

#include "gepi_count.h"
#include "gepi_topk.h"

namespace gepi {

/* One scan's band in a count round. */
struct BatchDesc {
    CountArgs g; /* tile_begin / n_tiles: the band                            */
};

/* Slot -> (scan, tile).  Workgroups are dispatched to the 8 XCDs round
 * robin (XCD = blockIdx.x % 8).  Every scan's band is split into 8 contiguous
 * chunks, one per XCD (so each XCD's L2 sees neighbouring tiles), and each
 * XCD runs its chunks scan after scan: slot s of XCD x (taken from the
 * XCD's work counter) is in scan i when xprefix[x][i] <= s < xprefix[x][i+1],
 * at tile tile_begin + xstart[x][i] + (s - xprefix[x][i]).  Chunk sizes
 * follow the round-robin dealing of the band, so every XCD gets the same
 * number of tiles of every scan to within one.                            */
struct BatchMap {
    const int64_t *xprefix; /* [8][n + 1] */
    const int64_t *xstart;  /* [8][n]     */
    uint32_t *wq;           /* 2 sets of 8 work counters                  */
    int32_t n;
    int32_t wq_parity;      /* set of this launch                          */
};

struct F4BatchSrc {
    const BatchDesc *d;
    BatchMap map;
    __device__ __forceinline__ F4Job grab(int32_t) const {
        const int32_t x = blockIdx.x & 7;
        const int64_t slot = atomicAdd(&map.wq[8 * map.wq_parity + x], 1u);
        const int64_t *pre = map.xprefix + (int64_t)x * (map.n + 1);
        F4Job j;
        j.ok = slot < pre[map.n];
        j.item = 0;
        j.tile = 0;
        if (!j.ok) return j;
        int32_t lo = 0, hi = map.n - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= slot)
                lo = mid;
            else
                hi = mid - 1;
        }
        j.item = lo;
        j.tile = d[lo].g.tile_begin + map.xstart[(int64_t)x * map.n + lo] + (slot - pre[lo]);
        return j;
    }
    __device__ __forceinline__ const CountArgs *argp(int32_t item) const { return &d[item].g; }
};

/* One (scan, tile) per workgroup: slot = blockIdx.x / 8 of XCD blockIdx.x % 8. */
struct F4BatchOnceSrc {
    const BatchDesc *d;
    BatchMap map;
    __device__ __forceinline__ F4Job grab(int32_t n) const {
        const int32_t x = blockIdx.x & 7;
        const int64_t slot = blockIdx.x >> 3;
        const int64_t *pre = map.xprefix + (int64_t)x * (map.n + 1);
        F4Job j;
        j.ok = n == 0 && slot < pre[map.n];
        j.item = 0;
        j.tile = 0;
        if (!j.ok) return j;
        int32_t lo = 0, hi = map.n - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= slot)
                lo = mid;
            else
                hi = mid - 1;
        }
        j.item = lo;
        j.tile = d[lo].g.tile_begin + map.xstart[(int64_t)x * map.n + lo] + (slot - pre[lo]);
        return j;
    }
    __device__ __forceinline__ const CountArgs *argp(int32_t item) const { return &d[item].g; }
};

/* PERSIST: workgroups take slots from the work counters (grid of a few
 * workgroups per CU); else one slot per workgroup (grid = 8 x the largest
 * XCD share). */
template <int FFUN, bool PERSIST>
__global__ void __launch_bounds__(F4_THREADS, 2)
    k_count_f4_batch(const BatchDesc *__restrict__ d, BatchMap map) {
    if (PERSIST) {
        zero_other_counters(map.wq, map.wq_parity);
        const F4BatchSrc src{d, map};
        count_f4_run<FFUN>(src);
    } else {
        const F4BatchOnceSrc src{d, map};
        count_f4_run<FFUN>(src);
    }
}

/* p[0..n) = v (one workgroup of 64 threads per 64 words). */
__global__ void __launch_bounds__(64) k_fill_u32(uint32_t *p, int32_t n, uint32_t v) {
    const int32_t i = (int32_t)(blockIdx.x * 64 + threadIdx.x);
    if (i < n) p[i] = v;
}

struct ExactDesc {
    ExactArgs e;
    int32_t ffun;
    int32_t pad;
};

/* blockIdx.y = descriptor; a zeroed descriptor (no scan: a captured round
 * launches one row per descriptor slot) does nothing. */
__global__ void __launch_bounds__(256) k_exact_batch(const ExactDesc *__restrict__ d) {
    const ExactDesc x = d[blockIdx.y];
    if (x.e.st == nullptr) return;
    const unsigned long long start = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    if (x.ffun == GEPI_FFUNC_ENTROPY)
        exact_range<GEPI_FFUNC_ENTROPY>(x.e, start, stride);
    else
        exact_range<GEPI_FFUNC_FISHERPV>(x.e, start, stride);
}

/* What the host needs after a round, per scan. */
struct ItemReport {
    unsigned long long raw_count;  /* candidates after k_exact, before top-k  */
    unsigned long long count;      /* candidates kept after the round         */
    unsigned long long pend_count; /* filter survivors of the band            */
    uint32_t pend_overflow;        /* survivors lost: band to be redone       */
    uint32_t cand_overflow;        /* candidates lost: band dropped, redone   */
};

struct TopkDesc {
    ScanState *st;
    Cand *cand;
    unsigned long long cap;
    int64_t drop_t0, drop_t1; /* this round's band (fp4 tile numbering)       */
    ItemReport *report;
    uint32_t k;
    int32_t ffun;
    int32_t n_a4;             /* A blocks of the scan (fp4 tile numbering)    */
    int32_t pad;
};

constexpr int TOPK_THREADS = 1024;

__device__ __forceinline__ bool key_le(uint64_t ah, uint64_t al, uint64_t bh, uint64_t bl) {
    return ah < bh || (ah == bh && al <= bl);
}

/* Block-wide exclusive prefix of one flag per thread (1024 threads). */
__device__ __forceinline__ uint32_t block_prefix(bool flag, uint32_t *wsum, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t b = __ballot(flag);
    const uint32_t below = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < TOPK_THREADS / 64; w++) {
        const uint32_t c = wsum[w];
        before += w < wave ? c : 0u;
        all += c;
    }
    total = all;
    return before + below;
}

/* Top-k maintenance of one scan by one workgroup: after a candidate
 * overflow, the k-th key among the records kept becomes the threshold and the
 * band's records are dropped (the band is redone); otherwise, with more than
 * k candidates, exactly the k smallest keys are kept.  Resets the pending
 * list for the next round.                                                 */
__global__ void __launch_bounds__(TOPK_THREADS) k_topk_wg(const TopkDesc *__restrict__ dd) {
    const TopkDesc d = dd[blockIdx.x];
    ScanState *st = d.st;
    if (st == nullptr) return; /* empty descriptor slot (the whole workgroup) */
    __shared__ uint32_t hist[256];
    __shared__ uint64_t kh[SEL_SMALL], kl[SEL_SMALL];
    __shared__ uint32_t wsum[TOPK_THREADS / 64];
    __shared__ uint64_t s_hi, s_lo;
    __shared__ uint32_t s_bits, s_rank, s_done, s_gather;
    __shared__ unsigned long long s_wptr;
    const int t = threadIdx.x;

    const unsigned long long raw = st->count;
    const uint32_t pend_ovf = st->pend_overflow, cand_ovf = st->overflow;
    const unsigned long long pend_n = st->pend_count;
    __syncthreads(); /* every thread has read the state before it changes */
    const unsigned long long n = raw < d.cap ? raw : d.cap;
    const bool work = !pend_ovf && (cand_ovf || n > d.k);
    if (work) {
        if (t == 0) {
            s_hi = 0;
            s_lo = 0;
            s_bits = 0;
            s_rank = d.k;
            s_done = 0;
        }
        /* radix select of the k-th smallest key, 8-bit digits MSB first */
        while (true) {
            if (t < 256) hist[t] = 0;
            __syncthreads();
            SelPrefix pre;
            pre.key_hi = s_hi;
            pre.key_lo = s_lo;
            pre.bits_done = s_bits;
            for (unsigned long long i = t; i < n; i += TOPK_THREADS) {
                const uint64_t h = order_bits(d.cand[i].score), l = d.cand[i].idx;
                if (key_match(h, l, pre)) atomicAdd(&hist[key_digit(h, l, pre.bits_done)], 1u);
            }
            __syncthreads();
            if (t == 0) {
                uint32_t cum = 0, dg = 0;
                for (dg = 0; dg < 255; dg++) {
                    if (cum + hist[dg] >= s_rank) break;
                    cum += hist[dg];
                }
                const uint32_t b = s_bits;
                if (b < 64)
                    s_hi |= (uint64_t)dg << (56 - b);
                else
                    s_lo |= (uint64_t)dg << (120 - b);
                s_rank -= cum;
                s_bits = b + 8;
                s_done = s_bits == 128 ? 1u : (hist[dg] <= (uint32_t)SEL_SMALL ? 2u : 0u);
                s_gather = 0;
            }
            __syncthreads();
            if (s_done) break;
        }
        if (s_done == 2) {
            /* the remaining matching keys fit in LDS: gather and sort them */
            SelPrefix pre;
            pre.key_hi = s_hi;
            pre.key_lo = s_lo;
            pre.bits_done = s_bits;
            for (unsigned long long i = t; i < n; i += TOPK_THREADS) {
                const uint64_t h = order_bits(d.cand[i].score), l = d.cand[i].idx;
                if (key_match(h, l, pre)) {
                    const uint32_t p = atomicAdd(&s_gather, 1u);
                    if (p < (uint32_t)SEL_SMALL) {
                        kh[p] = h;
                        kl[p] = l;
                    }
                }
            }
            __syncthreads();
            const uint32_t g = s_gather < (uint32_t)SEL_SMALL ? s_gather : (uint32_t)SEL_SMALL;
            if (t >= (int)g) {
                kh[t] = ~0ull;
                kl[t] = ~0ull;
            }
            __syncthreads();
            for (int size = 2; size <= SEL_SMALL; size <<= 1) {
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    const int partner = t ^ stride;
                    if (partner > t) {
                        const bool up = (t & size) == 0;
                        const bool gt = kh[t] > kh[partner] || (kh[t] == kh[partner] && kl[t] > kl[partner]);
                        if (gt == up) {
                            const uint64_t a = kh[t], b = kl[t];
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
                s_hi = kh[s_rank - 1];
                s_lo = kl[s_rank - 1];
            }
            __syncthreads();
        }
        /* in-place compaction: a chunk is read before any of it is written,
         * and writes never pass the read position */
        const uint64_t khi = s_hi, klo = s_lo;
        if (t == 0) s_wptr = 0;
        __syncthreads();
        for (unsigned long long base = 0; base < n; base += TOPK_THREADS) {
            const unsigned long long i = base + t;
            Cand c;
            bool keep = false;
            if (i < n) {
                c = d.cand[i];
                keep = key_le(order_bits(c.score), c.idx, khi, klo);
                if (keep && cand_ovf) {
                    const int64_t a = (int64_t)(c.idx >> 32) >> 1;
                    const int64_t b = (int64_t)((c.idx >> 3) & 0x1FFFFFFFull) >> 1;
                    const int64_t tl = tile_index(a, b, TM, d.n_a4);
                    keep = tl < d.drop_t0 || tl >= d.drop_t1;
                }
            }
            uint32_t total;
            const uint32_t off = block_prefix(keep, wsum, total);
            const unsigned long long w0 = s_wptr;
            __syncthreads();
            if (keep) d.cand[w0 + off] = c;
            if (t == 0) s_wptr = w0 + total;
            __syncthreads();
        }
        if (t == 0) {
            st->count = s_wptr;
            st->count2 = 0;
            st->overflow = 0;
            adopt_threshold(st, khi, klo, d.ffun);
        }
    }
    if (t == 0) {
        ItemReport r;
        r.raw_count = raw;
        r.count = work ? s_wptr : (pend_ovf ? raw : n);
        r.pend_count = pend_n;
        r.pend_overflow = pend_ovf;
        r.cand_overflow = pend_ovf ? 0u : cand_ovf;
        *d.report = r;
        st->pend_count = 0;
        st->pend_overflow = 0;
    }
}

} // namespace gepi

#endif
