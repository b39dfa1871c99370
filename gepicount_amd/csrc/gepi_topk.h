/*
 * gepi_topk.h -- exact scoring of filter survivors and the on-device
 * streaming top-k (replaces k1_entropy / k1_fisherpv, src/kernels/entropyD.cu:31-50,
 * src/kernels/fisher_d.cu:290-311, the per-tile thrust::sort_by_key of
 * kernel1.cu:448 and the host merge of misc1.cu:234-264).
 *
 * Keys are (order_bits(score), idx) as 128-bit unsigned values; idx orders
 * like (sA, sB, fun), so the key order is the canonical (score, sA, sB, fun).
 * The select is an MSB-first radix select with 8-bit digits that finishes
 * with a bitonic sort in LDS once at most SEL_SMALL keys share the prefix.
 */
#ifndef GEPI_TOPK_H
#define GEPI_TOPK_H

#include "gepi_device.h"

namespace gepi {

struct ExactArgs {
    ScanState *st;
    const Pend *pend;
    unsigned long long pcap;
    Cand *cand;
    unsigned long long cap;
    const double *lhi, *llo;
};

/* Exact score of pending evaluations [start, n) with the given stride;
 * admits keys <= the threshold.  Nothing when the pending list overflowed. */
template <int FFUN>
__device__ __forceinline__ void exact_range(const ExactArgs &e, unsigned long long start,
                                            unsigned long long stride) {
    if (e.st->pend_overflow) return;
    unsigned long long n = e.st->pend_count;
    if (n > e.pcap) n = e.pcap;
    const double thr_s = e.st->thr_score;
    const uint64_t thr_i = e.st->thr_idx;
    for (unsigned long long i = start; i < n; i += stride) {
        const Pend p = e.pend[i];
        const int32_t w = p.w, x = p.x, y = p.y, z = p.z;
        const double s = (FFUN == GEPI_FFUNC_ENTROPY) ? entropy_exact(w, x, y, z, e.lhi, e.llo)
                                                      : fisher_exact(w, x, y, z);
        if (s < thr_s || (s == thr_s && p.idx <= thr_i)) {
            unsigned long long pos = atomicAdd(&e.st->count, 1ull);
            if (pos < e.cap) {
                Cand c;
                c.score = s;
                c.idx = p.idx;
                c.w = (uint32_t)w;
                c.x = (uint32_t)x;
                c.y = (uint32_t)y;
                c.z = (uint32_t)z;
                e.cand[pos] = c;
            } else {
                atomicOr(&e.st->overflow, 1u);
            }
        }
    }
}

template <int FFUN>
__global__ void __launch_bounds__(256) k_exact(ExactArgs e) {
    exact_range<FFUN>(e, (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x,
                      (unsigned long long)gridDim.x * blockDim.x);
}

/* Clears the pending list at the start of a band. */
static __global__ void k_reset_pending(ScanState *st) {
    if (threadIdx.x != 0) return;
    st->pend_count = 0;
    st->pend_overflow = 0;
}

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

static __global__ void __launch_bounds__(256) k_sel_hist(const Cand *__restrict__ cand, ScanState *st) {
    if (st->done) return;
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const SelPrefix s = load_prefix(st);
    const unsigned long long n = st->count;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        uint64_t hi = order_bits(cand[i].score), lo = cand[i].idx;
        if (key_match(hi, lo, s)) atomicAdd(&h[key_digit(hi, lo, s.bits_done)], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&st->hist[threadIdx.x], h[threadIdx.x]);
}

static __global__ void k_sel_pick(ScanState *st) {
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

static __global__ void __launch_bounds__(256) k_sel_gather(const Cand *__restrict__ cand, ScanState *st,
                                                     uint64_t *small) {
    if (st->done != 2) return;
    const SelPrefix s = load_prefix(st);
    const unsigned long long n = st->count;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        uint64_t hi = order_bits(cand[i].score), lo = cand[i].idx;
        if (key_match(hi, lo, s)) {
            uint32_t p = atomicAdd(&st->gather_n, 1u);
            if (p < SEL_SMALL) {
                small[2 * p] = hi;
                small[2 * p + 1] = lo;
            }
        }
    }
}

static __global__ void __launch_bounds__(1024) k_sel_small(ScanState *st, const uint64_t *small) {
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

/* Keeps records with key <= the selected key; records of workgroup tiles in
 * [drop_t0, drop_t1) (geometry tv, n_a4 as in tile_index) are dropped too. */
static __global__ void __launch_bounds__(256) k_compact(const Cand *__restrict__ src, ScanState *st,
                                                  Cand *__restrict__ dst, int64_t drop_t0,
                                                  int64_t drop_t1, int32_t tv, int32_t n_a4) {
    const uint64_t khi = st->key_hi, klo = st->key_lo;
    const unsigned long long n = st->count;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        Cand c = src[i];
        uint64_t hi = order_bits(c.score);
        bool keep = hi < khi || (hi == khi && c.idx <= klo);
        if (keep && drop_t1 > drop_t0) {
            const int64_t a = (int64_t)(c.idx >> 32) >> 1;
            const int64_t b = (int64_t)((c.idx >> 3) & 0x1FFFFFFFull) >> 1;
            const int64_t t = tile_index(a, b, tv, n_a4);
            keep = t < drop_t0 || t >= drop_t1;
        }
        if (keep) {
            unsigned long long p = atomicAdd(&st->count2, 1ull);
            dst[p] = c;
        }
    }
}

/* Threshold score and filter coefficient of a selected key. */
__device__ __forceinline__ void adopt_threshold(ScanState *st, uint64_t key_hi, uint64_t key_lo,
                                                int32_t ffun) {
    const double thr = from_order_bits(key_hi);
    st->thr_score = thr;
    st->thr_idx = key_lo;
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

/* After a compaction: adopt the compacted buffer and the new threshold. */
static __global__ void k_finalize(ScanState *st, int32_t ffun) {
    if (threadIdx.x != 0) return;
    st->count = st->count2;
    st->count2 = 0;
    st->overflow = 0;
    adopt_threshold(st, st->key_hi, st->key_lo, ffun);
}

/* Resets the select state for a new select of the rank-th smallest key. */
static __global__ void k_sel_reset(ScanState *st, uint32_t rank) {
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

/* Top-k maintenance of one scan by one 1024-thread workgroup: after a
 * candidate overflow, the k-th key among the records kept becomes the
 * threshold and the band's records (fp4 tile numbering) are dropped (the
 * band is redone); otherwise, with more than k candidates, exactly the k
 * smallest keys are kept, compacted in place.  Resets the pending list for
 * the next round; writes a report when d.report is set.                    */
__device__ __forceinline__ void topk_wg_body(const TopkDesc &d) {
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
        if (d.report) {
        ItemReport r;
        r.raw_count = raw;
        r.count = work ? s_wptr : (pend_ovf ? raw : n);
        r.pend_count = pend_n;
        r.pend_overflow = pend_ovf;
        r.cand_overflow = pend_ovf ? 0u : cand_ovf;
        *d.report = r;
        }
        st->pend_count = 0;
        st->pend_overflow = 0;
    }
}

/* The same for one scan, its descriptor by value (single scans). */
static __global__ void __launch_bounds__(TOPK_THREADS) k_topk_one(TopkDesc d) { topk_wg_body(d); }

} // namespace gepi

#endif
