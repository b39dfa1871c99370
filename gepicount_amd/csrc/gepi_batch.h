/*
 * gepi_batch.h -- kernels of the batched scan engine.  Many scans (panels
 * and phenotype vectors of different sizes) advance together in rounds; a
 * round is three kinds of launch whatever the number of scans (one count
 * launch per objective and job shape present):
 *
 *   k_count_f4_batch  one band of fp4 count tiles per scan in one persistent
 *                     grid: workgroups take (scan, tile) slots from per-XCD
 *                     work counters through a descriptor table
 *   k_count_f4w_batch the same for scans on wave tiles (Fisher's, and long
 *                     entropy scans): one (scan, unit of two tiles) slot per
 *                     workgroup
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
 * XCD share).  HALF: the jobs are half tiles (32 x 32 variants; the
 * descriptors' tile_begin and n_tiles count half tiles, two per tile). */
template <int FFUN, bool PERSIST, bool HALF>
__global__ void __launch_bounds__(F4_THREADS, 2)
    k_count_f4_batch(const BatchDesc *__restrict__ d, BatchMap map) {
    if (PERSIST) {
        zero_other_counters(map.wq, map.wq_parity);
        const F4BatchSrc src{d, map};
        count_f4_run<FFUN, HALF>(src);
    } else {
        const F4BatchOnceSrc src{d, map};
        count_f4_run<FFUN, HALF>(src);
    }
}

/* Wave-tile jobs (the engine's job shape 2): one unit, two 64 x 32-variant
 * tiles as four 32 x 32 wave tiles (count_f4w_unit), per workgroup; slot
 * blockIdx.x / 8 of XCD blockIdx.x % 8 as in F4BatchOnceSrc, whose ranges
 * here count units (a descriptor's tile_begin and n_tiles still count
 * tiles).  The scan's layout is the grouped 3-step one (k_class_f4<3,
 * true>).  A slot past the round's units returns before any barrier. */
template <int FFUN>
__global__ void __launch_bounds__(F4W_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_count_f4w_batch(const BatchDesc *__restrict__ d, BatchMap map) {
    __shared__ F4WShared sh;
    const F4BatchOnceSrc src{d, map};
    const F4Job j = src.grab(0);
    if (!j.ok) return;
    const CountArgs g = d[j.item].g;
    const int64_t unit = j.tile - g.tile_begin;
    constexpr bool PARK = FFUN == GEPI_FFUNC_FISHERPV;
    if (g.f32)
        count_f4w_unit<FFUN, true, PARK>(g, unit, sh);
    else
        count_f4w_unit<FFUN, false, PARK>(g, unit, sh);
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
/* k_topk_wg: the top-k maintenance of every scan of a round, one workgroup
 * per descriptor (topk_wg_body in gepi_topk.h). */
__global__ void __launch_bounds__(TOPK_THREADS) k_topk_wg(const TopkDesc *__restrict__ dd) {
    topk_wg_body(dd[blockIdx.x]);
}

} // namespace gepi

#endif
