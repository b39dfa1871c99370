/*
 * orc_gold_ref.c -- the reference's CPU gold scan as it runs there: one
 * thread, 16-bit lookup-table popcounts, float32 scores.  Test and
 * measurement infrastructure only (see oracle.h): it times the CPU path the
 * reference ships, not the oracle's own exact ranking.
 *
 * Restated from src/proc_gold_1.h:68-178 and src/proc_gold.h:38-71:
 *   - per row pair (j, k <= j): a0 = mask0 & vA & vB, a1 = mask1 & vA & vB
 *     (proc_gold_1.h:109-113), then for each of the five functions
 *     binFunc recomputes w, x, y, z with bitsElement, which adds the 16-bit
 *     table counts of each 16-bit slice of a word (src/misc/bitops.h:53-62);
 *   - the score is entropyDelta4's host branch in float32 with log2f
 *     (src/kernels/entropy_inline.h:70-76; orc_entropy_ref_f32);
 *   - results go into a buffer of 2*topk entries while the score is below
 *     the running limit (strict '<'); a full buffer is sorted, the limit
 *     becomes the topk-th score and the insert is retried
 *     (proc_gold_1.h:119-135); a final sort keeps topk (:167-172).
 * Equal float scores keep scan order (a stable sort), which is one of the
 * orders the reference's unstable thrust::sort can produce.
 *
 * orc_gold_scan_ref_f32 scans the rows j < j_end only, so that a run can be
 * bounded in time and extrapolated the way the reference's own progress
 * estimate does (proc_gold_1.h:140-153: time x (R / rows done)^2).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>

static unsigned char g_bits16[65536];
static int g_bits16_ready = 0;

static void preload_bits16(void) {
    if (g_bits16_ready) return;
    for (int v = 0; v < 65536; v++) {
        int c = 0;
        for (int t = v; t; t &= t - 1) c++;
        g_bits16[v] = (unsigned char)c;
    }
    g_bits16_ready = 1;
}

static inline int bits_element(uint64_t a) {
    int c = 0;
    for (int j = 0; j < 4; j++) {
        c += g_bits16[a & 0xffffu];
        a >>= 16;
    }
    return c;
}

typedef struct {
    float ent;
    int64_t sA, sB;
    int32_t fun;
    int32_t w, x, y, z;
    int64_t order; /* scan position: ties keep scan order */
} ref_ent;

static int ref_cmp(const void *pa, const void *pb) {
    const ref_ent *a = (const ref_ent *)pa, *b = (const ref_ent *)pb;
    if (a->ent < b->ent) return -1;
    if (a->ent > b->ent) return 1;
    return (a->order > b->order) - (a->order < b->order);
}

int orc_gold_scan_ref_f32(const uint64_t *bits, const uint64_t *valid, const uint64_t *mask0,
                          const uint64_t *mask1, int64_t R, int32_t W, int64_t topk, int64_t j_end,
                          orc_result *out, int64_t *n_out, double *seconds) {
    *n_out = 0;
    if (R < 0 || W < 0 || topk <= 0 || j_end < 0 || j_end > R) return -1;
    preload_bits16();
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    const int64_t cap = 2 * topk;
    ref_ent *buf = (ref_ent *)calloc((size_t)cap, sizeof(ref_ent));
    uint64_t *a0 = (uint64_t *)calloc((size_t)(W ? W : 1), sizeof(uint64_t));
    uint64_t *a1 = (uint64_t *)calloc((size_t)(W ? W : 1), sizeof(uint64_t));
    int64_t used = 0, order = 0;
    float limit = 3.402823466e+38f;
    for (int64_t j = 0; j < j_end; j++) {
        const uint64_t *A = bits + j * W, *vA = valid + j * W;
        for (int64_t k = 0; k <= j; k++) {
            const uint64_t *B = bits + k * W, *vB = valid + k * W;
            for (int32_t m = 0; m < W; m++) {
                a0[m] = mask0[m] & vA[m] & vB[m];
                a1[m] = mask1[m] & vA[m] & vB[m];
            }
            for (int fun = 0; fun < ORC_NUM_FUNCS; fun++, order++) {
                int w = 0, x = 0, y = 0, z = 0;
                for (int32_t m = 0; m < W; m++) {
                    uint64_t c;
                    switch (fun) {
                    case ORC_AND: c = A[m] & B[m]; break;
                    case ORC_NAND: c = ~A[m] & B[m]; break;
                    case ORC_ANDN: c = A[m] & ~B[m]; break;
                    case ORC_NANDN: c = ~(A[m] | B[m]); break;
                    default: c = A[m] ^ B[m]; break;
                    }
                    w += bits_element(c & a0[m]);
                    x += bits_element(c & a1[m]);
                    y += bits_element(a0[m]);
                    z += bits_element(a1[m]);
                }
                float ee = orc_entropy_ref_f32(w, x, y, z);
                if (!(ee < limit)) continue;
                if (used == cap) {
                    qsort(buf, (size_t)cap, sizeof(ref_ent), ref_cmp);
                    used = topk;
                    limit = buf[topk - 1].ent;
                    if (!(ee < limit)) continue;
                }
                ref_ent *e = &buf[used++];
                e->ent = ee;
                e->sA = j;
                e->sB = k;
                e->fun = fun;
                e->w = w;
                e->x = x;
                e->y = y;
                e->z = z;
                e->order = order;
            }
        }
    }
    qsort(buf, (size_t)used, sizeof(ref_ent), ref_cmp);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    int64_t n = used < topk ? used : topk;
    for (int64_t i = 0; i < n; i++) {
        out[i].sA = buf[i].sA;
        out[i].sB = buf[i].sB;
        out[i].w = buf[i].w;
        out[i].x = buf[i].x;
        out[i].y = buf[i].y;
        out[i].z = buf[i].z;
        out[i].obj = (double)buf[i].ent;
        out[i].fun = buf[i].fun;
        out[i].reserved = 0;
    }
    *n_out = n;
    free(buf);
    free(a0);
    free(a1);
    return 0;
}
