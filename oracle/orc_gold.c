/*
 * orc_gold.c -- oracle restatement of the reference's CPU gold scan.
 * Test infrastructure only (see oracle.h).
 *
 * process_gold_1 (src/proc_gold_1.h:68-178) visits every row pair (j, k <= j)
 * of the allele-row panel, builds the per-class masks
 *     a0 = mask0 & vA & vB,  a1 = mask1 & vA & vB           (proc_gold_1.h:109-113)
 * and for each of the five Boolean functions (proc_gold.h:45-60)
 *     AND = A&B, NAND = ~A&B, ANDN = A&~B, NANDN = ~(A|B), XOR = A^B
 * counts  w = popc(f & a0), x = popc(f & a1), y = popc(a0), z = popc(a1)
 * (proc_gold.h:65-68).  Results are ranked ascending by objective; equal
 * objectives keep scan order (j, k, fun), which is what the reference gets
 * from inserting in scan order, its strict '<' threshold and a stable sort
 * (proc_gold_1.h:119-135, :167).  The scan is restated literally (no count
 * identities), parallelised over j with per-thread bounded heaps.
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* Far above the approximate score's error bound (orc_entropy_approx). */
#define ORC_APPROX_MARGIN 1e-9

typedef struct {
    double obj;
    uint64_t idx; /* sA << 32 | sB << 3 | fun : orders like (sA, sB, fun) */
    int32_t w, x, y, z;
} gold_ent;

static inline int ent_less(const gold_ent *a, const gold_ent *b) {
    if (a->obj < b->obj) return 1;
    if (a->obj > b->obj) return 0;
    return a->idx < b->idx;
}

static int ent_cmp(const void *pa, const void *pb) {
    const gold_ent *a = (const gold_ent *)pa, *b = (const gold_ent *)pb;
    if (ent_less(a, b)) return -1;
    if (ent_less(b, a)) return 1;
    return 0;
}

/* bounded max-heap keeping the k smallest keys */
typedef struct {
    gold_ent *v;
    int64_t n, k;
} heap_t;

static void heap_sift_down(heap_t *h, int64_t i) {
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < h->n && ent_less(&h->v[m], &h->v[l])) m = l;
        if (r < h->n && ent_less(&h->v[m], &h->v[r])) m = r;
        if (m == i) return;
        gold_ent t = h->v[i];
        h->v[i] = h->v[m];
        h->v[m] = t;
        i = m;
    }
}

static void heap_push(heap_t *h, const gold_ent *e) {
    if (h->k == 0) return;
    if (h->n < h->k) {
        int64_t i = h->n++;
        h->v[i] = *e;
        while (i > 0) {
            int64_t p = (i - 1) / 2;
            if (!ent_less(&h->v[p], &h->v[i])) break;
            gold_ent t = h->v[i];
            h->v[i] = h->v[p];
            h->v[p] = t;
            i = p;
        }
        return;
    }
    if (!ent_less(e, &h->v[0])) return;
    h->v[0] = *e;
    heap_sift_down(h, 0);
}

static inline uint64_t bfun(int fun, uint64_t a, uint64_t b) {
    switch (fun) {
    case ORC_AND: return a & b;
    case ORC_NAND: return ~a & b;
    case ORC_ANDN: return a & ~b;
    case ORC_NANDN: return ~(a | b);
    default: return a ^ b;
    }
}

static void pair_counts(const uint64_t *A, const uint64_t *vA, const uint64_t *B,
                        const uint64_t *vB, const uint64_t *mask0, const uint64_t *mask1,
                        int32_t W, int32_t cnt[ORC_NUM_FUNCS][4]) {
    memset(cnt, 0, sizeof(int32_t) * ORC_NUM_FUNCS * 4);
    for (int32_t m = 0; m < W; m++) {
        uint64_t a0 = mask0[m] & vA[m] & vB[m];
        uint64_t a1 = mask1[m] & vA[m] & vB[m];
        int y = __builtin_popcountll(a0), z = __builtin_popcountll(a1);
        for (int f = 0; f < ORC_NUM_FUNCS; f++) {
            uint64_t c = bfun(f, A[m], B[m]);
            cnt[f][0] += __builtin_popcountll(c & a0);
            cnt[f][1] += __builtin_popcountll(c & a1);
            cnt[f][2] += y;
            cnt[f][3] += z;
        }
    }
}

int orc_all_counts(const uint64_t *bits, const uint64_t *valid, const uint64_t *mask0,
                   const uint64_t *mask1, int64_t R, int32_t W, int32_t *out) {
    if (R < 0 || W < 0) return -1;
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t j = 0; j < R; j++) {
        int64_t base = j * (j + 1) / 2;
        for (int64_t k = 0; k <= j; k++) {
            int32_t cnt[ORC_NUM_FUNCS][4];
            pair_counts(bits + j * W, valid + j * W, bits + k * W, valid + k * W, mask0, mask1, W,
                        cnt);
            int32_t *o = out + (base + k) * ORC_NUM_FUNCS * 4;
            memcpy(o, cnt, sizeof cnt);
        }
    }
    return 0;
}

int orc_gold_scan(const uint64_t *bits, const uint64_t *valid, const uint64_t *mask0,
                  const uint64_t *mask1, int64_t R, int32_t W, int64_t topk, int ffun,
                  int nthreads, orc_result *out, int64_t *n_out) {
    return orc_gold_scan_rows(bits, valid, mask0, mask1, R, W, 0, R, topk, ffun, nthreads, out,
                              n_out);
}

/* The gold scan restricted to the rows j in [j0, j1) (all k <= j of each).
 * The top-k of a whole scan is the top-k of the union of the top-k lists of
 * any partition of [0, R) into row ranges, because the key is total.       */
int orc_gold_scan_rows(const uint64_t *bits, const uint64_t *valid, const uint64_t *mask0,
                       const uint64_t *mask1, int64_t R, int32_t W, int64_t j0, int64_t j1,
                       int64_t topk, int ffun, int nthreads, orc_result *out, int64_t *n_out) {
    *n_out = 0;
    if (R < 0 || W < 0 || topk < 0 || (ffun != ORC_FFUNC_ENTROPY && ffun != ORC_FFUNC_FISHERPV))
        return -1;
    if (R >= (1LL << 29)) return -1;
    if (j0 < 0 || j1 > R || j0 > j1) return -1;
    int64_t total = ORC_NUM_FUNCS * (j1 * (j1 + 1) / 2 - j0 * (j0 + 1) / 2);
    int64_t k = topk < total ? topk : total;
    if (k == 0) return 0;
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    /* warm the score tables before the parallel region (they grow lazily) */
    (void)orc_entropy_exact(0, 0, W * 64, 0);
    heap_t *heaps = (heap_t *)calloc((size_t)nt, sizeof(heap_t));
    for (int t = 0; t < nt; t++) {
        heaps[t].v = (gold_ent *)malloc(sizeof(gold_ent) * (size_t)k);
        heaps[t].k = k;
    }
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4)
    for (int64_t jj = 0; jj < j1 - j0; jj++) {
        int64_t j = j1 - 1 - jj; /* long rows first */
        heap_t *h = &heaps[omp_get_thread_num()];
        for (int64_t kk = 0; kk <= j; kk++) {
            int32_t cnt[ORC_NUM_FUNCS][4];
            pair_counts(bits + j * W, valid + j * W, bits + kk * W, valid + kk * W, mask0, mask1,
                        W, cnt);
            for (int f = 0; f < ORC_NUM_FUNCS; f++) {
                gold_ent e;
                e.w = cnt[f][0];
                e.x = cnt[f][1];
                e.y = cnt[f][2];
                e.z = cnt[f][3];
                /* A full heap only admits keys below its top; an entropy more
                 * than ORC_APPROX_MARGIN above the top's cannot be one. */
                if (ffun == ORC_FFUNC_ENTROPY && h->n == h->k &&
                    orc_entropy_approx(e.w, e.x, e.y, e.z) > h->v[0].obj + ORC_APPROX_MARGIN)
                    continue;
                e.obj = (ffun == ORC_FFUNC_ENTROPY) ? orc_entropy_exact(e.w, e.x, e.y, e.z)
                                                    : orc_fisher_rank(e.w, e.x, e.y, e.z);
                e.idx = ((uint64_t)j << 32) | ((uint64_t)kk << 3) | (uint64_t)f;
                heap_push(h, &e);
            }
        }
    }
    int64_t n = 0;
    for (int t = 0; t < nt; t++) n += heaps[t].n;
    gold_ent *all = (gold_ent *)malloc(sizeof(gold_ent) * (size_t)(n ? n : 1));
    int64_t p = 0;
    for (int t = 0; t < nt; t++) {
        memcpy(all + p, heaps[t].v, sizeof(gold_ent) * (size_t)heaps[t].n);
        p += heaps[t].n;
        free(heaps[t].v);
    }
    free(heaps);
    qsort(all, (size_t)n, sizeof(gold_ent), ent_cmp);
    int64_t m = n < k ? n : k;
    for (int64_t i = 0; i < m; i++) {
        out[i].sA = (int64_t)(all[i].idx >> 32);
        out[i].sB = (int64_t)((all[i].idx >> 3) & 0x1FFFFFFFull);
        out[i].fun = (int32_t)(all[i].idx & 7);
        out[i].w = all[i].w;
        out[i].x = all[i].x;
        out[i].y = all[i].y;
        out[i].z = all[i].z;
        out[i].obj = all[i].obj;
        out[i].reserved = 0;
    }
    free(all);
    *n_out = m;
    return 0;
}
