/*
 * oracle.h -- CPU oracle for the gEpiCount all-pairs Boolean-epistasis scan.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gepicount_amd/, the C-ABI
 * library, the CLI) links, loads or calls this code.  It is used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
 *
 * It restates, in plain C (+ OpenMP), the reference's CPU gold path:
 *   - PLINK .fam/.bim/.bed reading   src/reader/PlinkReader.h:130-149, :345-440
 *                                    src/reader/BinPlinkReader.h:58-244
 *   - the gold scan                  src/proc_gold_1.h:68-178, src/proc_gold.h:38-71
 *   - entropy objective              src/kernels/entropy_inline.h:58-79
 *   - host Fisher p (report column)  src/misc/fisher_p.c:52-245
 *   - the ranked report              src/misc/print_misc.h:56-78
 *
 * Deliberate, documented differences from the reference (all shared with the
 * product so that both produce the same bits):
 *   - the entropy objective is the exactly-rounded FP64 value of -I(F;C),
 *     summed in double-double from a t*log2(t) table, instead of float32 with
 *     log2f; ties are broken by (sA, sB, fun), which is the order the reference
 *     gold produces for equal scores (scan order + stable sort);
 *   - the Fisher ranking objective (-f) is log10 of the two-sided p with R's
 *     relative tie tolerance 1+1e-7, evaluated by an exact hypergeometric
 *     recurrence (IEEE +,-,*,/ only) so that CPU and GPU agree bit for bit;
 *   - exactly min(topmost, 5*P) real results are returned (no zero padding).
 */
#ifndef GEPI_ORACLE_H
#define GEPI_ORACLE_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NUM_FUNCS 5
enum { ORC_AND = 0, ORC_NAND = 1, ORC_ANDN = 2, ORC_NANDN = 3, ORC_XOR = 4 };
enum { ORC_FFUNC_ENTROPY = 0, ORC_FFUNC_FISHERPV = 1 };

/* One ranked interaction (the fields of InteractionResult, inc/gEpiCount.h:75-82). */
typedef struct orc_result {
    int64_t sA, sB;
    int32_t w, x, y, z;
    double obj;
    int32_t fun;
    int32_t reserved;
} orc_result;

/* ---- PLINK reader restatement ---------------------------------------- */
typedef struct orc_plink {
    int64_t n_rows;          /* 2 x .bim lines                                */
    int32_t n_samples;
    int32_t words;           /* ceil(n_samples / 64)                          */
    uint64_t *bits;          /* [n_rows][words] allele bits                   */
    uint64_t *valid;         /* [n_rows][words] valid (non-missing) bits      */
    uint64_t *mask1;         /* class mask, phenotype 1 (control); NULL if none */
    uint64_t *mask2;         /* class mask, phenotype 2 (case);    NULL if none */
    int32_t *pheno;          /* per sample phenotype code                     */
    char **sample_name;      /* per sample, FID                               */
    char **snp_id;           /* per row                                       */
    char *allele;            /* per row                                       */
    int32_t *chr, *pos;      /* per row                                       */
    int64_t total_missing;
} orc_plink;

/* Returns 0 on success, negative on error; message in orc_last_error(). */
int orc_plink_read(const char *root, orc_plink **out);
void orc_plink_free(orc_plink *p);
const char *orc_last_error(void);

/* ---- scoring ----------------------------------------------------------- */
/* Exactly-rounded FP64 -I(F;C) for counts (w, x, y, z); 0 when y+z == 0. */
double orc_entropy_exact(int w, int x, int y, int z);
/* The same sum in plain double from the rounded table: |approx - exact| is
 * below 9 * log2(m) * 2^-50 (< 1e-12 for m < 2^31).  The gold scan uses it
 * only to skip the exact score of evaluations that cannot enter a full heap. */
double orc_entropy_approx(int w, int x, int y, int z);
/* The reference's host float32 expression (entropy_inline.h:70-76). */
float orc_entropy_ref_f32(int w, int x, int y, int z);
/* The reference's host Fisher log10 p (fisher_p.c:52-76), same argument order. */
float orc_fisher_ref(int n11, int n12, int n21, int n22);
/* Ranking Fisher objective: log10 two-sided p of table [[w, y-w], [x, z-x]]. */
double orc_fisher_rank(int w, int x, int y, int z);
/* Own log2 built from IEEE +,-,*,/ only (bit-reproducible on CPU and GPU). */
double orc_log2(double v);

/* ---- scans ------------------------------------------------------------- */
/* Gold scan restatement.  Planes are [n_rows][words]; masks are [words].
 * Writes min(topk, 5*P) results sorted by (obj, sA, sB, fun) ascending.
 * nthreads <= 0 means all OpenMP threads.  Returns 0 or negative error.     */
int orc_gold_scan(const uint64_t *bits, const uint64_t *valid, const uint64_t *mask0,
                  const uint64_t *mask1, int64_t n_rows, int32_t words, int64_t topk,
                  int ffun, int nthreads, orc_result *out, int64_t *n_out);

/* The same scan over the rows j in [j0, j1) only (every k <= j of each). */
int orc_gold_scan_rows(const uint64_t *bits, const uint64_t *valid, const uint64_t *mask0,
                       const uint64_t *mask1, int64_t n_rows, int32_t words, int64_t j0,
                       int64_t j1, int64_t topk, int ffun, int nthreads, orc_result *out,
                       int64_t *n_out);

/* The reference's gold scan as it runs there (one thread, 16-bit table
 * popcounts, float32 scores, proc_gold_1.h:68-178) over rows j < j_end;
 * seconds = the scan's wall time.  For timing and its own pinning tests.   */
int orc_gold_scan_ref_f32(const uint64_t *bits, const uint64_t *valid, const uint64_t *mask0,
                          const uint64_t *mask1, int64_t n_rows, int32_t words, int64_t topk,
                          int64_t j_end, orc_result *out, int64_t *n_out, double *seconds);

/* Counts for every (row pair, function) in gold scan order: for j in [0,R),
 * k in [0,j], fun in [0,5): out[4*e + {0,1,2,3}] = {w, x, y, z}.           */
int orc_all_counts(const uint64_t *bits, const uint64_t *valid, const uint64_t *mask0,
                   const uint64_t *mask1, int64_t n_rows, int32_t words, int32_t *out);

/* ---- report ------------------------------------------------------------ */
/* Writes the ranked report in the print_misc.h:56-78 format. */
void orc_print_report(FILE *f, const orc_plink *p, const orc_result *tops, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
