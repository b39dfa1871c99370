/*
 * orc_report.c -- oracle report printer and gold command line.
 * Test infrastructure only (see oracle.h).
 *
 * orc_print_report restates printInteractions/printInteraction
 * (src/misc/print_misc.h:56-78): header line, then one line per result,
 * "%d:\t%d[%s %s %d:%d]\t%s\t%d[%s %s %d:%d]\t%g\t%g\t(%d,%d,%d,%d)\n" with the
 * objective printed as float and FPv = fisher_pval(w, y-w, x, z-x) recomputed
 * from the counts (print_misc.h:73).
 *
 * main() mirrors `gEpiCount -g` (src/gEpiCount.cpp:47-96, flags :118-153):
 * it reads the panel, runs the gold scan and prints the report to stderr.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

static const char *FUNC_NAME[6] = {"AND", "NAND", "ANDN", "NANDN", "XOR", "Invalid"};

void orc_print_report(FILE *f, const orc_plink *p, const orc_result *tops, int64_t n) {
    fprintf(f, "\nRank:\tidxA[id allele chr:pos]\tFunc\tidxB[id allele chr:pos]\tObj\tFPv\t(c00,c01,"
               "c10,c11)\n");
    for (int64_t i = 0; i < n; i++) {
        const orc_result *r = &tops[i];
        int64_t a = r->sA, b = r->sB;
        float pv = orc_fisher_ref(r->w, r->y - r->w, r->x, r->z - r->x);
        int fn = (r->fun >= 0 && r->fun < 5) ? r->fun : 5;
        fprintf(f, "%d:\t%d[%s %c %d:%d]\t%s\t%d[%s %c %d:%d]\t%g\t%g\t(%d,%d,%d,%d)\n", (int)i,
                (int)a, p->snp_id[a], p->allele[a], p->chr[a], p->pos[a], FUNC_NAME[fn], (int)b,
                p->snp_id[b], p->allele[b], p->chr[b], p->pos[b], (double)(float)r->obj,
                (double)pv, r->w, r->y - r->w, r->x, r->z - r->x);
    }
}

#ifdef ORC_GOLD_MAIN
static void usage(void) {
    fprintf(stderr,
            "gepicount_gold -i <input-root> [options]   (oracle CPU gold scan)\n"
            "   -i | --input-root  BASE  .bed/.bim/.fam root\n"
            "   -f | --fisher-pvals      rank by Fisher p instead of entropy\n"
            "   -N | --topmost     M     keep M results [1000000]\n"
            "   -T | --threads     T     OpenMP threads [all]\n");
}

int main(int argc, char **argv) {
    const char *root = NULL;
    int64_t topk = 1000000;
    int ffun = ORC_FFUNC_ENTROPY, nthreads = 0;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if ((!strcmp(a, "-i") || !strcmp(a, "--input-root")) && i + 1 < argc) root = argv[++i];
        else if ((!strcmp(a, "-N") || !strcmp(a, "--topmost")) && i + 1 < argc)
            topk = atoll(argv[++i]);
        else if (!strcmp(a, "-f") || !strcmp(a, "--fisher-pvals")) ffun = ORC_FFUNC_FISHERPV;
        else if (!strcmp(a, "-g") || !strcmp(a, "--gold")) continue;
        else if ((!strcmp(a, "-T") || !strcmp(a, "--threads")) && i + 1 < argc)
            nthreads = atoi(argv[++i]);
        else {
            fprintf(stderr, "Unknown option: %s\n", a);
            usage();
            return 3;
        }
    }
    if (!root) {
        usage();
        return 3;
    }
    orc_plink *p;
    if (orc_plink_read(root, &p) != 0) {
        fprintf(stderr, "Error in %s: %s\n", root, orc_last_error());
        return 1;
    }
    if (!p->mask1 || !p->mask2) {
        fprintf(stderr, "Error in class mask data\n");
        orc_plink_free(p);
        return 1;
    }
    orc_result *tops = (orc_result *)calloc((size_t)(topk > 0 ? topk : 1), sizeof(orc_result));
    int64_t n = 0;
    int rc = orc_gold_scan(p->bits, p->valid, p->mask1, p->mask2, p->n_rows, p->words, topk, ffun,
                           nthreads, tops, &n);
    if (rc != 0) {
        fprintf(stderr, "gold scan failed\n");
        return 1;
    }
    orc_print_report(stderr, p, tops, n);
    free(tops);
    orc_plink_free(p);
    return 0;
}
#endif
