/*
 * gepi_cli.cpp -- the `gepicount` command line (replaces src/gEpiCount.cpp).
 *
 * Flags and defaults follow getParams/printHelp (src/gEpiCount.cpp:98-187):
 *   -i/--input-root BASE, -f/--fisher-pvals, -G/--gpu-num N, -N/--topmost M,
 *   -t/--alternate-grid, -D/--print-diagonal, --block-sizes X Y,
 *   --max-grid-area N, --shared-memory, --miss-snps F, --miss-samples F.
 * -g/--gold is refused: the CPU gold scan is the test oracle (oracle/), and
 * this tool only runs on the GPU.  Output streams follow main() (:47-96): the
 * informational lines go to stdout, progress and the ranked report
 * (printInteractions, src/misc/print_misc.h:56-78) to stderr.
 * Exit codes: 0 ok, 3 bad arguments (as the reference), 1 any other error.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gepicount.h"

namespace {

const char *PROGNAME = "gEpiCount";

void print_help() {
    fprintf(stderr,
            "%s -i <input-root> [options]\n\n"
            "   -i | --input-root      BASE Use BASE as full path and filename for .bed,.bim,.fam\n\n"
            "   -f | --fisher-pvals         Computes Fisher P-value instead of Entropy [no]\n"
            "   -G | --gpu-num         N    Use GPU number N  [0]\n"
            "   -t | --alternate-grid       Accepted for compatibility (no effect)\n"
            "   -D | --print-diagonal       Print progress and a sample of the current ranking\n"
            "                               (stride 100000) after every band [no]\n"
            "   -N | --topmost         M    Maintain a list of M results  [%d]\n\n"
            "        --block-sizes     N M  Accepted for compatibility; N*M must be a multiple of 32 [16,16]\n"
            "        --max-grid-area   N    Cap evaluations per kernel launch at N x 1024 x 1024 [auto]\n"
            "        --shared-memory        Accepted for compatibility (tiles always use LDS)\n"
            "        --count-path      P    popcount | mfma | fp4 | fp4rows | auto  [auto]\n"
            "        --miss-snps       F    Max missing values fraction for SNPs [0.05]\n"
            "        --miss-samples    F    Max missing values fraction for samples  [0.05]\n"
            "        --qa-remove            Drop SNPs and samples beyond those fractions before\n"
            "                               scanning (QA on the GPU) [no]\n"
            "   -g | --gold                 Not available: the CPU gold scan is oracle/gepicount_gold\n\n",
            PROGNAME, GEPI_DEFAULT_TOPMOST);
}

bool need(int j, int argc, int extra) { return j + extra < argc; }

/* -D: progress and a strided sample of the current ranking after every band,
 * as the reference's tile loop prints them (src/kernels/kernel1.cu:420-438;
 * stride and context of src/misc/print_misc.h:30-31).                      */
struct ProgressCtx {
    const gepi_plink *pr;
    const int64_t *kept; /* --qa-remove: original variant of each scanned one */
};

void print_progress(const gepi_progress *pg, void *user) {
    const ProgressCtx *ctx = static_cast<const ProgressCtx *>(user);
    const double total_s = pg->ms_elapsed / 1000.0;
    const double pct = pg->tiles_total > 0 ? 100.0 * (double)pg->tiles_done / (double)pg->tiles_total
                                           : 100.0;
    double rem = pg->ms_remaining / 1000.0;
    const char *unit = "sec";
    if (rem > 3600.0) {
        rem /= 3600.0;
        unit = "hr";
    }
    fprintf(stderr, "K1: Total time %gsec (%g%%) Remaining: %g%s\n", total_s, pct, rem, unit);
    fprintf(stderr, "K1: Last processed: tile %lld of %lld\n", (long long)pg->tiles_done,
            (long long)pg->tiles_total);
    std::vector<gepi_result> tops(pg->tops, pg->tops + pg->n_tops);
    if (ctx->kept)
        for (gepi_result &r : tops) {
            r.sA = 2 * ctx->kept[r.sA / 2] + (r.sA & 1);
            r.sB = 2 * ctx->kept[r.sB / 2] + (r.sB & 1);
        }
    gepi_print_interactions(ctx->pr, tops.data(), (int64_t)tops.size(), 100000, 5);
}

} // namespace

int main(int argc, char **argv) {
    fprintf(stderr, "%s, version: %s\n\n", PROGNAME, gepi_version());
    gepi_params par;
    gepi_params_init(&par);
    std::string root;
    float miss_snps = 0.05f, miss_samp = 0.05f;
    bool ok = true, diag = false, qa_remove = false;
    for (int j = 1; j < argc; ++j) {
        std::string a(argv[j]);
        if (a == "-g" || a == "--gold") {
            fprintf(stderr, "The CPU gold scan is not part of this tool; use oracle/gepicount_gold.\n");
            ok = false;
        } else if (a == "-t" || a == "--alternate-grid" || a == "--shared-memory") {
        } else if ((a == "-G" || a == "--gpu-num") && need(j, argc, 1)) {
            par.gpu_num = atoi(argv[++j]);
        } else if (a == "-D" || a == "--print-diagonal") {
            diag = true;
            par.intermediate = 1;
        } else if (a == "-f" || a == "--fisher-pvals") {
            par.ffun = GEPI_FFUNC_FISHERPV;
        } else if ((a == "-i" || a == "--input-root") && need(j, argc, 1)) {
            root = argv[++j];
        } else if ((a == "-N" || a == "--topmost") && need(j, argc, 1)) {
            par.topmost = atoll(argv[++j]);
        } else if (a == "--block-sizes" && need(j, argc, 2)) {
            par.blk_x = atoi(argv[++j]);
            par.blk_y = atoi(argv[++j]);
        } else if (a == "--max-grid-area" && need(j, argc, 1)) {
            par.max_grid_area = atoi(argv[++j]);
        } else if (a == "--count-path" && need(j, argc, 1)) {
            std::string v(argv[++j]);
            if (v == "popcount") par.count_path = 1;
            else if (v == "mfma") par.count_path = 2;
            else if (v == "fp4") par.count_path = 3;
            else if (v == "fp4rows") par.count_path = 4;
            else if (v == "auto") par.count_path = 0;
            else {
                fprintf(stderr, "Unknown count path: %s\n", v.c_str());
                ok = false;
            }
        } else if (a == "--miss-snps" && need(j, argc, 1)) {
            miss_snps = (float)atof(argv[++j]);
        } else if (a == "--miss-samples" && need(j, argc, 1)) {
            miss_samp = (float)atof(argv[++j]);
        } else if (a == "--qa-remove") {
            qa_remove = true;
        } else {
            fprintf(stderr, "Unknown option: %s\n", a.c_str());
            ok = false;
        }
    }
    if ((par.blk_x * par.blk_y) % 32 != 0) {
        fprintf(stderr, "Block size not multiple of war size (32)\n");
        ok = false;
    }
    if (root.empty()) {
        fprintf(stderr, "Insufficient arguments\n");
        ok = false;
    }
    if (par.topmost < 0) {
        fprintf(stderr, "topmost must be >= 0\n");
        ok = false;
    }
    if (!ok) {
        print_help();
        return 3;
    }

    gepi_plink *pr = nullptr;
    if (gepi_plink_open(root.c_str(), &pr) != GEPI_OK) {
        printf("Error in %s\n", root.c_str());
        fprintf(stderr, "%s\n", gepi_last_error());
        return 1;
    }
    printf("\nElement Size: %d\n", 8);
    size_t dsz = (size_t)8 * pr->words * pr->n_rows;
    printf("Size of each Data Matrix: %zu (%zuMB)\n", dsz, (dsz + (1 << 20) - 1) >> 20);
    gepi_plink_check_quality(pr, miss_snps, miss_samp, nullptr, nullptr);
    if (!pr->mask_ctrl || !pr->mask_case) {
        fprintf(stderr, "Error in class mask data\n");
        gepi_plink_close(pr);
        return 1;
    }
    size_t tsz = (size_t)par.topmost * sizeof(gepi_result);
    printf("Size of array of %lld topmost interaction results: %zu (%zuMB)\n",
           (long long)par.topmost, tsz, (tsz + (1 << 20) - 1) >> 20);
    fflush(stdout);

    std::vector<gepi_result> tops((size_t)(par.topmost > 0 ? par.topmost : 1));
    int64_t n = 0;
    gepi_scan_stats st;
    ProgressCtx progress{pr, nullptr};
    if (diag) gepi_set_progress(print_progress, &progress);
    gepi_panel *panel = nullptr;
    int rc = gepi_panel_create_bed(pr->bed_payload, pr->n_rows / 2, pr->n_samples, par.gpu_num,
                                   &panel);
    if (rc == GEPI_OK && qa_remove) {
        /* checkQuality(remove = true) on the device: scan the conforming
         * variants over the conforming samples, then report original rows */
        gepi_panel *kept = nullptr;
        std::vector<int64_t> idx((size_t)(pr->n_rows / 2 > 0 ? pr->n_rows / 2 : 1));
        std::vector<uint64_t> keep((size_t)(pr->words > 0 ? pr->words : 1));
        int64_t nk = 0;
        rc = gepi_panel_filter(panel, miss_snps, miss_samp, &kept, idx.data(), &nk, keep.data());
        gepi_panel_destroy(panel);
        panel = nullptr;
        if (rc == GEPI_OK) {
            std::vector<uint64_t> m0((size_t)pr->words), m1((size_t)pr->words);
            for (int32_t w = 0; w < pr->words; w++) {
                m0[(size_t)w] = pr->mask_ctrl[w] & keep[(size_t)w];
                m1[(size_t)w] = pr->mask_case[w] & keep[(size_t)w];
            }
            printf("QA removal: %lld of %lld SNPs kept\n", (long long)nk,
                   (long long)(pr->n_rows / 2));
            progress.kept = idx.data();
            rc = gepi_scan(kept, m0.data(), m1.data(), &par, tops.data(), &n, &st);
            gepi_panel_destroy(kept);
            for (int64_t i = 0; i < n && rc == GEPI_OK; i++) {
                gepi_result &r = tops[(size_t)i];
                r.sA = 2 * idx[(size_t)(r.sA / 2)] + (r.sA & 1);
                r.sB = 2 * idx[(size_t)(r.sB / 2)] + (r.sB & 1);
            }
        }
    } else if (rc == GEPI_OK) {
        rc = gepi_scan(panel, pr->mask_ctrl, pr->mask_case, &par, tops.data(), &n, &st);
        gepi_panel_destroy(panel);
    }
    if (rc != GEPI_OK) {
        fprintf(stderr, "E[%d]: scan failed (%d): %s\n", par.gpu_num, rc, gepi_last_error());
        gepi_plink_close(pr);
        return 1;
    }
    fprintf(stderr,
            "K1: objective %s, %lld evaluations, %d+%d class samples, count path %d, %d launches\n"
            "K1: Time: scan %.3f ms (count kernels %.3f ms, layout %.3f ms, top-k %.3f ms), "
            "exact evaluations %lld\n",
            par.ffun == GEPI_FFUNC_ENTROPY ? "ENTROPY" : "FISHER PV", (long long)st.evals,
            st.n_class0, st.n_class1, st.count_path, st.bands, st.ms_scan, st.ms_count_kernels,
            st.ms_prepare, st.ms_select, (long long)st.exact_evals);
    gepi_print_interactions(pr, tops.data(), n, 1, 0);
    gepi_plink_close(pr);
    printf("Cleaned\n");
    return 0;
}
