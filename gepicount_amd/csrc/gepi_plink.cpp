/*
 * gepi_plink.cpp -- host-side input staging and report for libgepicount.
 *
 * Reader (replaces BinPlinkReader<ui8> / PlinkReader<ui8>):
 *   .fam  src/reader/PlinkReader.h:345-373: lines shorter than 5 characters
 *         are skipped; the sample name is the first token; sex and phenotype
 *         are the 5th and 6th fields (integer parse as "%d" does).
 *   .bim  PlinkReader.h:376-405: lines shorter than 9 characters are skipped;
 *         every line yields two rows, 2i for the allele in column 5 and 2i+1
 *         for the allele in column 6, sharing id, chromosome and position.
 *   .bed  src/reader/BinPlinkReader.h:58-98: magic 6c 1b and size
 *         N*ceil(n/4)+3 are checked; :160-224: code 0 -> row 2i + valid,
 *         2 -> both rows + valid, 3 -> row 2i+1 + valid, 1 -> missing;
 *         sample s is bit s%64 of word s/64.
 *   classes PlinkReader.h:408-440: a mask per phenotype value; -9 never sets
 *         a bit; classData(1) is the control mask, classData(2) the case mask.
 * Where the reference is undefined the reader fails closed: a mode byte other
 * than 1 (sample-major order, which the reference decodes as SNP-major), and
 * .bim lines whose numeric fields do not parse (the reference stream keeps its
 * fail bit and later rows get stale values).  Chromosome codes X, Y, XY and MT
 * map to 23..26 as in PLINK.
 *
 * checkQuality (PlinkReader.h:237-335), printInteractions/printInteraction
 * (src/misc/print_misc.h:56-78) and fisher_pval (src/misc/fisher_p.c:52-245)
 * are restated here for the report and the QA lines.
 */
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gepicount.h"
#include "gepi_internal.h"

namespace {

bool read_line(FILE *f, std::string &line) {
    line.clear();
    int c;
    bool any = false;
    while ((c = fgetc(f)) != EOF) {
        any = true;
        if (c == '\n') return true;
        line.push_back((char)c);
    }
    return any;
}

std::vector<std::string> tokens(const std::string &s) {
    std::vector<std::string> out;
    size_t i = 0, n = s.size();
    while (i < n) {
        while (i < n && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r')) i++;
        if (i >= n) break;
        size_t j = i;
        while (j < n && !(s[j] == ' ' || s[j] == '\t' || s[j] == '\r')) j++;
        out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

/* "%d" semantics: leading integer, anything after it ignored; false if none. */
bool scan_int(const std::string &t, long *v) {
    const char *s = t.c_str();
    char *end;
    errno = 0;
    long r = strtol(s, &end, 10);
    if (end == s) return false;
    *v = r;
    return true;
}

bool parse_chr(const std::string &t, int32_t *out) {
    char *end;
    long v = strtol(t.c_str(), &end, 10);
    if (end != t.c_str() && *end == 0) {
        *out = (int32_t)v;
        return true;
    }
    if (t == "X") { *out = 23; return true; }
    if (t == "Y") { *out = 24; return true; }
    if (t == "XY") { *out = 25; return true; }
    if (t == "MT") { *out = 26; return true; }
    return false;
}

char *dupstr(const std::string &s) {
    char *r = (char *)malloc(s.size() + 1);
    memcpy(r, s.c_str(), s.size() + 1);
    return r;
}

} // namespace

extern "C" void gepi_plink_close(gepi_plink *p) {
    if (!p) return;
    free(p->bit_data);
    free(p->valid_data);
    free(p->mask_ctrl);
    free(p->mask_case);
    free(p->pheno);
    if (p->sample_name)
        for (int32_t i = 0; i < p->n_samples; i++) free(p->sample_name[i]);
    free(p->sample_name);
    if (p->snp_id)
        for (int64_t r = 0; r < p->n_rows; r += 2) free(p->snp_id[r]);
    free(p->snp_id);
    free(p->allele);
    free(p->chr);
    free(p->pos);
    free(p->bed_payload);
    free(p);
}

extern "C" int gepi_plink_open(const char *root, gepi_plink **out) {
    if (!root || !out) return gepi_fail(GEPI_ERR_ARG, "gepi_plink_open: null argument");
    *out = nullptr;
    std::string base(root);
    gepi_plink *p = (gepi_plink *)calloc(1, sizeof(gepi_plink));

    /* .fam */
    {
        std::string fn = base + ".fam";
        FILE *f = fopen(fn.c_str(), "r");
        if (!f) {
            gepi_plink_close(p);
            return gepi_fail(GEPI_ERR_IO, "cannot open %s", fn.c_str());
        }
        std::vector<std::string> names;
        std::vector<int32_t> ph;
        std::string line;
        while (read_line(f, line)) {
            if (line.size() < 5) continue;
            size_t k = line.find_first_of(" \t");
            names.push_back(line.substr(0, k));
            std::vector<std::string> t = tokens(line);
            long sex = 0, phen = -9;
            if (t.size() < 6 || !scan_int(t[4], &sex) || !scan_int(t[5], &phen)) {
                sex = 0;
                phen = -9;
            }
            ph.push_back((int32_t)phen);
        }
        fclose(f);
        p->n_samples = (int32_t)names.size();
        p->pheno = (int32_t *)malloc(sizeof(int32_t) * (ph.size() + 1));
        p->sample_name = (char **)malloc(sizeof(char *) * (names.size() + 1));
        for (size_t i = 0; i < names.size(); i++) {
            p->pheno[i] = ph[i];
            p->sample_name[i] = dupstr(names[i]);
        }
    }
    /* .bim */
    {
        std::string fn = base + ".bim";
        FILE *f = fopen(fn.c_str(), "r");
        if (!f) {
            gepi_plink_close(p);
            return gepi_fail(GEPI_ERR_IO, "cannot open %s", fn.c_str());
        }
        std::vector<std::string> ids;
        std::vector<char> al;
        std::vector<int32_t> chr, pos;
        std::string line;
        int64_t lineno = 0;
        while (read_line(f, line)) {
            lineno++;
            if (line.size() < 9) continue;
            std::vector<std::string> t = tokens(line);
            int32_t c;
            char *end;
            bool ok = t.size() >= 6 && parse_chr(t[0], &c);
            if (ok) {
                strtod(t[2].c_str(), &end);
                ok = end != t[2].c_str() && *end == 0;
            }
            long pv = 0;
            if (ok) {
                pv = strtol(t[3].c_str(), &end, 10);
                ok = end != t[3].c_str() && *end == 0;
            }
            if (!ok) {
                fclose(f);
                gepi_plink_close(p);
                return gepi_fail(GEPI_ERR_FORMAT, "%s:%lld: malformed line", fn.c_str(),
                                 (long long)lineno);
            }
            ids.push_back(t[1]);
            al.push_back(t[4][0]);
            al.push_back(t[5][0]);
            chr.push_back(c);
            pos.push_back((int32_t)pv);
        }
        fclose(f);
        int64_t nv = (int64_t)ids.size();
        p->n_rows = 2 * nv;
        p->snp_id = (char **)malloc(sizeof(char *) * (p->n_rows + 1));
        p->allele = (char *)malloc(p->n_rows + 1);
        p->chr = (int32_t *)malloc(sizeof(int32_t) * (p->n_rows + 1));
        p->pos = (int32_t *)malloc(sizeof(int32_t) * (p->n_rows + 1));
        for (int64_t i = 0; i < nv; i++) {
            char *id = dupstr(ids[i]);
            p->snp_id[2 * i] = id;
            p->snp_id[2 * i + 1] = id;
            p->allele[2 * i] = al[2 * i];
            p->allele[2 * i + 1] = al[2 * i + 1];
            p->chr[2 * i] = p->chr[2 * i + 1] = chr[i];
            p->pos[2 * i] = p->pos[2 * i + 1] = pos[i];
        }
    }
    p->words = (p->n_samples + 63) / 64;
    /* .bed */
    {
        std::string fn = base + ".bed";
        FILE *f = fopen(fn.c_str(), "rb");
        if (!f) {
            gepi_plink_close(p);
            return gepi_fail(GEPI_ERR_IO, "cannot open %s", fn.c_str());
        }
        int64_t nv = p->n_rows / 2;
        int64_t bps = ((int64_t)p->n_samples + 3) / 4;
        int64_t expect = nv * bps + 3;
        fseek(f, 0, SEEK_END);
        int64_t size = ftell(f);
        fseek(f, 0, SEEK_SET);
        unsigned char hdr[3] = {0, 0, 0};
        if (size < 3 || fread(hdr, 1, 3, f) != 3) {
            fclose(f);
            gepi_plink_close(p);
            return gepi_fail(GEPI_ERR_FORMAT, "%s: too short", fn.c_str());
        }
        if (hdr[0] != 0x6c || hdr[1] != 0x1b) {
            fclose(f);
            gepi_plink_close(p);
            return gepi_fail(GEPI_ERR_FORMAT, "%s: incorrect magic %02x%02x", fn.c_str(), hdr[1],
                             hdr[0]);
        }
        if (size != expect) {
            fclose(f);
            gepi_plink_close(p);
            return gepi_fail(GEPI_ERR_FORMAT, "%s: size (%lld) different than expected (%lld)",
                             fn.c_str(), (long long)size, (long long)expect);
        }
        if (hdr[2] != 1) {
            fclose(f);
            gepi_plink_close(p);
            return gepi_fail(GEPI_ERR_FORMAT, "%s: sample-major .bed (mode byte %d) not supported",
                             fn.c_str(), hdr[2]);
        }
        p->bed_payload = (uint8_t *)malloc((size_t)(expect - 3) + 1);
        if (expect > 3 && fread(p->bed_payload, 1, (size_t)(expect - 3), f) != (size_t)(expect - 3)) {
            fclose(f);
            gepi_plink_close(p);
            return gepi_fail(GEPI_ERR_FORMAT, "%s: short read", fn.c_str());
        }
        fclose(f);
        const int32_t W = p->words;
        size_t tot = (size_t)p->n_rows * (size_t)W;
        p->bit_data = (uint64_t *)calloc(tot + 1, sizeof(uint64_t));
        p->valid_data = (uint64_t *)calloc(tot + 1, sizeof(uint64_t));
        int64_t missing = 0;
        const int32_t n = p->n_samples;
#pragma omp parallel for schedule(static) reduction(+ : missing)
        for (int64_t i = 0; i < nv; i++) {
            const uint8_t *src = p->bed_payload + i * bps;
            uint64_t *a1 = p->bit_data + (size_t)(2 * i) * W;
            uint64_t *a2 = a1 + W;
            uint64_t *v1 = p->valid_data + (size_t)(2 * i) * W;
            uint64_t *v2 = v1 + W;
            for (int32_t wd = 0; wd < W; wd++) {
                uint64_t b1 = 0, b2 = 0, vv = 0;
                int32_t s0 = wd * 64, s1 = s0 + 64 < n ? s0 + 64 : n;
                for (int32_t s = s0; s < s1; s++) {
                    unsigned code = (src[s >> 2] >> (2 * (s & 3))) & 3u;
                    uint64_t bit = 1ull << (s & 63);
                    if (code == 1u) {
                        missing++;
                        continue;
                    }
                    vv |= bit;
                    if (code != 3u) b1 |= bit;
                    if (code != 0u) b2 |= bit;
                }
                a1[wd] = b1;
                a2[wd] = b2;
                v1[wd] = vv;
                v2[wd] = vv;
            }
        }
        p->total_missing = missing;
    }
    /* classes */
    {
        const int32_t W = p->words;
        bool has1 = false, has2 = false;
        for (int32_t s = 0; s < p->n_samples; s++) {
            has1 |= p->pheno[s] == 1;
            has2 |= p->pheno[s] == 2;
        }
        if (has1) p->mask_ctrl = (uint64_t *)calloc((size_t)W + 1, sizeof(uint64_t));
        if (has2) p->mask_case = (uint64_t *)calloc((size_t)W + 1, sizeof(uint64_t));
        for (int32_t s = 0; s < p->n_samples; s++) {
            uint64_t bit = 1ull << (s & 63);
            if (p->pheno[s] == 1) p->mask_ctrl[s >> 6] |= bit;
            if (p->pheno[s] == 2) p->mask_case[s >> 6] |= bit;
        }
    }
    *out = p;
    return GEPI_OK;
}

extern "C" int gepi_plink_check_quality(const gepi_plink *p, float max_miss_snps,
                                        float max_miss_samples, int32_t *n_bad_samples,
                                        int32_t *n_bad_rows) {
    if (!p) return gepi_fail(GEPI_ERR_ARG, "gepi_plink_check_quality: null panel");
    const int32_t W = p->words;
    const int64_t R = p->n_rows;
    const int32_t n = p->n_samples;
    int32_t bad_s = 0, bad_r = 0;
    /* missing calls per sample in one pass over the rows: the set bits of
     * each row's ~valid within the n samples (a column walk per sample
     * reads the whole panel n times) */
    std::vector<int64_t> miss_s((size_t)n, 0);
    for (int64_t r = 0; r < R; r++) {
        const uint64_t *v = p->valid_data + (size_t)r * W;
        for (int32_t wd = 0; wd < W; wd++) {
            const int32_t s0 = wd * 64;
            uint64_t m = ~v[wd];
            if (n - s0 < 64) m &= (1ull << (n - s0)) - 1;
            for (; m; m &= m - 1) miss_s[(size_t)s0 + __builtin_ctzll(m)]++;
        }
    }
    for (int32_t s = 0; s < n; s++) {
        const int64_t valid = R - miss_s[(size_t)s];
        float miss = R ? (float)(R - valid) / (float)R : 0.0f;
        if (miss > max_miss_samples) {
            fprintf(stderr, "Sample %d:\t%s(%d)\tnMiss:%d\t(%4g)\n", s, p->sample_name[s],
                    p->pheno[s], (int)(R - valid), miss);
            bad_s++;
        }
    }
    fprintf(stderr, "\n%d Samples checked, %d with more than %4g missing values\n\n", n, bad_s,
            max_miss_samples);
    for (int64_t r = 0; r < R; r++) {
        int64_t valid = 0;
        for (int32_t wd = 0; wd < W; wd++)
            valid += __builtin_popcountll(p->valid_data[(size_t)r * W + wd]);
        float miss = n ? (float)(n - valid) / (float)n : 0.0f;
        if (miss > max_miss_snps) {
            fprintf(stderr, "SNP %lld:\t%s\tnMiss:%d\t(%4g)\n", (long long)r, p->snp_id[r],
                    (int)(n - valid), miss);
            bad_r++;
        }
    }
    fprintf(stderr, "\n%lld SNPs checked, %d with more than %4g missing values\n\n", (long long)R,
            bad_r, max_miss_snps);
    if (n_bad_samples) *n_bad_samples = bad_s;
    if (n_bad_rows) *n_bad_rows = bad_r;
    return GEPI_OK;
}

/* ---- fisher_pval: src/misc/fisher_p.c:52-245, operation for operation ----- */
namespace {

double fp_stirlerr(int n) {
    static const double sferr[16] = {
        0.0, 0.0810614667953272582196702, 0.0413406959554092940938221,
        0.02767792568499833914878929, 0.02079067210376509311152277,
        0.01664469118982119216319487, 0.01387612882307074799874573,
        0.01189670994589177009505572, 0.010411265261972096497478567,
        0.009255462182712732917728637, 0.008330563433362871256469318,
        0.007573675487951840794972024, 0.006942840107209529865664152,
        0.006408994188004207068439631, 0.005951370112758847735624416,
        0.005554733551962801371038690};
    const double S0 = 0.083333333333333333333, S1 = 0.00277777777777777777778,
                 S2 = 0.00079365079365079365079365, S3 = 0.000595238095238095238095238,
                 S4 = 0.0008417508417508417508417508;
    if (n <= 15) return sferr[n];
    double nn = (double)n * (double)n;
    if (n > 500) return (S0 - S1 / nn) / (double)n;
    if (n > 80) return (S0 - (S1 - S2 / nn) / nn) / (double)n;
    if (n > 35) return (S0 - (S1 - (S2 - S3 / nn) / nn) / nn) / (double)n;
    return (S0 - (S1 - (S2 - (S3 - S4 / nn) / nn) / nn) / nn) / (double)n;
}

double fp_bd0(double x, double np) {
    if (std::isinf(x) || std::isinf(np) || np == 0.0) return NAN;
    if (fabs(x - np) < 0.1 * (x + np)) {
        double v = (x - np) / (x + np);
        double s = (x - np) * v;
        double ej = 2 * x * v;
        v = v * v;
        for (int j = 1;; j++) {
            ej *= v;
            double s1 = s + ej / ((j << 1) + 1);
            if (s1 == s) return s1;
            s = s1;
        }
    }
    return x * log(x / np) + np - x;
}

double fp_dbinom_raw(int x, int n, double p, double q) {
    if (p == 0) return x == 0 ? 0.0 : -INFINITY;
    if (q == 0) return x == 0 ? 0.0 : -INFINITY;
    if (x == 0) {
        if (n == 0) return 0.0;
        return (p < 0.1) ? -fp_bd0((double)n, (double)n * q) - (double)n * p : (double)n * log(q);
    }
    if (x == n)
        return (q < 0.1) ? -fp_bd0((double)n, (double)n * p) - (double)n * q : (double)n * log(p);
    if (x < 0 || x > n) return -INFINITY;
    double lc = fp_stirlerr(n) - fp_stirlerr(x) - fp_stirlerr(n - x) -
                fp_bd0((double)x, (double)n * p) - fp_bd0((double)(n - x), (double)n * q);
    double lf = log(6.283185307179586476925286766559) + log((double)x) +
                log1p(-(double)x / (double)n);
    lc -= 0.5 * lf;
    return lc;
}

double fp_dhyper(int x, int r, int b, int n) {
    if (n < x || r < x || n - x > b) return -INFINITY;
    if (n == 0) return x == 0 ? 0.0 : -INFINITY;
    double p = (double)n / (double)(r + b);
    double q = (double)(r + b - n) / (double)(r + b);
    return fp_dbinom_raw(x, r, p, q) + fp_dbinom_raw(n - x, b, p, q) -
           fp_dbinom_raw(n, r + b, p, q);
}

} // namespace

extern "C" float gepi_fisher_pval(int32_t n11, int32_t n12, int32_t n21, int32_t n22) {
    if (n11 < 0 || n12 < 0 || n21 < 0 || n22 < 0) return NAN;
    int m = n11 + n21, n = n12 + n22, o = n11 + n12;
    int lo = o <= n ? 0 : o - n;
    int hi = o < m ? o : m;
    double dsumd = 0.0, dsumn = 0.0;
    double dthr = fp_dhyper(n11, m, n, o);
    for (int j = lo; j <= hi; ++j) {
        double daux = fp_dhyper(j, m, n, o);
        double daux2 = exp(daux);
        dsumd += daux2;
        if (daux <= dthr) dsumn += daux2;
    }
    return (float)(log10(dsumn) - log10(dsumd));
}

/* ---- report: src/misc/print_misc.h:56-78 ---------------------------------- */
static const char *k_func_name[6] = {"AND", "NAND", "ANDN", "NANDN", "XOR", "Invalid"};

static void print_one(int64_t num, const gepi_plink *p, const gepi_result &r) {
    int64_t a = r.sA, b = r.sB;
    float pv = gepi_fisher_pval(r.w, r.y - r.w, r.x, r.z - r.x);
    int fn = (r.fun >= 0 && r.fun < 5) ? r.fun : 5;
    fprintf(stderr, "%d:\t%d[%s %c %d:%d]\t%s\t%d[%s %c %d:%d]\t%g\t%g\t(%d,%d,%d,%d)\n",
            (int)num, (int)a, p->snp_id[a], p->allele[a], p->chr[a], p->pos[a], k_func_name[fn],
            (int)b, p->snp_id[b], p->allele[b], p->chr[b], p->pos[b], (double)(float)r.obj,
            (double)pv, r.w, r.y - r.w, r.x, r.z - r.x);
}

extern "C" int gepi_print_interactions(const gepi_plink *p, const gepi_result *tops, int64_t n,
                                       int32_t stride, int32_t context) {
    if (!p || (!tops && n > 0) || stride < 1)
        return gepi_fail(GEPI_ERR_ARG, "gepi_print_interactions: bad arguments");
    for (int64_t i = 0; i < n; i++)
        if (tops[i].sA < 0 || tops[i].sA >= p->n_rows || tops[i].sB < 0 || tops[i].sB >= p->n_rows)
            return gepi_fail(GEPI_ERR_ARG, "gepi_print_interactions: row index out of range");
    fprintf(stderr,
            "\nRank:\tidxA[id allele chr:pos]\tFunc\tidxB[id allele chr:pos]\tObj\tFPv\t(c00,c01,"
            "c10,c11)\n");
    if (stride - 2 * context < 1) context = stride / 2;
    for (int64_t j = 0; j < n; j += stride) {
        int64_t k0 = j - context < 0 ? 0 : j - context;
        int64_t k1 = j + context + 1 > n ? n : j + context + 1;
        for (int64_t k = k0; k < k1; ++k) print_one(k, p, tops[k]);
    }
    fflush(stderr);
    return GEPI_OK;
}
