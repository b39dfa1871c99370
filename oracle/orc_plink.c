/*
 * orc_plink.c -- oracle restatement of the reference's PLINK binary reader.
 * Test infrastructure only (see oracle.h).
 *
 *   .fam  PlinkReader.h:345-373  lines shorter than 5 chars skipped; sample name
 *         is the first token; sex and phenotype are the 5th and 6th fields read
 *         with "%*s %*s %*d %*d %d %d".
 *   .bim  PlinkReader.h:376-405  lines shorter than 9 chars skipped; each line
 *         becomes two rows: 2i carries allele column 5, 2i+1 allele column 6.
 *   .bed  BinPlinkReader.h:58-98 magic 0x1b6c and exact size N*ceil(n/4)+3;
 *         BinPlinkReader.h:160-224 decode: code 0 -> A1+valid, 2 -> A1+A2+valid,
 *         3 -> A2+valid, 1 -> missing (no bit); sample s -> bit s%64 of word s/64.
 *   masks PlinkReader.h:408-440  one mask per phenotype value; -9 contributes 0.
 *
 * Where the reference is undefined or silently wrong the oracle fails closed:
 * a .bed whose mode byte is not 1 (sample-major; the reference decodes it as
 * SNP-major), a .bim line whose numeric fields do not parse (the reference's
 * stream keeps its fail bit and later SNPs get stale values), and a .fam line
 * whose sscanf stops early (the reference reads uninitialised ints) gets
 * sex 0 / phenotype -9.  Chromosome codes X, Y, XY, MT map to 23..26 as PLINK
 * does; the reference only accepts integers there.
 */
#include "oracle.h"

#include <ctype.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[512];

const char *orc_last_error(void) { return g_err; }

static int fail(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

static char *xstrdup(const char *s) {
    size_t n = strlen(s);
    char *r = (char *)malloc(n + 1);
    memcpy(r, s, n + 1);
    return r;
}

/* Reads one line (any length) without its trailing newline; NULL at EOF. */
static char *read_line(FILE *f, char **buf, size_t *cap) {
    size_t n = 0;
    int c;
    if (*buf == NULL) {
        *cap = 256;
        *buf = (char *)malloc(*cap);
    }
    while ((c = fgetc(f)) != EOF) {
        if (c == '\n') break;
        if (n + 1 >= *cap) {
            *cap *= 2;
            *buf = (char *)realloc(*buf, *cap);
        }
        (*buf)[n++] = (char)c;
    }
    if (c == EOF && n == 0) return NULL;
    (*buf)[n] = 0;
    return *buf;
}

static int parse_chr(const char *tok, int32_t *out) {
    char *end;
    long v = strtol(tok, &end, 10);
    if (end != tok && *end == 0) {
        *out = (int32_t)v;
        return 0;
    }
    if (!strcmp(tok, "X")) { *out = 23; return 0; }
    if (!strcmp(tok, "Y")) { *out = 24; return 0; }
    if (!strcmp(tok, "XY")) { *out = 25; return 0; }
    if (!strcmp(tok, "MT")) { *out = 26; return 0; }
    return -1;
}

void orc_plink_free(orc_plink *p) {
    if (!p) return;
    free(p->bits);
    free(p->valid);
    free(p->mask1);
    free(p->mask2);
    free(p->pheno);
    if (p->sample_name)
        for (int32_t i = 0; i < p->n_samples; i++) free(p->sample_name[i]);
    free(p->sample_name);
    if (p->snp_id)
        for (int64_t i = 0; i < p->n_rows; i += 2) free(p->snp_id[i]);
    free(p->snp_id);
    free(p->allele);
    free(p->chr);
    free(p->pos);
    free(p);
}

static int read_fam(const char *root, orc_plink *p) {
    char fname[4096];
    snprintf(fname, sizeof fname, "%s.fam", root);
    FILE *f = fopen(fname, "r");
    if (!f) return fail("cannot open %s", fname);
    char *buf = NULL;
    size_t cap = 0;
    int32_t n = 0, ncap = 64;
    p->pheno = (int32_t *)malloc(sizeof(int32_t) * ncap);
    p->sample_name = (char **)malloc(sizeof(char *) * ncap);
    char *line;
    while ((line = read_line(f, &buf, &cap)) != NULL) {
        if (strlen(line) < 5) continue;
        if (n == ncap) {
            ncap *= 2;
            p->pheno = (int32_t *)realloc(p->pheno, sizeof(int32_t) * ncap);
            p->sample_name = (char **)realloc(p->sample_name, sizeof(char *) * ncap);
        }
        size_t k = strcspn(line, " \t");
        char save = line[k];
        line[k] = 0;
        p->sample_name[n] = xstrdup(line);
        line[k] = save;
        int d = 0, e = -9;
        if (sscanf(line, "%*s %*s %*d %*d %d %d", &d, &e) < 2) {
            /* the reference reads uninitialised ints here */
            d = 0;
            e = -9;
        }
        p->pheno[n] = e;
        n++;
    }
    free(buf);
    fclose(f);
    p->n_samples = n;
    return 0;
}

static int read_bim(const char *root, orc_plink *p) {
    char fname[4096];
    snprintf(fname, sizeof fname, "%s.bim", root);
    FILE *f = fopen(fname, "r");
    if (!f) return fail("cannot open %s", fname);
    char *buf = NULL;
    size_t cap = 0;
    int64_t r = 0, rcap = 128;
    p->snp_id = (char **)malloc(sizeof(char *) * rcap);
    p->allele = (char *)malloc(rcap);
    p->chr = (int32_t *)malloc(sizeof(int32_t) * rcap);
    p->pos = (int32_t *)malloc(sizeof(int32_t) * rcap);
    char *line;
    int64_t lineno = 0;
    while ((line = read_line(f, &buf, &cap)) != NULL) {
        lineno++;
        if (strlen(line) < 9) continue;
        char t[6][1024];
        int nt = sscanf(line, "%1023s %1023s %1023s %1023s %1023s %1023s", t[0], t[1], t[2], t[3],
                        t[4], t[5]);
        int32_t chr, pos;
        char *end;
        if (nt < 6 || parse_chr(t[0], &chr) != 0) {
            free(buf);
            fclose(f);
            return fail("%s:%lld: malformed line", fname, (long long)lineno);
        }
        strtod(t[2], &end);
        if (end == t[2] || *end) {
            free(buf);
            fclose(f);
            return fail("%s:%lld: malformed genetic distance", fname, (long long)lineno);
        }
        long pv = strtol(t[3], &end, 10);
        if (end == t[3] || *end) {
            free(buf);
            fclose(f);
            return fail("%s:%lld: malformed position", fname, (long long)lineno);
        }
        pos = (int32_t)pv;
        if (r + 2 > rcap) {
            rcap *= 2;
            p->snp_id = (char **)realloc(p->snp_id, sizeof(char *) * rcap);
            p->allele = (char *)realloc(p->allele, rcap);
            p->chr = (int32_t *)realloc(p->chr, sizeof(int32_t) * rcap);
            p->pos = (int32_t *)realloc(p->pos, sizeof(int32_t) * rcap);
        }
        char *id = xstrdup(t[1]);
        p->snp_id[r] = id;
        p->snp_id[r + 1] = id; /* shared; freed once via the even row */
        p->allele[r] = t[4][0];
        p->allele[r + 1] = t[5][0];
        p->chr[r] = p->chr[r + 1] = chr;
        p->pos[r] = p->pos[r + 1] = pos;
        r += 2;
    }
    free(buf);
    fclose(f);
    p->n_rows = r;
    return 0;
}

static int read_bed(const char *root, orc_plink *p) {
    char fname[4096];
    snprintf(fname, sizeof fname, "%s.bed", root);
    FILE *f = fopen(fname, "rb");
    if (!f) return fail("cannot open %s", fname);
    int64_t nvar = p->n_rows / 2;
    int64_t bps = (p->n_samples + 3) / 4;
    int64_t expect = nvar * bps + 3;
    fseek(f, 0, SEEK_END);
    int64_t size = ftell(f);
    fseek(f, 0, SEEK_SET);
    unsigned char hdr[3] = {0, 0, 0};
    if (size < 3 || fread(hdr, 1, 3, f) != 3) {
        fclose(f);
        return fail("%s: too short", fname);
    }
    if (hdr[0] != 0x6c || hdr[1] != 0x1b) {
        fclose(f);
        return fail("%s: incorrect magic %02x%02x", fname, hdr[1], hdr[0]);
    }
    if (size != expect) {
        fclose(f);
        return fail("%s: size (%lld) different than expected (%lld)", fname, (long long)size,
                    (long long)expect);
    }
    if (hdr[2] != 1) {
        fclose(f);
        return fail("%s: sample-major .bed (mode byte %d) is not supported", fname, hdr[2]);
    }
    unsigned char *data = (unsigned char *)malloc((size_t)(expect - 3) + 1);
    if (expect > 3 && fread(data, 1, (size_t)(expect - 3), f) != (size_t)(expect - 3)) {
        free(data);
        fclose(f);
        return fail("%s: short read", fname);
    }
    fclose(f);
    const int32_t W = p->words;
    size_t tot = (size_t)p->n_rows * (size_t)W;
    p->bits = (uint64_t *)calloc(tot ? tot : 1, sizeof(uint64_t));
    p->valid = (uint64_t *)calloc(tot ? tot : 1, sizeof(uint64_t));
    int64_t missing = 0;
#pragma omp parallel for schedule(static) reduction(+ : missing)
    for (int64_t i = 0; i < nvar; i++) {
        const unsigned char *src = data + i * bps;
        uint64_t *a1 = p->bits + (size_t)(2 * i) * W;
        uint64_t *a2 = p->bits + (size_t)(2 * i + 1) * W;
        uint64_t *v1 = p->valid + (size_t)(2 * i) * W;
        uint64_t *v2 = p->valid + (size_t)(2 * i + 1) * W;
        for (int32_t s = 0; s < p->n_samples; s++) {
            int code = (src[s >> 2] >> (2 * (s & 3))) & 3;
            uint64_t bit = 1ull << (s & 63);
            int wd = s >> 6;
            switch (code) {
            case 0: a1[wd] |= bit; v1[wd] |= bit; break;
            case 2: a1[wd] |= bit; a2[wd] |= bit; v1[wd] |= bit; break;
            case 3: a2[wd] |= bit; v1[wd] |= bit; break;
            default: missing++; break;
            }
        }
        for (int32_t wd = 0; wd < W; wd++) v2[wd] = v1[wd];
    }
    free(data);
    p->total_missing = missing;
    return 0;
}

static void classify(orc_plink *p) {
    const int32_t W = p->words;
    int has1 = 0, has2 = 0;
    for (int32_t s = 0; s < p->n_samples; s++) {
        if (p->pheno[s] == 1) has1 = 1;
        if (p->pheno[s] == 2) has2 = 1;
    }
    p->mask1 = has1 ? (uint64_t *)calloc(W ? W : 1, sizeof(uint64_t)) : NULL;
    p->mask2 = has2 ? (uint64_t *)calloc(W ? W : 1, sizeof(uint64_t)) : NULL;
    for (int32_t s = 0; s < p->n_samples; s++) {
        uint64_t bit = 1ull << (s & 63);
        if (p->pheno[s] == 1) p->mask1[s >> 6] |= bit;
        if (p->pheno[s] == 2) p->mask2[s >> 6] |= bit;
    }
}

int orc_plink_read(const char *root, orc_plink **out) {
    *out = NULL;
    orc_plink *p = (orc_plink *)calloc(1, sizeof(orc_plink));
    if (read_fam(root, p) != 0 || read_bim(root, p) != 0) {
        orc_plink_free(p);
        return -1;
    }
    p->words = (p->n_samples + 63) / 64;
    if (read_bed(root, p) != 0) {
        orc_plink_free(p);
        return -1;
    }
    classify(p);
    *out = p;
    return 0;
}
