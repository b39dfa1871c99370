/*
 * orc_score.c -- oracle scoring functions.  Test infrastructure only (see oracle.h).
 *
 * orc_entropy_exact  -I(F;C) = (L(y)+L(z) + L(xx)+L(m-xx) - L(w)-L(y-w)-L(x)-L(z-x)
 *                    - L(m)) / m with L(t) = t*log2(t), L(0) = 0, xx = w+x, m = y+z.
 *                    This is the expression of entropy_inline.h:71-76 with the
 *                    logs of ratios expanded; every guard there removes exactly
 *                    the terms that vanish with L(0) = 0, and m = 0 gives 0.
 *                    Arguments are sorted and common ones cancelled before a
 *                    double-double sum, so the result is the exactly rounded
 *                    double and symmetric tables give identical bits.
 * orc_entropy_ref_f32  entropy_inline.h:70-76 (host branch) in float with log2f.
 * orc_fisher_ref     fisher_p.c:52-245 restated operation for operation.
 * orc_fisher_rank    log10 of R's two-sided Fisher p (tolerance 1+1e-7) from
 *                    an exact hypergeometric recurrence; see below.
 */
#include "oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* t*log2(t) table in double-double, built from binary128.                   */
static double *g_lhi = NULL, *g_llo = NULL;
static int64_t g_lmax = -1;

static void ltable_ensure(int64_t tmax) {
    if (tmax <= g_lmax) return;
    int64_t n = tmax + 1;
    if (n < 1024) n = 1024;
    double *hi = (double *)malloc(sizeof(double) * (size_t)n);
    double *lo = (double *)malloc(sizeof(double) * (size_t)n);
    hi[0] = 0.0;
    lo[0] = 0.0;
    for (int64_t t = 1; t < n; t++) {
        __float128 q = (__float128)t * log2q((__float128)t);
        double h = (double)q;
        hi[t] = h;
        lo[t] = (double)(q - (__float128)h);
    }
    free(g_lhi);
    free(g_llo);
    g_lhi = hi;
    g_llo = lo;
    g_lmax = n - 1;
}

static inline void two_sum(double a, double b, double *s, double *e) {
    double x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    *s = x;
    *e = ar + br;
}

#define CSWAP(a, b)            \
    do {                       \
        if ((a) > (b)) {       \
            int64_t t_ = (a);  \
            (a) = (b);         \
            (b) = t_;          \
        }                      \
    } while (0)

double orc_entropy_exact(int w, int x, int y, int z) {
    int64_t m = (int64_t)y + z;
    if (m == 0) return 0.0;
    int64_t xx = (int64_t)w + x;
    ltable_ensure(m);
    int64_t P[4] = {y, z, xx, m - xx};
    int64_t N[5] = {w, (int64_t)y - w, x, (int64_t)z - x, m};
    /* sort both argument lists ascending */
    for (int i = 0; i < 4; i++)
        for (int j = 0; j + 1 < 4 - i; j++) CSWAP(P[j], P[j + 1]);
    for (int i = 0; i < 5; i++)
        for (int j = 0; j + 1 < 5 - i; j++) CSWAP(N[j], N[j + 1]);
    /* cancel common arguments (multiset difference) */
    int keepP[4] = {1, 1, 1, 1}, keepN[5] = {1, 1, 1, 1, 1};
    {
        int i = 0, j = 0;
        while (i < 4 && j < 5) {
            if (P[i] == N[j]) {
                keepP[i] = 0;
                keepN[j] = 0;
                i++;
                j++;
            } else if (P[i] < N[j]) {
                i++;
            } else {
                j++;
            }
        }
    }
    double s = 0.0, e = 0.0, err;
    for (int i = 0; i < 4; i++) {
        if (!keepP[i]) continue;
        two_sum(s, g_lhi[P[i]], &s, &err);
        double t = err + g_llo[P[i]];
        e = e + t;
    }
    for (int j = 0; j < 5; j++) {
        if (!keepN[j]) continue;
        two_sum(s, -g_lhi[N[j]], &s, &err);
        double t = err - g_llo[N[j]];
        e = e + t;
    }
    double num = s + e;
    double r = num / (double)m;
    if (r == 0.0) r = 0.0;
    return r;
}

double orc_entropy_approx(int w, int x, int y, int z) {
    int64_t m = (int64_t)y + z;
    if (m == 0) return 0.0;
    int64_t xx = (int64_t)w + x;
    ltable_ensure(m);
    double s = g_lhi[y] + g_lhi[z] + g_lhi[xx] + g_lhi[m - xx];
    s -= g_lhi[w] + g_lhi[y - w] + g_lhi[x] + g_lhi[z - x] + g_lhi[m];
    return s / (double)m;
}

float orc_entropy_ref_f32(int w, int x, int y, int z) {
    float m = (float)(y + z);
    float xx = (float)(w + x);
    float ret = 0.0f;
    if (xx > 0.0f && xx < m)
        ret += (xx / m) * log2f(xx / m) + ((m - xx) / m) * log2f((m - xx) / m);
    if (y > 0 && y > w && w > 0)
        ret -= ((float)w / m) * log2f((float)w / (float)y) +
               ((float)(y - w) / m) * log2f((float)(y - w) / (float)y);
    if (z > 0 && z > x && x > 0)
        ret -= ((float)x / m) * log2f((float)x / (float)z) +
               ((float)(z - x) / m) * log2f((float)(z - x) / (float)z);
    return ret;
}

/* ------------------------------------------------------------------------ */
/* fisher_p.c restated (R's dhyper/dbinom_raw/bd0/stirlerr, log domain).    */
static double ref_stirlerr(int n) {
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

static double ref_bd0(double x, double np) {
    if (isinf(x) || isinf(np) || np == 0.0) return NAN;
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

static double ref_dbinom_raw(int x, int n, double p, double q) {
    double lf, lc;
    if (p == 0) return x == 0 ? 0.0 : -INFINITY;
    if (q == 0) return x == 0 ? 0.0 : -INFINITY;
    if (x == 0) {
        if (n == 0) return 0.0;
        lc = (p < 0.1) ? -ref_bd0((double)n, (double)n * q) - (double)n * p : (double)n * log(q);
        return lc;
    }
    if (x == n) {
        lc = (q < 0.1) ? -ref_bd0((double)n, (double)n * p) - (double)n * q : (double)n * log(p);
        return lc;
    }
    if (x < 0 || x > n) return -INFINITY;
    lc = ref_stirlerr(n) - ref_stirlerr(x) - ref_stirlerr(n - x) -
         ref_bd0((double)x, (double)n * p) - ref_bd0((double)(n - x), (double)n * q);
    lf = log(6.283185307179586476925286766559) + log((double)x) +
         log1p(-(double)x / (double)n);
    lc -= 0.5 * lf;
    return lc;
}

static double ref_dhyper(int x, int r, int b, int n) {
    if (n < x || r < x || n - x > b) return -INFINITY;
    if (n == 0) return x == 0 ? 0.0 : -INFINITY;
    double p = (double)n / (double)(r + b);
    double q = (double)(r + b - n) / (double)(r + b);
    double p1 = ref_dbinom_raw(x, r, p, q);
    double p2 = ref_dbinom_raw(n - x, b, p, q);
    double p3 = ref_dbinom_raw(n, r + b, p, q);
    return p1 + p2 - p3;
}

float orc_fisher_ref(int n11, int n12, int n21, int n22) {
    if (n11 < 0 || n12 < 0 || n21 < 0 || n22 < 0) return NAN;
    int m = n11 + n21;
    int n = n12 + n22;
    int o = n11 + n12;
    int lo = o <= n ? 0 : o - n;
    int hi = o < m ? o : m;
    double dsumd = 0.0, dsumn = 0.0;
    double dthr = ref_dhyper(n11, m, n, o);
    for (int j = lo; j <= hi; ++j) {
        double daux = ref_dhyper(j, m, n, o);
        double daux2 = exp(daux);
        dsumd += daux2;
        if (daux <= dthr) dsumn += daux2;
    }
    return (float)(log10(dsumn) - log10(dsumd));
}

/* ------------------------------------------------------------------------ */
/* Own log2: v = f * 2^k with f in [1/sqrt2, sqrt2), t = (f-1)/(f+1),
 * log2(f) = (2/ln 2) * (t + t^3/3 + t^5/5 + ...), |t| <= 0.1716, 13 terms.   */
static const double LOG2_C[13] = {
    2.8853900817779268,  /* 2/(1*ln2)  */
    0.9617966939259756,  /* 2/(3*ln2)  */
    0.5770780163555853,  /* 2/(5*ln2)  */
    0.4121985831111324,  /* 2/(7*ln2)  */
    0.3205988979753252,  /* 2/(9*ln2)  */
    0.2623081892525388,  /* 2/(11*ln2) */
    0.22195308321368667, /* 2/(13*ln2) */
    0.19235933878519512, /* 2/(15*ln2) */
    0.16972882833987804, /* 2/(17*ln2) */
    0.15186263588304877, /* 2/(19*ln2) */
    0.1373995277037108,  /* 2/(21*ln2) */
    0.12545174268599682, /* 2/(23*ln2) */
    0.11541560327111708, /* 2/(25*ln2) */
};

double orc_log2(double v) {
    int k;
    double f = frexp(v, &k); /* f in [0.5, 1) */
    if (f < 0.7071067811865476) {
        f = f * 2.0;
        k = k - 1;
    }
    double t = (f - 1.0) / (f + 1.0);
    double t2 = t * t;
    double acc = LOG2_C[12];
    for (int i = 11; i >= 0; i--) {
        acc = acc * t2;
        acc = acc + LOG2_C[i];
    }
    double r = t * acc;
    return r + (double)k;
}

/* ------------------------------------------------------------------------ */
/* Fisher ranking objective.
 * Table [[a, b], [c, d]] = [[w, y-w], [x, z-x]] (rows: control, case; columns:
 * f = 1, f = 0).  The p value is invariant under row swap, column swap and
 * transpose, so the table is first mapped to the lexicographically smallest of
 * its 8 images.  With margins r0 = a+b, c0 = a+c, c1 = b+d, cell a ranges over
 * [max(0, r0-c1), min(r0, c0)] with P(j+1)/P(j) = (c0-j)(r0-j) / ((j+1)(c1-r0+j+1)).
 * T(j) = P(j)/P(a) is walked outward from j = a in a (mantissa, exponent)
 * form; p = sum_{T(j) <= 1+1e-7} T(j) / sum_j T(j)  (R's fisher.test rule).   */
typedef struct { double t; int e; } xf; /* value t * 2^e, t in [1, 2) or 0 */

static inline xf xf_norm(double t, int e) {
    xf r;
    if (t == 0.0) {
        r.t = 0.0;
        r.e = 0;
        return r;
    }
    int k;
    double f = frexp(t, &k);
    r.t = f * 2.0;
    r.e = e + k - 1;
    return r;
}

static inline double xf_to_double(xf v) {
    if (v.t == 0.0 || v.e < -1000) return 0.0;
    return ldexp(v.t, v.e);
}

static inline xf xf_add(xf a, xf b) {
    if (a.t == 0.0) return b;
    if (b.t == 0.0) return a;
    if (a.e < b.e) {
        xf t = a;
        a = b;
        b = t;
    }
    int d = b.e - a.e; /* <= 0 */
    double bs = (d < -1000) ? 0.0 : ldexp(b.t, d);
    return xf_norm(a.t + bs, a.e);
}

static inline int xf_le_tol(xf v) { /* v <= 1 + 1e-7 */
    if (v.t == 0.0) return 1;
    if (v.e < 0) return 1;
    if (v.e > 0) return 0;
    return v.t <= 1.0000001;
}

static void canon_table(int64_t *a, int64_t *b, int64_t *c, int64_t *d) {
    int64_t v[8][4] = {
        {*a, *b, *c, *d}, {*b, *a, *d, *c}, {*c, *d, *a, *b}, {*d, *c, *b, *a},
        {*a, *c, *b, *d}, {*c, *a, *d, *b}, {*b, *d, *a, *c}, {*d, *b, *c, *a}};
    int best = 0;
    for (int i = 1; i < 8; i++) {
        for (int k = 0; k < 4; k++) {
            if (v[i][k] < v[best][k]) {
                best = i;
                break;
            }
            if (v[i][k] > v[best][k]) break;
        }
    }
    *a = v[best][0];
    *b = v[best][1];
    *c = v[best][2];
    *d = v[best][3];
}

double orc_fisher_rank(int w, int x, int y, int z) {
    int64_t a = w, b = (int64_t)y - w, c = x, d = (int64_t)z - x;
    if (a + b + c + d == 0) return 0.0;
    canon_table(&a, &b, &c, &d);
    int64_t r0 = a + b, c0 = a + c, c1 = b + d;
    int64_t lo = r0 - c1 > 0 ? r0 - c1 : 0;
    int64_t hi = r0 < c0 ? r0 : c0;
    xf one = {1.0, 0};
    xf s_all = one;
    double s_tail = 1.0;
    int64_t n_in = 1;
    xf cur = one;
    for (int64_t j = a; j < hi; j++) { /* T(j+1) = T(j) * ratio(j) */
        double num = (double)(c0 - j) * (double)(r0 - j);
        double den = (double)(j + 1) * (double)(c1 - r0 + j + 1);
        cur = xf_norm(cur.t * (num / den), cur.e);
        s_all = xf_add(s_all, cur);
        if (xf_le_tol(cur)) {
            s_tail = s_tail + xf_to_double(cur);
            n_in++;
        }
    }
    cur = one;
    for (int64_t j = a; j > lo; j--) { /* T(j-1) = T(j) / ratio(j-1) */
        double num = (double)(c0 - j + 1) * (double)(r0 - j + 1);
        double den = (double)j * (double)(c1 - r0 + j);
        cur = xf_norm(cur.t * (den / num), cur.e);
        s_all = xf_add(s_all, cur);
        if (xf_le_tol(cur)) {
            s_tail = s_tail + xf_to_double(cur);
            n_in++;
        }
    }
    if (n_in == hi - lo + 1) return 0.0;
    double l2 = orc_log2(s_tail) - ((double)s_all.e + orc_log2(s_all.t));
    double r = l2 * 0.3010299956639812;
    if (r == 0.0) r = 0.0;
    return r;
}
