/*
 * gepi_score.h -- exact objective functions, compiled for host and gfx950.
 *
 * Both objectives use only IEEE-754 double +, -, *, / (the build passes
 * -ffp-contract=off), table lookups, frexp/ldexp and integer work, so the host
 * and the GPU produce identical bits for identical counts.  That makes the
 * top-k order reproducible across devices and scheduling orders.
 *
 * Entropy (src/kernels/entropy_inline.h:30-80).  The reference evaluates
 *   H(F|C) - H(F) = -I(F;C) <= 0
 * in float32 with guarded log terms.  Expanding every log of a ratio gives
 *   -I*m = L(y)+L(z) + L(xx)+L(m-xx) - L(w)-L(y-w)-L(x)-L(z-x) - L(m),
 * L(t) = t*log2(t), L(0) = 0, xx = w+x, m = y+z; each reference guard drops
 * exactly the terms that vanish under L(0) = 0, and m = 0 returns 0 as the
 * reference does.  Both argument lists are sorted, common arguments cancel,
 * and the rest is summed in double-double from a (hi, lo) table, giving the
 * exactly rounded value; symmetric tables (class swap, outcome swap,
 * transpose) give identical bits.
 *
 * Fisher (src/kernels/fisher_d.cu:249-285, src/misc/fisher_p.c:52-76).  The
 * reference sums exp(dhyper) over the support and keeps terms whose log
 * density is <= the observed one, in log space with libm/libdevice.  Here the
 * table [[w, y-w], [x, z-x]] is mapped to its canonical image under row swap,
 * column swap and transpose (the p value is invariant), and
 * T(j) = P(j)/P(obs) is walked outward from the observed cell with the exact
 * ratio P(j+1)/P(j) = (c0-j)(r0-j) / ((j+1)(c1-r0+j+1)), kept as (mantissa,
 * exponent) so it never overflows.  Terms with T(j) <= 1 + 1e-7 form the tail
 * (R's fisher.test tolerance), and log10 p = log2(tail/total) * log10(2) with
 * an own log2 (atanh series on the mantissa).
 */
#ifndef GEPI_SCORE_H
#define GEPI_SCORE_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GEPI_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define GEPI_HD static inline
#endif

namespace gepi {

GEPI_HD void two_sum(double a, double b, double &s, double &e) {
    double x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    s = x;
    e = ar + br;
}

GEPI_HD void cswap(int32_t &a, int32_t &b) {
    int32_t lo = a < b ? a : b;
    int32_t hi = a < b ? b : a;
    a = lo;
    b = hi;
}

/* Exactly rounded -I(F;C) in bits.  lhi/llo: t*log2(t) as double-double for
 * t in [0, y+z].  All small arrays are indexed with constants only (after
 * unrolling) so that they stay in registers on the GPU.                     */
GEPI_HD double entropy_exact(int32_t w, int32_t x, int32_t y, int32_t z, const double *lhi,
                             const double *llo) {
    int32_t m = y + z;
    if (m == 0) return 0.0;
    int32_t xx = w + x;
    int32_t P[4] = {y, z, xx, m - xx};
    int32_t N[5] = {w, y - w, x, z - x, m};
    /* sorting networks (4 and 5 inputs) */
    cswap(P[0], P[1]); cswap(P[2], P[3]); cswap(P[0], P[2]); cswap(P[1], P[3]); cswap(P[1], P[2]);
    cswap(N[0], N[1]); cswap(N[3], N[4]); cswap(N[2], N[4]); cswap(N[2], N[3]); cswap(N[1], N[4]);
    cswap(N[0], N[3]); cswap(N[0], N[2]); cswap(N[1], N[3]); cswap(N[1], N[2]);
    /* multiset difference: cancel equal arguments pairwise */
    bool kp[4], kn[5];
#pragma unroll
    for (int b = 0; b < 5; b++) kn[b] = true;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        kp[a] = true;
#pragma unroll
        for (int b = 0; b < 5; b++) {
            bool hit = kp[a] && kn[b] && P[a] == N[b];
            kp[a] = kp[a] && !hit;
            kn[b] = kn[b] && !hit;
        }
    }
    double s = 0.0, e = 0.0, err;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        if (!kp[a]) continue;
        two_sum(s, lhi[P[a]], s, err);
        double t = err + llo[P[a]];
        e = e + t;
    }
#pragma unroll
    for (int b = 0; b < 5; b++) {
        if (!kn[b]) continue;
        two_sum(s, -lhi[N[b]], s, err);
        double t = err - llo[N[b]];
        e = e + t;
    }
    double num = s + e;
    double r = num / (double)m;
    if (r == 0.0) r = 0.0;
    return r;
}

/* log2 from IEEE operations only: v = f*2^k, f in [1/sqrt2, sqrt2),
 * t = (f-1)/(f+1), log2 f = sum_i c_i t^(2i+1), c_i = 2/((2i+1) ln 2).       */
GEPI_HD double log2_ieee(double v) {
    const double C[13] = {2.8853900817779268,  0.9617966939259756,  0.5770780163555853,
                          0.4121985831111324,  0.3205988979753252,  0.2623081892525388,
                          0.22195308321368667, 0.19235933878519512, 0.16972882833987804,
                          0.15186263588304877, 0.1373995277037108,  0.12545174268599682,
                          0.11541560327111708};
    int k;
    double f = frexp(v, &k);
    if (f < 0.7071067811865476) {
        f = f * 2.0;
        k = k - 1;
    }
    double t = (f - 1.0) / (f + 1.0);
    double t2 = t * t;
    double acc = C[12];
#pragma unroll
    for (int i = 11; i >= 0; i--) {
        acc = acc * t2;
        acc = acc + C[i];
    }
    double r = t * acc;
    return r + (double)k;
}

struct XF {
    double t;
    int e;
}; /* t * 2^e, t in [1, 2) or t == 0 */

GEPI_HD XF xf_norm(double t, int e) {
    XF r;
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

GEPI_HD double xf_to_double(XF v) {
    if (v.t == 0.0 || v.e < -1000) return 0.0;
    return ldexp(v.t, v.e);
}

GEPI_HD XF xf_add(XF a, XF b) {
    if (a.t == 0.0) return b;
    if (b.t == 0.0) return a;
    if (a.e < b.e) {
        XF t = a;
        a = b;
        b = t;
    }
    int d = b.e - a.e;
    double bs = (d < -1000) ? 0.0 : ldexp(b.t, d);
    return xf_norm(a.t + bs, a.e);
}

GEPI_HD bool xf_le_tol(XF v) {
    if (v.t == 0.0) return true;
    if (v.e < 0) return true;
    if (v.e > 0) return false;
    return v.t <= 1.0000001;
}

GEPI_HD bool lex_less(int32_t a, int32_t b, int32_t c, int32_t d, int32_t A, int32_t B, int32_t C,
                      int32_t D) {
    if (a != A) return a < A;
    if (b != B) return b < B;
    if (c != C) return c < C;
    return d < D;
}

/* Lexicographically smallest of the 8 images of [[a,b],[c,d]] under row swap,
 * column swap and transpose.                                                 */
GEPI_HD void canon_table(int32_t &a, int32_t &b, int32_t &c, int32_t &d) {
    const int32_t a0 = a, b0 = b, c0 = c, d0 = d;
    int32_t ba = a0, bb = b0, bc = c0, bd = d0;
#define GEPI_CANON_TRY(p, q, r, s)                    \
    if (lex_less(p, q, r, s, ba, bb, bc, bd)) {       \
        ba = p;                                       \
        bb = q;                                       \
        bc = r;                                       \
        bd = s;                                       \
    }
    GEPI_CANON_TRY(b0, a0, d0, c0)
    GEPI_CANON_TRY(c0, d0, a0, b0)
    GEPI_CANON_TRY(d0, c0, b0, a0)
    GEPI_CANON_TRY(a0, c0, b0, d0)
    GEPI_CANON_TRY(c0, a0, d0, b0)
    GEPI_CANON_TRY(b0, d0, a0, c0)
    GEPI_CANON_TRY(d0, b0, c0, a0)
#undef GEPI_CANON_TRY
    a = ba;
    b = bb;
    c = bc;
    d = bd;
}

/* log10 of the two-sided Fisher p of [[w, y-w], [x, z-x]]. */
GEPI_HD double fisher_exact(int32_t w, int32_t x, int32_t y, int32_t z) {
    int32_t a = w, b = y - w, c = x, d = z - x;
    if (a + b + c + d == 0) return 0.0;
    canon_table(a, b, c, d);
    int32_t r0 = a + b, c0 = a + c, c1 = b + d;
    int32_t lo = r0 - c1 > 0 ? r0 - c1 : 0;
    int32_t hi = r0 < c0 ? r0 : c0;
    XF one;
    one.t = 1.0;
    one.e = 0;
    XF s_all = one;
    double s_tail = 1.0;
    int32_t n_in = 1;
    XF cur = one;
    /* Each walk may stop early without changing a bit of the result: past
     * the mode (step ratio <= 1, and the ratio only falls further along a
     * walk, so the computed terms never grow again) a term below 2^-55 is
     * under half an ulp of s_all (>= 1) and of s_tail (>= 1), so it and every
     * term after it leave both sums as they are, and each of them is in the
     * tail (T <= 1 + 1e-7): only their count is needed.                    */
    for (int32_t j = a; j < hi; j++) {
        double num = (double)(c0 - j) * (double)(r0 - j);
        double den = (double)(j + 1) * (double)(c1 - r0 + j + 1);
        cur = xf_norm(cur.t * (num / den), cur.e);
        if (cur.e <= -56 && num <= den) {
            n_in += hi - j; /* terms j+1 .. hi */
            break;
        }
        s_all = xf_add(s_all, cur);
        if (xf_le_tol(cur)) {
            s_tail = s_tail + xf_to_double(cur);
            n_in++;
        }
    }
    cur = one;
    for (int32_t j = a; j > lo; j--) {
        double num = (double)(c0 - j + 1) * (double)(r0 - j + 1);
        double den = (double)j * (double)(c1 - r0 + j);
        cur = xf_norm(cur.t * (den / num), cur.e);
        if (cur.e <= -56 && den <= num) {
            n_in += j - lo; /* terms j-1 .. lo */
            break;
        }
        s_all = xf_add(s_all, cur);
        if (xf_le_tol(cur)) {
            s_tail = s_tail + xf_to_double(cur);
            n_in++;
        }
    }
    if (n_in == hi - lo + 1) return 0.0;
    double l2 = log2_ieee(s_tail) - ((double)s_all.e + log2_ieee(s_all.t));
    double r = l2 * 0.3010299956639812;
    if (r == 0.0) r = 0.0;
    return r;
}

/* Orderable unsigned image of a double (no NaN): ascending order preserved. */
GEPI_HD uint64_t order_bits(double v) {
    union {
        double d;
        uint64_t u;
    } c;
    c.d = v;
    return (c.u & 0x8000000000000000ull) ? ~c.u : (c.u | 0x8000000000000000ull);
}

GEPI_HD double from_order_bits(uint64_t u) {
    union {
        double d;
        uint64_t u;
    } c;
    c.u = (u & 0x8000000000000000ull) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u;
    return c.d;
}

/* Packed tie-break index: orders like (sA, sB, fun). */
GEPI_HD uint64_t pack_idx(int64_t sA, int64_t sB, int32_t fun) {
    return ((uint64_t)sA << 32) | ((uint64_t)sB << 3) | (uint64_t)fun;
}

} // namespace gepi

#endif
