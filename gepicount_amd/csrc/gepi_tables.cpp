/* gepi_tables.cpp -- host construction of the score tables (binary128 math). */
#include <quadmath.h>

#include "gepi_internal.h"

void gepi_build_tables(int32_t n, double *lhi, double *llo, double *lfact) {
    if (n <= 0) return;
    __float128 acc = 0;
    lhi[0] = llo[0] = lfact[0] = 0.0;
    for (int32_t t = 1; t < n; t++) {
        __float128 q = (__float128)t * log2q((__float128)t);
        double h = (double)q;
        lhi[t] = h;
        llo[t] = (double)(q - (__float128)h);
        acc += logq((__float128)t);
        lfact[t] = (double)acc;
    }
}
