/* gepi_internal.h -- internal helpers shared by the library's translation units. */
#ifndef GEPI_INTERNAL_H
#define GEPI_INTERNAL_H

#include <stdint.h>

/* Records a printf-style message as the calling thread's last error and
 * returns `code` (a negative gepi_status).                                   */
int gepi_fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

/* Host tables for t in [0, n): t*log2(t) as double-double (hi, lo), exact to
 * binary128 before rounding, and ln(t!) rounded to double.                    */
void gepi_build_tables(int32_t n, double *lhi, double *llo, double *lfact);

#endif
