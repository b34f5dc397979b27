/* Vector distances (replaces reference src/util.h:18-22; the HDF5 helpers of
 * src/util.c:56-236 are file I/O of the drivers and are not part of this library). */
#pragma once
#include "config.h"

#ifdef __cplusplus
extern "C" {
#endif

double uniform_distance_real(long n, const double* x, const double* y);   /* src/util.c:8 */
double uniform_distance(long n, const numeric* x, const numeric* y);      /* src/util.c:23 */
double relative_distance(long n, const numeric* x, const numeric* y, double eps); /* src/util.c:38 */

#ifdef __cplusplus
}
#endif
