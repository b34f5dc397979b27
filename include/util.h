/* Vector distances and the drivers' HDF5 helpers (replaces reference src/util.h:18-34).  The
 * HDF5 helpers run on the library's own file-format reader and writer (mg_hdf5.h), with the
 * reference's names, arguments and return values (0 on success, negative on failure, a message
 * on stderr). */
#pragma once
#include <math.h>
#include "config.h"
#include "mg_hdf5.h"

/* relative distance |x - y| / max(|x| + |y|, eps) of two numbers (src/util.h:13-16) */
#ifdef __cplusplus
static inline double reldist(numeric x, numeric y, double eps)
{
    return std::abs(x - y) / fmax(std::abs(x) + std::abs(y), eps);
}
#else
static inline double reldist(numeric x, numeric y, double eps)
{
    return cabs(x - y) / fmax(cabs(x) + cabs(y), eps);
}
#endif

#ifdef __cplusplus
extern "C" {
#endif

double uniform_distance_real(long n, const double* x, const double* y);   /* src/util.c:8 */
double uniform_distance(long n, const numeric* x, const numeric* y);      /* src/util.c:24 */
double relative_distance(long n, const numeric* x, const numeric* y, double eps); /* src/util.c:40 */

/* number of dimensions of a dataset (src/util.c:56-79) */
herr_t get_hdf5_dataset_ndims(hid_t file, const char* name, int* ndims);
/* dimensions of a dataset (src/util.c:85-109) */
herr_t get_hdf5_dataset_dims(hid_t file, const char* name, hsize_t* dims);
/* whole dataset, converted to mem_type (src/util.c:115-134) */
herr_t read_hdf5_dataset(hid_t file, const char* name, hid_t mem_type, void* data);
/* attribute of the root group, converted to mem_type (src/util.c:142-161) */
herr_t read_hdf5_attribute(hid_t file, const char* name, hid_t mem_type, void* data);
/* new dataset stored as mem_type_store from data of mem_type_input (src/util.c:168-203) */
herr_t write_hdf5_dataset(hid_t file, const char* name, int degree, const hsize_t dims[], hid_t mem_type_store,
                          hid_t mem_type_input, const void* data);
/* new scalar attribute of the root group (src/util.c:210-236) */
herr_t write_hdf5_scalar_attribute(hid_t file, const char* name, hid_t mem_type_store, hid_t mem_type_input,
                                   const void* data);

#ifdef __cplusplus
}
#endif
