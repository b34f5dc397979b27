/*
 * Scalar and index types of the C-ABI (replaces reference src/config.h:6-55).
 * numeric is double _Complex for C callers (the reference always builds with
 * COMPLEX_CIRCUIT, CMakeLists.txt:16) and the layout-identical
 * std::complex<double> for C++ callers.
 */
#pragma once

#include <stdlib.h>

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> numeric;
#else
#include <complex.h>
typedef double complex numeric;
#endif

#define MEM_DATA_ALIGN 64

/* integer type for addressing entries of a statevector (reference src/config.h:55) */
typedef long int intqs;

static inline void aligned_free(void* memblock) { free(memblock); }
