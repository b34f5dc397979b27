/* HDF5-free file I/O for the reference drivers' input and output files.
 *
 * The reference drivers (examples/hubbard1d/hubbard1d_opt.c:81-162 reading, :238-310 writing) and
 * its C tests use exactly these HDF5 names: H5Fopen / H5Fcreate / H5Fclose, H5F_ACC_RDONLY,
 * H5F_ACC_TRUNC, H5P_DEFAULT, the types H5T_NATIVE_DOUBLE, H5T_NATIVE_INT, H5T_IEEE_F64LE,
 * H5T_STD_I32LE, the handle types hid_t / herr_t / hsize_t, and the helpers of src/util.h:24-34
 * (declared in util.h).  This header defines that subset over the library's own reader and
 * writer of the HDF5 file format, so the drivers build and run without libhdf5.
 *
 * Supported files: what the reference's files use, i.e. superblock version 0 or 1, groups as
 * symbol tables (B-tree version 1 + local heap), version-1 object headers (with continuation
 * blocks), contiguous or compact datasets, little-endian IEEE floats (4, 8 bytes) and signed or
 * unsigned integers (1, 2, 4, 8 bytes), simple or scalar dataspaces, and attributes on the root
 * group.  Anything else (chunked or filtered data, newer superblocks, shared messages) is
 * reported as an error.  Written files use the same subset and are read by libhdf5 and h5dump.
 * Names are looked up in the root group only (the reference files keep everything there).
 *
 * Do not include <hdf5.h> in the same translation unit, and do not pass these handles to
 * libhdf5: the handles and type ids are this library's own.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

#define H5P_DEFAULT ((hid_t)0)
#define H5F_ACC_RDONLY 0x0000u
#define H5F_ACC_RDWR 0x0001u
#define H5F_ACC_TRUNC 0x0002u

/* type ids: file storage types and the (little-endian, LP64) native memory types */
#define MG_H5T_ID(k) ((hid_t)(0x4d47485400000000LL + (k)))
#define H5T_IEEE_F64LE MG_H5T_ID(1)
#define H5T_IEEE_F32LE MG_H5T_ID(2)
#define H5T_STD_I32LE MG_H5T_ID(3)
#define H5T_STD_I64LE MG_H5T_ID(4)
#define H5T_STD_U8LE MG_H5T_ID(5)
#define H5T_STD_I8LE MG_H5T_ID(6)
#define H5T_STD_I16LE MG_H5T_ID(7)
#define H5T_STD_U32LE MG_H5T_ID(8)
#define H5T_STD_U64LE MG_H5T_ID(9)
#define H5T_NATIVE_DOUBLE H5T_IEEE_F64LE
#define H5T_NATIVE_FLOAT H5T_IEEE_F32LE
#define H5T_NATIVE_INT H5T_STD_I32LE
#define H5T_NATIVE_UINT H5T_STD_U32LE
#define H5T_NATIVE_LONG H5T_STD_I64LE
#define H5T_NATIVE_LLONG H5T_STD_I64LE
#define H5T_NATIVE_HSIZE H5T_STD_U64LE
#define H5T_NATIVE_UCHAR H5T_STD_U8LE
#define H5T_NATIVE_CHAR H5T_STD_I8LE
#define H5T_NATIVE_SHORT H5T_STD_I16LE

/* open an existing file (H5F_ACC_RDONLY); returns a handle > 0, or -1 */
hid_t mg_h5_fopen(const char* name, unsigned flags);
/* create (H5F_ACC_TRUNC: replace) a file; datasets and attributes are written at close */
hid_t mg_h5_fcreate(const char* name, unsigned flags);
/* close a handle; for a created file this writes it; 0 or -1 */
herr_t mg_h5_fclose(hid_t file);

#define H5Fopen(name, flags, fapl) mg_h5_fopen((name), (flags))
#define H5Fcreate(name, flags, fcpl, fapl) mg_h5_fcreate((name), (flags))
#define H5Fclose(file) mg_h5_fclose(file)

#ifdef __cplusplus
}
#endif
