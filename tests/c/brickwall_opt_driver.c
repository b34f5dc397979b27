/*
 * A plain C driver of the reference API, compiled against the headers in include/ and linked with
 * librqcopt_mg.so only: the call sequence and file I/O of the reference's brickwall
 * optimisation driver (examples/hubbard1d/hubbard1d_opt.c): the target unitary `expiH` and the
 * initial gates `vlist` / permutations `perm<i>` read from the driver's HDF5 input files
 * (hubbard1d_opt.c:81-151), a dense target passed as a linear_func callback (the zgemv ufunc of
 * hubbard1d_opt.c:41-52), set_rtr_default_params(16 * nlayers) and the 11-argument
 * optimize_matchgate_brickwall_circuit_hmat, then the HVP optimiser with the same inputs; each
 * result written in the driver's output layout (vlist_opt, f_iter and attributes,
 * hubbard1d_opt.c:238-310).  No libhdf5: the HDF5 calls are the library's (mg_hdf5.h).
 *
 * Usage: brickwall_opt_driver <unitary.hdf5> <init.hdf5> <niter> <out_hmat.hdf5> <out_hvp.hdf5>
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "matchgate.h"
#include "matchgate_target.h"
#include "mg_brickwall_opt.h"
#include "statevector.h"
#include "trust_region.h"
#include "util.h"

struct dense_target {
    const numeric* U;
    int n;
};

/* psi_out = U psi */
static int ufunc(const struct statevector* psi, void* udata, struct statevector* psi_out)
{
    const struct dense_target* t = (const struct dense_target*)udata;
    for (int a = 0; a < t->n; a++) {
        numeric s = 0;
        for (int b = 0; b < t->n; b++) s += t->U[(size_t)a * t->n + b] * psi->data[b];
        psi_out->data[a] = s;
    }
    return 0;
}

static int write_result(const char* path, const struct matchgate* vlist_opt, const double* f_iter, int nlayers,
                        int nqubits, int niter)
{
    hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return -1;
    hsize_t vdims[3] = {nlayers, 8, 2};
    hsize_t fdims[1] = {niter + 1};
    const double wtime = 0.0;
    const int full_target = 1;
    if (write_hdf5_dataset(file, "vlist_opt", 3, vdims, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, (const numeric*)vlist_opt) < 0 ||
        write_hdf5_dataset(file, "f_iter", 1, fdims, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, f_iter) < 0 ||
        write_hdf5_scalar_attribute(file, "nqubits", H5T_STD_I32LE, H5T_NATIVE_INT, &nqubits) ||
        write_hdf5_scalar_attribute(file, "nlayers", H5T_STD_I32LE, H5T_NATIVE_INT, &nlayers) ||
        write_hdf5_scalar_attribute(file, "Walltime", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &wtime) ||
        write_hdf5_scalar_attribute(file, "FULL_TARGET", H5T_STD_I32LE, H5T_NATIVE_INT, &full_target)) {
        H5Fclose(file);
        return -1;
    }
    return H5Fclose(file);
}

int main(int argc, char** argv)
{
    if (argc != 6) {
        fprintf(stderr, "usage: %s unitary.hdf5 init.hdf5 niter out_hmat.hdf5 out_hvp.hdf5\n", argv[0]);
        return 2;
    }
    const int niter = atoi(argv[3]);

    hid_t file = H5Fopen(argv[2], H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return 3;
    int nqubits = 0, nlayers = 0;
    if (read_hdf5_attribute(file, "nqubits", H5T_NATIVE_INT, &nqubits) < 0 ||
        read_hdf5_attribute(file, "nlayers", H5T_NATIVE_INT, &nlayers) < 0 || nqubits < 2 || nqubits > 12 ||
        nlayers < 1)
        return 3;
    struct matchgate* vstart = malloc(nlayers * sizeof(struct matchgate));
    if (read_hdf5_dataset(file, "vlist", H5T_NATIVE_DOUBLE, (numeric*)vstart) < 0) return 3;
    int* perm_data = malloc((size_t)nlayers * nqubits * sizeof(int));
    const int** perms = malloc(nlayers * sizeof(int*));
    for (int i = 0; i < nlayers; i++) {
        char varname[32];
        sprintf(varname, "perm%i", i);
        if (read_hdf5_dataset(file, varname, H5T_NATIVE_INT, perm_data + (size_t)i * nqubits) < 0) return 3;
        perms[i] = perm_data + (size_t)i * nqubits;
    }
    H5Fclose(file);

    const int S = 1 << nqubits;
    file = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return 3;
    int nd = 0;
    hsize_t dims[3] = {0, 0, 0};
    if (get_hdf5_dataset_ndims(file, "expiH", &nd) < 0 || nd != 3 || get_hdf5_dataset_dims(file, "expiH", dims) < 0 ||
        dims[0] != (hsize_t)S || dims[1] != (hsize_t)S || dims[2] != 2)
        return 3;
    numeric* U = malloc((size_t)S * S * sizeof(numeric));
    if (read_hdf5_dataset(file, "expiH", H5T_NATIVE_DOUBLE, U) < 0) return 3;
    H5Fclose(file);

    struct dense_target t = {U, S};
    double* f_iter = calloc(2 * (niter + 1), sizeof(double));
    numeric* f_citer = calloc(niter + 1, sizeof(numeric));
    struct matchgate* vopt = calloc(2 * nlayers, sizeof(struct matchgate));
    struct rtr_params params;
    set_rtr_default_params(16 * nlayers, &params);
    optimize_matchgate_brickwall_circuit_hmat(ufunc, &t, vstart, nlayers, nqubits, perms, &params, niter, f_iter,
                                              f_citer, vopt);
    set_rtr_default_params(16 * nlayers, &params);
    optimize_matchgate_brickwall_circuit_hvp(ufunc, &t, vstart, nlayers, nqubits, perms, &params, niter,
                                             f_iter + (niter + 1), f_citer, vopt + nlayers);

    if (write_result(argv[4], vopt, f_iter, nlayers, nqubits, niter) < 0 ||
        write_result(argv[5], vopt + nlayers, f_iter + (niter + 1), nlayers, nqubits, niter) < 0)
        return 4;
    printf("f_iter[%d] hmat %.15f hvp %.15f\n", niter, f_iter[niter], f_iter[2 * niter + 1]);
    free(f_iter); free(f_citer); free(vopt); free(perms); free(perm_data); free(vstart); free(U);
    return 0;
}
