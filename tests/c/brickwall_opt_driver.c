/*
 * A plain C driver of the reference API, compiled against include/*.h and linked with
 * librqcopt_mg.so only: the call sequence of the reference's brickwall optimisation
 * driver (examples/hubbard1d/hubbard1d_opt.c:160-200): a dense target passed as a
 * linear_func callback (the zgemv ufunc of hubbard1d_opt.c:41-52), tied brickwall gates
 * with per-layer permutations, set_rtr_default_params(16 * nlayers) and the 11-argument
 * optimize_matchgate_brickwall_circuit_hmat; then the HVP optimiser with the same inputs.
 *
 * Usage: brickwall_opt_driver <in.bin> <out.bin>
 *   in.bin : int32 nqubits, nlayers, niter; complex128 U[2^n][2^n] (row-major);
 *            complex128 vlist[nlayers][8]; int32 perms[nlayers][nqubits]
 *   out.bin: float64 f_iter_hmat[niter+1]; complex128 vlist_hmat[nlayers][8];
 *            float64 f_iter_hvp[niter+1];  complex128 vlist_hvp[nlayers][8]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "matchgate.h"
#include "matchgate_target.h"
#include "mg_brickwall_opt.h"
#include "statevector.h"
#include "trust_region.h"

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

static void* read_all(const char* path, size_t* n)
{
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    *n = (size_t)ftell(f);
    fseek(f, 0, SEEK_SET);
    void* buf = malloc(*n);
    if (buf && fread(buf, 1, *n, f) != *n) {
        free(buf);
        buf = NULL;
    }
    fclose(f);
    return buf;
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    size_t n = 0;
    unsigned char* in = read_all(argv[1], &n);
    if (!in || n < 12) return 3;
    int hdr[3];
    memcpy(hdr, in, sizeof(hdr));
    const int nqubits = hdr[0], nlayers = hdr[1], niter = hdr[2];
    const int S = 1 << nqubits;
    const numeric* U = (const numeric*)(in + 12);
    const struct matchgate* vstart = (const struct matchgate*)(in + 12 + (size_t)S * S * sizeof(numeric));
    const int* perm_data = (const int*)((const unsigned char*)vstart + (size_t)nlayers * sizeof(struct matchgate));
    const int** perms = malloc(nlayers * sizeof(int*));
    for (int i = 0; i < nlayers; i++) perms[i] = perm_data + (size_t)i * nqubits;

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

    FILE* out = fopen(argv[2], "wb");
    if (!out) return 4;
    fwrite(f_iter, sizeof(double), niter + 1, out);
    fwrite(vopt, sizeof(struct matchgate), nlayers, out);
    fwrite(f_iter + (niter + 1), sizeof(double), niter + 1, out);
    fwrite(vopt + nlayers, sizeof(struct matchgate), nlayers, out);
    fclose(out);
    printf("f_iter[%d] hmat %.15f hvp %.15f\n", niter, f_iter[niter], f_iter[2 * niter + 1]);
    free(f_iter); free(f_citer); free(vopt); free(perms); free(in);
    return 0;
}
