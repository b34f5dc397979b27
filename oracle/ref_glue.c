/*
 * Test and baseline infrastructure: linked with the reference's own kernel sources into
 * oracle/_ref/libmg_ref.so (oracle/Makefile, target `ref`).  The reference defines its
 * target callbacks only inside its example drivers (examples/hubbard1d/hubbard1d_opt.c:30-52),
 * which are not part of that build, so these are written here: a dense-matrix map (the
 * zgemv 'ufunc', without CBLAS) and a matrix-free brickwall map (the 'ufunc_matfree'); and a
 * timing harness that runs the reference's per-column target work over a column sample (the
 * CPU baseline of the benchmark).
 */
#include "matchgate_brickwall.h"
#include "matchgate_target.h"
#include "statevector.h"

/* udata of ref_dense_ufunc: row-major n x n matrix */
struct ref_dense_map {
    const numeric* U;
    long n;
};

/* psi_out = U psi (examples/hubbard1d/hubbard1d_opt.c:41-52) */
int ref_dense_ufunc(const struct statevector* psi, void* udata, struct statevector* psi_out)
{
    const struct ref_dense_map* m = udata;
    for (long a = 0; a < m->n; a++) {
        const numeric* row = m->U + a * m->n;
        numeric s = 0;
        for (long b = 0; b < m->n; b++) s += row[b] * psi->data[b];
        psi_out->data[a] = s;
    }
    return 0;
}

/* udata of ref_brickwall_ufunc (the struct u_splitting of examples/hubbard1d/hubbard1d_opt.c:23-28) */
struct ref_brickwall_map {
    const struct matchgate* ulist;
    int ulayers;
    const int** uperms;
};

/* psi_out = U psi for a brickwall U (examples/hubbard1d/hubbard1d_opt.c:30-39) */
int ref_brickwall_ufunc(const struct statevector* psi, void* udata, struct statevector* psi_out)
{
    const struct ref_brickwall_map* m = udata;
    return apply_matchgate_brickwall_unitary(m->ulist, m->ulayers, m->uperms, psi, psi_out);
}

#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* The reference's per-column work of the target loop (src/matchgate_target.c:746-864), timed
 * over the basis columns [0, ncols) with nthreads OpenMP threads, each accumulating its own
 * partial sums as the reference does: seed -conj(U e_b) from the map, brickwall forward sweep
 * (with the state cache when hessian != 0), the trace term, and with hessian != 0 the backward
 * sweep with the placeholder Hessian and the gradient and Hessian accumulation.  Returns the
 * wall time in seconds; *f_partial receives the trace over those columns.  Used to time a
 * bounded sample of a full evaluation. */
double ref_time_target_columns(linear_func ufunc, void* udata, const struct matchgate* vlist, int nlayers, int nqubits,
                               const int** perms, long ncols, int nthreads, int hessian, numeric* f_partial)
{
    const long n = 1L << nqubits;
    const int m = 8 * nlayers;
    numeric f_total = 0;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
#pragma omp parallel num_threads(nthreads) reduction(+ : f_total)
    {
        struct statevector psi = {0}, Upsi = {0}, Wpsi = {0};
        allocate_statevector(nqubits, &psi);
        allocate_statevector(nqubits, &Upsi);
        allocate_statevector(nqubits, &Wpsi);
        struct quantum_circuit_cache cache = {0};
        if (hessian) allocate_quantum_circuit_cache(nqubits, nlayers * (nqubits / 2), &cache);
        struct matchgate* dv_unit = calloc(nlayers, sizeof(struct matchgate));
        struct matchgate* dv_part = calloc(nlayers, sizeof(struct matchgate));
        numeric* h_unit = hessian ? calloc((size_t)m * m, sizeof(numeric)) : NULL;
        numeric* h_part = hessian ? calloc((size_t)m * m, sizeof(numeric)) : NULL;
        numeric fp = 0;
#pragma omp for schedule(static)
        for (long b = 0; b < ncols; b++) {
            memset(psi.data, 0, n * sizeof(numeric));
            psi.data[b] = 1;
            ufunc(&psi, udata, &Upsi);
            for (long a = 0; a < n; a++) Upsi.data[a] = -conj(Upsi.data[a]);
            if (hessian)
                matchgate_brickwall_unitary_forward(vlist, nlayers, perms, &psi, &cache, &Wpsi);
            else
                apply_matchgate_brickwall_unitary(vlist, nlayers, perms, &psi, &Wpsi);
            for (long a = 0; a < n; a++) fp += Upsi.data[a] * Wpsi.data[a];
            if (hessian) {
                matchgate_brickwall_unitary_backward_hessian(vlist, nlayers, perms, &cache, &Upsi, &psi, dv_unit, h_unit);
                for (int i = 0; i < nlayers; i++) add_matchgate(&dv_part[i], &dv_unit[i]);
                for (long i = 0; i < (long)m * m; i++) h_part[i] += h_unit[i];
            }
        }
        f_total += fp;
        free(h_part);
        free(h_unit);
        free(dv_part);
        free(dv_unit);
        if (hessian) free_quantum_circuit_cache(&cache);
        free_statevector(&Wpsi);
        free_statevector(&Upsi);
        free_statevector(&psi);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (f_partial) *f_partial = f_total;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
