/*
 * Test and baseline infrastructure: linked with the reference's own kernel sources into
 * oracle/_ref/libmg_ref.so (oracle/Makefile, target `ref`).  The reference defines its
 * target callbacks only inside its example drivers (examples/hubbard1d/hubbard1d_opt.c:30-52),
 * which are not part of that build, so these are written here: a dense-matrix map (the
 * zgemv 'ufunc', without CBLAS) and a matrix-free brickwall map (the 'ufunc_matfree').
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
