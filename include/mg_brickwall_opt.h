/* Brickwall optimisation driver (replaces reference src/mg_brickwall_opt.h:8-10,
 * the 11-argument form with f_citer). */
#pragma once
#include "matchgate_brickwall.h"
#include "matchgate_target.h"
#include "trust_region.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference src/mg_brickwall_opt.c:99-122: RTR with the explicit Riemannian Hessian */
void optimize_matchgate_brickwall_circuit_hmat(linear_func ufunc, void* udata, const struct matchgate vlist_start[],
                                               const int nlayers, const int nqubits, const int* perms[],
                                               struct rtr_params* params, const int niter, double* f_iter,
                                               numeric* f_citer, struct matchgate vlist_opt[]);

/* RTR driven by GPU Hessian-vector products (no reference counterpart: the reference
 * declares the HVP optimiser, trust_region.h:84-86, but has no brickwall caller).
 * Same arguments as the _hmat form. */
void optimize_matchgate_brickwall_circuit_hvp(linear_func ufunc, void* udata, const struct matchgate vlist_start[],
                                              const int nlayers, const int nqubits, const int* perms[],
                                              struct rtr_params* params, const int niter, double* f_iter,
                                              numeric* f_citer, struct matchgate vlist_opt[]);

#ifdef __cplusplus
}
#endif
