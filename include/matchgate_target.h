/*
 * Trace over basis columns: f = -Tr(U^dagger W), gradient and Hessian, evaluated
 * on the GPU (replaces reference src/matchgate_target.h:10-51; the MPI block
 * :54-75 is dropped: one GPU, no rank split).
 *
 * U is given, as in the reference, by a linear_func callback.  Because the GPU
 * needs U on the device, every call materialises the callback as dense rows on the
 * device (2^n host calls, from several host threads, so the callback must be
 * reentrant exactly as under the reference's OpenMP loop): like the reference, which
 * calls ufunc for every column on every evaluation (src/matchgate_target.c:795), a
 * changed udata takes effect at the next call.  A caller whose map does not change
 * may call mg_target_cache_reuse(1): materialised targets are then kept per
 * (ufunc, udata, nqubits) and reused until mg_target_cache_clear() (at most 8
 * targets and 64 GiB of device rows; the least recently used one is dropped first).
 * Callers that know their target should use the descriptor API of mg_gpu.h instead
 * (dense or brickwall targets on the device).
 *
 * Return codes as in the reference: 0 ok, -1 allocation/device error, -2 ufunc
 * failed (src/matchgate_target.c:400-487).
 */
#pragma once
#include "matchgate.h"
#include "quantum_circuit.h"
#include "statevector.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef int (*linear_func)(const struct statevector* psi, void* fdata, struct statevector* psi_out); /* :10 */

int matchgate_circuit_unitary_target(linear_func ufunc, void* udata, const struct matchgate gates[], const int ngates,
                                     const int wires[], const int nqubits, numeric* fval); /* :13 */
int parallel_matchgate_circuit_unitary_target(linear_func ufunc, void* udata, const struct matchgate gates[],
                                              const int ngates, const int wires[], const int nqubits, numeric* fval); /* :16 */
int matchgate_circuit_unitary_target_and_gradient(linear_func ufunc, void* udata, const struct matchgate gates[],
                                                  const int ngates, const int wires[], const int nqubits,
                                                  numeric* fval, struct matchgate dgates[]); /* :19 */
int matchgate_brickwall_unitary_target(linear_func ufunc, void* udata, const struct matchgate vlist[],
                                       const int nlayers, const int nqubits, const int* perms[], numeric* fval); /* :24 */
int matchgate_brickwall_unitary_target_and_gradient(linear_func ufunc, void* udata, const struct matchgate vlist[],
                                                    const int nlayers, const int nqubits, const int* perms[],
                                                    numeric* fval, struct matchgate dvlist[]); /* :29 */
/* hess: (8 nlayers)^2 complex, hess[((i*8+x)*nlayers + j)*8 + y] = d^2 f / dV_i[x] dV_j[y] */
int matchgate_brickwall_unitary_target_gradient_hessian(linear_func ufunc, void* udata, const struct matchgate vlist[],
                                                        const int nlayers, const int nqubits, const int* perms[],
                                                        numeric* fval, struct matchgate dvlist[], numeric* hess); /* :34 */
int matchgate_ti_brickwall_unitary_target_gradient_hessian(linear_func ufunc, void* udata,
                                                           const struct matchgate vlist[], const int nlayers,
                                                           const int nqubits, const int* perms[], numeric* fval,
                                                           struct matchgate dvlist[], numeric* hess); /* :39 */
/* grad_vec: 8 nlayers reals; H: (8 nlayers)^2 real Riemannian Hessian */
int matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix(linear_func ufunc, void* udata,
                                                                      const struct matchgate vlist[], const int nlayers,
                                                                      const int nqubits, const int* perms[],
                                                                      numeric* fval, double* grad_vec, double* H); /* :44 */
int parallel_matchgate_brickwall_unitary_target_gradient_hessian(linear_func ufunc, void* udata,
                                                                 const struct matchgate vlist[], const int nlayers,
                                                                 const int nqubits, const int* perms[], numeric* fval,
                                                                 struct matchgate dvlist[], numeric* hess); /* :49 */

/* drop every cached materialised linear_func target (no reference counterpart) */
void mg_target_cache_clear(void);
/* on != 0: keep materialised linear_func targets per (ufunc, udata, nqubits) and reuse them
 * until mg_target_cache_clear(); the caller promises that the map does not change meanwhile.
 * 0 (default): materialise on every call, as the reference evaluates ufunc on every call
 * (no reference counterpart) */
void mg_target_cache_reuse(int on);

#ifdef __cplusplus
}
#endif
