/* Per-column brickwall sweeps on host-memory states, executed on the GPU
 * (replaces reference src/matchgate_brickwall.h:8-21).  Layer i applies vlist[i]
 * to the pairs (perms[i][2k], perms[i][2k+1]), k = 0 .. nqubits/2 - 1. */
#pragma once
#include "matchgate.h"
#include "quantum_circuit.h"
#include "statevector.h"

#ifdef __cplusplus
extern "C" {
#endif

int apply_matchgate_brickwall_unitary(const struct matchgate vlist[], const int nlayers, const int* perms[],
                                      const struct statevector* psi, struct statevector* psi_out); /* src/matchgate_brickwall.c:14 */
int matchgate_brickwall_unitary_forward(const struct matchgate vlist[], int nlayers, const int* perms[],
                                        const struct statevector* psi, struct quantum_circuit_cache* cache,
                                        struct statevector* psi_out); /* :52 */
int matchgate_brickwall_unitary_backward(const struct matchgate vlist[], int nlayers, const int* perms[],
                                         const struct quantum_circuit_cache* cache, const struct statevector* dpsi_out,
                                         struct statevector* dpsi, struct matchgate dvlist[]); /* :87 */
/* hess: (8 nlayers)^2 complex, row-major, hess[((i*8+x)*nlayers + j)*8 + y] (:135) */
int matchgate_brickwall_unitary_backward_hessian(const struct matchgate vlist[], int nlayers, const int* perms[],
                                                 const struct quantum_circuit_cache* cache,
                                                 const struct statevector* dpsi_out, struct statevector* dpsi,
                                                 struct matchgate dvlist[], numeric* hess); /* :137 */
int matchgate_ti_brickwall_unitary_backward(const struct matchgate vlist[], int nlayers, const int* perms[],
                                            const struct quantum_circuit_cache* cache, const struct statevector* dpsi_out,
                                            struct statevector* dpsi, struct matchgate dvlist[]); /* :266 */
int matchgate_ti_brickwall_unitary_backward_hessian(const struct matchgate vlist[], int nlayers, const int* perms[],
                                                    const struct quantum_circuit_cache* cache,
                                                    const struct statevector* dpsi_out, struct statevector* dpsi,
                                                    struct matchgate dvlist[], numeric* hess); /* :308 */

#ifdef __cplusplus
}
#endif
