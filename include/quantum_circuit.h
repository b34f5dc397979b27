/* Sequential matchgate circuits on host-memory states, executed on the GPU
 * (replaces reference src/quantum_circuit.h:11-30). */
#pragma once
#include "matchgate.h"
#include "statevector.h"

#ifdef __cplusplus
extern "C" {
#endif

/* sequence of intermediate statevectors kept for the backward pass */
struct quantum_circuit_cache {
    struct statevector* psi_list;
    int nqubits;
    int ngates;
};

int allocate_quantum_circuit_cache(const int nqubits, const int ngates, struct quantum_circuit_cache* cache); /* :42 */
void free_quantum_circuit_cache(struct quantum_circuit_cache* cache);                                         /* :71 */
int apply_quantum_matchgate_circuit(const struct matchgate gates[], const int ngates, const int wires[],
                                    const struct statevector* psi, struct statevector* psi_out);               /* :13 */
int quantum_matchgate_circuit_forward(const struct matchgate gates[], const int ngates, const int wires[],
                                      const struct statevector* psi, struct quantum_circuit_cache* cache,
                                      struct statevector* psi_out);                                           /* :90 */
int quantum_matchgate_circuit_backward(const struct matchgate gates[], const int ngates, const int wires[],
                                       const struct quantum_circuit_cache* cache, const struct statevector* dpsi_out,
                                       struct statevector* dpsi, struct matchgate dgates[]);                  /* :115 */

#ifdef __cplusplus
}
#endif
