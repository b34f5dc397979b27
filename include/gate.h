/*
 * Matchgate statevector kernels on host-memory states, executed on the GPU
 * (replaces reference src/gate.h:7-22, src/gate.c:575-1143).
 */
#pragma once
#include "matchgate.h"
#include "statevector.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/gate.c:581 */
void apply_matchgate(const struct matchgate* gate, const int i, const int j, const struct statevector* psi,
                     struct statevector* psi_out);
/* src/gate.c:659 */
void apply_matchgate_backward(const struct matchgate* gate, const int i, const int j, const struct statevector* psi,
                              const struct statevector* dpsi_out, struct statevector* dpsi, struct matchgate* dgate);
/* src/gate.c:779 */
void apply_matchgate_to_array(const struct matchgate* gate, const int i, const int j,
                              const struct statevector_array* psi, struct statevector_array* psi_out);
/* src/gate.c:858 */
void apply_matchgate_backward_array(const struct matchgate* gate, const int i, const int j,
                                    const struct statevector_array* psi, const struct statevector* dpsi_out,
                                    struct matchgate* dgates);
/* src/gate.c:982 */
void apply_matchgate_placeholder(const int i, const int j, const struct statevector* psi,
                                 struct statevector_array* psi_out);
/* src/gate.c:1066 (general two-qubit gate: the 16 states d(G psi)/dG[r][c], psi_out->nstates == 16) */
void apply_gate_placeholder(const int i, const int j, const struct statevector* psi,
                            struct statevector_array* psi_out);

#ifdef __cplusplus
}
#endif
