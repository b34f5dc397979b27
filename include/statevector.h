/* Statevectors in host memory (replaces reference src/statevector.h:10-42). */
#pragma once
#include "config.h"

#ifdef __cplusplus
extern "C" {
#endif

struct statevector {
    numeric* data; /* 2^nqubits entries */
    int nqubits;
};

/* reference src/statevector.c:11 */
int allocate_statevector(int nqubits, struct statevector* psi);
/* reference src/statevector.c:32 */
void free_statevector(struct statevector* psi);
/* reference src/statevector.c:49 */
void transpose_statevector(const struct statevector* psi, const int* perm, struct statevector* psi_trans);
/* reference src/statevector.c:79 */
void add_statevectors(const struct statevector* psi1, const struct statevector* psi2, struct statevector* psi_sum);

/* interleaved array: data[idx * nstates + k] (reference src/statevector.h:28-37) */
struct statevector_array {
    numeric* data;
    int nqubits;
    int nstates;
};

/* reference src/statevector.c:95 */
int allocate_statevector_array(int nqubits, int nstates, struct statevector_array* psi_list);
/* reference src/statevector.c:116 */
void free_statevector_array(struct statevector_array* psi_list);

#ifdef __cplusplus
}
#endif
