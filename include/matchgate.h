/*
 * Matchgates and their 2x2-block algebra (replaces reference src/matchgate.h:5-73,
 * src/matchgate.c).  A matchgate is 128 bytes: center block (odd parity {01,10})
 * then corner block (even parity {00,11}), each 2x2 row-major.
 */
#pragma once
#include "config.h"

#ifdef __cplusplus
extern "C" {
#endif

struct mat2x2 {
    numeric data[4];
};

struct matchgate {
    struct mat2x2 center_block;
    struct mat2x2 corner_block;
};

/* tangent parameters per gate (complex build), reference src/matchgate.h:44-48 */
static const int num_tangent_params = 8;

void compare_matchgate_arrays(const struct matchgate* m, const struct matchgate* n, int d); /* src/matchgate.c:7 */
void identity_mat2x2(struct mat2x2* a);                                                      /* :40 */
void identity_matchgate(struct matchgate* a);                                                /* :65 */
void scale_matchgate(struct matchgate* a, const double x);                                   /* :80 */
void conjugate_matchgate(struct matchgate* a);                                               /* :93 */
void add_matchgate(struct matchgate* a, const struct matchgate* b);                          /* :109 */
void sub_matchgate(struct matchgate* a, const struct matchgate* b);                          /* :122 */
void add_matchgates(const struct matchgate* a, const struct matchgate* b, struct matchgate* c); /* :135 */
void sub_matchgates(const struct matchgate* a, const struct matchgate* b, struct matchgate* c); /* :148 */
void transpose_matchgate(const struct matchgate* a, struct matchgate* at);                   /* :165 */
void adjoint_matchgate(const struct matchgate* a, struct matchgate* ah);                     /* :182 */
void symm_matchgate(const struct matchgate* w, struct matchgate* z);                         /* :199 */
void antisymm_matchgate(const struct matchgate* w, struct matchgate* z);                     /* :222 */
void real_to_antisymm_matchgate(const double* r, struct matchgate* w);                       /* :251 */
void antisymm_matchgate_to_real(const struct matchgate* w, double* r);                       /* :277 */
void real_to_tangent_matchgate(const double* r, const struct matchgate* v, struct matchgate* z); /* :296 */
void tangent_matchgate_to_real(const struct matchgate* v, const struct matchgate* z, double* r); /* :309 */
void project_tangent_matchgate(const struct matchgate* u, const struct matchgate* z, struct matchgate* p); /* :323 */
void multiply_mat2x2(const struct mat2x2* a, const struct mat2x2* b, struct mat2x2* c);      /* :340 */
void multiply_matchgates(const struct matchgate* a, const struct matchgate* b, struct matchgate* c); /* :356 */
void symmetric_triple_matchgate_product(const struct matchgate* a, const struct matchgate* b,
                                        const struct matchgate* c, struct matchgate* ret);   /* :381 */
int inverse_mat2x2(const struct mat2x2* a, struct mat2x2* ainv);                             /* :414 */
int inverse_matchgate(const struct matchgate* a, struct matchgate* ainv);                    /* :494 */
void polar_matchgate_factor(const struct matchgate* a, struct matchgate* u);                 /* :517 */
void retract_matchgate(const struct matchgate* u, const double* eta, struct matchgate* v);   /* :548 */

#ifdef __cplusplus
}
#endif
