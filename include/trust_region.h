/*
 * Riemannian trust region with Steihaug-Toint truncated CG, host FP64
 * (replaces reference src/trust_region.h:11-90).  Same parameters, defaults and
 * callback types; riemannian_trust_region_optimize_hvp, declared but not
 * implemented by the reference (src/trust_region.c:319-381 is commented out), is
 * implemented here.
 */
#pragma once
#include <stdbool.h>
#include "config.h"

#ifdef __cplusplus
extern "C" {
#endif

struct truncated_cg_params {
    int maxiter;   /* maximum number of iterations */
    double abstol; /* absolute tolerance (early stopping) */
    double reltol; /* relative tolerance (early stopping) */
};

/* reference src/trust_region.h:24-29 */
static inline void set_truncated_cg_default_params(int n, struct truncated_cg_params* params)
{
    params->maxiter = 2 * n;
    params->abstol = 1e-8;
    params->reltol = 1e-6;
}

bool truncated_cg_hmat(const double* grad, const double* hess, int n, double radius,
                       const struct truncated_cg_params* params, double* z); /* :32 */

typedef void (*hessian_vector_product_func)(const double* x, void* fdata, double* vec, double* hvp); /* :35 */

bool truncated_cg_hvp(const double* x, const double* grad, const hessian_vector_product_func f_hvp, void* fdata,
                      const int n, const double radius, const struct truncated_cg_params* params, double* z); /* :37 */

typedef double (*target_func)(const double* x, void* fdata, numeric* f_complex);                    /* :40 */
typedef double (*target_gradient_func)(const double* x, void* fdata, double* grad);                /* :42 */
typedef double (*target_gradient_hessian_func)(const double* x, void* fdata, numeric* complex_f, double* grad,
                                               double* hess);                                        /* :44 */
typedef void (*retract_func)(const double* x, const double* eta, void* rdata, double* xs);          /* :46 */

struct rtr_params {
    struct truncated_cg_params tcg_params; /* parameters of the inner truncated CG */
    double rho_trust;                      /* acceptance threshold */
    double radius_init;                    /* initial radius */
    double maxradius;                      /* maximum radius */
    target_func g_func;                    /* optional function evaluated at each iteration */
    void* g_data;
    double* g_iter;                        /* niter + 1 evaluations of g_func */
};

/* reference src/trust_region.h:69-78 */
static inline void set_rtr_default_params(int n, struct rtr_params* params)
{
    set_truncated_cg_default_params(n, &params->tcg_params);
    params->rho_trust = 0.125;
    params->radius_init = 0.01;
    params->maxradius = 0.1;
    params->g_func = NULL;
    params->g_data = NULL;
    params->g_iter = NULL;
}

void riemannian_trust_region_optimize_hmat(target_func f, target_gradient_hessian_func f_deriv_hess, void* fdata,
                                           retract_func retract, void* rdata, const int n, const double* x_init,
                                           const int m, const struct rtr_params* params, const int niter,
                                           double* f_iter, numeric* f_citer, double* x_final); /* :80 */

void riemannian_trust_region_optimize_hvp(const target_func f, const target_gradient_func f_deriv,
                                          const hessian_vector_product_func f_hvp, void* fdata, retract_func retract,
                                          void* rdata, const int n, const double* x_init, const int m,
                                          const struct rtr_params* params, const int niter, double* f_iter,
                                          double* x_final); /* :84 */

#ifdef __cplusplus
}
#endif
