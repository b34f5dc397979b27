// Host-side helpers shared by the C-ABI files: projection of the sweep outputs
// onto the Riemannian gradient and Hessian-vector product.
#pragma once
#include <functional>
#include <memory>

#include "../../include/matchgate.h"
#include "../../include/trust_region.h"
#include "mg_runtime.hpp"

// the opaque handles of include/mg_gpu.h
struct mg_target {
    std::unique_ptr<mg::Target> t;
};
struct mg_program {
    mg::Program p;
};

namespace mg {

// grad_vec[8 i ..] = tangent_to_real(V_i, P_{V_i}(conj(dgrad_i)))  (reference src/matchgate_target.c:903-908)
void riemannian_gradient(const struct matchgate* V, int P, const zd* dgrad, double* grad_vec);

// one parameter gate's block of the Riemannian Hessian-vector product:
// tangent_to_real(V, P_V(conj(hz)) - (Z g^H V + V g^H Z) / 2), g = conj(dgrad)
// (reference src/matchgate_target.c:937-951, applied to a direction)
void weingarten_block(const struct matchgate* V, const zd* g, const struct matchgate* Z, const zd* hz, double* r);

// riemannian_trust_region_optimize_hvp; with heta_from_cg the model's H eta in the ratio test
// is the truncated CG's own sum of its steps times the products H d it already evaluated
// (linear in the steps, equal to a separate product up to rounding), which saves one
// Hessian-vector product per iteration
void rtr_optimize_hvp(const target_func f, const target_gradient_func f_deriv, const hessian_vector_product_func f_hvp,
                      void* fdata, retract_func retract, void* rdata, const int n, const double* x_init, const int m,
                      const struct rtr_params* params, const int niter, double* f_iter, double* x_final,
                      bool heta_from_cg, const std::function<bool()>* stop = nullptr);

}  // namespace mg
