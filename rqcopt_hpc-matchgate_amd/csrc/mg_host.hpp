// Host-side helpers shared by the C-ABI files: projection of the sweep outputs
// onto the Riemannian gradient and Hessian-vector product.
#pragma once
#include <memory>

#include "../../include/matchgate.h"
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

}  // namespace mg
