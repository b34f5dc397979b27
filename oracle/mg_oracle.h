/*
 * CPU oracle (test infrastructure only) for the matchgate-brickwall target,
 * gradient, Hessian and Hessian-vector product.
 *
 * This is a from-scratch FP64 restatement of the reference's algorithms, used
 * by tests/ (to check results), by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  The product library never links or calls it.
 *
 * Conventions (reference src/matchgate.h:5-14, src/gate.c:581-652):
 *   - a gate is 8 complex numbers: center[4] (odd parity {01,10}) followed by
 *     corner[4] (even parity {00,11}), each a row-major 2x2 block;
 *   - qubit 0 is the most significant bit of the amplitude index;
 *   - a gate on wires (i, j) with j < i uses the center block reversed.
 *
 * "Gate program": G gate instances, instance g acts with parameter gate
 * param[g] on wires (wires[2g], wires[2g+1]).  The reference's tied brickwall
 * (one gate per layer, L/2 instances per layer, src/matchgate_target.c:308-324)
 * is the program built by orc_brickwall_program().
 */
#pragma once
#include <complex.h>
#include <stdint.h>

typedef double complex ocplx;

/* ---------------- gate kernels (reference src/gate.c) ---------------- */
void orc_apply_gate(const ocplx* g, int i, int j, int L, const ocplx* psi, ocplx* out);
void orc_apply_gate_backward(const ocplx* g, int i, int j, int L, const ocplx* psi,
                             const ocplx* dout, ocplx* dpsi, ocplx* dg);
void orc_apply_gate_to_array(const ocplx* g, int i, int j, int L, int nstates, const ocplx* psi, ocplx* out);
void orc_apply_gate_backward_array(const ocplx* g, int i, int j, int L, int nstates, const ocplx* psi,
                                   const ocplx* dout, ocplx* dgs);
void orc_apply_gate_placeholder(int i, int j, int L, const ocplx* psi, ocplx* out);

/* ---------------- gate programs ---------------- */
void orc_brickwall_program(int L, int nlayers, const int* perms, int* param, int* wires);
void orc_apply_program(const ocplx* gates, int G, const int* param, const int* wires, int L,
                       const ocplx* psi, ocplx* out);
/* forward with cache: cache[g*S..] = state before instance g */
void orc_program_forward(const ocplx* gates, int G, const int* param, const int* wires, int L,
                         const ocplx* psi, ocplx* cache, ocplx* out);
/* backward: per-instance gradients dgi[G*8] (not summed), dpsi */
void orc_program_backward(const ocplx* gates, int G, const int* param, const int* wires, int L,
                          const ocplx* cache, const ocplx* dout, ocplx* dpsi, ocplx* dgi);
/* backward + complex Hessian over parameters (P params, hess m x m, m = 8P) */
void orc_program_backward_hessian(const ocplx* gates, int P, int G, const int* param, const int* wires, int L,
                                  const ocplx* cache, const ocplx* dout, ocplx* dpsi, ocplx* dgrad, ocplx* hess);
/* translation-invariant shortcut (reference src/matchgate_brickwall.c:308-435), tied brickwall only */
void orc_ti_brickwall_backward(const ocplx* vlist, int nlayers, const int* perms, int L, const ocplx* cache,
                               const ocplx* dout, ocplx* dpsi, ocplx* dgrad);
void orc_ti_brickwall_backward_hessian(const ocplx* vlist, int nlayers, const int* perms, int L,
                                       const ocplx* cache, const ocplx* dout, ocplx* dpsi, ocplx* dgrad, ocplx* hess);

/* ---------------- targets: trace over basis columns, dense target U ---------------- */
/* f = -Tr(U^dagger W); U row-major S x S */
void orc_target(const ocplx* U, const ocplx* gates, int G, const int* param, const int* wires, int L,
                int nthreads, ocplx* f);
void orc_target_gradient(const ocplx* U, const ocplx* gates, int P, int G, const int* param, const int* wires, int L,
                         int nthreads, ocplx* f, ocplx* dgrad);
void orc_target_gradient_hessian(const ocplx* U, const ocplx* gates, int P, int G, const int* param, const int* wires,
                                 int L, int ti, int nthreads, ocplx* f, ocplx* dgrad, ocplx* hess);
/* f, complex gradient and complex directional second derivative D^2 f[Z] (holomorphic);
 * column range [b0, b1) so that a bounded sample can be timed */
void orc_target_gradient_hvp(const ocplx* U, const ocplx* gates, const ocplx* Z, int P, int G, const int* param,
                             const int* wires, int L, long b0, long b1, int nthreads,
                             ocplx* f, ocplx* dgrad, ocplx* dhz);
/* the same with the target given as a gate program (gate k on (twires[2k], twires[2k+1])),
 * applied to e_b per column instead of a dense U */
void orc_target_gradient_hvp_program(const ocplx* tgates, int Gt, const int* twires, const ocplx* gates, const ocplx* Z,
                                     int P, int G, const int* param, const int* wires, int L, long b0, long b1,
                                     int nthreads, ocplx* f, ocplx* dgrad, ocplx* dhz);

/* ---------------- matchgate algebra (reference src/matchgate.c) ---------------- */
void orc_multiply(const ocplx* a, const ocplx* b, ocplx* c);
void orc_adjoint(const ocplx* a, ocplx* ah);
void orc_conjugate(ocplx* a);
void orc_symm(const ocplx* w, ocplx* z);
void orc_antisymm(const ocplx* w, ocplx* z);
void orc_real_to_antisymm(const double* r, ocplx* w);
void orc_antisymm_to_real(const ocplx* w, double* r);
void orc_real_to_tangent(const double* r, const ocplx* v, ocplx* z);
void orc_tangent_to_real(const ocplx* v, const ocplx* z, double* r);
void orc_project_tangent(const ocplx* u, const ocplx* z, ocplx* p);
void orc_symmetric_triple(const ocplx* a, const ocplx* b, const ocplx* c, ocplx* ret);
int  orc_inverse_mat2x2(const ocplx* a, ocplx* ainv);
void orc_polar(const ocplx* a, ocplx* u);
void orc_retract(const ocplx* u, const double* eta, ocplx* v);

/* ---------------- Riemannian projection ---------------- */
void orc_gradient_vector_hessian_matrix(const ocplx* vlist, int P, const ocplx* dgrad, const ocplx* hess,
                                        double* grad_vec, double* H);
void orc_riemannian_gradient(const ocplx* vlist, int P, const ocplx* dgrad, double* grad_vec);
void orc_real_to_directions(const ocplx* vlist, int P, const double* v, ocplx* Z);
void orc_riemannian_hvp(const ocplx* vlist, int P, const ocplx* dgrad, const ocplx* Z, const ocplx* dhz, double* hvp);

/* ---------------- trust region (reference src/trust_region.c) ---------------- */
int orc_truncated_cg_hmat(const double* grad, const double* hess, int n, double radius,
                          int maxiter, double abstol, double reltol, double* z);
/* Riemannian trust region on the gate program with dense target.
 * mode 0: explicit Hessian matrix (reference algorithm), mode 1: HVP */
void orc_rtr_optimize(const ocplx* U, const ocplx* vstart, int P, int G, const int* param, const int* wires, int L,
                      int mode, int niter, int tcg_maxiter, double rho_trust, double radius_init, double maxradius,
                      double abstol, double reltol, int nthreads, double* f_iter, ocplx* vopt, int* hvp_count);
