// Riemannian trust region with the truncated CG on the device, one HIP graph per iteration
// (mg_device_tr.hip).
#pragma once
#include "mg_runtime.hpp"

struct rtr_params;

namespace mg {

// the loop state in device memory (copied to pinned host memory at the end of each iteration)
struct DevTRState {
    double radius, maxradius, rho_trust, abstol, reltol;
    double fx, fxn, rsq, stoptol;
    int maxiter;    // truncated-CG steps at most
    int tcg_it;     // steps of the current truncated CG
    int done;       // the truncated CG has ended (its remaining steps are skipped)
    int skip_tail;  // the iteration's tail (retraction, ratio test) waits for the CG
    int bnd;        // the CG ended on the trust-region boundary
    int k;          // iterations done
    int nhvp;       // HVP sweeps
    int reuse;      // H eta from the CG's products, f + gradient at the trial point kept
    int kept;       // (reuse) grad and fx hold the gradient and f at the iterate: no head sweep
    double fx_im, fxn_im;
};

struct DevTRStats {
    int iterations = 0;
    long graph_launches = 0;  // graphs launched (one per iteration, plus continuations of long CGs)
};

// riemannian_trust_region_optimize_hvp on program p and dense target t, precision kFP64 or kFP32
// with the last precision_switch iterations in FP64.  reuse 0: the reference control flow (H eta by
// its own HVP sweep, f alone at the trial point); 1: H eta summed from the truncated CG's products
// and f with its gradient at the trial point, kept for the next iteration (as mg_serve's tr_reuse;
// the same iterates up to rounding, one HVP sweep and one f + gradient sweep fewer per iteration).
// Returns 1 when done, 0 when this configuration runs on the host loop instead (a circuit target, a
// per-iteration callback, a sweep without a plan), -1 on error.
int optimize_device(const Target& t, const Program& p, const zd* gates_start, const rtr_params* prm, int niter,
                    int precision, int precision_switch, int reuse, double* f_iter, zd* f_citer, zd* gates_opt,
                    int* nhvp, DevTRStats* stats = nullptr);

}  // namespace mg
