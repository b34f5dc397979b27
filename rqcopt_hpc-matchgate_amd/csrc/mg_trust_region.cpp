// Riemannian trust region (Absil, Mahony, Sepulchre 2008, Alg. 10) with the
// Steihaug-Toint truncated CG (Alg. 11), host FP64 (C-ABI of include/trust_region.h).
//
// Same control flow, parameters and return values as reference
// src/trust_region.c:19-306; the HVP variant implements the optimiser that the
// reference declares (src/trust_region.h:84-86) but leaves commented out
// (src/trust_region.c:319-381).  Dense level-1/2 kernels are plain loops (no CBLAS).
// mg::rtr_optimize_hvp adds the option to take H eta for the ratio test from the truncated
// CG's own products (one Hessian-vector product per iteration fewer).
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <vector>

#include "../../include/trust_region.h"
#include "mg_host.hpp"

namespace {

double dot(int n, const double* a, const double* b)
{
    double s = 0.0;
    for (int k = 0; k < n; k++) s += a[k] * b[k];
    return s;
}

// roots of x^2 + 2 p x + q = 0, ascending (reference src/trust_region.c:19-45)
void quadratic_roots(double p, double q, double x[2])
{
    assert(p * p - q >= 0);
    if (p == 0) {
        const double a = std::sqrt(-q);
        x[0] = -a;
        x[1] = a;
        return;
    }
    const double a = -(p + std::copysign(1.0, p) * std::sqrt(p * p - q));
    const double b = q / a;
    x[0] = std::fmin(a, b);
    x[1] = std::fmax(a, b);
}

// t > 0 with ||b + t d|| = radius (reference src/trust_region.c:54-72)
double boundary_step(const double* b, const double* d, int n, double radius)
{
    const double dsq = dot(n, d, d);
    if (dsq == 0) {
        std::fprintf(stderr, "input vector 'd' is zero\n");
        return 0;
    }
    double t[2];
    quadratic_roots(dot(n, b, d) / dsq, (dot(n, b, b) - radius * radius) / dsq, t);
    if (t[1] < 0) std::fprintf(stderr, "encountered t < 0: t = %g\n", t[1]);
    return t[1];
}

// minimise <grad, z> + 1/2 <z, H z> subject to ||z|| <= radius; returns true when
// the iterate was moved to the boundary.  Hz (nullable) receives H z, accumulated from the
// products H d of the CG directions with the same steps as z.
bool steihaug(const double* grad, const std::function<void(const double*, double*)>& hv, int n, double radius,
              const truncated_cg_params* prm, double* z, double* Hz = nullptr,
              const std::function<bool()>* stop = nullptr)
{
    std::vector<double> r(grad, grad + n), d(n), Hd(n);
    double rsq = dot(n, r.data(), r.data());
    const double stoptol = std::fmax(prm->abstol, prm->reltol * std::sqrt(rsq));
    std::memset(z, 0, n * sizeof(double));
    if (Hz) std::memset(Hz, 0, n * sizeof(double));
    for (int k = 0; k < n; k++) d[k] = -r[k];
    for (int it = 0; it < prm->maxiter; it++) {
        hv(d.data(), Hd.data());
        if (stop && (*stop)()) return false;
        const double dHd = dot(n, d.data(), Hd.data());
        const double t = boundary_step(z, d.data(), n, radius);
        const double alpha = rsq / dHd;
        if (dHd <= 0 || alpha > t) {
            for (int k = 0; k < n; k++) z[k] += t * d[k];
            if (Hz)
                for (int k = 0; k < n; k++) Hz[k] += t * Hd[k];
            return true;
        }
        for (int k = 0; k < n; k++) r[k] += alpha * Hd[k];
        for (int k = 0; k < n; k++) z[k] += alpha * d[k];
        if (Hz)
            for (int k = 0; k < n; k++) Hz[k] += alpha * Hd[k];
        const double rsq_next = dot(n, r.data(), r.data());
        if (std::sqrt(rsq_next) <= stoptol) break;
        const double beta = rsq_next / rsq;
        for (int k = 0; k < n; k++) d[k] = beta * d[k] - r[k];
        rsq = rsq_next;
    }
    return false;
}

void dense_mv(const double* H, int n, const double* x, double* y)
{
    for (int a = 0; a < n; a++) y[a] = dot(n, H + (size_t)a * n, x);
}

}  // namespace

extern "C" {

bool truncated_cg_hmat(const double* grad, const double* hess, int n, double radius,
                       const struct truncated_cg_params* params, double* z)
{
    return steihaug(grad, [&](const double* d, double* Hd) { dense_mv(hess, n, d, Hd); }, n, radius, params, z);
}

bool truncated_cg_hvp(const double* x, const double* grad, const hessian_vector_product_func f_hvp, void* fdata,
                      const int n, const double radius, const struct truncated_cg_params* params, double* z)
{
    return steihaug(grad, [&](const double* d, double* Hd) { f_hvp(x, fdata, const_cast<double*>(d), Hd); }, n, radius,
                    params, z);
}

// radius update and acceptance: reference src/trust_region.c:271-285
static void tr_update(double rho, bool on_boundary, const rtr_params* params, double* radius, bool* accept)
{
    if (rho < 0.25)
        *radius *= 0.25;
    else if (rho > 0.75 && on_boundary)
        *radius = std::fmin(2 * *radius, params->maxradius);
    *accept = rho > params->rho_trust;
}

void riemannian_trust_region_optimize_hmat(target_func f, target_gradient_hessian_func f_deriv_hess, void* fdata,
                                           retract_func retract, void* rdata, const int n, const double* x_init,
                                           const int m, const struct rtr_params* params, const int niter,
                                           double* f_iter, numeric* f_citer, double* x_final)
{
    assert(0 <= params->rho_trust && params->rho_trust < 0.25);
    std::vector<double> x(x_init, x_init + m), xn(m), grad(n), hess((size_t)n * n), eta(n), Heta(n);
    double radius = params->radius_init;
    numeric fc;
    if (params->g_func) params->g_iter[0] = params->g_func(x.data(), params->g_data, &fc);
    for (int k = 0; k < niter; k++) {
        const double fx = f_deriv_hess(x.data(), fdata, &fc, grad.data(), hess.data());
        const bool bnd = truncated_cg_hmat(grad.data(), hess.data(), n, radius, &params->tcg_params, eta.data());
        retract(x.data(), eta.data(), rdata, xn.data());
        f_iter[k] = fx;
        if (f_citer) f_citer[k] = fc;
        dense_mv(hess.data(), n, eta.data(), Heta.data());
        const double rho = (f(xn.data(), fdata, &fc) - fx) / (dot(n, grad.data(), eta.data()) + 0.5 * dot(n, eta.data(), Heta.data()));
        bool accept;
        tr_update(rho, bnd, params, &radius, &accept);
        if (accept) x.swap(xn);
        if (params->g_func) params->g_iter[k + 1] = params->g_func(x.data(), params->g_data, &fc);
    }
    std::memcpy(x_final, x.data(), m * sizeof(double));
    f_iter[niter] = f(x_final, fdata, &fc);
    if (f_citer) f_citer[niter] = fc;
}

}  // extern "C"

namespace mg {

void rtr_optimize_hvp(const target_func f, const target_gradient_func f_deriv, const hessian_vector_product_func f_hvp,
                      void* fdata, retract_func retract, void* rdata, const int n, const double* x_init, const int m,
                      const struct rtr_params* params, const int niter, double* f_iter, double* x_final,
                      bool heta_from_cg, const std::function<bool()>* stop)
{
    assert(0 <= params->rho_trust && params->rho_trust < 0.25);
    std::vector<double> x(x_init, x_init + m), xn(m), grad(n), eta(n), Heta(n);
    double radius = params->radius_init;
    numeric fc;
    if (params->g_func) params->g_iter[0] = params->g_func(x.data(), params->g_data, &fc);
    // a caller's stop request (an evaluation failed): the remaining f_iter are NaN, x_final the
    // current iterate
    auto stopped = [&](int k) {
        if (!stop || !(*stop)()) return false;
        for (int i = k; i <= niter; i++) f_iter[i] = NAN;
        std::memcpy(x_final, x.data(), m * sizeof(double));
        return true;
    };
    for (int k = 0; k < niter; k++) {
        const double fx = f_deriv(x.data(), fdata, grad.data());
        if (stopped(k)) return;
        const bool bnd = steihaug(
            grad.data(), [&](const double* d, double* Hd) { f_hvp(x.data(), fdata, const_cast<double*>(d), Hd); }, n,
            radius, &params->tcg_params, eta.data(), heta_from_cg ? Heta.data() : nullptr, stop);
        if (stopped(k)) return;
        retract(x.data(), eta.data(), rdata, xn.data());
        f_iter[k] = fx;
        if (!heta_from_cg) f_hvp(x.data(), fdata, eta.data(), Heta.data());
        const double fxn = f(xn.data(), fdata, &fc);
        if (stopped(k + 1)) return;
        const double rho = (fxn - fx) / (dot(n, grad.data(), eta.data()) + 0.5 * dot(n, eta.data(), Heta.data()));
        bool accept;
        tr_update(rho, bnd, params, &radius, &accept);
        if (accept) x.swap(xn);
        if (params->g_func) params->g_iter[k + 1] = params->g_func(x.data(), params->g_data, &fc);
    }
    std::memcpy(x_final, x.data(), m * sizeof(double));
    f_iter[niter] = f(x_final, fdata, &fc);
}

}  // namespace mg

extern "C" {

void riemannian_trust_region_optimize_hvp(const target_func f, const target_gradient_func f_deriv,
                                          const hessian_vector_product_func f_hvp, void* fdata, retract_func retract,
                                          void* rdata, const int n, const double* x_init, const int m,
                                          const struct rtr_params* params, const int niter, double* f_iter,
                                          double* x_final)
{
    mg::rtr_optimize_hvp(f, f_deriv, f_hvp, fdata, retract, rdata, n, x_init, m, params, niter, f_iter, x_final, false);
}

void mg_trust_region_optimize_hvp(const target_func f, const target_gradient_func f_deriv,
                                  const hessian_vector_product_func f_hvp, void* fdata, retract_func retract,
                                  void* rdata, const int n, const double* x_init, const int m,
                                  const struct rtr_params* params, const int niter, double* f_iter, double* x_final,
                                  int heta_from_cg)
{
    mg::rtr_optimize_hvp(f, f_deriv, f_hvp, fdata, retract, rdata, n, x_init, m, params, niter, f_iter, x_final,
                         heta_from_cg != 0);
}

}  // extern "C"
