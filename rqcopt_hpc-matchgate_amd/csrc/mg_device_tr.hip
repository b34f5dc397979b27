// Riemannian trust region with the truncated CG on the device, one HIP graph per iteration.
//
// The host form (mg_trust_region.cpp; reference src/trust_region.c:235-306 with the HVP solver of
// src/trust_region.h:84-86) stops after every evaluation to run the truncated CG, the retraction
// and the ratio test.  Here the iterate, the CG vectors and the radius live in device memory and
// one trust-region iteration is one graph launch:
//
//   head   gate coefficients at x, f + gradient sweep, Riemannian gradient, CG start
//   K x    coefficients at (x, d), HVP sweep, Riemannian H d, one CG step       (skipped once done)
//   tail   retraction, f at the trial point, H eta, ratio test, radius, accept  (skipped until done)
//
// followed by a copy of the loop state to pinned host memory.  A CG that needs more than K steps
// continues in a second graph (K more steps and the tail).  Every sweep runs from a plan
// (plan_sweep): the single-request path's own kernel and layout, with its coefficients rewritten
// on the device by k_coefs before each launch.  The small matchgate algebra (projections,
// Weingarten term, polar retraction) follows mg_algebra.cpp / mg_api.cpp, one thread per gate.
//
// Reference: trust region src/trust_region.c:255-291, truncated CG src/trust_region.c:74-170,
// retraction src/matchgate.c:517-562, Riemannian gradient and Hessian src/matchgate_target.c:903-967.
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>

#include "../../include/trust_region.h"
#include "mg_device_tr.hpp"
#include "mg_plan.hpp"

namespace mg {

namespace {

using C = cpx<double>;

// ---- complex 2x2 block algebra (row-major blocks; a matchgate is center then corner) ----------
__device__ inline C cadd(C a, C b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline C csub(C a, C b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline C cmuld(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline C cconj(C a) { return {a.re, -a.im}; }
__device__ inline C cscale(C a, double s) { return {a.re * s, a.im * s}; }
__device__ inline double cabsd(C a) { return hypot(a.re, a.im); }
// a / b, Smith's method (no overflow for large |b|)
__device__ inline C cdiv(C a, C b)
{
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

struct MG {
    C b[2][4];  // [0] center, [1] corner
};

__device__ inline MG load_mg(const double* x)
{
    MG m;
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < 4; k++) m.b[q][k] = {x[8 * q + 2 * k], x[8 * q + 2 * k + 1]};
    return m;
}
__device__ inline void store_mg(double* x, const MG& m)
{
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < 4; k++) {
            x[8 * q + 2 * k] = m.b[q][k].re;
            x[8 * q + 2 * k + 1] = m.b[q][k].im;
        }
}
// the 8 complex entries of parameter q in a reduced sweep output, at double offset `base`
__device__ inline MG load_out(const double* out, int q, int base)
{
    MG m;
    for (int e = 0; e < 8; e++) m.b[e >> 2][e & 3] = {out[2 + 32 * q + base + 2 * e], out[2 + 32 * q + base + 2 * e + 1]};
    return m;
}
__device__ inline MG mul(const MG& a, const MG& b)
{
    MG c;
    for (int q = 0; q < 2; q++) {
        const C* x = a.b[q];
        const C* y = b.b[q];
        c.b[q][0] = cadd(cmuld(x[0], y[0]), cmuld(x[1], y[2]));
        c.b[q][1] = cadd(cmuld(x[0], y[1]), cmuld(x[1], y[3]));
        c.b[q][2] = cadd(cmuld(x[2], y[0]), cmuld(x[3], y[2]));
        c.b[q][3] = cadd(cmuld(x[2], y[1]), cmuld(x[3], y[3]));
    }
    return c;
}
__device__ inline MG adjoint(const MG& a)
{
    MG c;
    for (int q = 0; q < 2; q++) {
        c.b[q][0] = cconj(a.b[q][0]);
        c.b[q][3] = cconj(a.b[q][3]);
        c.b[q][1] = cconj(a.b[q][2]);
        c.b[q][2] = cconj(a.b[q][1]);
    }
    return c;
}
__device__ inline MG conj_mg(const MG& a)
{
    MG c;
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < 4; k++) c.b[q][k] = cconj(a.b[q][k]);
    return c;
}
// skew-Hermitian blocks of the real coordinates r[8]
__device__ inline MG real_to_antisymm(const double* r)
{
    MG w;
    for (int q = 0; q < 2; q++) {
        const double* x = r + 4 * q;
        w.b[q][0] = {0.0, x[0]};
        w.b[q][3] = {0.0, x[3]};
        w.b[q][1] = {0.5 * (x[1] - x[2]), 0.5 * (x[1] + x[2])};
        w.b[q][2] = {-w.b[q][1].re, w.b[q][1].im};  // -conj
    }
    return w;
}
// coordinates of antisymm(v^dagger z): Re + Im of each entry of the skew-Hermitian part
__device__ inline void tangent_to_real(const MG& v, const MG& z, double* r)
{
    const MG t = mul(adjoint(v), z);
    for (int q = 0; q < 2; q++) {
        const C* s = t.b[q];
        const C off = cscale(csub(s[1], cconj(s[2])), 0.5);
        const C a[4] = {{0.0, s[0].im}, off, {-off.re, off.im}, {0.0, s[3].im}};
        for (int k = 0; k < 4; k++) r[4 * q + k] = a[k].re + a[k].im;
    }
}
// z - u symm(u^dagger z)
__device__ inline MG project_tangent(const MG& u, const MG& z)
{
    const MG w = mul(adjoint(u), z);
    MG s;
    for (int q = 0; q < 2; q++) {
        const C off = cscale(cadd(w.b[q][1], cconj(w.b[q][2])), 0.5);
        s.b[q][0] = {w.b[q][0].re, 0.0};
        s.b[q][3] = {w.b[q][3].re, 0.0};
        s.b[q][1] = off;
        s.b[q][2] = cconj(off);
    }
    const MG us = mul(u, s);
    MG p;
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < 4; k++) p.b[q][k] = csub(z.b[q][k], us.b[q][k]);
    return p;
}
// Riemannian Hessian-vector coordinates of one gate: P_V(conj(hz)) - (Z g^H V + V g^H Z) / 2 with
// g = conj of the Euclidean gradient entries (mg_api.cpp weingarten_block)
__device__ inline void weingarten(const MG& v, const MG& g, const MG& z, const MG& hz, double* r)
{
    const MG gp = project_tangent(v, conj_mg(hz));
    const MG gh = adjoint(conj_mg(g));
    const MG t1 = mul(mul(z, gh), v), t2 = mul(mul(v, gh), z);
    MG d;
    for (int q = 0; q < 2; q++)
        for (int k = 0; k < 4; k++) d.b[q][k] = csub(gp.b[q][k], cscale(cadd(t1.b[q][k], t2.b[q][k]), 0.5));
    tangent_to_real(v, d, r);
}
// Gauss-Jordan inverse of a 2x2 block with partial pivoting (mg_algebra.cpp inverse_mat2x2)
__device__ inline void inverse22(const C* a, C* out)
{
    C m[4] = {a[0], a[1], a[2], a[3]};
    C v[4] = {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {1.0, 0.0}};
    for (int k = 0; k < 2; k++) {
        int piv = k;
        double p = cabsd(m[2 * k + k]);
        if (k == 0 && cabsd(m[2]) > p) {
            piv = 1;
            p = cabsd(m[2]);
        }
        if (piv != k)
            for (int c = 0; c < 2; c++) {
                const C t0 = m[2 * k + c];
                m[2 * k + c] = m[2 * piv + c];
                m[2 * piv + c] = t0;
                const C t1 = v[2 * k + c];
                v[2 * k + c] = v[2 * piv + c];
                v[2 * piv + c] = t1;
            }
        for (int r = 0; r < 2; r++) {
            if (r == k) continue;
            const C s = cdiv(m[2 * r + k], m[2 * k + k]);
            for (int c = k + 1; c < 2; c++) m[2 * r + c] = csub(m[2 * r + c], cmuld(s, m[2 * k + c]));
            m[2 * r + k] = {0.0, 0.0};
            for (int c = 0; c < 2; c++) v[2 * r + c] = csub(v[2 * r + c], cmuld(s, v[2 * k + c]));
        }
        for (int c = k + 1; c < 2; c++) m[2 * k + c] = cdiv(m[2 * k + c], m[2 * k + k]);
        for (int c = 0; c < 2; c++) v[2 * k + c] = cdiv(v[2 * k + c], m[2 * k + k]);
        m[2 * k + k] = {1.0, 0.0};
    }
    for (int c = 0; c < 4; c++) out[c] = v[c];
}
// unitary polar factor, Newton u <- (u + u^{-dagger}) / 2 (mg_algebra.cpp polar_matchgate_factor)
__device__ inline MG polar(const MG& a)
{
    MG u = a;
    for (int k = 0; k < 14; k++) {
        MG inv;
        inverse22(u.b[0], inv.b[0]);
        inverse22(u.b[1], inv.b[1]);
        const MG w = adjoint(inv);
        double d0 = 0.0, d1 = 0.0;
        for (int q = 0; q < 2; q++)
            for (int e = 0; e < 4; e++) {
                u.b[q][e] = cscale(cadd(u.b[q][e], w.b[q][e]), 0.5);
                const double dd = cabsd(csub(u.b[q][e], w.b[q][e]));
                if (q == 0) d0 = fmax(d0, dd);
                else d1 = fmax(d1, dd);
            }
        if (k >= 4 && d0 < 1e-14 && d1 < 1e-14) break;
    }
    return u;
}
// 2x2 inverse by the closed form (the sweeps' V^{-1}; mg_runtime.hip build_instances)
__device__ inline void inverse2(const C* a, C* inv)
{
    const C det = csub(cmuld(a[0], a[3]), cmuld(a[1], a[2]));
    inv[0] = cdiv(a[3], det);
    inv[1] = cdiv({-a[1].re, -a[1].im}, det);
    inv[2] = cdiv({-a[2].re, -a[2].im}, det);
    inv[3] = cdiv(a[0], det);
}

constexpr int kBT = 256;  // threads of the single-workgroup kernels

// sum over the workgroup (fixed order: tree over the threads' partial sums)
__device__ inline double block_sum(double v, double* sh)
{
    const int t = threadIdx.x;
    __syncthreads();
    sh[t] = v;
    __syncthreads();
    for (int s = kBT / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}
__device__ inline double block_dot(const double* a, const double* b, int n, double* sh)
{
    double v = 0.0;
    for (int k = threadIdx.x; k < n; k += kBT) v += a[k] * b[k];
    return block_sum(v, sh);
}

// ---- kernels ----------------------------------------------------------------------------------
// per instance: V, V^{-1}, Z in the sorted basis (center reversed for flipped wires) from the gates
// x and the tangent coordinates v (null: Z = 0); the kernel's copy at its position kpos[g] in the
// sweep precision, and the FP64 program-order copy when prog64 is set
template <typename R>
__global__ void k_coefs(const double* x, const double* v, const int2* pm, const int* kpos, int G, cpx<R>* kcoef,
                        C* prog64, const int* skip)
{
    if (skip && *skip) return;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const int par = pm[g].x, flip = pm[g].y;
    const MG V = load_mg(x + 16 * (size_t)par);
    MG Z{};
    if (v) Z = mul(V, real_to_antisymm(v + 8 * (size_t)par));
    // V, V^{-1}, Z: center (reversed for flipped wires) then corner
    C c[24];
    for (int k = 0; k < 4; k++) {
        c[k] = V.b[0][flip ? 3 - k : k];
        c[4 + k] = V.b[1][k];
        c[16 + k] = Z.b[0][flip ? 3 - k : k];
        c[20 + k] = Z.b[1][k];
    }
    inverse2(c, c + 8);
    inverse2(c + 4, c + 12);
    cpx<R>* kd = kcoef + (size_t)kpos[g] * kCoefPerInst;
    for (int k = 0; k < 24; k++) kd[k] = {(R)c[k].re, (R)c[k].im};
    if (prog64)
        for (int k = 0; k < 24; k++) prog64[(size_t)g * kCoefPerInst + k] = c[k];
}

// f and the Riemannian gradient at x from an f + gradient sweep (with reuse, once kept: those of
// the last trial point or of the rejected step's iterate); the truncated CG's start (r = grad,
// d = -grad, z = 0, H z = 0); f_citer[k] = f
__global__ __launch_bounds__(kBT) void k_grad_init(const double* out, const double* x, int P, double* grad, double* r,
                                                   double* d, double* z, double* Hz, DevTRState* st, double* f_citer)
{
    __shared__ double sh[kBT];
    const int t = threadIdx.x, n = 8 * P;
    const bool have = st->reuse && st->kept;
    if (!have)
        for (int q = t; q < P; q += kBT) {
            const MG V = load_mg(x + 16 * (size_t)q);
            tangent_to_real(V, conj_mg(load_out(out, q, 0)), grad + 8 * (size_t)q);
        }
    __syncthreads();
    for (int k = t; k < n; k += kBT) {
        r[k] = grad[k];
        d[k] = -grad[k];
        z[k] = 0.0;
        Hz[k] = 0.0;
    }
    const double rsq = block_dot(grad, grad, n, sh);
    if (t == 0) {
        if (!have) {
            st->fx = out[0];
            st->fx_im = out[1];
        }
        if (f_citer) {
            f_citer[2 * st->k] = st->fx;
            f_citer[2 * st->k + 1] = st->fx_im;
        }
        st->rsq = rsq;
        st->stoptol = fmax(st->abstol, st->reltol * sqrt(rsq));
        st->tcg_it = 0;
        st->done = 0;
        st->skip_tail = 1;
        st->bnd = 0;
    }
}

// H d from an HVP sweep along d, then one step of the truncated CG (mg_trust_region.cpp
// steihaug; the roots as quadratic_roots / boundary_step)
__global__ __launch_bounds__(kBT) void k_hvp_tcg(const double* out, const double* x, int P, double* Hd, double* r,
                                                 double* d, double* z, double* Hz, DevTRState* st)
{
    __shared__ double sh[kBT];
    if (st->done) return;
    const int t = threadIdx.x, n = 8 * P;
    for (int q = t; q < P; q += kBT) {
        const MG V = load_mg(x + 16 * (size_t)q);
        const MG Z = mul(V, real_to_antisymm(d + 8 * (size_t)q));
        weingarten(V, load_out(out, q, 0), Z, load_out(out, q, 16), Hd + 8 * (size_t)q);
    }
    __syncthreads();
    const double dHd = block_dot(d, Hd, n, sh);
    const double dsq = block_dot(d, d, n, sh);
    const double zd_ = block_dot(z, d, n, sh);
    const double zz = block_dot(z, z, n, sh);
    const double radius = st->radius, rsq = st->rsq;
    // t > 0 with ||z + t d|| = radius: the larger root of t^2 + 2 p t + q = 0
    double tb = 0.0;
    if (dsq != 0.0) {
        const double p = zd_ / dsq, qq = (zz - radius * radius) / dsq;
        if (p == 0.0) {
            tb = sqrt(-qq);
        } else {
            const double a = -(p + copysign(1.0, p) * sqrt(p * p - qq));
            const double b = qq / a;
            tb = fmax(a, b);
        }
    }
    const double alpha = rsq / dHd;
    const bool boundary = dHd <= 0.0 || alpha > tb;
    if (boundary) {
        for (int k = t; k < n; k += kBT) {
            z[k] += tb * d[k];
            Hz[k] += tb * Hd[k];
        }
        if (t == 0) {
            st->bnd = 1;
            st->done = 1;
        }
    } else {
        for (int k = t; k < n; k += kBT) {
            r[k] += alpha * Hd[k];
            z[k] += alpha * d[k];
            Hz[k] += alpha * Hd[k];
        }
        const double rsq_next = block_dot(r, r, n, sh);
        const bool conv = sqrt(rsq_next) <= st->stoptol;
        if (!conv) {
            const double beta = rsq_next / rsq;
            for (int k = t; k < n; k += kBT) d[k] = beta * d[k] - r[k];
        }
        if (t == 0) {
            if (conv) st->done = 1;
            else st->rsq = rsq_next;
        }
    }
    __syncthreads();
    if (t == 0) {
        st->nhvp++;
        st->tcg_it++;
        if (!st->done && st->tcg_it >= st->maxiter) st->done = 1;
        st->skip_tail = st->done ? 0 : 1;
    }
}

// the trial point: per gate the polar factor of V (I + A(eta)) (mg_algebra.cpp retract_matchgate)
__global__ void k_retract(const double* x, const double* z, double* xn, int P, const int* skip)
{
    if (*skip) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    MG a = real_to_antisymm(z + 8 * (size_t)q);
    for (int b = 0; b < 2; b++) {
        a.b[b][0].re += 1.0;
        a.b[b][3].re += 1.0;
    }
    store_mg(xn + 16 * (size_t)q, polar(mul(load_mg(x + 16 * (size_t)q), a)));
}

__global__ void k_f_trial(const double* out, DevTRState* st, const int* skip)
{
    if (*skip) return;
    if (threadIdx.x == 0) st->fxn = out[0];
}

// H eta from an HVP sweep along eta, then the ratio test, the radius and the acceptance
// (mg_trust_region.cpp rtr_optimize_hvp / tr_update); f_iter[k] = f(x_k)
__global__ __launch_bounds__(kBT) void k_heta_update(const double* out, double* x, const double* xn, const double* z,
                                                     const double* grad, double* Heta, int P, DevTRState* st,
                                                     double* f_iter)
{
    __shared__ double sh[kBT];
    __shared__ int accept_sh;
    if (st->skip_tail) return;
    const int t = threadIdx.x, n = 8 * P;
    for (int q = t; q < P; q += kBT) {
        const MG V = load_mg(x + 16 * (size_t)q);
        const MG Z = mul(V, real_to_antisymm(z + 8 * (size_t)q));
        weingarten(V, load_out(out, q, 0), Z, load_out(out, q, 16), Heta + 8 * (size_t)q);
    }
    __syncthreads();
    const double gz = block_dot(grad, z, n, sh);
    const double zhz = block_dot(z, Heta, n, sh);
    if (t == 0) {
        const double rho = (st->fxn - st->fx) / (gz + 0.5 * zhz);
        if (rho < 0.25) st->radius *= 0.25;
        else if (rho > 0.75 && st->bnd) st->radius = fmin(2 * st->radius, st->maxradius);
        accept_sh = rho > st->rho_trust;
        f_iter[st->k] = st->fx;
        st->nhvp++;
        st->k++;
    }
    __syncthreads();
    if (accept_sh)
        for (int k = t; k < 16 * P; k += kBT) x[k] = xn[k];
}

// (reuse) f and the Riemannian gradient at the trial point from an f + gradient sweep there
__global__ __launch_bounds__(kBT) void k_trial_grad(const double* out, const double* xn, int P, double* grad_n,
                                                    DevTRState* st)
{
    if (st->skip_tail) return;
    const int t = threadIdx.x;
    for (int q = t; q < P; q += kBT) {
        const MG V = load_mg(xn + 16 * (size_t)q);
        tangent_to_real(V, conj_mg(load_out(out, q, 0)), grad_n + 8 * (size_t)q);
    }
    if (t == 0) {
        st->fxn = out[0];
        st->fxn_im = out[1];
    }
}

// (reuse) the ratio test with H eta = the CG's sum of its steps' products; on acceptance the
// trial point's f and gradient become the iterate's
__global__ __launch_bounds__(kBT) void k_update_reuse(double* x, const double* xn, const double* z, double* grad,
                                                      const double* grad_n, const double* Hz, int P, DevTRState* st,
                                                      double* f_iter)
{
    __shared__ double sh[kBT];
    __shared__ int accept_sh;
    if (st->skip_tail) return;
    const int t = threadIdx.x, n = 8 * P;
    const double gz = block_dot(grad, z, n, sh);
    const double zhz = block_dot(z, Hz, n, sh);
    if (t == 0) {
        const double rho = (st->fxn - st->fx) / (gz + 0.5 * zhz);
        if (rho < 0.25) st->radius *= 0.25;
        else if (rho > 0.75 && st->bnd) st->radius = fmin(2 * st->radius, st->maxradius);
        accept_sh = rho > st->rho_trust;
        f_iter[st->k] = st->fx;
        st->k++;
        st->kept = 1;
        if (accept_sh) {
            st->fx = st->fxn;
            st->fx_im = st->fxn_im;
        }
    }
    __syncthreads();
    if (accept_sh) {
        for (int k = t; k < 16 * P; k += kBT) x[k] = xn[k];
        for (int k = t; k < n; k += kBT) grad[k] = grad_n[k];
    }
}

// (reuse) f and gradient held from sweeps of the other precision: the next head sweeps again
__global__ void k_drop_kept(DevTRState* st)
{
    if (threadIdx.x == 0) st->kept = 0;
}

// every device buffer of one optimisation
struct Plans {
    SweepPlan p[3];  // modes f, f + gradient, f + gradient + HVP
    DevBuf kpos[3];
};

struct Run {
    int P = 0, G = 0;
    bool reuse = false;
    DevBuf x, xn, grad, grad_n, r, d, z, Hd, Hz, Heta, st, f_iter, f_citer, pm;
    std::unique_ptr<Plans> plans[2];       // [0] FP64, [1] FP32
    hipGraphExec_t step[2] = {nullptr, nullptr}, cont[2] = {nullptr, nullptr};
    DevTRState* host = nullptr;             // pinned mirror of the loop state
    ~Run()
    {
        for (int k = 0; k < 2; k++) {
            if (step[k]) (void)hipGraphExecDestroy(step[k]);
            if (cont[k]) (void)hipGraphExecDestroy(cont[k]);
        }
        std::lock_guard<std::recursive_mutex> g(capture_mutex());
        if (host) (void)hipHostFree(host);
    }
};

#define TR_HIP(call)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            set_error(std::string("device trust region: ") + #call + " failed: " + hipGetErrorString(e_)); \
            return -1;                                                                              \
        }                                                                                           \
    } while (0)

template <typename R>
hipError_t enqueue_coefs(const Run& run, const SweepPlan& pl, const DevBuf& kpos, const double* x, const double* v,
                         const int* skip, hipStream_t s)
{
    const int nb = (run.G + 63) / 64;
    hipLaunchKernelGGL(k_coefs<R>, dim3(nb), dim3(64), 0, s, x, v, (const int2*)run.pm.p, (const int*)kpos.p, run.G,
                       (cpx<R>*)pl.kcoef, pl.prog64, skip);
    return hipGetLastError();
}

hipError_t enqueue_coefs_any(const Run& run, const SweepPlan& pl, const DevBuf& kpos, const double* x, const double* v,
                             const int* skip, hipStream_t s)
{
    return pl.fp32 ? enqueue_coefs<float>(run, pl, kpos, x, v, skip, s) : enqueue_coefs<double>(run, pl, kpos, x, v, skip, s);
}

// one CG step: coefficients at (x, d), HVP sweep, H d and the step
hipError_t enqueue_cg(const Run& run, const Plans& pl, hipStream_t s)
{
    DevTRState* st = (DevTRState*)run.st.p;
    const int* done = &st->done;
    hipError_t e = enqueue_coefs_any(run, pl.p[2], pl.kpos[2], (const double*)run.x.p, (const double*)run.d.p, done, s);
    if (e == hipSuccess) e = launch_plan(pl.p[2], done, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_hvp_tcg, dim3(1), dim3(kBT), 0, s, (const double*)pl.p[2].out, (const double*)run.x.p,
                           run.P, (double*)run.Hd.p, (double*)run.r.p, (double*)run.d.p, (double*)run.z.p,
                           (double*)run.Hz.p, st);
        e = hipGetLastError();
    }
    return e;
}

// retraction, f at the trial point, H eta, ratio test; then the loop state to the host
hipError_t enqueue_tail(const Run& run, const Plans& pl, hipStream_t s)
{
    DevTRState* st = (DevTRState*)run.st.p;
    const int* skip = &st->skip_tail;
    hipLaunchKernelGGL(k_retract, dim3((run.P + 63) / 64), dim3(64), 0, s, (const double*)run.x.p, (const double*)run.z.p,
                       (double*)run.xn.p, run.P, skip);
    hipError_t e = hipGetLastError();
    if (run.reuse) {
        // f + gradient at the trial point; H eta from the CG's sum
        if (e == hipSuccess) e = enqueue_coefs_any(run, pl.p[1], pl.kpos[1], (const double*)run.xn.p, nullptr, skip, s);
        if (e == hipSuccess) e = launch_plan(pl.p[1], skip, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_trial_grad, dim3(1), dim3(kBT), 0, s, (const double*)pl.p[1].out, (const double*)run.xn.p,
                               run.P, (double*)run.grad_n.p, st);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_update_reuse, dim3(1), dim3(kBT), 0, s, (double*)run.x.p, (const double*)run.xn.p,
                               (const double*)run.z.p, (double*)run.grad.p, (const double*)run.grad_n.p,
                               (const double*)run.Hz.p, run.P, st, (double*)run.f_iter.p);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(run.host, run.st.p, sizeof(DevTRState), hipMemcpyDeviceToHost, s);
        return e;
    }
    if (e == hipSuccess) e = enqueue_coefs_any(run, pl.p[0], pl.kpos[0], (const double*)run.xn.p, nullptr, skip, s);
    if (e == hipSuccess) e = launch_plan(pl.p[0], skip, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_f_trial, dim3(1), dim3(64), 0, s, (const double*)pl.p[0].out, st, skip);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = enqueue_coefs_any(run, pl.p[2], pl.kpos[2], (const double*)run.x.p, (const double*)run.z.p, skip, s);
    if (e == hipSuccess) e = launch_plan(pl.p[2], skip, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_heta_update, dim3(1), dim3(kBT), 0, s, (const double*)pl.p[2].out, (double*)run.x.p,
                           (const double*)run.xn.p, (const double*)run.z.p, (const double*)run.grad.p,
                           (double*)run.Heta.p, run.P, st, (double*)run.f_iter.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(run.host, run.st.p, sizeof(DevTRState), hipMemcpyDeviceToHost, s);
    return e;
}

hipError_t enqueue_head(const Run& run, const Plans& pl, hipStream_t s)
{
    DevTRState* st = (DevTRState*)run.st.p;
    // with reuse the sweep runs only while nothing is kept (the first iteration, a precision switch)
    const int* skip = run.reuse ? &st->kept : nullptr;
    hipError_t e = enqueue_coefs_any(run, pl.p[1], pl.kpos[1], (const double*)run.x.p, nullptr, skip, s);
    if (e == hipSuccess) e = launch_plan(pl.p[1], skip, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_grad_init, dim3(1), dim3(kBT), 0, s, (const double*)pl.p[1].out, (const double*)run.x.p,
                           run.P, (double*)run.grad.p, (double*)run.r.p, (double*)run.d.p, (double*)run.z.p,
                           (double*)run.Hz.p, st, (double*)run.f_citer.p);
        e = hipGetLastError();
    }
    return e;
}

int capture(hipStream_t s, hipGraphExec_t* exec, const std::function<hipError_t()>& body)
{
    hipGraph_t g = nullptr;
    {
        std::lock_guard<std::recursive_mutex> lk(capture_mutex());
        TR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const hipError_t e = body();
        const hipError_t ee = hipStreamEndCapture(s, &g);
        if (e != hipSuccess || ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            set_error(std::string("device trust region: graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : ee));
            return -1;
        }
    }
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    TR_HIP(ei);
    return 0;
}

constexpr int kCgPerGraph = 2;  // truncated-CG steps in a step graph before a continuation graph

// plans, kernel positions and the two graphs of one precision
int prepare(Run& run, int pi, const Target& t, const Program& p, const zd* gates, hipStream_t s)
{
    auto pl = std::make_unique<Plans>();
    for (int m = 0; m < 3; m++) {
        if (plan_sweep(t, p, gates, m, pi ? kFP32 : kFP64, &pl->p[m]) < 0) return -1;
        if (pl->p[m].kind == 0) return 0;
        std::vector<int> kpos(p.G);
        for (int k = 0; k < p.G; k++) kpos[pl->p[m].korder[k]] = k;
        TR_HIP(pl->kpos[m].ensure((size_t)p.G * sizeof(int)));
        TR_HIP(hipMemcpyAsync(pl->kpos[m].p, kpos.data(), (size_t)p.G * sizeof(int), hipMemcpyHostToDevice, s));
    }
    TR_HIP(hipStreamSynchronize(s));
    // one launch of each plan outside the captures (kernel attributes; the skip flag below is set)
    DevTRState* st = (DevTRState*)run.st.p;
    for (int m = 0; m < 3; m++) TR_HIP(launch_plan(pl->p[m], &st->skip_tail, s));
    TR_HIP(hipStreamSynchronize(s));
    const Plans& P = *pl;
    if (capture(s, &run.step[pi], [&]() -> hipError_t {
            hipError_t e = enqueue_head(run, P, s);
            for (int k = 0; k < kCgPerGraph && e == hipSuccess; k++) e = enqueue_cg(run, P, s);
            if (e == hipSuccess) e = enqueue_tail(run, P, s);
            return e;
        }) < 0)
        return -1;
    if (capture(s, &run.cont[pi], [&]() -> hipError_t {
            hipError_t e = hipSuccess;
            for (int k = 0; k < kCgPerGraph && e == hipSuccess; k++) e = enqueue_cg(run, P, s);
            if (e == hipSuccess) e = enqueue_tail(run, P, s);
            return e;
        }) < 0)
        return -1;
    run.plans[pi] = std::move(pl);
    return 1;
}

}  // namespace

int optimize_device(const Target& t, const Program& p, const zd* gates_start, const rtr_params* prm, int niter,
                    int precision, int precision_switch, int reuse, double* f_iter, zd* f_citer, zd* gates_opt,
                    int* nhvp, DevTRStats* stats)
{
    if (prm->g_func || t.kind != 0 || niter < 1) return 0;
    const int P = p.P, n = 8 * P;
    const hipStream_t s = stream();
    Run run;
    run.P = P;
    run.G = p.G;
    run.reuse = reuse != 0;
    TR_HIP(run.x.ensure(16 * (size_t)P * sizeof(double)));
    TR_HIP(run.xn.ensure(16 * (size_t)P * sizeof(double)));
    for (DevBuf* b : {&run.grad, &run.grad_n, &run.r, &run.d, &run.z, &run.Hd, &run.Hz, &run.Heta})
        TR_HIP(b->ensure((size_t)n * sizeof(double)));
    TR_HIP(run.st.ensure(sizeof(DevTRState)));
    TR_HIP(run.f_iter.ensure((size_t)(niter + 1) * sizeof(double)));
    TR_HIP(run.f_citer.ensure((size_t)(niter + 1) * 2 * sizeof(double)));
    TR_HIP(run.pm.ensure((size_t)p.G * sizeof(int2)));
    {
        std::lock_guard<std::recursive_mutex> g(capture_mutex());
        TR_HIP(hipHostMalloc((void**)&run.host, sizeof(DevTRState), hipHostMallocDefault));
    }
    // per instance: its parameter gate and whether its wires are reversed
    std::vector<int2> pm(p.G);
    for (int g = 0; g < p.G; g++) pm[g] = {p.param[g], p.wires[2 * g + 1] < p.wires[2 * g] ? 1 : 0};
    DevTRState st0{};
    st0.radius = prm->radius_init;
    st0.maxradius = prm->maxradius;
    st0.rho_trust = prm->rho_trust;
    st0.abstol = prm->tcg_params.abstol;
    st0.reltol = prm->tcg_params.reltol;
    st0.maxiter = prm->tcg_params.maxiter;
    st0.skip_tail = 1;
    st0.reuse = run.reuse ? 1 : 0;
    TR_HIP(hipMemcpyAsync(run.pm.p, pm.data(), pm.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    TR_HIP(hipMemcpyAsync(run.x.p, gates_start, 16 * (size_t)P * sizeof(double), hipMemcpyHostToDevice, s));
    TR_HIP(hipMemcpyAsync(run.st.p, &st0, sizeof(DevTRState), hipMemcpyHostToDevice, s));
    TR_HIP(hipStreamSynchronize(s));
    const int sw = std::max(0, precision_switch);
    const bool need32 = precision == kFP32 && niter > sw, need64 = precision == kFP64 || sw > 0;
    for (int pi = 0; pi < 2; pi++) {
        if ((pi == 1 && !need32) || (pi == 0 && !need64)) continue;
        const int r = prepare(run, pi, t, p, gates_start, s);
        if (r <= 0) return r;
    }
    long launches = 0;
    for (int k = 0; k < niter; k++) {
        const int pi = (precision == kFP64 || k >= niter - sw) ? 0 : 1;
        // kept f and gradient of FP32 sweeps do not start the FP64 iterations
        if (run.reuse && k > 0 && pi == 0 && precision == kFP32 && k == niter - sw) {
            hipLaunchKernelGGL(k_drop_kept, dim3(1), dim3(64), 0, s, (DevTRState*)run.st.p);
            TR_HIP(hipGetLastError());
        }
        TR_HIP(hipGraphLaunch(run.step[pi], s));
        TR_HIP(hipStreamSynchronize(s));
        launches++;
        // a truncated CG longer than the step graph's steps: continuation graphs (each K more steps)
        for (int c = 0; !run.host->done && c <= prm->tcg_params.maxiter / kCgPerGraph + 1; c++) {
            TR_HIP(hipGraphLaunch(run.cont[pi], s));
            TR_HIP(hipStreamSynchronize(s));
            launches++;
        }
        if (!run.host->done || run.host->k != k + 1) {
            set_error("device trust region: the truncated CG did not finish");
            return -1;
        }
    }
    std::vector<double> fi(niter + 1), fc(2 * (niter + 1));
    TR_HIP(hipMemcpyAsync(gates_opt, run.x.p, 16 * (size_t)P * sizeof(double), hipMemcpyDeviceToHost, s));
    TR_HIP(hipMemcpyAsync(fi.data(), run.f_iter.p, niter * sizeof(double), hipMemcpyDeviceToHost, s));
    TR_HIP(hipMemcpyAsync(fc.data(), run.f_citer.p, 2 * niter * sizeof(double), hipMemcpyDeviceToHost, s));
    TR_HIP(hipStreamSynchronize(s));
    // the final f at the last iterate: the precision of the last iteration, and FP64 for f_citer
    EvalOut out;
    const int plast = (precision == kFP64 || sw >= 1) ? kFP64 : kFP32;
    if (evaluate(t, p, gates_opt, nullptr, kF, plast, &out) < 0) return -1;
    fi[niter] = out.f.real();
    if (f_citer) {
        if (plast != kFP64 && evaluate(t, p, gates_opt, nullptr, kF, kFP64, &out) < 0) return -1;
        fc[2 * niter] = out.f.real();
        fc[2 * niter + 1] = out.f.imag();
        for (int k = 0; k <= niter; k++) f_citer[k] = zd(fc[2 * k], fc[2 * k + 1]);
    }
    std::memcpy(f_iter, fi.data(), (niter + 1) * sizeof(double));
    if (nhvp) *nhvp = run.host->nhvp;
    if (stats) {
        stats->iterations = niter;
        stats->graph_launches = launches;
    }
    return 1;
}

}  // namespace mg
