// Device helpers of the gate-window sweep kernels (mg_batch.hip, mg_sweep_win.hip):
// scalar reads of read-only descriptors, window tuples held in registers, and the
// per-gate forward and reverse operations on a tuple's quadruples.
#pragma once
#include "mg_batch.hpp"
#include "mg_quad.hpp"
#include "mg_window.hpp"

namespace mg {

// Read-only descriptor data (gate instances, coefficients, window blocks) is read
// through the constant address space: with a wave-uniform address the loads become
// scalar loads into SGPRs, which the gate kernels then use as scalar operands.
#define MG_CONST __attribute__((address_space(4)))
template <typename T>
__device__ inline const MG_CONST T* cptr(const T* p)
{
    return (const MG_CONST T*)p;
}

template <typename R, int N>
__device__ inline void load_coef_c(const cpx<R>* src, cpx<R>* dst)
{
    const MG_CONST R* p = cptr(reinterpret_cast<const R*>(src));
#pragma unroll
    for (int k = 0; k < N; k++) dst[k] = cpx<R>{p[2 * k], p[2 * k + 1]};
}

// a window block's positions and amplitude offsets, in scalar registers
template <int W>
struct BlockRegs {
    int g0, g1;
    int pos[W];
    uint32_t off[1 << W];
};

template <int W>
__device__ inline BlockRegs<W> load_block(const WBlock* b)
{
    const MG_CONST WBlock* p = cptr(b);
    BlockRegs<W> r;
    r.g0 = p->g0;
    r.g1 = p->g1;
#pragma unroll
    for (int i = 0; i < W; i++) r.pos[i] = p->pos[i];
#pragma unroll
    for (int j = 0; j < (1 << W); j++) r.off[j] = p->off[j];
    return r;
}

template <typename R, bool SWZ>
__device__ inline uint32_t lin(uint32_t i)
{
    if constexpr (SWZ) return win_swz<batch_bank_bits<R>()>(i);
    else return i;
}

// tuple u -> address with zero bits at the W window positions (ascending)
template <int W>
__device__ inline uint32_t deposit(uint32_t u, const int* pos)
{
#pragma unroll
    for (int i = 0; i < W; i++) {
        const int p = pos[i];
        u = ((u >> p) << (p + 1)) | (u & ((1u << p) - 1));
    }
    return u;
}

// compile-time quadruples of a window-local gate: local bits LA > LB hold the gate's
// bit positions hi > lo; the quadruple order is that of load_quad (00, lo set, hi set, both)
template <int W, int LA, int LB>
struct WQuads {
    static constexpr int N = 1 << (W - 2);
    static constexpr int idx(int r)
    {
        int x = ((r >> LB) << (LB + 1)) | (r & ((1 << LB) - 1));
        return ((x >> LA) << (LA + 1)) | (x & ((1 << LA) - 1));
    }
};

template <typename R, int W>
struct Win {
    cpx<R> a[1 << W];
};

template <typename R, int W, int LA, int LB>
__device__ inline Quad<R> wq_get(const Win<R, W>& s, int r)
{
    const int x = WQuads<W, LA, LB>::idx(r);
    return {s.a[x], s.a[x | (1 << LB)], s.a[x | (1 << LA)], s.a[x | (1 << LA) | (1 << LB)]};
}

template <typename R, int W, int LA, int LB>
__device__ inline void wq_put(Win<R, W>& s, int r, const Quad<R>& q)
{
    const int x = WQuads<W, LA, LB>::idx(r);
    s.a[x] = q.x;
    s.a[x | (1 << LB)] = q.y;
    s.a[x | (1 << LA)] = q.z;
    s.a[x | (1 << LA) | (1 << LB)] = q.w;
}

// ---- per-gate operations on a window tuple ----------------------------------------------
// forward: P <- V P  (and D <- V D + Z P)
template <typename R, int W, int LA, int LB, int MODE>
__device__ inline void wg_forward(const cpx<R>* cf, Win<R, W>& P, Win<R, W>& D)
{
#pragma unroll
    for (int r = 0; r < WQuads<W, LA, LB>::N; r++) {
        const Quad<R> p = wq_get<R, W, LA, LB>(P, r);
        if constexpr (MODE == 2) wq_put<R, W, LA, LB>(D, r, gapply_add(cf, wq_get<R, W, LA, LB>(D, r), gapply(cf + 16, p)));
        wq_put<R, W, LA, LB>(P, r, gapply(cf, p));
    }
}

// reverse: psi_g = V^{-1} psi_{g+1} (and dpsi), outer products, lam <- V^T lam (and dlam)
template <typename R, int W, int LA, int LB, int MODE>
__device__ inline void wg_backward(const cpx<R>* cf, Win<R, W>& Lm, Win<R, W>& P, Win<R, W>& DL, Win<R, W>& D,
                                   cpx<R>* ga, cpx<R>* ha)
{
#pragma unroll
    for (int r = 0; r < WQuads<W, LA, LB>::N; r++) {
        const Quad<R> l = wq_get<R, W, LA, LB>(Lm, r);
        const Quad<R> p = gapply(cf + 8, wq_get<R, W, LA, LB>(P, r));
        wq_put<R, W, LA, LB>(P, r, p);
        outer_acc(ga, l, p);
        if constexpr (MODE == 2) {
            const Quad<R> dp = gapply(cf + 8, gapply_sub(cf + 16, p, wq_get<R, W, LA, LB>(D, r)));
            wq_put<R, W, LA, LB>(D, r, dp);
            const Quad<R> dl = wq_get<R, W, LA, LB>(DL, r);
            outer_acc(ha, dl, p);
            outer_acc(ha, l, dp);
            wq_put<R, W, LA, LB>(DL, r, gapplyT_add(cf, dl, gapplyT(cf + 16, l)));
        }
        wq_put<R, W, LA, LB>(Lm, r, gapplyT(cf, l));
    }
}

// dispatch on the window-local pair index (la, lb): (1,0) (2,0) (2,1) (3,0) (3,1) (3,2).
// The host only emits pairs inside the window, so the last pair of a W-bit window is the
// default: every path applies the gate, and no path passes the states through unchanged
// (a pass-through path makes the compiler copy every state register before the switch).
#define MG_WCASES(W, CALL)                                                  \
    if constexpr (W == 2) {                                                 \
        CALL(W, 1, 0);                                                      \
    } else if constexpr (W == 3) {                                          \
        switch (wcase) {                                                    \
            case 0: CALL(W, 1, 0); break;                                   \
            case 1: CALL(W, 2, 0); break;                                   \
            default: CALL(W, 2, 1); break;                                  \
        }                                                                   \
    } else {                                                                \
        switch (wcase) {                                                    \
            case 0: CALL(W, 1, 0); break;                                   \
            case 1: CALL(W, 2, 0); break;                                   \
            case 2: CALL(W, 2, 1); break;                                   \
            case 3: CALL(W, 3, 0); break;                                   \
            case 4: CALL(W, 3, 1); break;                                   \
            default: CALL(W, 3, 2); break;                                  \
        }                                                                   \
    }

template <typename R, int W, bool SWZ>
__device__ inline void win_load(const cpx<R>* st, uint32_t sb, const uint32_t* off, Win<R, W>& s)
{
#pragma unroll
    for (int j = 0; j < (1 << W); j++) s.a[j] = st[sb ^ off[j]];
}

template <typename R, int W, bool SWZ>
__device__ inline void win_store(cpx<R>* st, uint32_t sb, const uint32_t* off, const Win<R, W>& s)
{
#pragma unroll
    for (int j = 0; j < (1 << W); j++) st[sb ^ off[j]] = s.a[j];
}

}  // namespace mg
