// Register-resident column sweep: the same f / gradient / D^2 f[Z] trace as
// sweep_kernel (mg_sweep.hip), for columns whose states do not fit the LDS
// (FP64 at L=12: 4 states x 64 KB; FP32 at L=13).
//
// A workgroup of T threads owns one column; thread t keeps, for every live state,
// the QPT quadruples q = t + k*T of the current gate in VGPRs (T * QPT * 4 = 2^L),
// i.e. the column's four states are spread over the CU's 512 KB register file.
// Between two gates whose wire pairs differ the states are re-laid out through LDS
// (write in the old gate's quadruple order, barrier, read in the new order), NST
// states per round.  The per-gate FP64 reductions and the accumulator layout are
// those of sweep_kernel.  The reverse sweep recovers psi_g with V^dagger, so the
// host selects this kernel only for gates unitary to 1e-12.
//
// Reference: per-column loop src/matchgate_target.c:788-837, forward
// src/matchgate_brickwall.c:52-78, backward kernel src/gate.c:659-772.
#include "mg_device.hpp"
#include "mg_quad.hpp"
#include "mg_sweep.hpp"

namespace mg {

template <typename R, int QPT>
struct RegState {
    Quad<R> q[QPT];
};

// store the QPT quadruples of one state at the (lo, hi) quadruple positions
// (SWZ: swizzled LDS layout, else plain global memory)
template <typename R, int QPT, int T, bool SWZ = true>
__device__ inline void put_state(cpx<R>* buf, const RegState<R, QPT>& s, int lo, int hi, int tid)
{
#pragma unroll
    for (int k = 0; k < QPT; k++) store_quad<R, SWZ>(buf, quad_base(tid + k * T, lo, hi), lo, hi, s.q[k]);
}

template <typename R, int QPT, int T, bool SWZ = true>
__device__ inline void get_state(const cpx<R>* buf, RegState<R, QPT>& s, int lo, int hi, int tid)
{
#pragma unroll
    for (int k = 0; k < QPT; k++) s.q[k] = load_quad<R, SWZ>(buf, quad_base(tid + k * T, lo, hi), lo, hi);
}

// Re-lay the first NS states out from quadruple order (lo, hi) to (lo2, hi2),
// NST states per LDS round.  Ends with a barrier after the last read.  NST is a
// compile-time constant so that every state register is overwritten unconditionally.
template <typename R, int QPT, int T, int NS, int NST>
__device__ inline void relayout(RegState<R, QPT> (&st)[4], cpx<R>* buf, uint32_t S, int lo, int hi, int lo2, int hi2,
                                int tid)
{
#pragma unroll
    for (int s0 = 0; s0 < NS; s0 += NST) {
#pragma unroll
        for (int s = s0; s < NS && s < s0 + NST; s++) put_state<R, QPT, T>(buf + (size_t)(s - s0) * S, st[s], lo, hi, tid);
        __syncthreads();
#pragma unroll
        for (int s = s0; s < NS && s < s0 + NST; s++) get_state<R, QPT, T>(buf + (size_t)(s - s0) * S, st[s], lo2, hi2, tid);
        __syncthreads();
    }
}

template <typename R>
__device__ inline Quad<R> basis_quad(uint32_t base, int lo, int hi, uint32_t b)
{
    const R one = R(1), zero = R(0);
    return {{base == b ? one : zero, zero},
            {(base | (1u << lo)) == b ? one : zero, zero},
            {(base | (1u << hi)) == b ? one : zero, zero},
            {(base | (1u << lo) | (1u << hi)) == b ? one : zero, zero}};
}

// MODE 0: f; 1: f + gradient; 2: f + gradient + D^2 f[Z].  State slots: 0 lam, 1 psi, 2 dlam, 3 dpsi.
template <typename R, int MODE, int T, int QPT, int NST>
__global__ __launch_bounds__(T) void sweep_reg_kernel(SweepArgs<R> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NW = T / 64;
    constexpr int NV = MODE == 2 ? 32 : 16;  // doubles staged per gate (gradient, then D^2 f[Z])
    constexpr int NS = MODE == 2 ? 4 : 2;
    const uint32_t S = 1u << a.L;
    const int tid0 = threadIdx.x, lane = tid0 & 63, wv = tid0 >> 6;

    double* acc = reinterpret_cast<double*>(smem);
    double* stage = acc + 32 * a.P;  // [2][NW][32]: gradient values at 0..15, D^2 f[Z] at 16..31
    double* fst = stage + 2 * NW * 32;
    cpx<R>* buf = reinterpret_cast<cpx<R>*>(smem + a.lds_state_offset);

    RegState<R, QPT> st[4];
    for (int k = tid0; k < 32 * a.P; k += T) acc[k] = 0.0;
    double fre = 0.0, fim = 0.0;
    __syncthreads();

    const Inst last = a.inst[a.G - 1];
    for (long b = a.b0 + blockIdx.x; b < a.b1; b += gridDim.x) {
        // an opaque copy of the thread index: keeps the compiler from hoisting per-thread
        // addresses out of the column loop, which would pin VGPRs for the whole kernel
        int tid = tid0;
        asm volatile("" : "+v"(tid));
        // ---- seed lam = -conj(U e_b), in the quadruple order of the last gate ----------
        if (a.target_kind == 0) {
            const cpx<R>* row = a.seed + (size_t)b * S;
            get_state<R, QPT, T, false>(row, st[0], last.lo, last.hi, tid);
        } else {
            Inst prev = a.tinst[0];
#pragma unroll
            for (int k = 0; k < QPT; k++)
                st[0].q[k] = basis_quad<R>(quad_base(tid + k * T, prev.lo, prev.hi), prev.lo, prev.hi, (uint32_t)b);
            for (int g = 0; g < a.Gt; g++) {
                const Inst in = a.tinst[g];
                if (in.lo != prev.lo || in.hi != prev.hi)
                    relayout<R, QPT, T, 1, NST>(st, buf, S, prev.lo, prev.hi, in.lo, in.hi, tid);
                cpx<R> V[8];
                load_coef<R, 8>(a.tcoef + (size_t)g * kCoefPerInst, V);
#pragma unroll
                for (int k = 0; k < QPT; k++) st[0].q[k] = gapply(V, st[0].q[k]);
                prev = in;
            }
            if (prev.lo != last.lo || prev.hi != last.hi)
                relayout<R, QPT, T, 1, NST>(st, buf, S, prev.lo, prev.hi, last.lo, last.hi, tid);
#pragma unroll
            for (int k = 0; k < QPT; k++) {
                Quad<R>& q = st[0].q[k];
                q.x = {-q.x.re, q.x.im};
                q.y = {-q.y.re, q.y.im};
                q.z = {-q.z.re, q.z.im};
                q.w = {-q.w.re, q.w.im};
            }
        }
        // ---- forward: psi (slot 1), dpsi (slot 3) ------------------------------------
        {
            const Inst i0 = a.inst[0];
#pragma unroll
            for (int k = 0; k < QPT; k++) {
                st[1].q[k] = basis_quad<R>(quad_base(tid + k * T, i0.lo, i0.hi), i0.lo, i0.hi, (uint32_t)b);
                if constexpr (MODE == 2) st[3].q[k] = Quad<R>{{0, 0}, {0, 0}, {0, 0}, {0, 0}};
            }
        }
        for (int g = 0; g < a.G; g++) {
            cpx<R> cf[24];
            load_coef<R, 8>(a.coef + (size_t)g * kCoefPerInst, cf);
            if constexpr (MODE == 2) load_coef<R, 8>(a.coef + (size_t)g * kCoefPerInst + 16, cf + 16);
#pragma unroll
            for (int k = 0; k < QPT; k++) {
                const Quad<R> p = st[1].q[k];
                if constexpr (MODE == 2) st[3].q[k] = gapply_add(cf, st[3].q[k], gapply(cf + 16, p));
                st[1].q[k] = gapply(cf, p);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (g + 1 < a.G) {
                const Inst in = a.inst[g], nx = a.inst[g + 1];
                if (in.lo != nx.lo || in.hi != nx.hi) {
                    // move psi (and dpsi) into slots 0.. of a temporary view: relayout works on
                    // the first NS slots, so shuffle slots to keep lam (slot 0) untouched
                    RegState<R, QPT> tmp[4];
                    tmp[0] = st[1];
                    if constexpr (MODE == 2) tmp[1] = st[3];
                    relayout<R, QPT, T, MODE == 2 ? 2 : 1, NST>(tmp, buf, S, in.lo, in.hi, nx.lo, nx.hi, tid);
                    st[1] = tmp[0];
                    if constexpr (MODE == 2) st[3] = tmp[1];
                }
            }
        }
        // ---- f_b = lam . psi_G (both in the last gate's order) ---------------------------
        {
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int k = 0; k < QPT; k++) {
                const Quad<R>& l = st[0].q[k];
                const Quad<R>& p = st[1].q[k];
                const cpx<R> t0 = cmul(l.x, p.x), t1 = cmul(l.y, p.y), t2 = cmul(l.z, p.z), t3 = cmul(l.w, p.w);
                sr += (double)t0.re + (double)t1.re + (double)t2.re + (double)t3.re;
                si += (double)t0.im + (double)t1.im + (double)t2.im + (double)t3.im;
            }
            sr = wave_sum(sr);
            si = wave_sum(si);
            if (lane == 0) {
                fst[2 * wv] = sr;
                fst[2 * wv + 1] = si;
            }
            __syncthreads();
            if (tid == 0) {
                double tr = 0.0, ti = 0.0;
                for (int w = 0; w < NW; w++) {
                    tr += fst[2 * w];
                    ti += fst[2 * w + 1];
                }
                fre += tr;
                fim += ti;
            }
        }
        if constexpr (MODE >= 1) {
            if constexpr (MODE == 2) {
#pragma unroll
                for (int k = 0; k < QPT; k++) st[2].q[k] = Quad<R>{{0, 0}, {0, 0}, {0, 0}, {0, 0}};
            }
            for (int g = a.G - 1; g >= 0; g--) {
                const Inst in = a.inst[g];
                cpx<R> cf[24];
                load_coef<R, 8>(a.coef + (size_t)g * kCoefPerInst, cf);
                if constexpr (MODE == 2) load_coef<R, 8>(a.coef + (size_t)g * kCoefPerInst + 16, cf + 16);
                const int bufsel = g & 1;
                // recover psi_g (and dpsi_g) with V^dagger = V^{-1}, then the gradient outer products
                {
                    cpx<R> ga[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) ga[e] = {R(0), R(0)};
#pragma unroll
                    for (int k = 0; k < QPT; k++) {
                        const Quad<R> p = gapplyH(cf, st[1].q[k]);
                        st[1].q[k] = p;
                        if constexpr (MODE == 2) st[3].q[k] = gapplyH(cf, gapply_sub(cf + 16, p, st[3].q[k]));
                        outer_acc(ga, st[0].q[k], p);
                        // keep the quadruples' temporaries from overlapping (VGPR pressure)
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    R v[16];
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        v[2 * e] = ga[e].re;
                        v[2 * e + 1] = ga[e].im;
                    }
                    wave_halve<16, 32, R>(v, lane);
                    if ((lane & 3) == 0) stage[(bufsel * NW + wv) * 32 + (lane >> 2)] = (double)v[0];
                }
                if constexpr (MODE == 2) {
                    cpx<R> ha[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) ha[e] = {R(0), R(0)};
#pragma unroll
                    for (int k = 0; k < QPT; k++) {
                        outer_acc(ha, st[2].q[k], st[1].q[k]);
                        outer_acc(ha, st[0].q[k], st[3].q[k]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    R v[16];
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        v[2 * e] = ha[e].re;
                        v[2 * e + 1] = ha[e].im;
                    }
                    wave_halve<16, 32, R>(v, lane);
                    if ((lane & 3) == 0) stage[(bufsel * NW + wv) * 32 + 16 + (lane >> 2)] = (double)v[0];
                }
                // propagate lam (and dlam)
#pragma unroll
                for (int k = 0; k < QPT; k++) {
                    const Quad<R> l = st[0].q[k];
                    if constexpr (MODE == 2) st[2].q[k] = gapplyT_add(cf, st[2].q[k], gapplyT(cf + 16, l));
                    st[0].q[k] = gapplyT(cf, l);
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (g > 0 && (a.inst[g - 1].lo != in.lo || a.inst[g - 1].hi != in.hi)) {
                    const Inst nx = a.inst[g - 1];
                    relayout<R, QPT, T, NS, NST>(st, buf, S, in.lo, in.hi, nx.lo, nx.hi, tid);
                } else {
                    __syncthreads();
                }
                if (wv == g % NW && lane < NV) {
                    double s = 0.0;
                    for (int w = 0; w < NW; w++) s += stage[(bufsel * NW + w) * 32 + lane];
                    const int ci = lane >> 1;
                    int e = ci & 7;
                    if (in.flip && e < 4) e = 3 - e;
                    acc[32 * in.param + 2 * ((ci & 8) | e) + (lane & 1)] += s;
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    double* out = a.partial + (size_t)blockIdx.x * (2 + 32 * a.P);
    if (tid0 == 0) {
        out[0] = fre;
        out[1] = fim;
    }
    for (int k = tid0; k < 32 * a.P; k += T) out[2 + k] = acc[k];
}

template <typename R, int MODE, int T, int QPT, int NST>
static hipError_t launch_reg_t(const SweepArgs<R>& a, int grid, size_t lds, hipStream_t s)
{
    auto k = sweep_reg_kernel<R, MODE, T, QPT, NST>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(T), lds, s, a);
    return hipGetLastError();
}

template <typename R, int T, int QPT, int NST>
static hipError_t launch_reg_m(const SweepArgs<R>& a, int mode, int grid, size_t lds, hipStream_t s)
{
    if (mode == 0) return launch_reg_t<R, 0, T, QPT, NST>(a, grid, lds, s);
    if (mode == 1) return launch_reg_t<R, 1, T, QPT, NST>(a, grid, lds, s);
    if (mode == 2) return launch_reg_t<R, 2, T, QPT, NST>(a, grid, lds, s);
    return hipErrorInvalidValue;
}

// a.nstate = states per LDS re-layout round (1, 2 or 4)
template <typename R>
hipError_t launch_sweep_reg(const SweepArgs<R>& a, int mode, int threads, int qpt, int grid, size_t lds, hipStream_t s)
{
#define MG_REG(TT, QQ)                                                                   \
    if (threads == TT && qpt == QQ) {                                                    \
        if (a.nstate >= 4) return launch_reg_m<R, TT, QQ, 4>(a, mode, grid, lds, s);     \
        if (a.nstate >= 2) return launch_reg_m<R, TT, QQ, 2>(a, mode, grid, lds, s);     \
        return launch_reg_m<R, TT, QQ, 1>(a, mode, grid, lds, s);                        \
    }
    MG_REG(512, 1) MG_REG(512, 2) MG_REG(1024, 1) MG_REG(1024, 2)
#undef MG_REG
    return hipErrorInvalidValue;
}

template hipError_t launch_sweep_reg<double>(const SweepArgs<double>&, int, int, int, int, size_t, hipStream_t);
template hipError_t launch_sweep_reg<float>(const SweepArgs<float>&, int, int, int, int, size_t, hipStream_t);

}  // namespace mg
