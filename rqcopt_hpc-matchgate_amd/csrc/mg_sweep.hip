// Column-sweep kernels: the trace over basis columns b of
//   f = -Tr(U^dagger W) = sum_b  s_b . (W e_b),   s_b = -conj(U e_b)
// together with the holomorphic gradient dF/dV and the directional second
// derivative D^2 f[Z] (forward-over-reverse), accumulated in FP64.
//
// Reference: the per-column loop of src/matchgate_target.c:788-837 (seed :795-805,
// forward :807, dot :812-815, backward :821), the brickwall forward
// src/matchgate_brickwall.c:52-78 and the backward kernel src/gate.c:659-772.
// The reference caches every intermediate state per column; here a workgroup keeps
// one column's states on chip and recovers psi_g on the way back either by
// uncomputing (psi_g = V_g^{-1} psi_{g+1}) or from a checkpoint buffer.
//
// Work split: workgroup w handles columns b0 + w, b0 + w + gridDim.x, ...;
// thread t handles quadruples t, t + T, ... of each gate.  Per gate, each wave
// reduces its partial outer products with a transposing butterfly (32 values in
// 32 shuffles), stages them in LDS, and after the gate's barrier one wave adds the
// staged values, in wave order, into the workgroup's per-parameter accumulator.
// Workgroup partials are summed in block order by reduce_partials: results are
// bitwise reproducible run to run.
#include "mg_device.hpp"
#include "mg_fwd_mfma.hpp"
#include "mg_quad.hpp"
#include "mg_sweep.hpp"

namespace mg {

// ---- the sweep kernel ----------------------------------------------------------------
// MODE 0: f;  MODE 1: f + gradient;  MODE 2: f + gradient + D^2 f[Z]
// LDS:   states in LDS (else in a per-workgroup global scratch slice)
// CKPT:  forward states kept in a per-workgroup checkpoint buffer (no uncompute)
template <typename R, int MODE, bool LDS, bool CKPT, int T>
__global__ __launch_bounds__(T) void sweep_kernel(SweepArgs<R> a)
{
    if (a.skip && *a.skip) return;
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NW = T / 64;
    constexpr int NV = MODE == 2 ? 32 : 16;  // doubles reduced per gate (and kept per parameter)
    constexpr int SHIFT = 6 - log2c<NV>();
    const int L = a.L;
    const uint32_t S = 1u << L, Q = S >> 2;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    double* acc = reinterpret_cast<double*>(smem);    // [P][NV]
    double* stage = acc + NV * a.P;                   // [2][NW][NV]
    double* fst = stage + 2 * NW * NV;                // [NW][2]
    cpx<R>* st;
    if constexpr (LDS) {
        st = reinterpret_cast<cpx<R>*>(smem + a.lds_state_offset);
    } else {
        st = a.scratch + (size_t)blockIdx.x * a.nstate * S;
    }
    // state slots: lam, then psi (if not CKPT), then dlam, dpsi (MODE 2)
    cpx<R>* lam = st;
    cpx<R>* psi = CKPT ? nullptr : st + S;
    cpx<R>* dlam = st + (CKPT ? 1 : 2) * S;
    cpx<R>* dpsi = CKPT ? nullptr : st + 3 * S;
    cpx<R>* ck = nullptr;
    if constexpr (CKPT) ck = a.ckpt + (size_t)blockIdx.x * 2 * (a.G + 1) * S;  // [G+1][psi | dpsi]

    for (int k = tid; k < NV * a.P; k += T) acc[k] = 0.0;
    double fre = 0.0, fim = 0.0;
    __syncthreads();

    for (long b = a.b0 + blockIdx.x; b < a.b1; b += gridDim.x) {
        // ---- seed: lam = -conj(U e_b) --------------------------------------------------
        if (a.target_kind == 0) {
            const cpx<R>* row = a.seed + (size_t)b * S;
            for (uint32_t i = tid; i < S; i += T) sel<R, LDS>(lam, i) = row[i];
        } else {
            for (uint32_t i = tid; i < S; i += T) sel<R, LDS>(lam, i) = {R(i == (uint32_t)b), R(0)};
            __syncthreads();
            for (int g = 0; g < a.Gt; g++) {
                const Inst in = a.tinst[g];
                cpx<R> V[8];
                load_coef<R, 8>(a.tcoef + (size_t)g * kCoefPerInst, V);
                const uint32_t olo = swz_idx<R, LDS>(1u << in.lo), ohi = swz_idx<R, LDS>(1u << in.hi);
                for (uint32_t q = tid; q < Q; q += T) {
                    const uint32_t sb = swz_idx<R, LDS>(quad_base(q, in.lo, in.hi));
                    store_quad_o(lam, sb, olo, ohi, gapply(V, load_quad_o(lam, sb, olo, ohi)));
                }
                __syncthreads();
            }
            for (uint32_t i = tid; i < S; i += T) {
                cpx<R>& v = sel<R, LDS>(lam, i);
                v = {-v.re, v.im};
            }
        }
        // ---- forward sweep ---------------------------------------------------------------
        cpx<R>* p0 = CKPT ? ck : psi;
        cpx<R>* d0 = CKPT ? ck + S : dpsi;
        for (uint32_t i = tid; i < S; i += T) {
            sel<R, LDS && !CKPT>(p0, i) = {R(i == (uint32_t)b), R(0)};
            if constexpr (MODE == 2) sel<R, LDS && !CKPT>(d0, i) = {R(0), R(0)};
        }
        __syncthreads();
        if constexpr (sizeof(R) == 4 && LDS && !CKPT) {
            if (a.fwin) {
                if constexpr (MODE == 2)
                    fwd_windows_hvp<NW>(reinterpret_cast<const FWin*>(a.fwin), a.fwaf, 0, a.nwin, psi, dpsi, lane, wv);
                else
                    fwd_windows<NW>(reinterpret_cast<const FWin*>(a.fwin), a.fwaf, 0, a.nwin, psi, lane, wv);
                goto forward_done;
            }
        }
        for (int g = 0; g < a.G; g++) {
            const Inst in = a.inst[g];
            cpx<R> cf[24];
            load_coef<R, 8>(a.coef + (size_t)g * kCoefPerInst, cf);
            if constexpr (MODE == 2) load_coef<R, 8>(a.coef + (size_t)g * kCoefPerInst + 16, cf + 16);
            const cpx<R>* pin = CKPT ? ck + (size_t)2 * g * S : psi;
            cpx<R>* pout = CKPT ? ck + (size_t)2 * (g + 1) * S : psi;
            const cpx<R>* din = CKPT ? ck + (size_t)(2 * g + 1) * S : dpsi;
            cpx<R>* dout = CKPT ? ck + (size_t)(2 * g + 3) * S : dpsi;
            // one swizzle per quadruple: the wire bits' images are uniform
            const uint32_t olo = swz_idx<R, LDS && !CKPT>(1u << in.lo), ohi = swz_idx<R, LDS && !CKPT>(1u << in.hi);
            // with two states, two quadruples at a time (pin and pout are one array, or two
            // checkpoint slots); with one state the single-quadruple loop measured faster
            uint32_t q = tid;
            if constexpr (MODE == 2)
                for (; q + T < Q; q += 2 * T)
                    fwd_quads<R, MODE, LDS && !CKPT, 2>(q, T, in.lo, in.hi, olo, ohi, cf, pin, pout, din, dout);
            for (; q < Q; q += T)
                fwd_quads<R, MODE, LDS && !CKPT, 1>(q, T, in.lo, in.hi, olo, ohi, cf, pin, pout, din, dout);
            __syncthreads();
        }
    forward_done:
        // ---- f_b = s_b . psi_G -----------------------------------------------------------
        {
            const cpx<R>* pG = CKPT ? ck + (size_t)2 * a.G * S : psi;
            double sr = 0.0, si = 0.0;
            for (uint32_t i = tid; i < S; i += T) {
                const cpx<R> t = cmul(sel<R, LDS>(lam, i), sel<R, LDS && !CKPT>(pG, i));
                sr += (double)t.re;
                si += (double)t.im;
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
            // ---- reverse sweep ---------------------------------------------------------
            if constexpr (MODE == 2) {
                for (uint32_t i = tid; i < S; i += T) sel<R, LDS>(dlam, i) = {R(0), R(0)};
            }
            __syncthreads();
            for (int g = a.G - 1; g >= 0; g--) {
                const Inst in = a.inst[g];
                cpx<R> cf[24];
                load_coef<R, MODE == 2 ? 24 : 16>(a.coef + (size_t)g * kCoefPerInst, cf);
                cpx<R> ga[8], ha[8];
                if constexpr (CKPT) {
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        ga[k] = {R(0), R(0)};
                        ha[k] = {R(0), R(0)};
                    }
                }
                const uint32_t olo = swz_idx<R, LDS>(1u << in.lo), ohi = swz_idx<R, LDS>(1u << in.hi);
                if constexpr (!CKPT)
                    rev_gate<R, MODE, LDS, true>(tid, T, Q, in.lo, in.hi, olo, ohi, cf, lam, psi, dlam, dpsi, ga, ha);
                for (uint32_t q = tid; CKPT && q < Q; q += T) {
                    const uint32_t base = quad_base(q, in.lo, in.hi);
                    const uint32_t sb = swz_idx<R, LDS>(base);
                    const Quad<R> l = load_quad_o(lam, sb, olo, ohi);
                    Quad<R> p, dp, dl;
                    if constexpr (CKPT) {
                        p = load_quad(ck + (size_t)2 * g * S, base, in.lo, in.hi);
                        if constexpr (MODE == 2) dp = load_quad(ck + (size_t)(2 * g + 1) * S, base, in.lo, in.hi);
                    } else {
                        p = gapply(cf + 8, load_quad_o(psi, sb, olo, ohi));
                        store_quad_o(psi, sb, olo, ohi, p);
                        if constexpr (MODE == 2) {
                            const Quad<R> dn = load_quad_o(dpsi, sb, olo, ohi);
                            dp = gapply(cf + 8, gapply_sub(cf + 16, p, dn));
                            store_quad_o(dpsi, sb, olo, ohi, dp);
                        }
                    }
                    outer_acc(ga, l, p);
                    if constexpr (MODE == 2) {
                        dl = load_quad_o(dlam, sb, olo, ohi);
                        outer_acc(ha, dl, p);
                        outer_acc(ha, l, dp);
                        store_quad_o(dlam, sb, olo, ohi, gapplyT_add(cf, dl, gapplyT(cf + 16, l)));
                    }
                    store_quad_o(lam, sb, olo, ohi, gapplyT(cf, l));
                }
                // lane sums in the sweep precision, then FP64 across waves, columns and workgroups
                R v[NV];
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    v[2 * k] = ga[k].re;
                    v[2 * k + 1] = ga[k].im;
                    if constexpr (MODE == 2) {
                        v[16 + 2 * k] = ha[k].re;
                        v[16 + 2 * k + 1] = ha[k].im;
                    }
                }
                wave_halve<NV, 32, R>(v, lane);
                const int buf = g & 1;
                if ((lane & ((1 << SHIFT) - 1)) == 0) stage[(buf * NW + wv) * NV + (lane >> SHIFT)] = (double)v[0];
                __syncthreads();
                if (wv == g % NW && lane < NV) {
                    double s = 0.0;
                    for (int w = 0; w < NW; w++) s += stage[(buf * NW + w) * NV + lane];
                    const int ci = lane >> 1;
                    int e = ci & 7;
                    if (in.flip && e < 4) e = 3 - e;
                    acc[NV * in.param + 2 * ((ci & 8) | e) + (lane & 1)] += s;
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    double* out = a.partial + (size_t)blockIdx.x * (2 + 32 * a.P);
    if (tid == 0) {
        out[0] = fre;
        out[1] = fim;
    }
    for (int k = tid; k < 32 * a.P; k += T) out[2 + k] = (k & 31) < NV ? acc[(k >> 5) * NV + (k & 31)] : 0.0;
}

// out[k] = sum over workgroups w of partial[w][k], in a fixed order: thread row ty
// sums w = ty, ty + 16, ... and row 0 then adds the 16 row sums in order.
__global__ __launch_bounds__(1024) void reduce_partials(const double* __restrict__ partial, int nblk, int stride,
                                                        double* __restrict__ out, const int* skip)
{
    if (skip && *skip) return;
    __shared__ double rows[16][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + tx;
    double s = 0.0;
    if (k < stride)
        for (int w = ty; w < nblk; w += 16) s += partial[(size_t)w * stride + k];
    rows[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && k < stride) {
        double t = 0.0;
        for (int r = 0; r < 16; r++) t += rows[r][tx];
        out[k] = t;
    }
}

// ---- host-side launch ----------------------------------------------------------------
template <typename R, int MODE, bool LDS, bool CKPT, int T>
static hipError_t launch_t(const SweepArgs<R>& a, int grid, size_t lds, hipStream_t s)
{
    auto k = sweep_kernel<R, MODE, LDS, CKPT, T>;
    if (lds > 65536) {
        hipError_t e = ensure_lds_limit((const void*)k, lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(T), lds, s, a);
    return hipGetLastError();
}

template <typename R, int T>
static hipError_t launch_T(const SweepArgs<R>& a, int mode, bool lds_states, bool ckpt, int grid, size_t lds, hipStream_t s)
{
#define MG_CASE(M, LD, CK) \
    if (mode == M && lds_states == LD && ckpt == CK) return launch_t<R, M, LD, CK, T>(a, grid, lds, s);
    MG_CASE(0, true, false) MG_CASE(0, false, false)
    MG_CASE(1, true, false) MG_CASE(1, false, false) MG_CASE(1, true, true) MG_CASE(1, false, true)
    MG_CASE(2, true, false) MG_CASE(2, false, false) MG_CASE(2, true, true) MG_CASE(2, false, true)
#undef MG_CASE
    return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_sweep(const SweepArgs<R>& a, int mode, bool lds_states, bool ckpt, int threads, int grid, size_t lds,
                        hipStream_t s)
{
    if (threads == 256) return launch_T<R, 256>(a, mode, lds_states, ckpt, grid, lds, s);
    if (threads == 512) return launch_T<R, 512>(a, mode, lds_states, ckpt, grid, lds, s);
    if (threads == 1024) return launch_T<R, 1024>(a, mode, lds_states, ckpt, grid, lds, s);
    return hipErrorInvalidValue;
}

template hipError_t launch_sweep<double>(const SweepArgs<double>&, int, bool, bool, int, int, size_t, hipStream_t);
template hipError_t launch_sweep<float>(const SweepArgs<float>&, int, bool, bool, int, int, size_t, hipStream_t);

hipError_t launch_reduce(const double* partial, int nblk, int stride, double* out, hipStream_t s, const int* skip)
{
    hipLaunchKernelGGL(reduce_partials, dim3((stride + 63) / 64), dim3(1024), 0, s, partial, nblk, stride, out, skip);
    return hipGetLastError();
}

}  // namespace mg
