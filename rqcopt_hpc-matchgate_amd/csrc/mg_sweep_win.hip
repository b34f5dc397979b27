// Register-resident column sweep with gate windows at L=12 (FP64, whose states,
// 4 x 64 KB, do not fit the LDS, and FP32).
//
// The program's gates are regrouped into blocks acting inside W-qubit windows
// (mg_window.cpp).  Each thread keeps, for every live state, the 2^W amplitudes of one
// window tuple in VGPRs (FP64: 512 threads x 8, FP32: 256 x 16 = 2^12); a block's gates are
// applied there, and only between blocks are the states re-laid out through LDS
// (write in the block's tuple order, barrier, read in the next block's order),
// two FP64 (one when the accumulator of many parameter gates leaves no room for two) or four FP32
// states per round.  Compared with a re-layout per gate this halves the LDS
// traffic and the barriers of the brickwall sweeps.
//
// Reverse sweep per gate, in three passes over the tuple's quadruples so that one
// accumulator set is live at a time: (A) psi_g = V^dagger psi_{g+1} (and dpsi) with
// the gradient outer products, (B) the D^2 f[Z] outer products, (C) lam <- V^T lam
// (and dlam).  Each set is reduced by the transposing wave butterfly and staged per
// gate; after the block's first re-layout barrier one wave per gate adds the staged
// sums in wave order into the accumulator (as sweep_kernel, mg_sweep.hip).  The host
// selects this kernel for gates unitary to 1e-12 (it uncomputes with V^dagger).
//
// Reference: per-column loop src/matchgate_target.c:788-837, forward
// src/matchgate_brickwall.c:52-78, backward kernel src/gate.c:659-772.
#include "mg_fwd_mfma.hpp"
#include "mg_sweep_win.hpp"
#include "mg_win.hpp"

namespace mg {

namespace {

// FP64: 512 threads x 3-qubit windows (8 amplitudes per state); FP32: 256 threads x 4-qubit
// windows (16 amplitudes per state, one relayout per ~3.6 brickwall gates)
template <typename R>
struct WinShape;
template <>
struct WinShape<double> {
    static constexpr int T = 512, W = 3;
};
template <>
struct WinShape<float> {
    static constexpr int T = 256, W = 4;
};

// (A) uncompute psi (and dpsi), gradient outer products
template <typename R, int kW, int LA, int LB, int MODE>
__device__ inline void pass_a(const cpx<R>* V, const cpx<R>* Z, const Win<R, kW>& Lm, Win<R, kW>& P, Win<R, kW>& D,
                              cpx<R>* ga)
{
#pragma unroll
    for (int r = 0; r < WQuads<kW, LA, LB>::N; r++) {
        const Quad<R> p = gapplyH(V, wq_get<R, kW, LA, LB>(P, r));
        wq_put<R, kW, LA, LB>(P, r, p);
        if constexpr (MODE == 2)
            wq_put<R, kW, LA, LB>(D, r, gapplyH(V, gapply_sub(Z, p, wq_get<R, kW, LA, LB>(D, r))));
        // the first quadruple starts the sums
        if (r == 0) outer_acc<R, true>(ga, wq_get<R, kW, LA, LB>(Lm, r), p);
        else outer_acc(ga, wq_get<R, kW, LA, LB>(Lm, r), p);
        __builtin_amdgcn_sched_barrier(0);  // one quadruple's temporaries at a time
    }
}

// (B) D^2 f[Z] outer products
template <typename R, int kW, int LA, int LB>
__device__ inline void pass_b(const Win<R, kW>& Lm, const Win<R, kW>& P, const Win<R, kW>& DL, const Win<R, kW>& D,
                              cpx<R>* ha)
{
#pragma unroll
    for (int r = 0; r < WQuads<kW, LA, LB>::N; r++) {
        if (r == 0) outer_acc<R, true>(ha, wq_get<R, kW, LA, LB>(DL, r), wq_get<R, kW, LA, LB>(P, r));
        else outer_acc(ha, wq_get<R, kW, LA, LB>(DL, r), wq_get<R, kW, LA, LB>(P, r));
        outer_acc(ha, wq_get<R, kW, LA, LB>(Lm, r), wq_get<R, kW, LA, LB>(D, r));
    }
}

// (C) lam <- V^T lam, dlam <- V^T dlam + Z^T lam
template <typename R, int kW, int LA, int LB, int MODE>
__device__ inline void pass_c(const cpx<R>* V, const cpx<R>* Z, Win<R, kW>& Lm, Win<R, kW>& DL)
{
#pragma unroll
    for (int r = 0; r < WQuads<kW, LA, LB>::N; r++) {
        const Quad<R> l = wq_get<R, kW, LA, LB>(Lm, r);
        if constexpr (MODE == 2)
            wq_put<R, kW, LA, LB>(DL, r, gapplyT_add(V, wq_get<R, kW, LA, LB>(DL, r), gapplyT(Z, l)));
        wq_put<R, kW, LA, LB>(Lm, r, gapplyT(V, l));
        __builtin_amdgcn_sched_barrier(0);
    }
}

// reduce 8 complex per-thread sums over the wave and stage them in FP64
template <typename R>
__device__ inline void stage8(const cpx<R>* s, double* slot, int lane)
{
    R v[16];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        v[2 * e] = s[e].re;
        v[2 * e + 1] = s[e].im;
    }
    wave_halve<16, 32, R>(v, lane);
    if ((lane & 3) == 0) slot[lane >> 2] = (double)v[0];
}

// after the barrier that follows a window block's reverse sweep, the wave that owns a gate's
// parameter adds the gate's staged sums: gates of one block that share a parameter (tied layers)
// are added by one wave, in order
template <int kNW, int NV>
__device__ inline void win_rev_adds(const WinSweepArgs& a, int g0, int g1, const double* stage, double* acc, int wv,
                                    int lane)
{
    for (int g = g0; g < g1; g++) {
        const int gparam = cptr(a.inst)[g].param, gflip = cptr(a.inst)[g].flip;
        if (wv == gparam % kNW && lane < NV) {
            const double* sg = stage + (size_t)(g - g0) * kNW * 32;
            double s = 0.0;
            for (int w = 0; w < kNW; w++) s += sg[w * 32 + lane];
            const int ci = lane >> 1;
            int e = ci & 7;
            if (gflip && e < 4) e = 3 - e;
            acc[NV * gparam + 2 * ((ci & 8) | e) + (lane & 1)] += s;
        }
    }
}

}  // namespace

template <typename R, int MODE, int RS>
__global__ __launch_bounds__(WinShape<R>::T) void sweep_win_kernel(WinSweepArgs a)
{
    if (a.skip && *a.skip) return;
    constexpr int kT = WinShape<R>::T, kW = WinShape<R>::W, kNW = kT / 64;
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NV = MODE == 2 ? 32 : 16;
    // RS states per re-layout round (the buffer holds RS states): FP32 4, FP64 2 or 1.  With two
    // per round the reverse sweep stores psi for the re-layout while the block's last lam update
    // runs: with the HVP psi and dpsi form the first round and lam and dlam the second, with the
    // gradient psi and lam share the round
    constexpr bool kEarly = RS == 2 && MODE >= 1;
    const uint32_t S = 1u << a.L;
    const int tid0 = threadIdx.x, lane = tid0 & 63, wv = tid0 >> 6;

    double* stage = reinterpret_cast<double*>(smem);  // [a.slots][kNW][32]
    double* fst = stage + a.slots * kNW * 32;         // [kNW][2]
    double* acc = reinterpret_cast<double*>(smem + a.acc_off);
    cpx<R>* buf = reinterpret_cast<cpx<R>*>(smem + a.buf_off);  // [2 or 4][S] re-layout buffer

    for (int k = tid0; k < NV * a.P; k += kT) acc[k] = 0.0;
    double fre = 0.0, fim = 0.0;
    __syncthreads();

    const int nb = a.nblocks;
    const cpx<R>* coef = reinterpret_cast<const cpx<R>*>(a.coef);
    for (long b = a.b0 + blockIdx.x; b < a.b1; b += gridDim.x) {
        // an opaque copy of the thread index: keeps per-thread addresses from being
        // hoisted out of the column loop (and pinned in VGPRs for the whole kernel)
        int tid = tid0;
        asm volatile("" : "+v"(tid));
        Win<R, kW> Lm, P, DL, D;
        // ---- forward: psi, dpsi ------------------------------------------------------------
        if constexpr (sizeof(R) == 8 && (MODE < 2 || RS >= 2)) {
            if (a.fwin) {
                // on the matrix cores, the column (and its derivative) in the re-layout buffer, then
                // into registers in the last window block's tuple order
                cpx<R>* cpsi = buf;
                cpx<R>* cdpsi = buf + S;
                for (uint32_t j = (uint32_t)tid; j < S; j += kT) {
                    cpsi[mat_swz(j)] = {j == (uint32_t)b ? R(1) : R(0), R(0)};
                    if constexpr (MODE == 2) cdpsi[mat_swz(j)] = {R(0), R(0)};
                }
                __syncthreads();
                fwd_windows64<kNW, MODE == 2>(reinterpret_cast<const DWin*>(a.fwin), a.fwaf, a.nfwin, cpsi, cdpsi, lane,
                                              wv);
                const BlockRegs<kW> bl = load_block<kW>(a.blk_glob + nb - 1);
                const uint32_t base = deposit<kW>((uint32_t)tid, bl.pos);
#pragma unroll
                for (int j = 0; j < (1 << kW); j++) {
                    const uint32_t o = mat_swz(base | bl.off[j]);
                    P.a[j] = cpsi[o];
                    if constexpr (MODE == 2) D.a[j] = cdpsi[o];
                }
                // the reverse sweep's re-layouts reuse the buffer
                __syncthreads();
                goto forward_done;
            }
        }
        {
            const BlockRegs<kW> b0 = load_block<kW>(a.blk_glob);
            const uint32_t base = deposit<kW>((uint32_t)tid, b0.pos);
#pragma unroll
            for (int j = 0; j < (1 << kW); j++) {
                P.a[j] = {(base | b0.off[j]) == (uint32_t)b ? R(1) : R(0), R(0)};
                D.a[j] = {R(0), R(0)};
            }
        }
        for (int k = 0; k < nb; k++) {
            const BlockRegs<kW> bk = load_block<kW>(a.blk_lds + k);
            for (int g = bk.g0; g < bk.g1; g++) {
                const int wcase = cptr(a.inst)[g].lo;
                cpx<R> cf[24];
                load_coef_c<R, 8>(coef + (size_t)g * kCoefPerInst, cf);
                if constexpr (MODE == 2) load_coef_c<R, 8>(coef + (size_t)g * kCoefPerInst + 16, cf + 16);
#define MG_FWD(WW, A, B) wg_forward<R, WW, A, B, MODE == 2 ? 2 : 0>(cf, P, D)
                MG_WCASES(kW, MG_FWD)
#undef MG_FWD
            }
            if (k + 1 < nb) {
                const BlockRegs<kW> nx = load_block<kW>(a.blk_lds + k + 1);
                const uint32_t sb = lin<R, true>(deposit<kW>((uint32_t)tid, bk.pos));
                const uint32_t sn = lin<R, true>(deposit<kW>((uint32_t)tid, nx.pos));
                win_store<R, kW, true>(buf, sb, bk.off, P);
                if constexpr (MODE == 2 && RS >= 2) win_store<R, kW, true>(buf + S, sb, bk.off, D);
                __syncthreads();
                win_load<R, kW, true>(buf, sn, nx.off, P);
                if constexpr (MODE == 2 && RS >= 2) win_load<R, kW, true>(buf + S, sn, nx.off, D);
                __syncthreads();
                if constexpr (MODE == 2 && RS == 1) {
                    win_store<R, kW, true>(buf, sb, bk.off, D);
                    __syncthreads();
                    win_load<R, kW, true>(buf, sn, nx.off, D);
                    __syncthreads();
                }
            }
        }
    forward_done:
        // ---- seed lam = -conj(U e_b) in the last block's tuple order (after the forward sweep,
        //      which does not need it) ---------------------------------------------------------
        {
            const BlockRegs<kW> bl = load_block<kW>(a.blk_glob + nb - 1);
            const uint32_t base = deposit<kW>((uint32_t)tid, bl.pos);
            const cpx<R>* row = reinterpret_cast<const cpx<R>*>(a.seed) + (size_t)b * S;
#pragma unroll
            for (int j = 0; j < (1 << kW); j++) Lm.a[j] = row[base | bl.off[j]];
        }
        // ---- f_b = lam . psi_G (both in the last block's order) ------------------------------
        {
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int j = 0; j < (1 << kW); j++) {
                const cpx<R> t = cmul(Lm.a[j], P.a[j]);
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
                for (int w = 0; w < kNW; w++) {
                    tr += fst[2 * w];
                    ti += fst[2 * w + 1];
                }
                fre += tr;
                fim += ti;
            }
        }
        if constexpr (MODE >= 1) {
#pragma unroll
            for (int j = 0; j < (1 << kW); j++) DL.a[j] = {R(0), R(0)};
            for (int k = nb - 1; k >= 0; k--) {
                const BlockRegs<kW> bk = load_block<kW>(a.blk_lds + k);
                const uint32_t sb = lin<R, true>(deposit<kW>((uint32_t)tid, bk.pos));
                for (int g = bk.g1 - 1; g >= bk.g0; g--) {
                    const int wcase = cptr(a.inst)[g].lo;
                    double* slot = stage + ((g - bk.g0) * kNW + wv) * 32;
                    cpx<R> V[8], Z[8];
                    load_coef_c<R, 8>(coef + (size_t)g * kCoefPerInst, V);
                    if constexpr (MODE == 2) load_coef_c<R, 8>(coef + (size_t)g * kCoefPerInst + 16, Z);
                    {
                        cpx<R> ga[8];  // set by the gate's first quadruple
#define MG_PA(WW, A, B) pass_a<R, WW, A, B, MODE>(V, Z, Lm, P, D, ga)
                        MG_WCASES(kW, MG_PA)
#undef MG_PA
                        stage8(ga, slot, lane);
                    }
                    if constexpr (MODE == 2) {
                        cpx<R> ha[8];
#define MG_PB(WW, A, B) pass_b<R, WW, A, B>(Lm, P, DL, D, ha)
                        MG_WCASES(kW, MG_PB)
#undef MG_PB
                        stage8(ha, slot + 16, lane);
                    }
                    if constexpr (kEarly) {
                        // psi (and dpsi) are final for this block once the block's first gate has
                        // uncomputed them: their re-layout stores overlap the last lam update
                        if (g == bk.g0 && k > 0) {
                            if constexpr (MODE == 2) {
                                win_store<R, kW, true>(buf, sb, bk.off, P);
                                win_store<R, kW, true>(buf + S, sb, bk.off, D);
                            } else {
                                win_store<R, kW, true>(buf + S, sb, bk.off, P);
                            }
                        }
                    }
#define MG_PC(WW, A, B) pass_c<R, WW, A, B, MODE>(V, Z, Lm, DL)
                    MG_WCASES(kW, MG_PC)
#undef MG_PC
                }
                // re-layout to the previous block's order (RS states per round); the staged
                // sums are consumed after the first barrier
                const bool move = k > 0;
                BlockRegs<kW> pv = bk;
                if (move) pv = load_block<kW>(a.blk_lds + k - 1);
                const uint32_t sn = lin<R, true>(deposit<kW>((uint32_t)tid, pv.pos));
                if constexpr (kEarly && MODE == 2) {
                    __syncthreads();
                    win_rev_adds<kNW, NV>(a, bk.g0, bk.g1, stage, acc, wv, lane);
                    if (move) {
                        win_load<R, kW, true>(buf, sn, pv.off, P);
                        if constexpr (MODE == 2) win_load<R, kW, true>(buf + S, sn, pv.off, D);
                        __syncthreads();
                        win_store<R, kW, true>(buf, sb, bk.off, Lm);
                        if constexpr (MODE == 2) win_store<R, kW, true>(buf + S, sb, bk.off, DL);
                        __syncthreads();
                        win_load<R, kW, true>(buf, sn, pv.off, Lm);
                        if constexpr (MODE == 2) win_load<R, kW, true>(buf + S, sn, pv.off, DL);
                    }
                    __syncthreads();
                    continue;
                }
                if (move) {
                    win_store<R, kW, true>(buf, sb, bk.off, Lm);
                    if constexpr (RS >= 2 && !kEarly) win_store<R, kW, true>(buf + S, sb, bk.off, P);
                    if constexpr (MODE == 2 && RS >= 4) {
                        win_store<R, kW, true>(buf + 2 * S, sb, bk.off, DL);
                        win_store<R, kW, true>(buf + 3 * S, sb, bk.off, D);
                    }
                }
                __syncthreads();
                win_rev_adds<kNW, NV>(a, bk.g0, bk.g1, stage, acc, wv, lane);
                if (move) {
                    win_load<R, kW, true>(buf, sn, pv.off, Lm);
                    if constexpr (RS >= 2) win_load<R, kW, true>(buf + S, sn, pv.off, P);
                    if constexpr (MODE == 2 && RS >= 4) {
                        win_load<R, kW, true>(buf + 2 * S, sn, pv.off, DL);
                        win_load<R, kW, true>(buf + 3 * S, sn, pv.off, D);
                    }
                    __syncthreads();
                    if constexpr (RS == 1) {
                        win_store<R, kW, true>(buf, sb, bk.off, P);
                        __syncthreads();
                        win_load<R, kW, true>(buf, sn, pv.off, P);
                        __syncthreads();
                    }
                    if constexpr (MODE == 2 && RS == 2) {
                        win_store<R, kW, true>(buf, sb, bk.off, DL);
                        win_store<R, kW, true>(buf + S, sb, bk.off, D);
                        __syncthreads();
                        win_load<R, kW, true>(buf, sn, pv.off, DL);
                        win_load<R, kW, true>(buf + S, sn, pv.off, D);
                        __syncthreads();
                    }
                    if constexpr (MODE == 2 && RS == 1) {
                        win_store<R, kW, true>(buf, sb, bk.off, DL);
                        __syncthreads();
                        win_load<R, kW, true>(buf, sn, pv.off, DL);
                        __syncthreads();
                        win_store<R, kW, true>(buf, sb, bk.off, D);
                        __syncthreads();
                        win_load<R, kW, true>(buf, sn, pv.off, D);
                        __syncthreads();
                    }
                } else {
                    __syncthreads();
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
    for (int k = tid0; k < 32 * a.P; k += kT) out[2 + k] = (k & 31) < NV ? acc[(k >> 5) * NV + (k & 31)] : 0.0;
}

// ---- tiled variant (FP64, L > 12) ------------------------------------------------------------
namespace {

// bits of x scattered to the positions p[0..n) (wave-uniform positions, scalar reads)
__device__ inline uint32_t scatter_c(uint32_t x, const int* p, int n)
{
    const MG_CONST int* q = cptr(p);
    uint32_t r = 0;
    for (int i = 0; i < n; i++) r |= ((x >> i) & 1u) << q[i];
    return r;
}

}  // namespace

// the staged sums of a window block's gates into the accumulators (as win_rev_adds)
template <int kNW, int NV>
__device__ inline void win_tile_rev_adds(const WinTileArgs& a, int g0, int g1, const double* stage, double* acc, int wv,
                                         int lane)
{
    for (int g = g0; g < g1; g++) {
        const int gparam = cptr(a.inst)[g].param, gflip = cptr(a.inst)[g].flip;
        if (wv == gparam % kNW && lane < NV) {
            const double* sg = stage + (size_t)(g - g0) * kNW * 32;
            double s = 0.0;
            for (int w = 0; w < kNW; w++) s += sg[w * 32 + lane];
            const int ci = lane >> 1;
            int e = ci & 7;
            if (gflip && e < 4) e = 3 - e;
            acc[NV * gparam + 2 * ((ci & 8) | e) + (lane & 1)] += s;
        }
    }
}

// Columns of 2^L > 2^12 amplitudes in 12-qubit tiles (512 threads x 8 amplitudes in
// registers), each tile block's gates applied in 3-qubit register windows.  MODE 0/1/2: f,
// + gradient, + HVP against the dense target rows (states between tile blocks in the
// per-workgroup scratch).  MODE 3: materialise a target program, row b of a.rows =
// -conj(W e_b), each row doubling as its column's state between tile blocks.
template <int MODE, int RS>
__global__ __launch_bounds__(512) void sweep_win_tile_kernel(WinTileArgs a)
{
    using R = double;
    constexpr int kT = 512, kW = 3, kNW = kT / 64, TB = 12;
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NV = MODE == 2 ? 32 : 16;
    // psi stored for the reverse re-layout while the block's last lam update runs (as in
    // sweep_win_kernel) only for the gradient: for the HVP the longer live ranges raise this
    // kernel's spills (L=16 dense rows 2.17 -> 2.36 s measured)
    constexpr bool kEarly = RS == 2 && MODE == 1;
    constexpr uint32_t TSZ = 1u << TB;  // tile amplitudes (the re-layout buffer's index space)
    const int L = a.L;
    const uint32_t S = 1u << L, ntiles = 1u << (L - TB);
    const int tid0 = threadIdx.x, lane = tid0 & 63, wv = tid0 >> 6;

    double* stage = reinterpret_cast<double*>(smem);  // [a.slots][kNW][32]
    double* fst = stage + a.slots * kNW * 32;         // [kNW][2]
    double* acc = reinterpret_cast<double*>(smem + a.acc_off);
    cpx<R>* buf = reinterpret_cast<cpx<R>*>(smem + a.buf_off);  // [RS][TSZ] re-layout buffer
    // the column's states in global memory: psi, dpsi (MODE 2), lam (MODE >= 1), dlam (MODE 2)
    constexpr int LAM = MODE == 2 ? 2 : 1;
    cpx<R>* sc = reinterpret_cast<cpx<R>*>(a.scratch) + (size_t)blockIdx.x * a.nstate * S;
    cpx<R>* g_psi = sc;
    cpx<R>* g_dpsi = sc + S;
    cpx<R>* g_lam = sc + (size_t)LAM * S;
    cpx<R>* g_dlam = sc + 3 * (size_t)S;
    const cpx<R>* coef = reinterpret_cast<const cpx<R>*>(a.coef);
    const cpx<R>* seed = reinterpret_cast<const cpx<R>*>(a.seed);

    for (int k = tid0; k < NV * a.P; k += kT) acc[k] = 0.0;
    double fre = 0.0, fim = 0.0;
    __syncthreads();

    const int ntb = a.ntb;
    for (long b = a.b0 + blockIdx.x; b < a.b1; b += gridDim.x) {
        int tid = tid0;
        asm volatile("" : "+v"(tid));
        Win<R, kW> Lm, P, DL, D;
        double fr = 0.0, fi = 0.0;
        const cpx<R>* row = MODE == 3 ? nullptr : seed + (size_t)(b - a.row0) * S;
        cpx<R>* f_psi = MODE == 3 ? reinterpret_cast<cpx<R>*>(a.rows) + (size_t)(b - a.row0) * S : g_psi;
        // ---- passes over the tile blocks: forward 0 .. ntb-2 (psi, dpsi); the last tile block
        // forward and at once backwards, tile by tile (f_b = s_b . psi_G, lam seeded from the target
        // row read once for both, the states handed over in registers); then reverse ntb-2 .. 0 (lam,
        // psi uncomputed, dlam, dpsi).  One loop, so that the forward and the reverse code each
        // appear once.
        constexpr bool kRev = MODE == 1 || MODE == 2;
        const int npass = kRev ? 2 * ntb - 1 : ntb;
        for (int ps = 0; ps < npass; ps++) {
            const bool fwd = ps < ntb;
            const int tb = fwd ? ps : 2 * ntb - 2 - ps;
            const WTBlock* tbk = a.tblocks + tb;
            const int k0 = cptr(tbk)->k0, k1 = cptr(tbk)->k1;
            for (uint32_t t = 0; t < ntiles; t++) {
                const uint32_t ot = scatter_c(t, tbk->other, L - TB);
                if (fwd) {
                    {
                        const BlockRegs<kW> bf = load_block<kW>(a.blk_glob + k0);
                        const uint32_t gb = ot | scatter_c(deposit<kW>((uint32_t)tid, bf.pos), tbk->tpos, TB);
                        if (tb == 0) {
#pragma unroll
                            for (int j = 0; j < (1 << kW); j++) {
                                P.a[j] = {(gb ^ bf.off[j]) == (uint32_t)b ? R(1) : R(0), R(0)};
                                D.a[j] = {R(0), R(0)};
                            }
                        } else {
                            win_load<R, kW, false>(f_psi, gb, bf.off, P);
                            if constexpr (MODE == 2) win_load<R, kW, false>(g_dpsi, gb, bf.off, D);
                        }
                    }
                    for (int k = k0; k < k1; k++) {
                        const BlockRegs<kW> bk = load_block<kW>(a.blk_lds + k);
                        for (int g = bk.g0; g < bk.g1; g++) {
                            const int wcase = cptr(a.inst)[g].lo;
                            cpx<R> cf[24];
                            load_coef_c<R, 8>(coef + (size_t)g * kCoefPerInst, cf);
                            if constexpr (MODE == 2) load_coef_c<R, 8>(coef + (size_t)g * kCoefPerInst + 16, cf + 16);
#define MG_FWD(WW, A, B) wg_forward<R, WW, A, B, MODE == 2 ? 2 : 0>(cf, P, D)
                            MG_WCASES(kW, MG_FWD)
#undef MG_FWD
                        }
                        if (k + 1 < k1) {
                            const BlockRegs<kW> nx = load_block<kW>(a.blk_lds + k + 1);
                            const uint32_t sb = lin<R, true>(deposit<kW>((uint32_t)tid, bk.pos));
                            const uint32_t sn = lin<R, true>(deposit<kW>((uint32_t)tid, nx.pos));
                            win_store<R, kW, true>(buf, sb, bk.off, P);
                            if constexpr (MODE == 2 && RS >= 2) win_store<R, kW, true>(buf + TSZ, sb, bk.off, D);
                            __syncthreads();
                            win_load<R, kW, true>(buf, sn, nx.off, P);
                            if constexpr (MODE == 2 && RS >= 2) win_load<R, kW, true>(buf + TSZ, sn, nx.off, D);
                            __syncthreads();
                            if constexpr (MODE == 2 && RS == 1) {
                                win_store<R, kW, true>(buf, sb, bk.off, D);
                                __syncthreads();
                                win_load<R, kW, true>(buf, sn, nx.off, D);
                                __syncthreads();
                            }
                        }
                    }
                    // the last window block's layout: back to global memory, or (last tile block)
                    // f_b = s_b . psi_G and the reverse sweep's seed lam = s_b, dlam = 0
                    const BlockRegs<kW> bl = load_block<kW>(a.blk_glob + k1 - 1);
                    const uint32_t gb = ot | scatter_c(deposit<kW>((uint32_t)tid, bl.pos), tbk->tpos, TB);
                    if constexpr (MODE == 3) {
                        if (tb == ntb - 1) {
#pragma unroll
                            for (int j = 0; j < (1 << kW); j++) P.a[j] = {-P.a[j].re, P.a[j].im};
                        }
                        win_store<R, kW, false>(f_psi, gb, bl.off, P);
                        continue;
                    } else {
                        if (tb < ntb - 1) {
                            win_store<R, kW, false>(g_psi, gb, bl.off, P);
                            if constexpr (MODE == 2) win_store<R, kW, false>(g_dpsi, gb, bl.off, D);
                            continue;
                        }
#pragma unroll
                        for (int j = 0; j < (1 << kW); j++) {
                            const cpx<R> r = row[gb ^ bl.off[j]];
                            const cpx<R> tt = cmul(r, P.a[j]);
                            fr += tt.re;
                            fi += tt.im;
                            Lm.a[j] = r;
                            DL.a[j] = {R(0), R(0)};
                        }
                        if constexpr (!kRev) continue;
                    }
                } else {
                    const BlockRegs<kW> bl = load_block<kW>(a.blk_glob + k1 - 1);
                    const uint32_t gb = ot | scatter_c(deposit<kW>((uint32_t)tid, bl.pos), tbk->tpos, TB);
                    win_load<R, kW, false>(g_psi, gb, bl.off, P);
                    win_load<R, kW, false>(g_lam, gb, bl.off, Lm);
                    if constexpr (MODE == 2) {
                        win_load<R, kW, false>(g_dpsi, gb, bl.off, D);
                        win_load<R, kW, false>(g_dlam, gb, bl.off, DL);
                    }
                }
                if constexpr (kRev) {
                    for (int k = k1 - 1; k >= k0; k--) {
                        const BlockRegs<kW> bk = load_block<kW>(a.blk_lds + k);
                        const uint32_t sb = lin<R, true>(deposit<kW>((uint32_t)tid, bk.pos));
                        for (int g = bk.g1 - 1; g >= bk.g0; g--) {
                            const int wcase = cptr(a.inst)[g].lo;
                            double* slot = stage + ((g - bk.g0) * kNW + wv) * 32;
                            cpx<R> V[8], Z[8];
                            load_coef_c<R, 8>(coef + (size_t)g * kCoefPerInst, V);
                            if constexpr (MODE == 2) load_coef_c<R, 8>(coef + (size_t)g * kCoefPerInst + 16, Z);
                            {
                                cpx<R> ga[8];  // set by the gate's first quadruple
#define MG_PA(WW, A, B) pass_a<R, WW, A, B, MODE>(V, Z, Lm, P, D, ga)
                                MG_WCASES(kW, MG_PA)
#undef MG_PA
                                stage8(ga, slot, lane);
                            }
                            if constexpr (MODE == 2) {
                                cpx<R> ha[8];
#define MG_PB(WW, A, B) pass_b<R, WW, A, B>(Lm, P, DL, D, ha)
                                MG_WCASES(kW, MG_PB)
#undef MG_PB
                                stage8(ha, slot + 16, lane);
                            }
                            if constexpr (kEarly) {
                                // psi (and dpsi) final for this window block: stored for the
                                // re-layout while the last lam update runs
                                if (g == bk.g0 && k > k0) {
                                    if constexpr (MODE == 2) {
                                        win_store<R, kW, true>(buf, sb, bk.off, P);
                                        win_store<R, kW, true>(buf + TSZ, sb, bk.off, D);
                                    } else {
                                        win_store<R, kW, true>(buf + TSZ, sb, bk.off, P);
                                    }
                                }
                            }
#define MG_PC(WW, A, B) pass_c<R, WW, A, B, MODE>(V, Z, Lm, DL)
                            MG_WCASES(kW, MG_PC)
#undef MG_PC
                        }
                        // re-layout to the previous window block's order; the staged sums are
                        // consumed after the first barrier
                        const bool move = k > k0;
                        BlockRegs<kW> pv = bk;
                        if (move) pv = load_block<kW>(a.blk_lds + k - 1);
                        const uint32_t sn = lin<R, true>(deposit<kW>((uint32_t)tid, pv.pos));
                        if constexpr (kEarly && MODE == 2) {
                            __syncthreads();
                            win_tile_rev_adds<kNW, NV>(a, bk.g0, bk.g1, stage, acc, wv, lane);
                            if (move) {
                                win_load<R, kW, true>(buf, sn, pv.off, P);
                                win_load<R, kW, true>(buf + TSZ, sn, pv.off, D);
                                __syncthreads();
                                win_store<R, kW, true>(buf, sb, bk.off, Lm);
                                win_store<R, kW, true>(buf + TSZ, sb, bk.off, DL);
                                __syncthreads();
                                win_load<R, kW, true>(buf, sn, pv.off, Lm);
                                win_load<R, kW, true>(buf + TSZ, sn, pv.off, DL);
                            }
                            __syncthreads();
                            continue;
                        }
                        if (move) {
                            win_store<R, kW, true>(buf, sb, bk.off, Lm);
                            if constexpr (RS >= 2 && !kEarly) win_store<R, kW, true>(buf + TSZ, sb, bk.off, P);
                        }
                        __syncthreads();
                        win_tile_rev_adds<kNW, NV>(a, bk.g0, bk.g1, stage, acc, wv, lane);
                        if (move) {
                            win_load<R, kW, true>(buf, sn, pv.off, Lm);
                            if constexpr (RS >= 2) win_load<R, kW, true>(buf + TSZ, sn, pv.off, P);
                            __syncthreads();
                            if constexpr (RS == 1) {
                                win_store<R, kW, true>(buf, sb, bk.off, P);
                                __syncthreads();
                                win_load<R, kW, true>(buf, sn, pv.off, P);
                                __syncthreads();
                            }
                            if constexpr (MODE == 2 && RS == 2) {
                                win_store<R, kW, true>(buf, sb, bk.off, DL);
                                win_store<R, kW, true>(buf + TSZ, sb, bk.off, D);
                                __syncthreads();
                                win_load<R, kW, true>(buf, sn, pv.off, DL);
                                win_load<R, kW, true>(buf + TSZ, sn, pv.off, D);
                                __syncthreads();
                            }
                            if constexpr (MODE == 2 && RS == 1) {
                                win_store<R, kW, true>(buf, sb, bk.off, DL);
                                __syncthreads();
                                win_load<R, kW, true>(buf, sn, pv.off, DL);
                                __syncthreads();
                                win_store<R, kW, true>(buf, sb, bk.off, D);
                                __syncthreads();
                                win_load<R, kW, true>(buf, sn, pv.off, D);
                                __syncthreads();
                            }
                        } else {
                            __syncthreads();
                        }
                    }
                    // back to global memory in the first window block's order (the next,
                    // lower tile block reads it)
                    if (tb > 0) {
                        const BlockRegs<kW> bf = load_block<kW>(a.blk_glob + k0);
                        const uint32_t gb = ot | scatter_c(deposit<kW>((uint32_t)tid, bf.pos), tbk->tpos, TB);
                        win_store<R, kW, false>(g_psi, gb, bf.off, P);
                        win_store<R, kW, false>(g_lam, gb, bf.off, Lm);
                        if constexpr (MODE == 2) {
                            win_store<R, kW, false>(g_dpsi, gb, bf.off, D);
                            win_store<R, kW, false>(g_dlam, gb, bf.off, DL);
                        }
                    }
                }
            }
            // the next tile block reads, in its own mapping, what other threads stored here
            __syncthreads();
        }
        if constexpr (MODE == 3) continue;
        {
            const double sr = wave_sum(fr), si = wave_sum(fi);
            if (lane == 0) {
                fst[2 * wv] = sr;
                fst[2 * wv + 1] = si;
            }
            __syncthreads();
            if (tid == 0) {
                double tr = 0.0, ti = 0.0;
                for (int w = 0; w < kNW; w++) {
                    tr += fst[2 * w];
                    ti += fst[2 * w + 1];
                }
                fre += tr;
                fim += ti;
            }
        }
        __syncthreads();
    }
    if constexpr (MODE == 3) return;
    __syncthreads();
    double* out = a.partial + (size_t)blockIdx.x * (2 + 32 * a.P);
    if (tid0 == 0) {
        out[0] = fre;
        out[1] = fim;
    }
    for (int k = tid0; k < 32 * a.P; k += kT) out[2 + k] = (k & 31) < NV ? acc[(k >> 5) * NV + (k & 31)] : 0.0;
}

hipError_t launch_materialize_win_tile(const WinTileArgs& a, int grid, size_t lds, hipStream_t s)
{
    if (a.L <= 12 || a.L - 12 > 20 || a.ntb < 1 || a.rs != 1 || !a.rows || a.slots < 1 || a.slots > kWinRegGates)
        return hipErrorInvalidValue;
    const void* k = (const void*)sweep_win_tile_kernel<3, 1>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((sweep_win_tile_kernel<3, 1>), dim3(grid), dim3(512), lds, s, a);
    return hipGetLastError();
}

template <int RS>
static hipError_t launch_wtile_rs(const WinTileArgs& a, int mode, int grid, size_t lds, hipStream_t s)
{
    const void* k = mode == 2 ? (const void*)sweep_win_tile_kernel<2, RS>
                  : mode == 1 ? (const void*)sweep_win_tile_kernel<1, RS> : (const void*)sweep_win_tile_kernel<0, RS>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (mode == 2) hipLaunchKernelGGL((sweep_win_tile_kernel<2, RS>), dim3(grid), dim3(512), lds, s, a);
    else if (mode == 1) hipLaunchKernelGGL((sweep_win_tile_kernel<1, RS>), dim3(grid), dim3(512), lds, s, a);
    else hipLaunchKernelGGL((sweep_win_tile_kernel<0, RS>), dim3(grid), dim3(512), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_sweep_win_tile(const WinTileArgs& a, int mode, int grid, size_t lds, hipStream_t s)
{
    // 12-qubit tiles of a larger column, the other bits listed in WTBlock::other
    if (a.L <= 12 || a.L - 12 > 20 || a.ntb < 1 || a.slots < 1 || a.slots > kWinRegGates) return hipErrorInvalidValue;
    if (a.rs == 2) return launch_wtile_rs<2>(a, mode, grid, lds, s);
    if (a.rs == 1) return launch_wtile_rs<1>(a, mode, grid, lds, s);
    return hipErrorInvalidValue;
}

size_t win_sweep_lds_bytes(int L, int P, int fp32, size_t limit, size_t* acc_off, size_t* buf_off, int* rs, int mode,
                           int slots)
{
    const int nw = (fp32 ? WinShape<float>::T : WinShape<double>::T) / 64;
    // staging: one slot per gate of the largest window block
    const size_t fixed = sizeof(double) * ((size_t)slots * nw * 32 + 2 * nw);
    const size_t a0 = (fixed + 255) & ~(size_t)255;
    // per parameter gate 16 complex sums (gradient, D^2 f[Z]); 8 below MODE 2
    const size_t b0 = (a0 + sizeof(double) * (mode == 2 ? 32 : 16) * (size_t)P + 255) & ~(size_t)255;
    if (acc_off) *acc_off = a0;
    if (buf_off) *buf_off = b0;
    const size_t state = (fp32 ? sizeof(cpx<float>) : sizeof(cpx<double>)) * ((size_t)1 << L);
    // four FP32 states per re-layout round; FP64 two, or one when two do not fit
    int n = fp32 ? 4 : 2;
    if (!fp32 && b0 + 2 * state > limit) n = 1;
    if (rs) *rs = n;
    return b0 + (size_t)n * state;
}

bool win_sweep_fits(int L, int fp32)
{
    return fp32 ? ((size_t)1 << L) == ((size_t)WinShape<float>::T << WinShape<float>::W)
                : ((size_t)1 << L) == ((size_t)WinShape<double>::T << WinShape<double>::W);
}

int win_sweep_window_bits(int fp32) { return fp32 ? WinShape<float>::W : WinShape<double>::W; }

template <typename R, int RS>
static hipError_t launch_win_r(const WinSweepArgs& a, int mode, int grid, size_t lds, hipStream_t s)
{
    const void* k = mode == 2 ? (const void*)sweep_win_kernel<R, 2, RS>
                  : mode == 1 ? (const void*)sweep_win_kernel<R, 1, RS> : (const void*)sweep_win_kernel<R, 0, RS>;
    if (lds > 65536) {
        hipError_t e = ensure_lds_limit(k, lds);
        if (e != hipSuccess) return e;
    }
    constexpr int T = WinShape<R>::T;
    if (mode == 2) hipLaunchKernelGGL((sweep_win_kernel<R, 2, RS>), dim3(grid), dim3(T), lds, s, a);
    else if (mode == 1) hipLaunchKernelGGL((sweep_win_kernel<R, 1, RS>), dim3(grid), dim3(T), lds, s, a);
    else hipLaunchKernelGGL((sweep_win_kernel<R, 0, RS>), dim3(grid), dim3(T), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_sweep_win(const WinSweepArgs& a, int mode, int fp32, int grid, size_t lds, hipStream_t s)
{
    if (a.slots < 1 || a.slots > kWinRegGates) return hipErrorInvalidValue;
    if (fp32) return a.rs == 4 ? launch_win_r<float, 4>(a, mode, grid, lds, s) : hipErrorInvalidValue;
    if (a.rs == 2) return launch_win_r<double, 2>(a, mode, grid, lds, s);
    if (a.rs == 1) return launch_win_r<double, 1>(a, mode, grid, lds, s);
    return hipErrorInvalidValue;
}

}  // namespace mg
