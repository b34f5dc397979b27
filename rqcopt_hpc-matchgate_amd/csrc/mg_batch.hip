// Batched column sweep: one persistent launch over the basis columns of many
// evaluation requests of different qubit counts, depths, modes and targets, with
// the gates of each request fused into W-qubit windows.
//
// Work split.  Each request ("instance") is cut into work items of a fixed number
// of columns.  Workgroups take items from an atomic counter (largest items first,
// as ordered by the host) until the list is exhausted.  Within an item, the T
// threads form T / cgw column groups of cgw threads; every group sweeps its own
// column of the same instance, so all groups pass the same gate sequence and share
// the workgroup barriers (a group may be narrower than a wave).
//
// Gate windows.  The host regroups the instance's gates into blocks whose gates all
// act inside one window of W qubits (mg_window.cpp).  A column's state splits into
// 2^(L-W) tuples of 2^W amplitudes that differ only in the window bits; per block a
// thread loads its tuples' amplitudes of every live state into registers, applies
// all of the block's gates there and stores them back: one LDS (or memory) pass and
// one barrier per block instead of per gate.  The per-gate outer products of the
// reverse sweep are reduced per tuple by a transposing wave butterfly and added
// into per-gate LDS slots (double-buffered by block parity); after the block's
// barrier one wave per gate adds the slots, in wave order, into the accumulator.
//
// The per-column arithmetic is that of sweep_kernel (mg_sweep.hip): seed
// lam = -conj(U e_b), forward psi (and dpsi), f_b = lam . psi_G, reverse sweep
// uncomputing psi_g with V^{-1}.  An item's partial sums depend only on its instance,
// its columns and its batch shape, and each instance's items are summed in item
// order by batch_reduce: a request evaluates to the same bits alone and in any mix.
//
// Reference: per-column loop src/matchgate_target.c:788-837, forward
// src/matchgate_brickwall.c:52-78, backward kernel src/gate.c:659-772.
#include "mg_batch.hpp"
#include "mg_quad.hpp"
#include "mg_win.hpp"

namespace mg {

// apply gates [g0, g1) of a forward-only program (target seeding) to one state
template <typename R, int W, bool SWZ>
__device__ inline void win_apply_program(cpx<R>* st, const WBlock* blocks, int nb, const Inst* inst, const cpx<R>* coef,
                                         uint32_t ntup, int tl, int cgw, bool on)
{
    for (int k = 0; k < nb; k++) {
        const BlockRegs<W> bk = load_block<W>(blocks + k);
        if (on)
            for (uint32_t u = tl; u < ntup; u += cgw) {
                const uint32_t base = deposit<W>(u, bk.pos);
                const uint32_t sb = lin<R, SWZ>(base);
                Win<R, W> P, D;
                win_load<R, W, SWZ>(st, sb, bk.off, P);
                for (int g = bk.g0; g < bk.g1; g++) {
                    const int wcase = cptr(inst)[g].lo;
                    cpx<R> cf[8];
                    load_coef_c<R, 8>(coef + (size_t)g * kCoefPerInst, cf);
#define MG_FWD0(WW, A, B) wg_forward<R, WW, A, B, 0>(cf, P, D)
                    MG_WCASES(W, MG_FWD0)
#undef MG_FWD0
                }
                win_store<R, W, SWZ>(st, sb, bk.off, P);
            }
        __syncthreads();
    }
}

template <typename R, int T, int W, int MODE, bool LDS>
__device__ void batch_item(const BatchInst<R>& in, const BatchItem& it, unsigned char* smem, cpx<R>* gslice,
                           double* out)
{
    constexpr int NW = T / 64;
    constexpr int NV = MODE == 2 ? 32 : 16;
    constexpr int SHIFT = 6 - log2c<NV>();
    constexpr int NSL = MODE == 2 ? 4 : 2;
    constexpr bool SWZ = LDS;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t S = 1u << in.L, ntup = S >> W;
    const int cgw = in.cgw, ncc = T / cgw, grp = tid / cgw, tl = tid - grp * cgw;

    double* stage = reinterpret_cast<double*>(smem);        // [2][kWinMaxGates][NW][32]
    double* fst = stage + 2 * kWinMaxGates * NW * 32;       // [NW][2]
    double* acc = reinterpret_cast<double*>(smem + in.acc_off);
    cpx<R>* st = (LDS ? reinterpret_cast<cpx<R>*>(smem + in.state_off) : gslice) + (size_t)grp * NSL * S;
    cpx<R>* lam = st;
    cpx<R>* psi = st + S;
    cpx<R>* dlam = st + 2 * S;
    cpx<R>* dpsi = st + 3 * S;

    for (int k = tid; k < 32 * in.P; k += T) acc[k] = 0.0;
    double fre = 0.0, fim = 0.0;
    __syncthreads();

    for (int r = it.b0; r < it.b1; r += ncc) {
        const uint32_t b = (uint32_t)(r + grp);
        const bool on = (int)b < it.b1;
        // ---- seed: lam = -conj(U e_b) -------------------------------------------------
        if (in.target_kind == 0) {
            if (on) {
                const cpx<R>* row = in.seed + (size_t)b * S;
                for (uint32_t i = tl; i < S; i += cgw) lam[lin<R, SWZ>(i)] = row[i];
            }
        } else {
            if (on)
                for (uint32_t i = tl; i < S; i += cgw) lam[lin<R, SWZ>(i)] = {R(i == b), R(0)};
            __syncthreads();
            win_apply_program<R, W, SWZ>(lam, in.tblocks, in.ntblocks, in.tinst, in.tcoef, ntup, tl, cgw, on);
            if (on)
                for (uint32_t i = tl; i < S; i += cgw) {
                    cpx<R>& v = lam[lin<R, SWZ>(i)];
                    v = {-v.re, v.im};
                }
        }
        // ---- forward sweep -------------------------------------------------------------
        if (on)
            for (uint32_t i = tl; i < S; i += cgw) {
                psi[lin<R, SWZ>(i)] = {R(i == b), R(0)};
                if constexpr (MODE == 2) dpsi[lin<R, SWZ>(i)] = {R(0), R(0)};
            }
        __syncthreads();
        for (int k = 0; k < in.nblocks; k++) {
            const BlockRegs<W> bk = load_block<W>(in.blocks + k);
            if (on)
                for (uint32_t u = tl; u < ntup; u += cgw) {
                    const uint32_t sb = lin<R, SWZ>(deposit<W>(u, bk.pos));
                    Win<R, W> P, D;
                    win_load<R, W, SWZ>(psi, sb, bk.off, P);
                    if constexpr (MODE == 2) win_load<R, W, SWZ>(dpsi, sb, bk.off, D);
                    for (int g = bk.g0; g < bk.g1; g++) {
                        const int wcase = cptr(in.inst)[g].lo;
                        cpx<R> cf[24];
                        load_coef_c<R, 8>(in.coef + (size_t)g * kCoefPerInst, cf);
                        if constexpr (MODE == 2) load_coef_c<R, 8>(in.coef + (size_t)g * kCoefPerInst + 16, cf + 16);
#define MG_FWD(WW, A, B) wg_forward<R, WW, A, B, MODE == 2 ? 2 : 0>(cf, P, D)
                        MG_WCASES(W, MG_FWD)
#undef MG_FWD
                    }
                    win_store<R, W, SWZ>(psi, sb, bk.off, P);
                    if constexpr (MODE == 2) win_store<R, W, SWZ>(dpsi, sb, bk.off, D);
                }
            __syncthreads();
        }
        // ---- f_b = lam . psi_G ------------------------------------------------------------
        {
            double sr = 0.0, si = 0.0;
            if (on)
                for (uint32_t i = tl; i < S; i += cgw) {
                    const cpx<R> t = cmul(lam[lin<R, SWZ>(i)], psi[lin<R, SWZ>(i)]);
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
            // ---- reverse sweep -------------------------------------------------------------
            if constexpr (MODE == 2) {
                if (on)
                    for (uint32_t i = tl; i < S; i += cgw) dlam[lin<R, SWZ>(i)] = {R(0), R(0)};
            }
            __syncthreads();
            for (int k = in.nblocks - 1; k >= 0; k--) {
                const BlockRegs<W> bk = load_block<W>(in.blocks + k);
                const int buf = k & 1;
                double* stg = stage + (size_t)buf * kWinMaxGates * NW * 32;
                for (uint32_t u = tl; u < ntup; u += cgw) {
                    // inactive groups and idle lanes still take part in the wave butterflies
                    const bool act = on;
                    const uint32_t sb = lin<R, SWZ>(deposit<W>(act ? u : 0u, bk.pos));
                    Win<R, W> Lm, P, DL, D;
                    if (act) {
                        win_load<R, W, SWZ>(lam, sb, bk.off, Lm);
                        win_load<R, W, SWZ>(psi, sb, bk.off, P);
                        if constexpr (MODE == 2) {
                            win_load<R, W, SWZ>(dlam, sb, bk.off, DL);
                            win_load<R, W, SWZ>(dpsi, sb, bk.off, D);
                        }
                    }
                    for (int g = bk.g1 - 1; g >= bk.g0; g--) {
                        const int wcase = cptr(in.inst)[g].lo;
                        cpx<R> cf[24];
                        load_coef_c<R, MODE == 2 ? 24 : 16>(in.coef + (size_t)g * kCoefPerInst, cf);
                        cpx<R> ga[8], ha[8];
#pragma unroll
                        for (int e = 0; e < 8; e++) {
                            ga[e] = {R(0), R(0)};
                            ha[e] = {R(0), R(0)};
                        }
                        if (act) {
#define MG_BWD(WW, A, B) wg_backward<R, WW, A, B, MODE>(cf, Lm, P, DL, D, ga, ha)
                            MG_WCASES(W, MG_BWD)
#undef MG_BWD
                        }
                        R v[NV];
#pragma unroll
                        for (int e = 0; e < 8; e++) {
                            v[2 * e] = ga[e].re;
                            v[2 * e + 1] = ga[e].im;
                            if constexpr (MODE == 2) {
                                v[16 + 2 * e] = ha[e].re;
                                v[16 + 2 * e + 1] = ha[e].im;
                            }
                        }
                        wave_halve<NV, 32, R>(v, lane);
                        if ((lane & ((1 << SHIFT) - 1)) == 0)
                            stg[((g - bk.g0) * NW + wv) * 32 + (lane >> SHIFT)] += (double)v[0];
                    }
                    if (act) {
                        win_store<R, W, SWZ>(lam, sb, bk.off, Lm);
                        win_store<R, W, SWZ>(psi, sb, bk.off, P);
                        if constexpr (MODE == 2) {
                            win_store<R, W, SWZ>(dlam, sb, bk.off, DL);
                            win_store<R, W, SWZ>(dpsi, sb, bk.off, D);
                        }
                    }
                }
                __syncthreads();
                // one wave per gate adds the staged sums in wave order, then clears its slots;
                // the wave is the one that owns the gate's parameter, so gates of one block
                // that share a parameter (tied layers) are added by one wave, in order
                for (int g = bk.g0; g < bk.g1; g++) {
                    const int gparam = cptr(in.inst)[g].param, gflip = cptr(in.inst)[g].flip;
                    if (wv == gparam % NW && lane < NV) {
                        double* sg = stg + (size_t)(g - bk.g0) * NW * 32;
                        double s = 0.0;
                        for (int w = 0; w < NW; w++) {
                            s += sg[w * 32 + lane];
                            sg[w * 32 + lane] = 0.0;
                        }
                        const int ci = lane >> 1;
                        int e = ci & 7;
                        if (gflip && e < 4) e = 3 - e;
                        acc[32 * gparam + 2 * ((ci & 8) | e) + (lane & 1)] += s;
                    }
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0) {
        out[0] = fre;
        out[1] = fim;
    }
    for (int k = tid; k < 32 * in.P; k += T) out[2 + k] = acc[k];
}

template <typename R, int T, int W>
__global__ __launch_bounds__(T) void batch_kernel(const BatchInst<R>* __restrict__ insts,
                                                  const BatchItem* __restrict__ items, int nitems, int* counter,
                                                  double* partial, cpx<R>* scratch, long long scratch_stride)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int next;
    cpx<R>* gslice = scratch ? scratch + (size_t)blockIdx.x * scratch_stride : nullptr;
    {
        double* stage = reinterpret_cast<double*>(smem);
        for (int k = threadIdx.x; k < 2 * kWinMaxGates * (T / 64) * 32; k += T) stage[k] = 0.0;
    }
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int k = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (k >= nitems) break;
        const BatchItem it = items[k];
        const BatchInst<R> in = insts[it.inst];
        double* out = partial + it.out;
        const int path = in.mode * 2 + (in.lds ? 1 : 0);
        switch (path) {
            case 0: batch_item<R, T, W, 0, false>(in, it, smem, gslice, out); break;
            case 1: batch_item<R, T, W, 0, true>(in, it, smem, gslice, out); break;
            case 2: batch_item<R, T, W, 1, false>(in, it, smem, gslice, out); break;
            case 3: batch_item<R, T, W, 1, true>(in, it, smem, gslice, out); break;
            case 4: batch_item<R, T, W, 2, false>(in, it, smem, gslice, out); break;
            default: batch_item<R, T, W, 2, true>(in, it, smem, gslice, out); break;
        }
    }
}

// seed[b][a] = -conj((U e_b)[a]) for the columns b of a target program, FP64: one
// column per group of cgw threads, the state in LDS (window form as batch_item)
template <int T, int W>
__global__ __launch_bounds__(T) void materialize_kernel(int L, const Inst* tinst, const cpx<double>* tcoef,
                                                        const WBlock* tblocks, int ntblocks, int cgw,
                                                        cpx<double>* seed)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t S = 1u << L, ntup = S >> W;
    const int ncc = T / cgw, grp = threadIdx.x / cgw, tl = threadIdx.x - grp * cgw;
    cpx<double>* st = reinterpret_cast<cpx<double>*>(smem) + (size_t)grp * S;
    for (uint32_t b0 = blockIdx.x * ncc; b0 < S; b0 += gridDim.x * ncc) {
        const uint32_t b = b0 + grp;
        const bool on = b < S;
        if (on)
            for (uint32_t i = tl; i < S; i += cgw) st[lin<double, true>(i)] = {double(i == b), 0.0};
        __syncthreads();
        win_apply_program<double, W, true>(st, tblocks, ntblocks, tinst, tcoef, ntup, tl, cgw, on);
        if (on)
            for (uint32_t i = tl; i < S; i += cgw) {
                const cpx<double> v = st[lin<double, true>(i)];
                seed[(size_t)b * S + i] = {-v.re, v.im};
            }
        __syncthreads();
    }
}

// the same for columns too large for LDS: the state is the seed row itself (global memory,
// plain offsets), one column per workgroup
template <int T, int W>
__global__ __launch_bounds__(T) void materialize_global_kernel(int L, const Inst* tinst, const cpx<double>* tcoef,
                                                               const WBlock* tblocks, int ntblocks, cpx<double>* seed)
{
    const uint32_t S = 1u << L, ntup = S >> W;
    const int tl = threadIdx.x;
    for (uint32_t b = blockIdx.x; b < S; b += gridDim.x) {
        cpx<double>* row = seed + (size_t)b * S;
        for (uint32_t i = tl; i < S; i += T) row[i] = {double(i == b), 0.0};
        __syncthreads();
        win_apply_program<double, W, false>(row, tblocks, ntblocks, tinst, tcoef, ntup, tl, T, true);
        for (uint32_t i = tl; i < S; i += T) {
            const cpx<double> v = row[i];
            row[i] = {-v.re, v.im};
        }
        __syncthreads();
    }
}

hipError_t launch_materialize_global(int L, const Inst* tinst, const cpx<double>* tcoef, const WBlock* tblocks,
                                     int ntblocks, cpx<double>* seed, int grid, hipStream_t s)
{
    constexpr int T = 256, W = 3;
    hipLaunchKernelGGL((materialize_global_kernel<T, W>), dim3(grid), dim3(T), 0, s, L, tinst, tcoef, tblocks, ntblocks,
                       seed);
    return hipGetLastError();
}

hipError_t launch_materialize(int L, const Inst* tinst, const cpx<double>* tcoef, const WBlock* tblocks, int ntblocks,
                              cpx<double>* seed, int grid, size_t lds, hipStream_t s)
{
    constexpr int T = 256, W = 3;
    const int cgw = (int)std::min<long>(T, std::max<long>(1, (1L << L) >> W));
    auto k = materialize_kernel<T, W>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(T), lds, s, L, tinst, tcoef, tblocks, ntblocks, cgw, seed);
    return hipGetLastError();
}

__global__ void batch_reset(int* counter) { *counter = 0; }

// the dynamic-LDS attribute is set before a capture, not while one is recorded
static bool hip_stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}

// out[seg.dst + k] = sum over items j of partial[seg.src + j * seg.stride + k], in j order
// (thread row ty sums j = ty, ty + 16, ...; row 0 adds the 16 row sums in order)
__global__ __launch_bounds__(1024) void batch_reduce(const ReduceSeg* __restrict__ segs, const int2* __restrict__ blocks,
                                                     const double* __restrict__ partial, double* __restrict__ out)
{
    __shared__ double rows[16][64];
    const int2 bk = blocks[blockIdx.x];
    const ReduceSeg sg = segs[bk.x];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int k = bk.y + tx;
    double s = 0.0;
    if (k < sg.stride)
        for (int j = ty; j < sg.n; j += 16) s += partial[sg.src + (long long)j * sg.stride + k];
    rows[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && k < sg.stride) {
        double t = 0.0;
        for (int r = 0; r < 16; r++) t += rows[r][tx];
        out[sg.dst + k] = t;
    }
}

template <typename R, int T, int W>
static hipError_t launch_bt(const BatchInst<R>* insts, const BatchItem* items, int nitems, int* counter,
                            double* partial, cpx<R>* scratch, long long scratch_stride, int grid, size_t lds,
                            hipStream_t s)
{
    auto k = batch_kernel<R, T, W>;
    // a negative grid only sets the kernel's LDS attribute (before a stream capture)
    if (grid < 0) return lds > 65536 ? hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)
                                     : hipSuccess;
    if (lds > 65536 && !hip_stream_capturing(s)) {
        hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(1), 0, s, counter);
    hipLaunchKernelGGL(k, dim3(grid), dim3(T), lds, s, insts, items, nitems, counter, partial, scratch, scratch_stride);
    return hipGetLastError();
}

bool batch_variant_exists(int precision_fp32, int threads, int wbits)
{
    if (precision_fp32)
        return (threads == 256 && wbits >= 2 && wbits <= 4) || (threads == 512 && (wbits == 2 || wbits == 3));
    return (threads == 256 && (wbits == 2 || wbits == 3)) || (threads == 512 && wbits == 2);
}

template <>
hipError_t launch_batch<float>(const BatchInst<float>* insts, const BatchItem* items, int nitems, int* counter,
                               double* partial, cpx<float>* scratch, long long scratch_stride, int threads, int wbits,
                               int grid, size_t lds, hipStream_t s)
{
#define MG_BV(TT, WW)                                                                                          \
    if (threads == TT && wbits == WW)                                                                          \
        return launch_bt<float, TT, WW>(insts, items, nitems, counter, partial, scratch, scratch_stride, grid, lds, s);
    MG_BV(256, 4) MG_BV(256, 3) MG_BV(256, 2) MG_BV(512, 3) MG_BV(512, 2)
    return hipErrorInvalidValue;
}

template <>
hipError_t launch_batch<double>(const BatchInst<double>* insts, const BatchItem* items, int nitems, int* counter,
                                double* partial, cpx<double>* scratch, long long scratch_stride, int threads,
                                int wbits, int grid, size_t lds, hipStream_t s)
{
#define MG_BVD(TT, WW)                                                                                         \
    if (threads == TT && wbits == WW)                                                                          \
        return launch_bt<double, TT, WW>(insts, items, nitems, counter, partial, scratch, scratch_stride, grid, lds, s);
    MG_BVD(256, 3) MG_BVD(256, 2) MG_BVD(512, 2)
#undef MG_BVD
#undef MG_BV
    return hipErrorInvalidValue;
}

hipError_t launch_batch_reduce(const ReduceSeg* segs, const int2* blocks, int nblocks, const double* partial, double* out,
                               hipStream_t s)
{
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_reduce, dim3(nblocks), dim3(1024), 0, s, segs, blocks, partial, out);
    return hipGetLastError();
}

}  // namespace mg
