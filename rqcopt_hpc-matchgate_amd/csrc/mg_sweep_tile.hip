// Tiled column sweep for columns larger than the LDS (L >= 13; FP64 from L = 12 when the
// register window kernel does not apply).
//
// The gate program is regrouped on the host (tile_blocks) into blocks of consecutive gates that
// all act inside one window of kw qubits (kw = 11 FP64, 12 FP32), a window that always holds the 3
// lowest index bits so that a tile's amplitudes come in runs of 8 contiguous elements.  A column's
// states live in a per-workgroup global slice; per block, the workgroup walks the 2^(L-kw) tiles
// of the window (the amplitudes sharing the bits outside it): load the tile of every live state
// into LDS (2^kw amplitudes each, 128 KB in all), apply the block's gates there one after the
// other as the LDS-resident kernel does (quadruples, one barrier per gate, per-gate outer
// products through the transposing wave butterfly into the workgroup's FP64 accumulator), store
// the tile back.  A brickwall needs a handful of such passes (4 for L=14, depth 11, FP64)
// where one-gate or 3/4-qubit-window passes need one per gate or per 2-4 gates.
//
// Per column: forward blocks (psi, dpsi); each tile of the last forward block goes on at once
// through the first reverse block (the same gates backwards), with f = seed . psi_G and lam seeded
// from the dense target row read once for both; then the other reverse blocks (psi and dpsi
// uncomputed with V^{-1}).  Partial sums per workgroup as the other sweep kernels.
//
// Reference: per-column loop src/matchgate_target.c:788-837, forward
// src/matchgate_brickwall.c:52-78, backward kernel src/gate.c:659-772.
#include <algorithm>

#include "mg_fwd_mfma.hpp"
#include "mg_mat_mfma.hpp"
#include "mg_quad.hpp"
#include "mg_sweep_tile.hpp"
#include "mg_win.hpp"


namespace mg {

namespace {

// tile bits: FP64 11 (2^11 amplitudes x 4 states x 16 bytes = 128 KB of LDS), FP32 12
template <typename R>
constexpr int tile_kw() { return sizeof(R) == 8 ? 11 : 12; }

// threads per workgroup (log2): 512.  FP64: 256 measured the same.  FP32 HVP: its forward pass
// runs on the matrix cores, where 512 threads give each wave two column blocks (L=16 1.49 -> 1.47 s,
// L=14 77.4 -> 75.6 ms against 1024 threads with the forward pass on the VALU, which was the faster
// of the two VALU forms: 1.52 s against 1.63 s at 512).  FP32 f and f + gradient: two workgroups
// (two tiles of one or two states) share a CU.
template <typename R, int MODE>
constexpr int tile_lt() { return 9; }
// waves per SIMD the register allocation must allow: FP32 f + gradient 4 (two 512-thread
// workgroups per CU, at most 128 VGPRs)
template <typename R, int MODE>
constexpr int tile_min_waves() { return sizeof(R) == 4 && MODE == 1 ? 4 : 1; }

__device__ inline uint32_t scatter_bits(uint32_t x, const int* pos, int n)
{
    uint32_t r = 0;
    for (int i = 0; i < n; i++) r |= ((x >> i) & 1u) << pos[i];
    return r;
}

// the amplitude index of tile position tid + m T of a block's tiles: win_tid | wm[m] | tile base
template <int PER>
struct TileMap {
    uint32_t win_tid;
    uint32_t wm[PER];
};

template <int KW, int LT, int PER>
__device__ inline void tile_map(const TBlock* blk, int tid, TileMap<PER>* m)
{
    m->win_tid = scatter_bits((uint32_t)tid, blk->pos, LT);  // T = 2^LT threads
#pragma unroll
    for (int q = 0; q < PER; q++) m->wm[q] = scatter_bits((uint32_t)q << LT, blk->pos, KW);
}

}  // namespace

// The forward pass of one FP32 tile over the block's 4-qubit windows on the matrix cores: psi
// (fwd_windows) for f and f + gradient, psi and dpsi <- M dpsi + N psi (fwd_windows_hvp) for the
// HVP; mg_fwd_mfma.hpp.

template <typename R, int MODE>
__global__ __launch_bounds__((1 << tile_lt<R, MODE>()), (tile_min_waves<R, MODE>())) void sweep_tile_kernel(TileArgs<R> a)
{
    constexpr int LT = tile_lt<R, MODE>(), T = 1 << LT, NW = T / 64;
    constexpr int NV = MODE == 2 ? 32 : 16;
    constexpr int SHIFT = 6 - log2c<NV>();
    constexpr int KW = tile_kw<R>();
    constexpr uint32_t TS = 1u << KW, Qt = TS >> 2;
    constexpr int PER = (int)(TS / T);
    extern __shared__ __align__(16) unsigned char smem[];
    const int L = a.L;
    const uint32_t S = 1u << L;
    const uint32_t ntiles = 1u << (L - KW);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

    const int nacc = a.slot ? a.nacc : a.P;
    double* acc = reinterpret_cast<double*>(smem);  // [nacc][32]
    double* stage = acc + NV * nacc;                // [2][NW][NV]
    double* fst = stage + 2 * NW * NV;              // [NW][2]
    cpx<R>* tl = reinterpret_cast<cpx<R>*>(smem + a.tile_off);
    // tile slots and global slices: psi, dpsi (MODE 2), lam (MODE >= 1), dlam (MODE 2)
    constexpr int LAM = MODE == 2 ? 2 : 1;
    cpx<R>* t_psi = tl;
    cpx<R>* t_dpsi = tl + TS;
    cpx<R>* t_lam = tl + LAM * TS;
    cpx<R>* t_dlam = tl + 3 * TS;
    cpx<R>* sc = a.scratch + (size_t)blockIdx.x * a.nstate * S;
    cpx<R>* g_psi = sc;
    cpx<R>* g_dpsi = sc + S;
    cpx<R>* g_lam = sc + (size_t)LAM * S;
    cpx<R>* g_dlam = sc + 3 * (size_t)S;

    double* out = a.partial + (size_t)blockIdx.x * (2 + 32 * a.P);
    for (int k = tid; k < NV * nacc; k += T) acc[k] = 0.0;
    if (a.slot)
        for (int k = tid; k < 32 * a.P; k += T) out[2 + k] = 0.0;
    double fre = 0.0, fim = 0.0;
    __syncthreads();

    const int nb = a.nblocks;
    for (long b = a.b0 + blockIdx.x; b < a.b1; b += gridDim.x) {
        const size_t rb = (size_t)(b - a.row0) * S;
        // seed row entry (FP32 sweeps may read the FP64 rows and round on load: the same values
        // as an FP32 copy of the rows)
        auto row = [&](uint32_t i) -> cpx<R> {
            if constexpr (sizeof(R) == 4) {
                if (a.seed64) {
                    const cpx<double> v = a.seed64[rb + i];
                    return {(float)v.re, (float)v.im};
                }
            }
            return a.seed[rb + i];
        };
        double fr = 0.0, fi = 0.0;
        // the reverse sweep of one tile through block bk's gates (lam, psi, dlam, dpsi in LDS)
        auto reverse_tile = [&](const TBlock* bk) {
            for (int g = bk->g1 - 1; g >= bk->g0; g--) {
                const Inst in = a.inst[g];
                cpx<R> cf[24];
                load_coef_c<R, MODE == 2 ? 24 : 16>(a.coef + (size_t)g * kCoefPerInst, cf);
                cpx<R> ga[8], ha[8];
                const uint32_t olo = swz_idx<R, true>(1u << in.lo), ohi = swz_idx<R, true>(1u << in.hi);
                rev_gate<R, MODE, true, (2 * T <= Qt)>(tid, T, Qt, in.lo, in.hi, olo, ohi, cf, t_lam, t_psi, t_dlam,
                                                       t_dpsi, ga, ha);
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
                const int buf = g & 1;
                if ((lane & ((1 << SHIFT) - 1)) == 0) stage[(buf * NW + wv) * NV + (lane >> SHIFT)] = (double)v[0];
                __syncthreads();
                if (wv == g % NW && lane < NV) {
                    double s = 0.0;
                    for (int w = 0; w < NW; w++) s += stage[(buf * NW + w) * NV + lane];
                    const int ci = lane >> 1;
                    int e = ci & 7;
                    if (in.flip && e < 4) e = 3 - e;
                    acc[NV * (a.slot ? a.slot[g] : in.param) + 2 * ((ci & 8) | e) + (lane & 1)] += s;
                }
            }
        };
        // the tile's states back to the column's slices (each thread its own positions)
        auto store_reverse = [&](const TileMap<PER>& mp, uint32_t ot) {
#pragma unroll
            for (int m = 0; m < PER; m++) {
                const uint32_t j = tid + m * T, gi = ot | mp.win_tid | mp.wm[m];
                g_psi[gi] = sel<R, true>(t_psi, j);
                g_lam[gi] = sel<R, true>(t_lam, j);
                if constexpr (MODE == 2) {
                    g_dpsi[gi] = sel<R, true>(t_dpsi, j);
                    g_dlam[gi] = sel<R, true>(t_dlam, j);
                }
            }
        };
        // after a reverse block (per-block slots): this block's sums of this column into the
        // workgroup's partial sums
        auto flush_slots = [&](const TBlock* bk) {
            if (a.slot) {
                const int* bp = a.bparam + bk->s0;
                for (int k2 = tid; k2 < NV * bk->ns; k2 += T) {
                    out[2 + 32 * bp[k2 / NV] + (k2 % NV)] += acc[k2];
                    acc[k2] = 0.0;
                }
                __syncthreads();
            }
        };
        // ---- passes over the blocks: forward 0 .. nb-2 (psi, dpsi), then the last block forward
        // and at once backwards, tile by tile (f = s_b . psi_G, lam seeded from the target row read
        // once for both, no store and reload of psi between the two sweeps), then reverse nb-2 .. 0
        // (lam, psi uncomputed, dlam, dpsi).  One loop, so that the forward and the reverse gate
        // code each appear once (their registers are not live together).
        const int npass = MODE >= 1 ? 2 * nb - 1 : nb;
        for (int ps = 0; ps < npass; ps++) {
            const int k = ps < nb ? ps : 2 * nb - 2 - ps;
            const bool fwd = ps < nb, rev = MODE >= 1 && ps >= nb - 1;
            const TBlock* bk = a.blocks + k;
            TileMap<PER> mp;
            tile_map<KW, LT, PER>(bk, tid, &mp);
            for (uint32_t t = 0; t < ntiles; t++) {
                const uint32_t ot = scatter_bits(t, bk->other, L - KW);
#pragma unroll
                for (int m = 0; m < PER; m++) {
                    const uint32_t j = tid + m * T, gi = ot | mp.win_tid | mp.wm[m];
                    if (!fwd) {
                        sel<R, true>(t_psi, j) = g_psi[gi];
                        if constexpr (MODE >= 1) sel<R, true>(t_lam, j) = g_lam[gi];
                        if constexpr (MODE == 2) {
                            sel<R, true>(t_dpsi, j) = g_dpsi[gi];
                            sel<R, true>(t_dlam, j) = g_dlam[gi];
                        }
                    } else if (k == 0) {
                        sel<R, true>(t_psi, j) = {R(gi == (uint32_t)b), R(0)};
                        if constexpr (MODE == 2) sel<R, true>(t_dpsi, j) = {R(0), R(0)};
                    } else {
                        sel<R, true>(t_psi, j) = g_psi[gi];
                        if constexpr (MODE == 2) sel<R, true>(t_dpsi, j) = g_dpsi[gi];
                    }
                }
                __syncthreads();
                if (fwd) {
                    if constexpr (sizeof(R) == 4) {
                        if (a.fwin) {
                            if constexpr (MODE == 2)
                                fwd_windows_hvp<NW>(reinterpret_cast<const FWin*>(a.fwin), a.fwaf, bk->w0, bk->w1, t_psi,
                                                    t_dpsi, lane, wv);
                            else
                                fwd_windows<NW>(reinterpret_cast<const FWin*>(a.fwin), a.fwaf, bk->w0, bk->w1, t_psi, lane,
                                                wv);
                            goto forward_done;
                        }
                    }
                    for (int g = bk->g0; g < bk->g1; g++) {
                        const Inst in = a.inst[g];
                        cpx<R> cf[24];
                        load_coef_c<R, 8>(a.coef + (size_t)g * kCoefPerInst, cf);
                        if constexpr (MODE == 2) load_coef_c<R, 8>(a.coef + (size_t)g * kCoefPerInst + 16, cf + 16);
                        const uint32_t olo = swz_idx<R, true>(1u << in.lo), ohi = swz_idx<R, true>(1u << in.hi);
                        uint32_t q = tid;
                        if constexpr (MODE == 2 && 2 * T <= Qt)
                            for (; q + T < Qt; q += 2 * T)
                                fwd_quads<R, MODE, true, 2>(q, T, in.lo, in.hi, olo, ohi, cf, t_psi, t_psi, t_dpsi, t_dpsi);
                        for (; q < Qt; q += T)
                            fwd_quads<R, MODE, true, 1>(q, T, in.lo, in.hi, olo, ohi, cf, t_psi, t_psi, t_dpsi, t_dpsi);
                        __syncthreads();
                    }
                forward_done:
                    if (k < nb - 1) {
                        // store (each thread its own positions, so the next tile's load needs no barrier)
#pragma unroll
                        for (int m = 0; m < PER; m++) {
                            const uint32_t j = tid + m * T, gi = ot | mp.win_tid | mp.wm[m];
                            g_psi[gi] = sel<R, true>(t_psi, j);
                            if constexpr (MODE == 2) g_dpsi[gi] = sel<R, true>(t_dpsi, j);
                        }
                        continue;
                    }
                    // the last block: f_b = s_b . psi_G, and lam = s_b, dlam = 0 for the reverse sweep
#pragma unroll
                    for (int m = 0; m < PER; m++) {
                        const uint32_t j = tid + m * T, gi = ot | mp.win_tid | mp.wm[m];
                        const cpx<R> r = row(gi);
                        const cpx<R> tt = cmul(r, sel<R, true>(t_psi, j));
                        fr += (double)tt.re;
                        fi += (double)tt.im;
                        if constexpr (MODE >= 1) sel<R, true>(t_lam, j) = r;
                        if constexpr (MODE == 2) sel<R, true>(t_dlam, j) = {R(0), R(0)};
                    }
                    if constexpr (MODE == 0) continue;
                    __syncthreads();
                }
                if constexpr (MODE >= 1) {
                    if (rev) {
                        reverse_tile(bk);
                        if (k > 0) store_reverse(mp, ot);
                    }
                }
            }
            // the next block reads, in its own mapping, what other threads stored here
            __syncthreads();
            if constexpr (MODE >= 1) {
                if (rev) flush_slots(bk);
            }
        }
        {
            const double sr = wave_sum(fr), si = wave_sum(fi);
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
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0) {
        out[0] = fre;
        out[1] = fim;
    }
    if (!a.slot)
        for (int k = tid; k < 32 * a.P; k += T) out[2 + k] = (k & 31) < NV ? acc[(k >> 5) * NV + (k & 31)] : 0.0;
}

int tile_window_bits(int fp32) { return fp32 ? tile_kw<float>() : tile_kw<double>(); }

size_t tile_sweep_lds_bytes(int P, int kw, int fp32, int nstate, size_t* tile_off)
{
    // nstate 4: the HVP mode (32 accumulated doubles per parameter), else 16
    const int hvp = nstate == 4;
    const int NW = (1 << (fp32 ? (hvp ? tile_lt<float, 2>() : tile_lt<float, 1>()) : tile_lt<double, 2>())) / 64;
    const int NV = hvp ? 32 : 16;
    const size_t head = sizeof(double) * (NV * (size_t)P + 2 * NW * NV + 2 * NW);
    const size_t off = (head + 255) & ~(size_t)255;
    if (tile_off) *tile_off = off;
    return off + (size_t)nstate * ((size_t)1 << kw) * (fp32 ? sizeof(cpx<float>) : sizeof(cpx<double>));
}

template <typename R, int MODE>
static hipError_t launch_tile_m(const TileArgs<R>& a, int grid, size_t lds, hipStream_t s)
{
    auto k = sweep_tile_kernel<R, MODE>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(1 << tile_lt<R, MODE>()), lds, s, a);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_sweep_tile(const TileArgs<R>& a, int mode, int grid, size_t lds, hipStream_t s)
{
    // the kernel's tile is 2^kw amplitudes, kw fixed per precision, and the column larger
    if (a.kw != tile_kw<R>() || a.kw >= a.L || a.L > 30 || a.L - a.kw > 32 - kTileMaxBits) return hipErrorInvalidValue;
    if (mode == 2) return launch_tile_m<R, 2>(a, grid, lds, s);
    if (mode == 1) return launch_tile_m<R, 1>(a, grid, lds, s);
    if (mode == 0) return launch_tile_m<R, 0>(a, grid, lds, s);
    return hipErrorInvalidValue;
}

template hipError_t launch_sweep_tile<double>(const TileArgs<double>&, int, int, size_t, hipStream_t);
template hipError_t launch_sweep_tile<float>(const TileArgs<float>&, int, int, size_t, hipStream_t);

}  // namespace mg
