// Target materialisation on the matrix cores: rows[b] = -conj(W e_b) for a target program W
// beyond 12 qubits.
//
// The program is cut into 12-qubit tile blocks (tile_blocks, as the tiled sweeps) and, inside
// each, into 4-qubit windows (window_blocks).  Every gate is a matchgate: it keeps the parity of
// its two bits, so a window's 16 x 16 product keeps the parity of the window's four bits and is
// two 8 x 8 blocks.  In real arithmetic an 8 x 8 complex block is one 16 x 16 matrix, applied to
// 16 columns (the amplitudes sharing the other 8 bits of the tile) by four
// v_mfma_f64_16x16x4f64: 8 complex MACs per amplitude and window where the 2 x 2 blocks of its
// ~3.6 gates take ~7.2, at the FP64 matrix rate.
//
// Per row: per tile block, per tile (2^(L-12) of them): load the tile of the row (the basis
// vector e_b in the first block) into LDS, apply the block's windows (one barrier each; every
// wave owns 4 of the tile's 16 column blocks, so a window updates its entries in place), store
// the tile back (with -conj in the last block).  Reference: the target map the reference calls
// per column, src/matchgate_target.c:795, for a brickwall target
// examples/hubbard1d/hubbard1d_opt.c:30-39 over src/matchgate_brickwall.c:14-43.
#include <algorithm>
#include <complex>

#include "mg_mat_mfma.hpp"
#include "mg_sweep_tile.hpp"
#include "mg_window.hpp"

namespace mg {

namespace {

// one 16x16x4 MFMA per precision: D += A B with lane l supplying A[l % 16][l / 16] and
// B[l / 16][l % 16]; FP64 returns rows 4 v + l / 16 of column l % 16 as results v = 0..3, FP32
// rows 4 (l / 16) + v (both checked on the device; the host orders the A rows to match)
template <typename R>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mac(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mac(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

__device__ inline uint32_t scatter(uint32_t x, const int* pos, int n)
{
    uint32_t r = 0;
    for (int i = 0; i < n; i++) r |= ((x >> i) & 1u) << pos[i];
    return r;
}

}  // namespace

template <typename R>
__global__ __launch_bounds__(256) void materialize_mfma_kernel(MatArgs a)
{
    typedef typename Mfma<R>::v4 v4;
    extern __shared__ __align__(16) unsigned char smem[];
    cpx<R>* tl = reinterpret_cast<cpx<R>*>(smem);  // [4096], index mat_swz(j)
    const R* afrag = reinterpret_cast<const R*>(a.afrag);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int L = a.L;
    const uint32_t S = 1u << L, ntiles = 1u << (L - 12);
    const int n0 = lane >> 4, col = lane & 15;
    cpx<R>* rows = reinterpret_cast<cpx<R>*>(a.rows);

    for (long b = a.b0 + blockIdx.x; b < a.b1; b += gridDim.x) {
        cpx<R>* row = rows + (size_t)(b - a.row0) * S;
        for (int k = 0; k < a.ntb; k++) {
            const WTBlock* tb = a.tblocks + k;
            const uint32_t lo = scatter((uint32_t)tid, tb->tpos, 8);
            const bool last = k == a.ntb - 1;
            for (uint32_t t = 0; t < ntiles; t++) {
                const uint32_t ot = scatter(t, tb->other, L - 12);
#pragma unroll 4
                for (uint32_t m = 0; m < 16; m++) {
                    const uint32_t gi = ot | lo | scatter(m, tb->tpos + 8, 4);
                    const uint32_t j = (uint32_t)tid | (m << 8);
                    tl[mat_swz(j)] = k == 0 ? cpx<R>{R(gi == (uint32_t)b ? 1 : 0), R(0)} : row[gi];
                }
                __syncthreads();
                for (int w = tb->k0; w < tb->k1; w++) {
                    const DWin* dw = a.wins + w;
                    R af[8];
#pragma unroll
                    for (int f = 0; f < 8; f++) af[f] = afrag[((size_t)w * 8 + f) * 64 + lane];
                    const uint32_t oj = (uint32_t)dw->offj[col];
                    // the wave's 8 products (4 column blocks x 2 parities): every read first, so
                    // that the in-place writes follow them, then 8 independent MFMA chains
                    uint32_t base[4];
                    cpx<R> x[8][2];
#pragma unroll
                    for (int c4 = 0; c4 < 4; c4++) {
                        base[c4] = oj ^ (uint32_t)dw->offcb[wv * 4 + c4];
#pragma unroll
                        for (int p = 0; p < 2; p++) {
                            x[2 * c4 + p][0] = tl[base[c4] ^ (uint32_t)dw->offp[p][n0]];
                            x[2 * c4 + p][1] = tl[base[c4] ^ (uint32_t)dw->offp[p][4 + n0]];
                        }
                    }
                    v4 acc[8];
#pragma unroll
                    for (int c = 0; c < 8; c++) acc[c] = v4{R(0), R(0), R(0), R(0)};
#pragma unroll
                    for (int kk = 0; kk < 4; kk++)
#pragma unroll
                        for (int c = 0; c < 8; c++) {
                            const cpx<R>& xv = x[c][kk >> 1];
                            acc[c] = Mfma<R>::mac(af[4 * (c & 1) + kk], (kk & 1) ? xv.im : xv.re, acc[c]);
                        }
                    // this lane's results: entries 2 n0 and 2 n0 + 1 of its column
#pragma unroll
                    for (int c = 0; c < 8; c++) {
                        const int p = c & 1;
                        tl[base[c >> 1] ^ (uint32_t)dw->offp[p][2 * n0]] = {acc[c][0], acc[c][1]};
                        tl[base[c >> 1] ^ (uint32_t)dw->offp[p][2 * n0 + 1]] = {acc[c][2], acc[c][3]};
                    }
                    __syncthreads();
                }
#pragma unroll 4
                for (uint32_t m = 0; m < 16; m++) {
                    const uint32_t gi = ot | lo | scatter(m, tb->tpos + 8, 4);
                    const cpx<R> v = tl[mat_swz((uint32_t)tid | (m << 8))];
                    row[gi] = last ? cpx<R>{-v.re, v.im} : v;
                }
                // the next tile's load reuses the LDS tile
                __syncthreads();
            }
        }
    }
}

hipError_t launch_materialize_mfma(const MatArgs& a, int fp32, int grid, hipStream_t s)
{
    // 32-bit amplitude indices: at most 30 qubits (the tiled sweeps' limit too)
    if (a.L <= 12 || a.L > 30) return hipErrorInvalidValue;
    if (fp32)
        hipLaunchKernelGGL(materialize_mfma_kernel<float>, dim3(grid), dim3(256), mat_mfma_lds(1), s, a);
    else
        hipLaunchKernelGGL(materialize_mfma_kernel<double>, dim3(grid), dim3(256), mat_mfma_lds(0), s, a);
    return hipGetLastError();
}

// ---- window fragments on the device -------------------------------------------------------

// Block w forms window w's matrices: thread t < 16 owns column n = t % 8 of parity block p = t / 8
// (entry (m, n) of block p is the window entry (kParIdx[p][m], kParIdx[p][n]); the position of
// window index i in its parity block is i / 2).  Left-multiplying by a gate mixes a column's
// entries only among themselves, so each thread applies the window's gates to its own column,
// M <- V M and N <- V N + Z M (with the M before the gate), then writes the A fragments that
// read its column.  Fragment slot (4 p + kk, lane) holds row r = lane % 16, column c = 4 kk +
// lane / 16 of block p's real 16 x 16 matrix: column c is entry n = 4 (c / 8) + c % 4 of the input,
// real (c / 4 even) or imaginary part, and row r is the real or imaginary part of output entry m
// (FP64 result layout: m = 2 (r % 4) + (r / 4) / 2, part (r / 4) % 2; FP32: m = 2 (r / 4) +
// (r % 4) / 2, part (r % 4) % 2).
__global__ __launch_bounds__(64) void window_frags_kernel(const int* wstart, const WinGate* gates,
                                                          const cpx<double>* coef, int hvp, double* out64,
                                                          float* out32, const int* skip)
{
    __shared__ double col[16][4][8];  // per thread: M re, M im, N re, N im of its column
    if (skip && *skip) return;
    const int w = blockIdx.x, t = threadIdx.x;
    if (t >= 16) return;  // no barrier below: every thread works on its own column
    const int p = t >> 3, n = t & 7;
    double* mr = col[t][0];
    double* mi = col[t][1];
    double* nr = col[t][2];
    double* ni = col[t][3];
    for (int m = 0; m < 8; m++) {
        mr[m] = m == n ? 1.0 : 0.0;
        mi[m] = nr[m] = ni[m] = 0.0;
    }
    for (int j = wstart[w]; j < wstart[w + 1]; j++) {
        const WinGate g = gates[j];
        const cpx<double>* v = coef + (size_t)g.inst * kCoefPerInst;
        const cpx<double>* z = v + 16;
        const int bl = 1 << g.la, bh = 1 << g.lb;
        for (int r = 0; r < 16; r++) {
            if (r & (bl | bh)) continue;
            // rows r, r | both bits keep the parity of r and mix by the corner block; rows with one
            // of the bits set have the other parity and mix by the center block
            const bool same = (__builtin_popcount(r) & 1) == p;
            const int a = (same ? r : (r | bl)) >> 1, b = (same ? (r | bl | bh) : (r | bh)) >> 1;
            const cpx<double>* c = same ? v + 4 : v;
            const cpx<double>* zz = same ? z + 4 : z;
            const double xr = mr[a], xi = mi[a], wr = mr[b], wi = mi[b];
            if (hvp) {
                const double yr = nr[a], yi = ni[a], ur = nr[b], ui = ni[b];
                nr[a] = ((c[0].re * yr - c[0].im * yi) + (c[1].re * ur - c[1].im * ui)) +
                        ((zz[0].re * xr - zz[0].im * xi) + (zz[1].re * wr - zz[1].im * wi));
                ni[a] = ((c[0].re * yi + c[0].im * yr) + (c[1].re * ui + c[1].im * ur)) +
                        ((zz[0].re * xi + zz[0].im * xr) + (zz[1].re * wi + zz[1].im * wr));
                nr[b] = ((c[2].re * yr - c[2].im * yi) + (c[3].re * ur - c[3].im * ui)) +
                        ((zz[2].re * xr - zz[2].im * xi) + (zz[3].re * wr - zz[3].im * wi));
                ni[b] = ((c[2].re * yi + c[2].im * yr) + (c[3].re * ui + c[3].im * ur)) +
                        ((zz[2].re * xi + zz[2].im * xr) + (zz[3].re * wi + zz[3].im * wr));
            }
            mr[a] = (c[0].re * xr - c[0].im * xi) + (c[1].re * wr - c[1].im * wi);
            mi[a] = (c[0].re * xi + c[0].im * xr) + (c[1].re * wi + c[1].im * wr);
            mr[b] = (c[2].re * xr - c[2].im * xi) + (c[3].re * wr - c[3].im * wi);
            mi[b] = (c[2].re * xi + c[2].im * xr) + (c[3].re * wi + c[3].im * wr);
        }
    }
    const int NF = hvp ? 16 : 8;
    for (int X = 0; X <= (hvp ? 1 : 0); X++) {
        const double* re = X ? nr : mr;
        const double* im = X ? ni : mi;
        for (int q = 0; q < 2; q++) {
            const int cc = 8 * (n >> 2) + 4 * q + (n & 3), kk = cc >> 2;
            for (int r = 0; r < 16; r++) {
                const int ln = 16 * (cc & 3) + r;
                const int m = out64 ? 2 * (r & 3) + ((r >> 2) >> 1) : 2 * (r >> 2) + ((r & 3) >> 1);
                const int part = out64 ? (r >> 2) & 1 : (r & 3) & 1;
                const double val = part == 0 ? (q == 0 ? re[m] : -im[m]) : (q == 0 ? im[m] : re[m]);
                const size_t idx = ((size_t)w * NF + 8 * X + 4 * p + kk) * 64 + ln;
                if (out64) out64[idx] = val;
                else out32[idx] = (float)val;
            }
        }
    }
}

hipError_t launch_window_frags(int nwin, const int* wstart, const WinGate* gates, const cpx<double>* coef, int hvp,
                               double* out64, float* out32, hipStream_t s, const int* skip)
{
    if (nwin < 1 || (!out64) == (!out32)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(window_frags_kernel, dim3(nwin), dim3(64), 0, s, wstart, gates, coef, hvp, out64, out32, skip);
    return hipGetLastError();
}

// ---- host: the window form ----------------------------------------------------------------

namespace {

// window indices of each parity, ascending
constexpr int kParIdx[2][8] = {{0, 3, 5, 6, 9, 10, 12, 15}, {1, 2, 4, 7, 8, 11, 13, 14}};

using zc = std::complex<double>;

// M <- G M for the gate with coefficients g (V: center block g0..g3 on {lo set, hi set}, corner
// g4..g7 on {none, both}) on window-local bits la (its lower bit) and lb
// a b + c d, written out (std::complex products call the library's IEEE-special-case routine,
// which would dominate the host time of the per-evaluation window matrices)
inline zc mul2(const cpx<double>& a, const zc& b, const cpx<double>& c, const zc& d)
{
    return {(a.re * b.real() - a.im * b.imag()) + (c.re * d.real() - c.im * d.imag()),
            (a.re * b.imag() + a.im * b.real()) + (c.re * d.imag() + c.im * d.real())};
}

void apply_gate_left(zc* M, const cpx<double>* g, int la, int lb)
{
    for (int r = 0; r < 16; r++) {
        if (r & ((1 << la) | (1 << lb))) continue;
        const int i00 = r, i01 = r | (1 << la), i10 = r | (1 << lb), i11 = r | (1 << la) | (1 << lb);
        for (int c = 0; c < 16; c++) {
            const zc x = M[i00 * 16 + c], y = M[i01 * 16 + c], z = M[i10 * 16 + c], w = M[i11 * 16 + c];
            M[i00 * 16 + c] = mul2(g[4], x, g[5], w);
            M[i01 * 16 + c] = mul2(g[0], y, g[1], z);
            M[i10 * 16 + c] = mul2(g[2], y, g[3], z);
            M[i11 * 16 + c] = mul2(g[6], x, g[7], w);
        }
    }
}

// A window's 16 x 16 matrix as its two 8 x 8 parity blocks, real and imaginary parts apart: entry
// (m, n) of block p is the entry (kParIdx[p][m], kParIdx[p][n]); the entries across parities are
// zero for a product of matchgates and are not stored
struct ParMat {
    double re[2][8][8], im[2][8][8];
    void set(bool identity)
    {
        for (int p = 0; p < 2; p++)
            for (int m = 0; m < 8; m++)
                for (int n = 0; n < 8; n++) {
                    re[p][m][n] = identity && m == n ? 1.0 : 0.0;
                    im[p][m][n] = 0.0;
                }
    }
};

// position of window index i in its parity block (the inverse of kParIdx)
struct ParPos {
    int pos[16];
    ParPos()
    {
        for (int p = 0; p < 2; p++)
            for (int m = 0; m < 8; m++) pos[kParIdx[p][m]] = m;
    }
};

// rows a, b of block p: M <- (c0 a + c1 b, c2 a + c3 b) and, with N, N <- the same of N's rows
// plus (z0 a + z1 b, z2 a + z3 b) of M's rows before the update
void mix_rows(ParMat& M, ParMat* N, int p, int a, int b, const cpx<double>* c, const cpx<double>* z)
{
    double* __restrict__ ar = M.re[p][a];
    double* __restrict__ ai = M.im[p][a];
    double* __restrict__ br = M.re[p][b];
    double* __restrict__ bi = M.im[p][b];
    if (N) {
        double* __restrict__ nar = N->re[p][a];
        double* __restrict__ nai = N->im[p][a];
        double* __restrict__ nbr = N->re[p][b];
        double* __restrict__ nbi = N->im[p][b];
        for (int n = 0; n < 8; n++) {
            const double xr = ar[n], xi = ai[n], wr = br[n], wi = bi[n];
            const double yr = nar[n], yi = nai[n], vr = nbr[n], vi = nbi[n];
            nar[n] = ((c[0].re * yr - c[0].im * yi) + (c[1].re * vr - c[1].im * vi)) +
                     ((z[0].re * xr - z[0].im * xi) + (z[1].re * wr - z[1].im * wi));
            nai[n] = ((c[0].re * yi + c[0].im * yr) + (c[1].re * vi + c[1].im * vr)) +
                     ((z[0].re * xi + z[0].im * xr) + (z[1].re * wi + z[1].im * wr));
            nbr[n] = ((c[2].re * yr - c[2].im * yi) + (c[3].re * vr - c[3].im * vi)) +
                     ((z[2].re * xr - z[2].im * xi) + (z[3].re * wr - z[3].im * wi));
            nbi[n] = ((c[2].re * yi + c[2].im * yr) + (c[3].re * vi + c[3].im * vr)) +
                     ((z[2].re * xi + z[2].im * xr) + (z[3].re * wi + z[3].im * wr));
        }
    }
    for (int n = 0; n < 8; n++) {
        const double xr = ar[n], xi = ai[n], wr = br[n], wi = bi[n];
        ar[n] = (c[0].re * xr - c[0].im * xi) + (c[1].re * wr - c[1].im * wi);
        ai[n] = (c[0].re * xi + c[0].im * xr) + (c[1].re * wi + c[1].im * wr);
        br[n] = (c[2].re * xr - c[2].im * xi) + (c[3].re * wr - c[3].im * wi);
        bi[n] = (c[2].re * xi + c[2].im * xr) + (c[3].re * wi + c[3].im * wr);
    }
}

// M <- V M and, with N, N <- V N + Z M (the M before the gate): v, z the gate's coefficients
// (center block 0-3 on the window indices with exactly one of the gate's bits set, corner block
// 4-7 on those with none or both), la, lb the window bits of the gate's lo and hi wires
void apply_gate_pair(ParMat& M, ParMat* N, const cpx<double>* v, const cpx<double>* z, int la, int lb)
{
    static const ParPos pp;
    for (int r = 0; r < 16; r++) {
        if (r & ((1 << la) | (1 << lb))) continue;
        const int i00 = r, i01 = r | (1 << la), i10 = r | (1 << lb), i11 = r | (1 << la) | (1 << lb);
        const int pr = __builtin_popcount(r) & 1;
        mix_rows(M, N, pr, pp.pos[i00], pp.pos[i11], v + 4, z + 4);
        mix_rows(M, N, 1 - pr, pp.pos[i01], pp.pos[i10], v, z);
    }
}

}  // namespace

namespace {

// the FP32 tile kernel's LDS index (lds_swz of cpx<float>: bits 4 and 5 folded into bits 0-3)
int tile_swz32(uint32_t a) { return (int)(a ^ (((a >> 4) & 1u) * 5u) ^ (((a >> 5) & 1u) * 10u)); }

// lane fragments of one parity block of a 16 x 16 window matrix, FP32 result layout (rows 4 g ..
// 4 g + 3: real and imaginary parts of entries 2 g and 2 g + 1)
// the 512 A-fragment values of a window matrix (2 parities x 4 kk x 64 lanes): matrix entry and
// which part (0: re, 1: -im, 2: im), for the FP64 (w = 0) and FP32 (w = 1) result layouts
struct FragTable {
    short entry[2][512];
    unsigned char part[2][512];
    FragTable()
    {
        for (int w = 0; w < 2; w++)
            for (int p = 0; p < 2; p++)
                for (int kk = 0; kk < 4; kk++)
                    for (int ln = 0; ln < 64; ln++) {
                        const int i = (p * 4 + kk) * 64 + ln;
                        const int r = ln % 16, c = 4 * kk + ln / 16;
                        const int n = (c / 8) * 4 + c % 4, q = (c / 4) % 2;
                        const int m = w == 0 ? 2 * (r % 4) + (r / 4) / 2 : 2 * (r / 4) + (r % 4) / 2;
                        const int pt = w == 0 ? (r / 4) % 2 : (r % 4) % 2;
                        entry[w][i] = (short)(p * 64 + m * 8 + n);
                        part[w][i] = (unsigned char)(pt == 0 ? (q == 0 ? 0 : 1) : (q == 0 ? 2 : 0));
                    }
    }
};
const FragTable& frag_table()
{
    static const FragTable t;
    return t;
}

template <typename T>
void push_fragments(const ParMat& M, int w, std::vector<T>* af)
{
    const FragTable& t = frag_table();
    const double* re = &M.re[0][0][0];
    const double* im = &M.im[0][0][0];
    const size_t n0 = af->size();
    af->resize(n0 + 512);
    T* o = af->data() + n0;
    for (int i = 0; i < 512; i++) {
        const int e = t.entry[w][i], pt = t.part[w][i];
        o[i] = (T)(pt == 0 ? re[e] : pt == 1 ? -im[e] : im[e]);
    }
}

void push_fragments32(const ParMat& M, std::vector<float>* af) { push_fragments<float>(M, 1, af); }

}  // namespace

int forward_windows(const std::vector<Inst>& inst, const std::vector<cpx<double>>& coef, std::vector<int>* order,
                    std::vector<FWin>* wins, std::vector<float>* af, bool hvp)
{
    constexpr int TB = 12, W = 4;
    std::vector<WBlock> wbs;
    window_blocks(inst, TB, W, 1 << 20, order, &wbs);
    if (order->size() != inst.size()) return -1;
    wins->reserve(wins->size() + wbs.size());
    af->reserve(af->size() + wbs.size() * (hvp ? 1024 : 512));
    for (const WBlock& wb : wbs) {
        ParMat M, N;
        M.set(true);
        N.set(false);
        for (int k = wb.g0; k < wb.g1; k++) {
            const Inst& l = inst[(*order)[k]];
            int la = -1, lb = -1;
            for (int i = 0; i < W; i++) {
                if (wb.pos[i] == l.lo) la = i;
                if (wb.pos[i] == l.hi) lb = i;
            }
            if (la < 0 || lb < 0) return -1;
            const cpx<double>* g = &coef[(size_t)(*order)[k] * kCoefPerInst];
            apply_gate_pair(M, hvp ? &N : nullptr, g, g + 16, la, lb);
        }
        FWin fw{};
        int oth[8], no = 0;
        for (int q = 0; q < TB; q++)
            if (std::find(wb.pos, wb.pos + W, q) == wb.pos + W) oth[no++] = q;
        auto put = [&](uint32_t x, const int* pos, int n) {
            uint32_t r = 0;
            for (int i = 0; i < n; i++) r |= ((x >> i) & 1u) << pos[i];
            return tile_swz32(r);
        };
        for (int p = 0; p < 2; p++)
            for (int n = 0; n < 8; n++) fw.offp[p][n] = put((uint32_t)kParIdx[p][n], wb.pos, W);
        for (int j = 0; j < 16; j++) {
            fw.offj[j] = put((uint32_t)j, oth, 4);
            fw.offcb[j] = put((uint32_t)j, oth + 4, 4);
        }
        wins->push_back(fw);
        push_fragments32(M, af);
        if (hvp) push_fragments32(N, af);
    }
    return (int)wbs.size();
}

int forward_window_form(const std::vector<Inst>& inst, bool fp64, std::vector<DWin>* wins, std::vector<int>* wstart,
                        std::vector<WinGate>* gates, std::vector<int>* order)
{
    constexpr int TB = 12, W = 4;
    std::vector<WBlock> wbs;
    window_blocks(inst, TB, W, 1 << 20, order, &wbs);
    if (order->size() != inst.size()) return -1;
    wins->clear();
    wstart->clear();
    gates->clear();
    for (const WBlock& wb : wbs) {
        wstart->push_back((int)gates->size());
        for (int k = wb.g0; k < wb.g1; k++) {
            const Inst& l = inst[(*order)[k]];
            WinGate g{(*order)[k], -1, -1, 0};
            for (int i = 0; i < W; i++) {
                if (wb.pos[i] == l.lo) g.la = i;
                if (wb.pos[i] == l.hi) g.lb = i;
            }
            if (g.la < 0 || g.lb < 0) return -1;
            gates->push_back(g);
        }
        DWin dw{};
        int oth[8], no = 0;
        for (int q = 0; q < TB; q++)
            if (std::find(wb.pos, wb.pos + W, q) == wb.pos + W) oth[no++] = q;
        auto put = [&](uint32_t x, const int* pos, int n) {
            uint32_t r = 0;
            for (int i = 0; i < n; i++) r |= ((x >> i) & 1u) << pos[i];
            return fp64 ? (int)mat_swz(r) : tile_swz32(r);
        };
        for (int p = 0; p < 2; p++)
            for (int n = 0; n < 8; n++) dw.offp[p][n] = put((uint32_t)kParIdx[p][n], wb.pos, W);
        for (int j = 0; j < 16; j++) {
            dw.offj[j] = put((uint32_t)j, oth, 4);
            dw.offcb[j] = put((uint32_t)j, oth + 4, 4);
        }
        wins->push_back(dw);
    }
    wstart->push_back((int)gates->size());
    return (int)wbs.size();
}

bool make_dense_windows(const std::vector<Inst>& inst, const std::vector<cpx<double>>& coef, int L, DenseWindows* d)
{
    constexpr int TB = 12, W = 4;
    d->tblocks.clear();
    d->wins.clear();
    d->afrag.clear();
    d->afrag32.clear();
    d->ngates = (int)inst.size();
    if (L <= TB || L - TB > 20) return false;
    std::vector<int> order1;
    std::vector<TBlock> tbs;
    tile_blocks(inst, L, TB, &order1, &tbs);
    if (order1.size() != inst.size()) return false;
    for (const TBlock& tb : tbs) {
        auto local = [&](int q) { return (int)(std::find(tb.pos, tb.pos + TB, q) - tb.pos); };
        std::vector<Inst> li;
        for (int k = tb.g0; k < tb.g1; k++) {
            Inst in = inst[order1[k]];
            in.lo = local(in.lo);
            in.hi = local(in.hi);
            if (in.lo >= TB || in.hi >= TB) return false;
            li.push_back(in);
        }
        std::vector<int> order2;
        std::vector<WBlock> wbs;
        window_blocks(li, TB, W, 1 << 20, &order2, &wbs);
        if (order2.size() != li.size()) return false;
        WTBlock t{};
        t.k0 = (int)d->wins.size();
        for (int i = 0; i < TB; i++) t.tpos[i] = tb.pos[i];
        for (int i = 0; i < L - TB; i++) t.other[i] = tb.other[i];
        for (const WBlock& wb : wbs) {
            // the window's dense matrix, gates in program order
            zc M[256];
            for (int i = 0; i < 256; i++) M[i] = (i / 16 == i % 16) ? 1.0 : 0.0;
            for (int k = wb.g0; k < wb.g1; k++) {
                const Inst& l = li[order2[k]];
                int la = -1, lb = -1;
                for (int i = 0; i < W; i++) {
                    if (wb.pos[i] == l.lo) la = i;
                    if (wb.pos[i] == l.hi) lb = i;
                }
                if (la < 0 || lb < 0) return false;
                apply_gate_left(M, &coef[(size_t)order1[tb.g0 + order2[k]] * kCoefPerInst], la, lb);
            }
            // entries across parities must vanish: the gates are matchgates
            for (int r = 0; r < 16; r++)
                for (int c = 0; c < 16; c++)
                    if ((__builtin_popcount(r) ^ __builtin_popcount(c)) & 1)
                        if (M[r * 16 + c] != 0.0) return false;
            // LDS offsets: window bit i is tile bit pos[i]; the other 8 tile bits ascending
            DWin dw{};
            int oth[8], no = 0;
            for (int q = 0; q < TB; q++)
                if (std::find(wb.pos, wb.pos + W, q) == wb.pos + W) oth[no++] = q;
            auto put = [&](uint32_t x, const int* pos, int n) {
                uint32_t r = 0;
                for (int i = 0; i < n; i++) r |= ((x >> i) & 1u) << pos[i];
                return (int)mat_swz(r);
            };
            for (int p = 0; p < 2; p++)
                for (int n = 0; n < 8; n++) dw.offp[p][n] = put((uint32_t)kParIdx[p][n], wb.pos, W);
            for (int j = 0; j < 16; j++) {
                dw.offj[j] = put((uint32_t)j, oth, 4);
                dw.offcb[j] = put((uint32_t)j, oth + 4, 4);
            }
            d->wins.push_back(dw);
            // A fragments: lane l holds row l % 16, column 4 kk + l / 16 of block p's real matrix.
            // Rows are ordered so that lane group g's four results are the real and imaginary
            // parts of entries 2 g and 2 g + 1: FP64 returns rows g, g + 4, g + 8, g + 12 there,
            // FP32 rows 4 g .. 4 g + 3.
            for (int p = 0; p < 2; p++)
                for (int kk = 0; kk < 4; kk++)
                    for (int ln = 0; ln < 64; ln++) {
                        const int r = ln % 16, c = 4 * kk + ln / 16;
                        const int n = (c / 8) * 4 + c % 4, q = (c / 4) % 2;  // input entry, re/im
                        auto value = [&](int m, int part) {
                            const zc e = M[kParIdx[p][m] * 16 + kParIdx[p][n]];
                            if (part == 0) return q == 0 ? e.real() : -e.imag();
                            return q == 0 ? e.imag() : e.real();
                        };
                        d->afrag.push_back(value(2 * (r % 4) + (r / 4) / 2, (r / 4) % 2));
                        d->afrag32.push_back((float)value(2 * (r / 4) + (r % 4) / 2, (r % 4) % 2));
                    }
        }
        t.k1 = (int)d->wins.size();
        d->tblocks.push_back(t);
    }
    return true;
}

}  // namespace mg
