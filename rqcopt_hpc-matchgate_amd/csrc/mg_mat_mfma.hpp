// Target materialisation on the matrix cores (mg_mat_mfma.hip).
#pragma once
#include <vector>

#include "mg_device.hpp"
#include "mg_sweep_win.hpp"

namespace mg {

// One 4-qubit window of a tile block.  A product of matchgates keeps the parity of the window's
// bits, so the window's 16 x 16 matrix is two 8 x 8 blocks, one per parity; each block, written
// as a real 16 x 16 matrix over (real, imaginary) parts, is one FP64 MFMA tile.  Offsets are
// indices into the LDS tile (12 bits), already swizzled (the swizzle is linear over XOR, so the
// index of window entry n of column j of column block cb is offp ^ offj ^ offcb).
// LDS index of tile entry j: the bits above the lowest four folded onto them, so that the
// columns of a block (any 4 of the tile's bits) fall in different banks; linear over XOR
__host__ __device__ inline uint32_t mat_swz(uint32_t j) { return j ^ ((j >> 4) & 15u) ^ ((j >> 8) & 15u); }

struct DWin {
    int offp[2][8];  // parity block p, parity index n (window indices ascending within a parity)
    int offj[16];    // column j of a column block (the 4 lowest of the 8 bits outside the window)
    int offcb[16];   // column block cb (the 4 highest)
};

// the target program in that form: tile blocks (WTBlock: window range [k0, k1), the tile's 12
// positions and the other L - 12), windows, and the MFMA A-operand fragments of every window,
// [window][4 p + kk][lane] doubles (row lane % 16, column 4 kk + lane / 16 of parity block p's
// real matrix; columns 0-3: real parts of entries 0-3, 4-7 their imaginary parts, 8-15 the same
// for entries 4-7; rows g, g + 4, g + 8, g + 12 (g < 4): real and imaginary parts of entries 2 g and
// 2 g + 1, which the FP64 MFMA returns as results 0-3 of lanes 16 g .. 16 g + 15)
struct DenseWindows {
    std::vector<WTBlock> tblocks;
    std::vector<DWin> wins;
    std::vector<double> afrag;   // FP64 MFMA
    std::vector<float> afrag32;  // FP32 MFMA (its own result layout, rows 4 g .. 4 g + 3)
    int ngates = 0;
};

// fuses the target's instances (program order, global bit positions, coefficients V first) into
// 4-qubit windows inside 12-qubit tile blocks; false when the blocking fails (L <= 12)
bool make_dense_windows(const std::vector<Inst>& inst, const std::vector<cpx<double>>& coef, int L, DenseWindows* d);

// The forward pass of an f or f + gradient sweep (psi <- V psi per gate) over 4-qubit windows of a
// tile block: per window psi_out = M psi_in with M the product of the window's V, which keeps parity
// and is two 8 x 8 blocks, FP32 MFMA tiles as for the materialisation.  Offsets are LDS indices of
// the FP32 tile kernel (lds_swz, linear over XOR).
using FWin = DWin;  // the same offsets, into the FP32 tile kernel's LDS layout (lds_swz)

// `inst` (tile-local bit positions) and `coef` (V first in each instance's coefficients) are one
// tile block's gates in program order; fills `order` with a window order of them, appends the
// windows and their FP32 fragments [window][8][lane] and returns the number of windows.  With
// `hvp`, the forward pass of the Hessian-vector product, whose direction dpsi follows
// dpsi <- V dpsi + Z psi per gate (Z the gate's direction coefficients, 16..23): per window
// dpsi_out = M dpsi_in + N psi_in with N = sum_i V_k..V_{i+1} Z_i V_{i-1}..V_1, which keeps parity
// like M; the fragments are then [window][16][lane], M's 8 then N's 8.
int forward_windows(const std::vector<Inst>& inst, const std::vector<cpx<double>>& coef, std::vector<int>* order,
                    std::vector<FWin>* wins, std::vector<float>* af, bool hvp = false);

// one gate of a forward-pass window: program instance, window bits of its lo and hi wires
struct WinGate {
    int inst, la, lb, pad;
};

// The window form of the forward pass of a 12-qubit column (L = 12) without the matrices: the
// program's gates cut into 4-qubit windows (window_blocks), per window its gates in order
// (gates[wstart[w] .. wstart[w + 1])), the LDS offsets of each window (`fp64`: the FP64 MFMA
// layout of the materialisation, mat_swz; else the FP32 tile kernel's lds_swz) and the window
// order of the instances (`order`, new -> old).  Returns the number of windows, or -1.
int forward_window_form(const std::vector<Inst>& inst, bool fp64, std::vector<DWin>* wins, std::vector<int>* wstart,
                        std::vector<WinGate>* gates, std::vector<int>* order);

// The windows' A fragments on the device, from this evaluation's coefficients (`coef`: FP64,
// [program instance][kCoefPerInst]): block w forms window w's M (and with `hvp` N) column by
// column and writes [window][8 or 16][lane] fragments in the FP64 (out64) or FP32 (out32) MFMA
// result layout, the same values forward_windows makes on the host.  Nonzero *skip: the launch
// returns at once (a launch inside a captured trust-region step whose truncated CG has ended)
hipError_t launch_window_frags(int nwin, const int* wstart, const WinGate* gates, const cpx<double>* coef, int hvp,
                               double* out64, float* out32, hipStream_t s, const int* skip = nullptr);

struct MatArgs {
    int L, ntb;
    const WTBlock* tblocks;
    const DWin* wins;
    const void* afrag;  // double (FP64 rows) or float (FP32 rows) fragments
    long b0, b1;        // rows [b0, b1)
    void* rows;         // cpx<double> or cpx<float>: row b at rows + (b - row0) S, -conj(W e_b)
    long row0;
};
// one 12-qubit tile of the rows' precision in LDS
inline size_t mat_mfma_lds(int fp32) { return ((size_t)1 << 12) * (fp32 ? sizeof(cpx<float>) : sizeof(cpx<double>)); }
hipError_t launch_materialize_mfma(const MatArgs& a, int fp32, int grid, hipStream_t s);

}  // namespace mg
