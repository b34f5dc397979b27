// Launch interface of the register-resident gate-window sweep (mg_sweep_win.hip).
#pragma once
#include "mg_device.hpp"
#include "mg_window.hpp"

namespace mg {

constexpr int kWinRegGates = 6;  // gates per window block at most (staging slots)

struct WinSweepArgs {
    int L, G, P, nblocks;
    const Inst* inst;          // [G] in block order; lo = window-local pair index
    const void* coef;          // cpx<R> [G][kCoefPerInst] in block order
    const WBlock* blk_lds;     // [nblocks] offsets swizzled for the LDS re-layout buffer
    const WBlock* blk_glob;    // [nblocks] plain offsets (seed rows, basis vectors)
    const void* seed;          // cpx<R> [S][S]: row b = -conj(U e_b)
    long b0, b1;               // column range
    double* partial;           // [gridDim.x][2 + 32 P]
    int acc_off, buf_off;      // byte offsets in dynamic LDS
    int rs;                    // states per re-layout round (FP32: 4; FP64: 2, or 1 for many parameters)
    int slots;                 // staging slots: gates of the largest window block (<= kWinRegGates)
    // FP64: the forward pass on the matrix cores over nfwin 4-qubit windows of the column held in
    // the re-layout buffer (struct DWin, fragments [window][8 or 16][lane], forward_windows64;
    // psi, and dpsi for the HVP, which needs two states per round), or none (fwin null)
    const void* fwin;
    const double* fwaf;
    int nfwin;
    const int* skip;  // nonzero *skip: the launch returns at once (as SweepArgs::skip); null: run
};

// Tiled variant (FP64, L > 12): the program is cut into tile blocks acting inside 12-qubit
// windows (tile_blocks); a tile (the 2^12 amplitudes sharing the bits outside the window) is
// loaded from the column's global slice into registers and swept through the tile block's own
// 3-qubit window blocks exactly as the L=12 kernel sweeps a whole column, then stored back.
struct WTBlock {
    int k0, k1;                         // window blocks [k0, k1) of this tile block
    int tpos[12];                       // the tile window's bit positions, ascending
    int other[20];                      // the other L - 12 bit positions, ascending
};

struct WinTileArgs {
    int L, G, P, ntb;
    const Inst* inst;          // [G] in block order; lo = window-local pair index
    const void* coef;          // cpx<double> [G][kCoefPerInst] in block order
    const WTBlock* tblocks;    // [ntb]
    const WBlock* blk_lds;     // window blocks: tile-local positions, offsets swizzled for the LDS buffer
    const WBlock* blk_glob;    // the same blocks with the amplitude-index images of the offsets
    const void* seed;          // cpx<double> [S][S]: row b = -conj(U e_b)
    long b0, b1;               // column range
    double* partial;           // [gridDim.x][2 + 32 P]
    void* scratch;             // cpx<double> [gridDim.x][nstate][S]
    int nstate;                // 1 (f), 2 (+ gradient), 4 (+ HVP)
    int acc_off, buf_off, rs;  // as WinSweepArgs (12-qubit tiles)
    int slots;                 // as WinSweepArgs
    void* rows;                // materialisation: cpx<double> [S][S] output rows (column b's state)
    long row0;                 // column b's seed (or materialised) row is row b - row0 of seed (rows)
};
hipError_t launch_sweep_win_tile(const WinTileArgs& a, int mode, int grid, size_t lds, hipStream_t s);
// rows[b] = -conj(W e_b) for the columns [b0, b1) of the program in `a` (the target's gates,
// P = 0, one state per re-layout round: rs = 1); seed, partial and scratch are not used
hipError_t launch_materialize_win_tile(const WinTileArgs& a, int grid, size_t lds, hipStream_t s);

// dynamic LDS of the kernel within `limit` bytes (and the accumulator and buffer offsets, and the
// states per re-layout round that fit: FP64 re-lays two states out per round, or one when the
// accumulator of many parameter gates leaves no room for two)
size_t win_sweep_lds_bytes(int L, int P, int fp32, size_t limit, size_t* acc_off, size_t* buf_off, int* rs,
                           int mode = 2, int slots = kWinRegGates);
// the kernel's threads x window amplitudes cover the column (L=12 in both precisions)
bool win_sweep_fits(int L, int fp32);
int win_sweep_window_bits(int fp32);
hipError_t launch_sweep_win(const WinSweepArgs& a, int mode, int fp32, int grid, size_t lds, hipStream_t s);

}  // namespace mg
