// Launch interface of the tiled column sweep for columns larger than the LDS (mg_sweep_tile.hip).
#pragma once
#include <vector>

#include "mg_device.hpp"

namespace mg {

constexpr int kTileMaxBits = 12;  // qubits of a tile at most (2^12 amplitudes per state)
constexpr int kTileLowBits = 3;   // every tile holds the 3 lowest index bits: 8 contiguous amplitudes

// a block of consecutive gates acting inside one tile window: the window's kw bit positions
// (ascending; local bit i of a tile position is bit pos[i] of the amplitude index) and the
// L - kw others (ascending; bit i of a tile number is bit other[i] of the amplitude index)
struct TBlock {
    int g0, g1;
    int pos[kTileMaxBits];
    int other[32 - kTileMaxBits];
    int s0, ns;  // per-block accumulator slots: the block's parameters are bparam[s0, s0 + ns)
    int w0, w1;  // forward-pass windows of the block on the matrix cores (TileArgs.fwin)
};

template <typename R>
struct TileArgs {
    int L, G, P, kw, nblocks;
    const Inst* inst;       // [G] in block order; lo/hi = bit positions inside the block's window
    const cpx<R>* coef;     // [G][kCoefPerInst] in block order: V, V^{-1}, Z
    const TBlock* blocks;   // [nblocks]
    const cpx<R>* seed;     // [S][S]: row b = -conj(U e_b)
    const cpx<double>* seed64;  // FP32 sweeps without an FP32 copy of the rows: the FP64 rows,
                                // rounded to FP32 on load (seed unused); null otherwise
    long b0, b1;            // column range
    long row0;              // column b's seed row is row b - row0 of seed (seed64)
    double* partial;        // [gridDim.x][2 + 32 P]
    cpx<R>* scratch;        // [gridDim.x][nstate][S]: psi, dpsi, lam, dlam (as the mode needs)
    int nstate;             // states per column: 1 (f), 2 (f + gradient), 4 (+ HVP)
    int tile_off;           // byte offset of the tile states in dynamic LDS
    // accumulator: null slot = one [32] slot per parameter gate (P slots), kept over the columns;
    // else slot[g] is instance g's slot among its block's parameters (bparam), nacc slots, added
    // to the workgroup's partial sums after each block of each column
    const int* slot;
    const int* bparam;
    int nacc;
    // FP32 f and f + gradient: the forward pass on the matrix cores, 4-qubit windows [w0, w1) per block (struct FWin
    // of mg_mat_mfma.hpp; the block's instances in window order); null: gate by gate
    const void* fwin;
    const float* fwaf;  // [window][8][lane] A fragments of the windows
};

// Regroups the instances (program order; bit positions hi > lo) into blocks whose gates act
// inside windows of W bit positions that include the kTileLowBits lowest ones.  `order` gets the
// new order (a topological order of the gate dependencies) and `blocks` the blocks.
void tile_blocks(const std::vector<Inst>& inst, int L, int W, std::vector<int>* order, std::vector<TBlock>* blocks);

int tile_window_bits(int fp32);                                  // 11 FP64, 12 FP32 (the kernel's tile)
size_t tile_sweep_lds_bytes(int P, int kw, int fp32, int nstate, size_t* tile_off);
template <typename R>
hipError_t launch_sweep_tile(const TileArgs<R>& a, int mode, int grid, size_t lds, hipStream_t s);

}  // namespace mg
