// Gate windows: a gate program regrouped into blocks of consecutive gates that
// all act inside one window of W qubits, so that a sweep can apply a whole block
// to registers holding the window's 2^W amplitudes (one LDS or memory pass per
// block instead of one per gate).
#pragma once
#include <cstdint>
#include <vector>

#include "mg_device.hpp"

namespace mg {

constexpr int kWinMaxBits = 4;    // largest window: 16 amplitudes per thread and state
constexpr int kWinMaxGates = 6;   // gates per block at most (per-gate staging slots in LDS)

struct WBlock {
    int g0, g1;                      // gates [g0, g1) of the reordered instance list
    int pos[kWinMaxBits];            // window bit positions, ascending (W used)
    uint32_t off[1 << kWinMaxBits];  // address offset of window amplitude j (swizzled in LDS)
};

// Reorders the instances (program order; bit positions hi > lo) into blocks:
// `order` receives the new order (old indices) and `blocks` the block list.  The
// order is a topological order of the program's gate dependencies, so the product
// of the gates is unchanged.  Every block's gates act inside its window.
void window_blocks(const std::vector<Inst>& inst, int L, int W, int kmax, std::vector<int>* order,
                   std::vector<WBlock>* blocks);

// window-local pair index of a gate with bit positions (lo, hi) in a window with
// ascending positions pos[0..W): la > lb are the local indices of hi and lo, and the
// index enumerates (la, lb) = (1,0), (2,0), (2,1), (3,0), (3,1), (3,2)
int window_case(const WBlock& b, int W, int lo, int hi);

// LDS address swizzle of the window kernels: the address bits above the bank group
// (B = log2 of the lanes one 128-byte LDS cycle serves) are XOR-folded into the low
// B bits.  Linear in GF(2), so swz(base | off) = swz(base) ^ swz(off) for disjoint bits.
template <int B>
__host__ __device__ inline uint32_t win_swz(uint32_t a)
{
    constexpr uint32_t m = (1u << B) - 1;
    uint32_t r = a;
#pragma unroll
    for (int s = B; s < 24; s += B) r ^= (a >> s) & m;  // addresses below 2^24 elements
    return r;
}

}  // namespace mg
