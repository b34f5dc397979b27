// Batched column sweep: many evaluation requests ("instances") of different
// sizes, modes and targets in one persistent launch (mg_batch.hip).
#pragma once
#include "mg_device.hpp"
#include "mg_window.hpp"

namespace mg {

template <typename R>
struct BatchInst {
    int L, G, P, mode;      // mode 0: f, 1: f + gradient, 2: f + gradient + D^2 f[Z]
    int target_kind, Gt;    // 0: dense seed rows, 1: target program applied in-kernel
    int cgw;                // threads per column: a power of two <= min(T, 2^(L-W)); T / cgw columns at once
    int lds;                // 1: states in LDS at state_off, 0: in the workgroup's global scratch slice
    int acc_off;            // byte offset of the [P][32] accumulator in dynamic LDS
    int state_off;          // byte offset of the states in dynamic LDS
    int nblocks, ntblocks;  // gate-window blocks of the program and of the target program
    const Inst* inst;       // [G] in block order; lo = window-local pair index (mg_window.hpp)
    const cpx<R>* coef;     // [G][kCoefPerInst] in block order
    const WBlock* blocks;   // [nblocks]
    const cpx<R>* seed;     // [S][S] (target_kind 0)
    const Inst* tinst;      // [Gt] in block order (target_kind 1)
    const cpx<R>* tcoef;    // [Gt][kCoefPerInst] in block order
    const WBlock* tblocks;  // [ntblocks]
};

// LDS swizzle fold width: log2 of the lanes one 128-byte LDS cycle serves
template <typename R>
constexpr int batch_bank_bits() { return sizeof(R) == 4 ? 4 : 3; }

// one work item: columns [b0, b1) of instance `inst`; its partial sums (2 + 32 P
// doubles: f, then per parameter 8 gradient and 8 D^2 f[Z] complex values) go to
// partial + out
struct BatchItem {
    int inst;
    int b0, b1;
    int pad;
    long long out;
};

// per instance: sum n partials of `stride` doubles starting at src, in item order, into dst
struct ReduceSeg {
    long long src, dst;
    int n, stride;
};

// fixed LDS bookkeeping in front of the accumulator: per-gate staging
// [2][kWinMaxGates][NW][32] doubles, f [NW][2]
inline size_t batch_lds_fixed_bytes(int T)
{
    const int nw = T / 64;
    const size_t b = sizeof(double) * (2 * (size_t)kWinMaxGates * nw * 32 + 2 * nw);
    return (b + 255) & ~(size_t)255;
}

// compiled (threads, window bits) variants: FP32 (256, 2..4) and (512, 2..3); FP64 (256, 2..3) and (512, 2).
// launch_batch with grid < 0 only sets the kernel's dynamic-LDS attribute (done before a stream
// capture; during a capture the launch itself does not set it).
bool batch_variant_exists(int precision_fp32, int threads, int wbits);

template <typename R>
hipError_t launch_batch(const BatchInst<R>* insts, const BatchItem* items, int nitems, int* counter, double* partial,
                        cpx<R>* scratch, long long scratch_stride, int threads, int wbits, int grid, size_t lds,
                        hipStream_t s);

// dense seed rows of a target program (FP64, window bits 3, 256 threads; blocks in LDS
// swizzled form): lds = (256 / cgw) * 2^L * 16 bytes with cgw = min(256, 2^(L-3))
hipError_t launch_materialize(int L, const Inst* tinst, const cpx<double>* tcoef, const WBlock* tblocks, int ntblocks,
                              cpx<double>* seed, int grid, size_t lds, hipStream_t s);

// the same for 2^L columns whose states do not fit LDS (in place in the seed rows; plain offsets)
hipError_t launch_materialize_global(int L, const Inst* tinst, const cpx<double>* tcoef, const WBlock* tblocks,
                                     int ntblocks, cpx<double>* seed, int grid, hipStream_t s);

hipError_t launch_batch_reduce(const ReduceSeg* segs, const int2* blocks, int nblocks, const double* partial, double* out,
                               hipStream_t s);

}  // namespace mg
