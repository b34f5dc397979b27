// Launch interface of the column-sweep kernels (mg_sweep.hip).
#pragma once
#include "mg_device.hpp"

namespace mg {

template <typename R>
struct SweepArgs {
    int L;                  // qubits; S = 2^L amplitudes per column
    int G;                  // gate instances of the program
    int P;                  // parameter gates (gradient slots)
    const Inst* inst;       // [G]
    const cpx<R>* coef;     // [G][kCoefPerInst]: V, V^{-1}, Z (sorted basis)
    int target_kind;        // 0: dense seed rows, 1: target program applied in-kernel
    const cpx<R>* seed;     // [S][S]: row b = -conj(U e_b)
    int Gt;                 // target program instances
    const Inst* tinst;      // [Gt]
    const cpx<R>* tcoef;    // [Gt][kCoefPerInst] (V used)
    long b0, b1;            // column range [b0, b1)
    double* partial;        // [gridDim.x][2 + 32 P]: f (re, im), then per parameter 16 complex
    cpx<R>* scratch;        // global state slices [gridDim.x][nstate][S] (when states are not in LDS)
    cpx<R>* ckpt;           // checkpoints [gridDim.x][G+1][2][S] (checkpoint mode)
    int nstate;             // state vectors per workgroup
    int lds_state_offset;   // byte offset of the states inside dynamic LDS (LDS mode)
    // FP32 sweeps with the states in LDS at L = 12: the forward pass on the matrix cores over nwin
    // 4-qubit windows (struct FWin; fragments [window][8][lane], or [window][16][lane] for the HVP;
    // the instances in window order), or none (fwin null: gate by gate)
    const void* fwin;
    const float* fwaf;
    int nwin;
    // a launch inside a captured trust-region step: nonzero *skip makes the launch return at once
    // (the step's truncated CG has already ended); null: always run
    const int* skip;
};

// doubles of LDS bookkeeping in front of the states: acc [P][32] (sweep_kernel: [P][16] below
// MODE 2, which has no D^2 f[Z] half), stage [2][NW][32], f [NW][2]
inline size_t sweep_lds_head_bytes(int P, int nwaves, int mode = 2)
{
    size_t b = sizeof(double) * ((mode == 2 ? 32 : 16) * (size_t)P + 2 * nwaves * 32 + 2 * nwaves);
    return (b + 255) & ~(size_t)255;
}

template <typename R>
hipError_t launch_sweep(const SweepArgs<R>& a, int mode, bool lds_states, bool ckpt, int threads, int grid, size_t lds,
                        hipStream_t s);

// register-resident variant (mg_sweep_reg.hip): threads x qpt quadruples = 2^L / 4;
// a.nstate = states per LDS re-layout round
template <typename R>
hipError_t launch_sweep_reg(const SweepArgs<R>& a, int mode, int threads, int qpt, int grid, size_t lds, hipStream_t s);

// skip as SweepArgs::skip
hipError_t launch_reduce(const double* partial, int nblk, int stride, double* out, hipStream_t s,
                         const int* skip = nullptr);


}  // namespace mg
