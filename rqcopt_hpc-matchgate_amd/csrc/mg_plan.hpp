// A single-request sweep laid out once and launched again with coefficients rewritten on the
// device (mg_runtime.hip: plan_sweep, launch_plan): the evaluations inside a trust-region step
// captured as one HIP graph (mg_device_tr.hip).
#pragma once
#include <vector>

#include "mg_mat_mfma.hpp"
#include "mg_runtime.hpp"
#include "mg_sweep.hpp"
#include "mg_sweep_win.hpp"

namespace mg {

struct SweepPlan {
    int kind = 0;  // 0: no plan for this configuration, 1: sweep_kernel (states in LDS), 2: sweep_win_kernel
    int mode = 0, fp32 = 0;
    int grid = 0, threads = 0, stride = 0;
    size_t lds = 0;
    SweepArgs<double> a64{};  // kind 1, FP64
    SweepArgs<float> a32{};   // kind 1, FP32
    WinSweepArgs wa{};        // kind 2
    // forward windows on the matrix cores: fragments formed from the FP64 program-order
    // coefficients (nwin 0: none)
    int nwin = 0;
    const int* wstart = nullptr;
    const WinGate* wgate = nullptr;
    double* frag64 = nullptr;
    float* frag32 = nullptr;
    // the coefficients a launch reads, to be rewritten before it: the kernel's copy (instance k of
    // the kernel's order is program instance korder[k], in the sweep precision) and the FP64
    // program-order copy of the fragments (null: none)
    void* kcoef = nullptr;
    std::vector<int> korder;
    cpx<double>* prog64 = nullptr;
    double* partial = nullptr;
    double* out = nullptr;  // [stride]: f (re, im), then per parameter 32 doubles (gradient, D^2 f[Z])
    DevBuf mem;             // the device memory of everything above
};

// lays out the evaluation (mode, precision) of program p against target t as the single-request
// path would run it, with `gates` (and zero directions) as the coefficients; plan->kind 0 when the
// path would use a kernel without a plan (register, tiled or checkpointing sweeps)
int plan_sweep(const Target& t, const Program& p, const zd* gates, int mode, int precision, SweepPlan* plan);
// frags (if any), sweep and reduction of a plan on stream s; nonzero *skip makes each return at once
hipError_t launch_plan(const SweepPlan& plan, const int* skip, hipStream_t s);

}  // namespace mg
