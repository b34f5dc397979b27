// Host runtime of the GPU evaluation path: device context, device-resident
// targets, gate programs and the f / gradient / D^2 f[Z] evaluator.
#pragma once
#include <complex>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "mg_device.hpp"
#include "../../include/config.h"

namespace mg {

using zd = std::complex<double>;

void set_error(const std::string& msg);
const char* last_error();

// grow-only device allocation owned by the context
// held by HIP graph captures and by the calls that may not overlap one on another thread
// (allocations, frees, synchronous copies)
std::recursive_mutex& capture_mutex();

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes);
    void release();
    ~DevBuf() { release(); }
};

// A gate program: G instances; instance g applies parameter gate param[g]
// to wires (wires[2g], wires[2g+1]).
struct Program {
    int L = 0, P = 0, G = 0;
    std::vector<int> param, wires;
    bool validate(std::string* why) const;
};

// tied brickwall of the reference: layer i applies gate i to the pairs
// (perms[i][2k], perms[i][2k+1]) (reference src/matchgate_target.c:308-324)
Program brickwall_program(int L, int nlayers, const int* const* perms);

// Device-resident target U.  Dense: the seed matrix row b = -conj(U e_b) in FP64
// (and FP32 on first use).  Program: a gate program (with its gates) applied to
// e_b inside the sweep kernel.
struct Target {
    int kind = 0;  // 0 dense, 1 program
    int L = 0;
    DevBuf seed64, seed32;
    int Gt = 0;
    DevBuf tinst, tcoef64, tcoef32;
    std::vector<Inst> tinst_h;
    std::vector<cpx<double>> tcoef_h;
    bool seed32_ready = false, tcoef32_ready = false;
    // target program in gate-window form for the batched kernel, per precision (FP64, FP32)
    struct WinCache {
        DevBuf inst, coef, blk_lds, blk_glob;
        int nblocks = 0;
        int wbits = 0;  // window bits it was built for (0: not built)
    };
    WinCache wc[2];
    // target program in 12-qubit tile form (FP64 register windows) for materialising its rows
    // beyond 12 qubits: all at once (target_densify) or one column chunk per pass (evaluations of
    // program targets with the tiled sweeps)
    struct TileForm {
        DevBuf inst, coef, blk;
        int G = 0, ntb = 0, nblk = 0, slots = 0;
        size_t acc_off = 0, buf_off = 0, lds = 0;
        bool ready = false;
    };
    TileForm tf;
    // the target program in 4-qubit windows for the matrix-core materialisation (L > 12)
    struct MatForm {
        DevBuf tb, win, af, af32;
        int ntb = 0;
        bool ready = false;
    };
    MatForm mf;
};

// U row-major (S x S)
Target* target_dense(const zd* U, int L);
// rows already in seed layout: seed[b*S + a] = -conj(U[a][b])
Target* target_dense_seed(std::vector<cpx<double>>&& seed, int L);
Target* target_program(const Program& prog, const zd* gates);
// turn a target program into dense seed rows on the device (FP64, materialised once on
// the GPU); later evaluations read the rows instead of applying the program per column
int target_densify(Target& t);

enum Mode { kF = 0, kFGrad = 1, kFGradHvp = 2 };
enum Precision { kFP64 = 0, kFP32 = 1 };

struct EvalOut {
    zd f;
    std::vector<zd> grad;  // [P*8]
    std::vector<zd> dhz;   // [P*8]
};

struct EvalStats {
    bool lds_states = false, ckpt = false;
    int reg_threads = 0;  // > 0: register-resident kernel with this many threads per workgroup
    int threads = 0;      // threads per workgroup of the launched sweep
    int grid = 0;
    size_t lds_bytes = 0;
    float kernel_ms = 0.f;  // sweep kernel duration (HIP events on the launch stream)
    const char* kernel = "";  // name of the sweep kernel
};

// bytes of the static device scratch that allocation-free evaluations draw from
constexpr size_t kNoallocScratchBytes = (size_t)32 << 20;

// gates, Z: P*8 complex each (Z may be null when mode < 2).  Column range [b0, b1).
// noalloc: every device buffer of the sweep comes from the static scratch (no hipMalloc);
// fails with an error when it does not fit there
int evaluate(const Target& t, const Program& p, const zd* gates, const zd* Z, int mode, int precision,
             EvalOut* out, long b0 = 0, long b1 = -1, EvalStats* stats = nullptr, bool noalloc = false);

// configuration knobs (mostly for tests and tuning)
struct Knobs {
    int force_ckpt = -1;      // -1 auto, 0 never, 1 always (when mode >= 1)
    int force_lds = -1;       // -1 auto, 0 global-scratch states, 1 LDS states (if they fit)
    int grid_per_cu = 0;      // 0 auto
    int force_reg = -1;       // -1 auto, 0 never, 1 whenever a register-resident configuration exists
    int reg_threads = 0;      // 0 auto, else 512 or 1024 threads for the register-resident kernel
    int lds_threads = 0;      // 0 auto, else 256, 512 or 1024 threads for the LDS/global kernel
    int force_win = -1;       // -1 auto (register gate-window kernel where it applies), 0 never
    int force_tile = -1;      // -1 auto (tiled kernel for dense targets whose states exceed the LDS), 0 never
    bool time_kernels = false;
    size_t row_chunk_bytes = 0;  // materialised target rows per column chunk (0: 4 GiB)
    int mat_kernel = -1;         // target rows beyond 12 qubits: -1 auto, 0 register tiles, 1 matrix cores
    int fwd_mfma = -1;           // FP32 forward passes on the matrix cores (tiled f, f + gradient; L=12 HVP in LDS): -1 auto, 0 gate by gate
    int device_tr = -1;          // mg_optimize with HVPs: -1 auto (trust region on the device, one graph per
                                 // iteration, where its sweeps have plans), 0 the host loop, 2 on the
                                 // device with reuse (H eta from the CG, f + gradient at the trial kept)
};
Knobs& knobs();

hipStream_t stream();
int device_cu_count();

// ---- batched evaluation (one persistent launch per precision over many requests) ----
struct BatchReq {
    Target* t = nullptr;
    const Program* p = nullptr;
    const zd* gates = nullptr;  // [P*8]
    const zd* Z = nullptr;      // [P*8] (mode 2)
    int mode = 0;
    int precision = 0;
    EvalOut* out = nullptr;
    long b0 = 0, b1 = 0;        // basis columns [b0, b1); b1 <= 0: all 2^L
};

struct BatchStats {
    int launches = 0;
    int items = 0;
    int grid = 0;
    size_t lds_bytes = 0;
    float kernel_ms = 0.f;
    double prep_ms = 0.0;  // host time preparing the launches (instances, window blocks, upload staging)
    int graph_replays = 0;   // launches replayed from a captured graph
    int graph_captures = 0;  // graphs captured
};

// Per (precision, L) shape of the batched kernel: column-group width cgw (threads per
// column, a power of two up to the 2^(L-W) window tuples) and columns per work item;
// 0 = default.  Results depend on these (summation
// order), never on which other requests share the launch.
struct BatchTuning {
    int threads[2] = {256, 256};  // per precision (FP64, FP32): threads per workgroup
    int wbits[2] = {3, 4};        // per precision: gate-window bits
    int grid_per_cu = 0;
    int cgw[2][31] = {};
    int cpi[2][31] = {};
    // requests with dense targets and single_min <= L <= single_max fill the GPU on their own:
    // they run on the single-request kernels (register gate-window / LDS / tiled sweeps) in
    // their own launches instead of inside the persistent batch (0 / 0: never)
    int single_min = 12, single_max = 30;
    // replay repeated launches (same arguments, buffers and sizes) from captured HIP graphs
    bool graphs = true;
};
BatchTuning& batch_tuning();
void batch_shape(int L, int precision, int* cgw, int* cpi);
// a batched kernel is compiled for (precision, threads, window bits)
bool batch_variant_exists(int precision_fp32, int threads, int wbits);

int evaluate_batch(BatchReq* reqs, int n, BatchStats* st);
// true if evaluate_batch runs a request of this target and program on the single-request
// kernels (dense target, or a program target beyond 12 qubits, and batch_tuning().single_min <= L
// <= single_max): it fills the GPU alone
bool batch_single_request(const Target& t, const Program& p);
BatchStats& last_batch_stats();

}  // namespace mg
