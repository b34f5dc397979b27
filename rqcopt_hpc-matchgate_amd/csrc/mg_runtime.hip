// Host runtime: device context, targets, programs and the sweep orchestrator.
#include "mg_runtime.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>

#include "mg_sweep.hpp"

namespace mg {

// ---- errors ----------------------------------------------------------------------------
static thread_local std::string g_err;
void set_error(const std::string& msg)
{
    g_err = msg;
    std::fprintf(stderr, "rqcopt_mg: %s\n", msg.c_str());
}
const char* last_error() { return g_err.c_str(); }

#define MG_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            set_error(std::string(#call) + " failed: " + hipGetErrorString(e_));             \
            return -1;                                                                       \
        }                                                                                    \
    } while (0)

hipError_t DevBuf::ensure(size_t bytes)
{
    if (bytes <= n && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, bytes ? bytes : 256);
    if (e != hipSuccess) {
        p = nullptr;
        n = 0;
        return e;
    }
    n = bytes;
    return hipSuccess;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}

// ---- context (one per process, for the device current at first use) ------------------
struct Context {
    int device = -1;
    hipStream_t s = nullptr;
    int ncu = 0;
    size_t lds_limit = 0;
    DevBuf inst, coef, partial, out, scratch, ckpt;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<double> out_h;
};

static Context* ctx()
{
    static std::unique_ptr<Context> c;
    if (!c) {
        auto n = std::make_unique<Context>();
        if (hipGetDevice(&n->device) != hipSuccess) return nullptr;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, n->device) != hipSuccess) return nullptr;
        n->ncu = prop.multiProcessorCount;
        n->lds_limit = prop.sharedMemPerBlockOptin ? prop.sharedMemPerBlockOptin : prop.sharedMemPerBlock;
        if (hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking) != hipSuccess) return nullptr;
        if (hipEventCreate(&n->ev0) != hipSuccess || hipEventCreate(&n->ev1) != hipSuccess) return nullptr;
        c = std::move(n);
    }
    return c.get();
}

hipStream_t stream()
{
    Context* c = ctx();
    return c ? c->s : nullptr;
}

int device_cu_count()
{
    Context* c = ctx();
    return c ? c->ncu : 0;
}

Knobs& knobs()
{
    static Knobs k;
    return k;
}

// ---- programs ----------------------------------------------------------------------------
bool Program::validate(std::string* why) const
{
    if (L < 2 || L > 30) { *why = "nqubits out of range"; return false; }
    if (G < 1 || P < 1) { *why = "empty program"; return false; }
    if ((int)param.size() != G || (int)wires.size() != 2 * G) { *why = "program arrays have wrong sizes"; return false; }
    for (int g = 0; g < G; g++) {
        const int i = wires[2 * g], j = wires[2 * g + 1];
        if (i < 0 || j < 0 || i >= L || j >= L || i == j) { *why = "invalid wire pair at instance " + std::to_string(g); return false; }
        if (param[g] < 0 || param[g] >= P) { *why = "invalid parameter index at instance " + std::to_string(g); return false; }
    }
    return true;
}

Program brickwall_program(int L, int nlayers, const int* const* perms)
{
    Program p;
    p.L = L;
    p.P = nlayers;
    const int per = L / 2;
    p.G = nlayers * per;
    p.param.resize(p.G);
    p.wires.resize(2 * p.G);
    for (int i = 0; i < nlayers; i++)
        for (int k = 0; k < per; k++) {
            p.param[i * per + k] = i;
            p.wires[2 * (i * per + k)] = perms[i][2 * k];
            p.wires[2 * (i * per + k) + 1] = perms[i][2 * k + 1];
        }
    return p;
}

// ---- instance coefficients -------------------------------------------------------------
// gate in the sorted basis: center reversed when the caller's wires are reversed
static void sorted_gate(const zd* g, bool flip, zd* out)
{
    if (flip) {
        out[0] = g[3]; out[1] = g[2]; out[2] = g[1]; out[3] = g[0];
    } else {
        for (int k = 0; k < 4; k++) out[k] = g[k];
    }
    for (int k = 4; k < 8; k++) out[k] = g[k];
}

static bool inverse2(const zd* a, zd* inv)
{
    const zd det = a[0] * a[3] - a[1] * a[2];
    if (std::abs(det) == 0.0) return false;
    inv[0] = a[3] / det;
    inv[1] = -a[1] / det;
    inv[2] = -a[2] / det;
    inv[3] = a[0] / det;
    return true;
}

static double unitarity_defect(const zd* b)
{
    // || B^dagger B - I ||_max
    const zd m00 = std::conj(b[0]) * b[0] + std::conj(b[2]) * b[2];
    const zd m01 = std::conj(b[0]) * b[1] + std::conj(b[2]) * b[3];
    const zd m11 = std::conj(b[1]) * b[1] + std::conj(b[3]) * b[3];
    return std::max({std::abs(m00 - 1.0), std::abs(m01), std::abs(m11 - 1.0)});
}

struct Built {
    std::vector<Inst> inst;
    std::vector<cpx<double>> coef;
    bool invertible = true;
    double defect = 0.0;
};

static void build_instances(const Program& p, const zd* gates, const zd* Z, Built* b)
{
    b->inst.resize(p.G);
    b->coef.assign((size_t)p.G * kCoefPerInst, cpx<double>{0.0, 0.0});
    b->invertible = true;
    b->defect = 0.0;
    for (int q = 0; q < p.P; q++) {
        b->defect = std::max(b->defect, unitarity_defect(gates + 8 * q));
        b->defect = std::max(b->defect, unitarity_defect(gates + 8 * q + 4));
    }
    for (int g = 0; g < p.G; g++) {
        int i = p.wires[2 * g], j = p.wires[2 * g + 1];
        const bool flip = j < i;
        if (flip) std::swap(i, j);
        b->inst[g] = Inst{p.L - 1 - i, p.L - 1 - j, p.param[g], flip ? 1 : 0};
        zd V[8], Vi[8], Zs[8];
        sorted_gate(gates + 8 * p.param[g], flip, V);
        if (!inverse2(V, Vi) || !inverse2(V + 4, Vi + 4)) {
            b->invertible = false;
            for (int k = 0; k < 8; k++) Vi[k] = 0.0;
        }
        if (Z) sorted_gate(Z + 8 * p.param[g], flip, Zs);
        else
            for (int k = 0; k < 8; k++) Zs[k] = 0.0;
        cpx<double>* c = &b->coef[(size_t)g * kCoefPerInst];
        for (int k = 0; k < 8; k++) {
            c[k] = {V[k].real(), V[k].imag()};
            c[8 + k] = {Vi[k].real(), Vi[k].imag()};
            c[16 + k] = {Zs[k].real(), Zs[k].imag()};
        }
    }
}

// ---- targets ------------------------------------------------------------------------------
__global__ void to_fp32(const cpx<double>* in, cpx<float>* out, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        out[k] = {(float)in[k].re, (float)in[k].im};
}

Target* target_dense_seed(std::vector<cpx<double>>&& seed, int L)
{
    Context* c = ctx();
    if (!c) { set_error("no HIP device"); return nullptr; }
    const size_t S = (size_t)1 << L;
    if (seed.size() != S * S) { set_error("dense target has the wrong size"); return nullptr; }
    auto t = std::make_unique<Target>();
    t->kind = 0;
    t->L = L;
    if (t->seed64.ensure(S * S * sizeof(cpx<double>)) != hipSuccess) { set_error("out of device memory for the target"); return nullptr; }
    if (hipMemcpy(t->seed64.p, seed.data(), S * S * sizeof(cpx<double>), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("copying the target to the device failed");
        return nullptr;
    }
    return t.release();
}

Target* target_dense(const zd* U, int L)
{
    const size_t S = (size_t)1 << L;
    std::vector<cpx<double>> seed(S * S);
    // seed[b][a] = -conj(U[a][b]): blocked transpose
    const size_t B = 64;
    for (size_t a0 = 0; a0 < S; a0 += B)
        for (size_t b0 = 0; b0 < S; b0 += B)
            for (size_t a = a0; a < std::min(S, a0 + B); a++)
                for (size_t b = b0; b < std::min(S, b0 + B); b++) {
                    const zd u = U[a * S + b];
                    seed[b * S + a] = {-u.real(), u.imag()};
                }
    return target_dense_seed(std::move(seed), L);
}

Target* target_program(const Program& prog, const zd* gates)
{
    Context* c = ctx();
    if (!c) { set_error("no HIP device"); return nullptr; }
    std::string why;
    if (!prog.validate(&why)) { set_error("target program: " + why); return nullptr; }
    auto t = std::make_unique<Target>();
    t->kind = 1;
    t->L = prog.L;
    t->Gt = prog.G;
    Built b;
    build_instances(prog, gates, nullptr, &b);
    t->tinst_h = b.inst;
    t->tcoef_h = b.coef;
    if (t->tinst.ensure(b.inst.size() * sizeof(Inst)) != hipSuccess ||
        t->tcoef64.ensure(b.coef.size() * sizeof(cpx<double>)) != hipSuccess) {
        set_error("out of device memory for the target program");
        return nullptr;
    }
    if (hipMemcpy(t->tinst.p, b.inst.data(), b.inst.size() * sizeof(Inst), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->tcoef64.p, b.coef.data(), b.coef.size() * sizeof(cpx<double>), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("copying the target program failed");
        return nullptr;
    }
    return t.release();
}

static int ensure_fp32_target(Target& t, hipStream_t s)
{
    if (t.kind == 0 && !t.seed32_ready) {
        const size_t S = (size_t)1 << t.L;
        MG_HIP(t.seed32.ensure(S * S * sizeof(cpx<float>)));
        hipLaunchKernelGGL(to_fp32, dim3(1024), dim3(256), 0, s, (const cpx<double>*)t.seed64.p, (cpx<float>*)t.seed32.p, S * S);
        MG_HIP(hipGetLastError());
        t.seed32_ready = true;
    }
    if (t.kind == 1 && !t.tcoef32_ready) {
        std::vector<cpx<float>> c32(t.tcoef_h.size());
        for (size_t k = 0; k < c32.size(); k++) c32[k] = {(float)t.tcoef_h[k].re, (float)t.tcoef_h[k].im};
        MG_HIP(t.tcoef32.ensure(c32.size() * sizeof(cpx<float>)));
        MG_HIP(hipMemcpyAsync(t.tcoef32.p, c32.data(), c32.size() * sizeof(cpx<float>), hipMemcpyHostToDevice, s));
        MG_HIP(hipStreamSynchronize(s));
        t.tcoef32_ready = true;
    }
    return 0;
}

// ---- evaluation -------------------------------------------------------------------------------
template <typename R>
static int run_sweep(Context* c, Target& t, const Program& p, const Built& b, int mode, long b0, long b1, EvalOut* out,
                     EvalStats* st)
{
    hipStream_t s = c->s;
    const int L = p.L;
    const size_t S = (size_t)1 << L;
    const Knobs& kn = knobs();
    // threads of the LDS/global kernel: more waves per CU when one workgroup fills the LDS
    int lds_T = kn.lds_threads;
    {
        // 512 threads when the states fill the LDS (one workgroup per CU), else 256
        const int nst0 = (mode == 2 ? 4 : 2);
        const bool fits = sweep_lds_head_bytes(p.P, 8) + (size_t)nst0 * S * sizeof(cpx<R>) <= c->lds_limit;
        if (lds_T == 0) lds_T = fits && S >= 2048 ? 512 : 256;
    }
    const int NW = lds_T / 64;

    // coefficients in precision R
    std::vector<cpx<R>> coefR(b.coef.size());
    for (size_t k = 0; k < coefR.size(); k++) coefR[k] = {(R)b.coef[k].re, (R)b.coef[k].im};
    MG_HIP(c->inst.ensure(b.inst.size() * sizeof(Inst)));
    MG_HIP(c->coef.ensure(coefR.size() * sizeof(cpx<R>)));
    MG_HIP(hipMemcpyAsync(c->inst.p, b.inst.data(), b.inst.size() * sizeof(Inst), hipMemcpyHostToDevice, s));
    MG_HIP(hipMemcpyAsync(c->coef.p, coefR.data(), coefR.size() * sizeof(cpx<R>), hipMemcpyHostToDevice, s));

    // checkpointing: needed when psi_g cannot be recovered by V^{-1} (non-unitary gates)
    bool ckpt = false;
    if (mode >= 1) {
        if (kn.force_ckpt == 1) ckpt = true;
        else if (kn.force_ckpt == 0) ckpt = !b.invertible;
        else ckpt = !b.invertible || b.defect > 1e-10;
    }
    const int nstate = ckpt ? (mode == 2 ? 2 : 1) : (mode == 2 ? 4 : 2);
    const size_t head = sweep_lds_head_bytes(p.P, NW);
    const size_t state_bytes = (size_t)nstate * S * sizeof(cpx<R>);
    if (head > c->lds_limit) { set_error("too many parameter gates for the LDS accumulator"); return -1; }
    bool lds_states = head + state_bytes <= c->lds_limit;
    if (kn.force_lds == 0) lds_states = false;

    // register-resident kernel: states in VGPRs, LDS as the re-layout buffer; needs gates
    // unitary to 1e-12 (it uncomputes with V^dagger) and threads x quads = 2^L / 4
    int reg_T = 0, reg_qpt = 0;
    if (!ckpt && b.defect <= 1e-12 && kn.force_reg != 0 && (!lds_states || kn.force_reg == 1)) {
        const size_t Q = S / 4;
        const int prefT = kn.reg_threads ? kn.reg_threads : (Q <= 1024 ? 512 : 1024);
        for (int T : {prefT, 512, 1024}) {
            if (Q % T) continue;
            const size_t qpt = Q / T;
            if ((T == 512 && (qpt == 1 || qpt == 2)) || (T == 1024 && (qpt == 1 || qpt == 2))) {
                reg_T = T;
                reg_qpt = (int)qpt;
                break;
            }
        }
        if (reg_T && head + S * sizeof(cpx<R>) > c->lds_limit) reg_T = 0;
    }
    const int nw_reg = reg_T / 64;
    const size_t head_reg = reg_T ? sweep_lds_head_bytes(p.P, nw_reg) : 0;
    int reg_nst = 0;
    if (reg_T) {
        reg_nst = (int)std::min<size_t>(mode == 2 ? 4 : 2, (c->lds_limit - head_reg) / (S * sizeof(cpx<R>)));
        if (reg_nst == 3) reg_nst = 2;  // instantiated: 1, 2, 4 states per round
        if (reg_nst < 1) reg_T = 0;
    }
    const size_t lds = reg_T ? head_reg + (size_t)reg_nst * S * sizeof(cpx<R>) : (lds_states ? head + state_bytes : head);

    const long ncols = b1 - b0;
    int per_cu = kn.grid_per_cu > 0 ? kn.grid_per_cu : std::max<int>(1, std::min<size_t>(4, (160 * 1024) / std::max<size_t>(lds, 1)));
    if (reg_T && kn.grid_per_cu <= 0) per_cu = std::max<int>(1, std::min<size_t>(2, (160 * 1024) / lds));
    long grid = std::min<long>(ncols, (long)c->ncu * per_cu);
    // bound the global scratch and checkpoint footprint (2 GiB each)
    const size_t lim = (size_t)2 << 30;
    if (!lds_states && !reg_T) grid = std::min<long>(grid, std::max<long>(1, lim / state_bytes));
    const size_t ck_per = ckpt ? (size_t)2 * (p.G + 1) * S * sizeof(cpx<R>) : 0;
    if (ckpt) grid = std::min<long>(grid, std::max<long>(1, lim / ck_per));
    grid = std::max<long>(grid, 1);

    const int stride = 2 + 32 * p.P;
    MG_HIP(c->partial.ensure((size_t)grid * stride * sizeof(double)));
    MG_HIP(c->out.ensure((size_t)stride * sizeof(double)));
    if (!lds_states && !reg_T) MG_HIP(c->scratch.ensure((size_t)grid * state_bytes));
    if (ckpt) MG_HIP(c->ckpt.ensure((size_t)grid * ck_per));

    SweepArgs<R> a{};
    a.L = L;
    a.G = p.G;
    a.P = p.P;
    a.inst = (const Inst*)c->inst.p;
    a.coef = (const cpx<R>*)c->coef.p;
    a.target_kind = t.kind;
    if (t.kind == 0) {
        a.seed = std::is_same<R, double>::value ? (const cpx<R>*)t.seed64.p : (const cpx<R>*)t.seed32.p;
    } else {
        a.Gt = t.Gt;
        a.tinst = (const Inst*)t.tinst.p;
        a.tcoef = std::is_same<R, double>::value ? (const cpx<R>*)t.tcoef64.p : (const cpx<R>*)t.tcoef32.p;
    }
    a.b0 = b0;
    a.b1 = b1;
    a.partial = (double*)c->partial.p;
    a.scratch = (cpx<R>*)c->scratch.p;
    a.ckpt = (cpx<R>*)c->ckpt.p;
    a.nstate = reg_T ? reg_nst : nstate;
    a.lds_state_offset = (int)(reg_T ? head_reg : head);

    if (kn.time_kernels) MG_HIP(hipEventRecord(c->ev0, s));
    if (reg_T)
        MG_HIP(launch_sweep_reg<R>(a, mode, reg_T, reg_qpt, (int)grid, lds, s));
    else
        MG_HIP(launch_sweep<R>(a, mode, lds_states, ckpt, lds_T, (int)grid, lds, s));
    if (kn.time_kernels) MG_HIP(hipEventRecord(c->ev1, s));
    MG_HIP(launch_reduce((const double*)c->partial.p, (int)grid, stride, (double*)c->out.p, s));
    c->out_h.resize(stride);
    MG_HIP(hipMemcpyAsync(c->out_h.data(), c->out.p, stride * sizeof(double), hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));

    const double* o = c->out_h.data();
    out->f = zd(o[0], o[1]);
    out->grad.assign((size_t)8 * p.P, zd(0.0));
    out->dhz.assign((size_t)8 * p.P, zd(0.0));
    for (int q = 0; q < p.P; q++)
        for (int e = 0; e < 8; e++) {
            out->grad[8 * q + e] = zd(o[2 + 32 * q + 2 * e], o[2 + 32 * q + 2 * e + 1]);
            out->dhz[8 * q + e] = zd(o[2 + 32 * q + 16 + 2 * e], o[2 + 32 * q + 16 + 2 * e + 1]);
        }
    if (st) {
        st->lds_states = lds_states && !reg_T;
        st->ckpt = ckpt;
        st->reg_threads = reg_T;
        st->threads = reg_T ? reg_T : lds_T;
        st->grid = (int)grid;
        st->lds_bytes = lds;
        st->kernel_ms = 0.f;
        if (kn.time_kernels) MG_HIP(hipEventElapsedTime(&st->kernel_ms, c->ev0, c->ev1));
    }
    return 0;
}

int evaluate(const Target& tc, const Program& p, const zd* gates, const zd* Z, int mode, int precision, EvalOut* out,
             long b0, long b1, EvalStats* stats)
{
    Context* c = ctx();
    if (!c) { set_error("no HIP device available"); return -1; }
    Target& t = const_cast<Target&>(tc);
    std::string why;
    if (!p.validate(&why)) { set_error("program: " + why); return -1; }
    if (t.L != p.L) { set_error("target and program have different qubit counts"); return -1; }
    if (mode < 0 || mode > 2) { set_error("invalid evaluation mode"); return -1; }
    if (mode == 2 && !Z) { set_error("Hessian-vector product needs a direction"); return -1; }
    const long S = 1L << p.L;
    if (b1 < 0) b1 = S;
    if (b0 < 0 || b1 > S || b0 >= b1) { set_error("invalid column range"); return -1; }
    Built b;
    build_instances(p, gates, Z, &b);
    if (mode >= 1 && !b.invertible && knobs().force_ckpt == 0) {
        set_error("singular gate block: cannot uncompute");
        return -1;
    }
    if (precision == kFP32) {
        if (ensure_fp32_target(t, c->s) < 0) return -1;
        return run_sweep<float>(c, t, p, b, mode, b0, b1, out, stats);
    }
    return run_sweep<double>(c, t, p, b, mode, b0, b1, out, stats);
}

}  // namespace mg
