// Host runtime: device context, targets, programs and the sweep orchestrator.
#include "mg_runtime.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "mg_batch.hpp"
#include "mg_mat_mfma.hpp"
#include "mg_plan.hpp"
#include "mg_sweep.hpp"
#include "mg_sweep_tile.hpp"
#include "mg_sweep_win.hpp"

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

// ---- graph capture and the calls that may not overlap it ---------------------------------
// Several host threads use the device at once (the server's lanes, each with its own stream).
// While one of them captures a HIP graph, another thread's allocation, free or synchronous copy
// (which orders itself against every stream of the device) fails the capture or itself; such
// calls and the capture sections hold this lock.
std::recursive_mutex& capture_mutex()
{
    static std::recursive_mutex m;
    return m;
}

hipError_t ensure_lds_limit(const void* kernel, size_t bytes)
{
    static std::mutex m;
    static std::map<const void*, size_t> set;
    std::lock_guard<std::mutex> g(m);
    size_t& cur = set[kernel];
    if (bytes <= cur) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) cur = bytes;
    return e;
}

// synchronous host-to-device copy, outside any other thread's graph capture
static hipError_t h2d(void* dst, const void* src, size_t bytes)
{
    std::lock_guard<std::recursive_mutex> g(capture_mutex());
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
}

hipError_t DevBuf::ensure(size_t bytes)
{
    if (bytes <= n && p) return hipSuccess;
    std::lock_guard<std::recursive_mutex> g(capture_mutex());
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
    std::lock_guard<std::recursive_mutex> g(capture_mutex());
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}

// ---- static device scratch of the allocation-free evaluation ----------------------------
// mg_evaluate_noalloc draws every device buffer of one single-request sweep (descriptors,
// workgroup partial sums, the reduced result) from this array instead of allocating.  Each
// device holds its own copy for the life of the process; host code reaches it only through
// the address hipGetSymbolAddress returns with that device current (Context::noalloc).
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

// ---- context (one per host thread, for the device current at its first use): a stream, the
// grow-only buffers and the captured graphs of that thread's evaluations, so that two threads
// (the server's lanes) can evaluate at the same time ------------------------------------------
struct Context {
    int device = -1;
    unsigned char* noalloc = nullptr;  // device address of `scratch` on this device
    hipStream_t s = nullptr;
    int ncu = 0;
    size_t lds_limit = 0;
    DevBuf partial, out, scratch, ckpt;
    DevBuf rows;    // materialised target rows of one column chunk (program targets beyond 12 qubits)
    DevBuf w_fwaf;  // L=12 forward-pass windows' MFMA fragments, formed on the device
    // host staging of the L=12 forward-pass window form, kept from one evaluation to the next
    std::vector<DWin> h_dwin;
    std::vector<int> h_wstart;
    std::vector<WinGate> h_wgate;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<double> out_h;
    // single-request evaluation: every descriptor of one evaluation in one device image, sent in
    // one copy from pinned staging; the event marks that copy's completion (the staging is reused
    // by the next evaluation); the reduced result comes back into pinned memory
    DevBuf e_up;
    void* e_up_h = nullptr;
    size_t e_up_hn = 0;
    hipEvent_t e_up_ev = nullptr;
    bool e_up_pending = false;
    void* e_out_h = nullptr;
    size_t e_out_hn = 0;
    // batched evaluation: packed descriptor upload, partials, outputs, scratch, counter
    DevBuf b_up, b_partial, b_out, b_scratch, b_counter;
    void* b_up_h = nullptr;   // pinned staging of the upload
    size_t b_up_hn = 0;
    void* b_out_h = nullptr;  // pinned staging of the outputs
    size_t b_out_hn = 0;

    // captured batched launches, least recently used evicted; keys launched once (eagerly)
    using GraphKey = std::array<long long, 21>;
    struct Graph {
        GraphKey key;
        hipGraphExec_t exec;
        long long used;
    };
    static constexpr size_t kGraphs = 16, kSeen = 64;
    std::vector<Graph> graphs;
    std::vector<GraphKey> seen;
    long long graph_tick = 0;

    hipGraphExec_t graph_find(const GraphKey& k)
    {
        for (Graph& g : graphs)
            if (g.key == k) {
                g.used = ++graph_tick;
                return g.exec;
            }
        return nullptr;
    }
    // true if k was launched before (and forgets it); otherwise remembers it
    bool graph_seen(const GraphKey& k)
    {
        auto it = std::find(seen.begin(), seen.end(), k);
        if (it != seen.end()) {
            seen.erase(it);
            return true;
        }
        if (seen.size() == kSeen) seen.erase(seen.begin());
        seen.push_back(k);
        return false;
    }
    void graph_insert(const GraphKey& k, hipGraphExec_t e)
    {
        if (graphs.size() == kGraphs) {
            auto lru = std::min_element(graphs.begin(), graphs.end(),
                                        [](const Graph& a, const Graph& b) { return a.used < b.used; });
            (void)hipGraphExecDestroy(lru->exec);
            graphs.erase(lru);
        }
        graphs.push_back({k, e, ++graph_tick});
    }
    ~Context()
    {
        for (Graph& g : graphs) (void)hipGraphExecDestroy(g.exec);
        std::lock_guard<std::recursive_mutex> lk(capture_mutex());
        for (void* h : {b_up_h, b_out_h, e_up_h, e_out_h})
            if (h) (void)hipHostFree(h);
    }
};

static Context* ctx()
{
    static thread_local std::unique_ptr<Context> c;
    if (!c) {
        auto n = std::make_unique<Context>();
        if (hipGetDevice(&n->device) != hipSuccess) return nullptr;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, n->device) != hipSuccess) return nullptr;
        n->ncu = prop.multiProcessorCount;
        n->lds_limit = prop.sharedMemPerBlockOptin ? prop.sharedMemPerBlockOptin : prop.sharedMemPerBlock;
        if (hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking) != hipSuccess) return nullptr;
        if (hipEventCreate(&n->ev0) != hipSuccess || hipEventCreate(&n->ev1) != hipSuccess ||
            hipEventCreateWithFlags(&n->e_up_ev, hipEventDisableTiming) != hipSuccess)
            return nullptr;
        void* sp = nullptr;
        const hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
            return nullptr;
        }
        n->noalloc = (unsigned char*)sp;
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
    if (h2d(t->seed64.p, seed.data(), S * S * sizeof(cpx<double>)) != hipSuccess) {
        set_error("copying the target to the device failed");
        return nullptr;
    }
    return t.release();
}

// in place: m holds U (row-major S x S) and receives the seed rows m[b][a] = -conj(U[a][b]); each
// pair of mirrored entries is swapped by the thread of its upper-triangle entry
__global__ void seed_rows_in_place(cpx<double>* m, size_t S)
{
    const size_t n = S * S;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        const size_t a = k / S, b = k % S;
        if (b < a) continue;
        const cpx<double> u = m[a * S + b], v = m[b * S + a];
        m[b * S + a] = {-u.re, u.im};
        if (b != a) m[a * S + b] = {-v.re, v.im};
    }
}

Target* target_dense(const zd* U, int L)
{
    Context* c = ctx();
    if (!c) { set_error("no HIP device"); return nullptr; }
    const size_t S = (size_t)1 << L;
    auto t = std::make_unique<Target>();
    t->kind = 0;
    t->L = L;
    // the matrix as given, then its seed rows formed on the device (a host transpose of a 2^12
    // target took longer than the upload)
    if (t->seed64.ensure(S * S * sizeof(cpx<double>)) != hipSuccess) { set_error("out of device memory for the target"); return nullptr; }
    if (hipMemcpyAsync(t->seed64.p, U, S * S * sizeof(cpx<double>), hipMemcpyHostToDevice, c->s) != hipSuccess) {
        set_error("copying the target to the device failed");
        return nullptr;
    }
    const long grid = std::min<long>((long)c->ncu * 8, std::max<long>(1, (long)((S * S + 255) / 256)));
    hipLaunchKernelGGL(seed_rows_in_place, dim3((unsigned)grid), dim3(256), 0, c->s, (cpx<double>*)t->seed64.p, S);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->s) != hipSuccess) {
        set_error("forming the target's seed rows failed");
        return nullptr;
    }
    return t.release();
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
    if (h2d(t->tinst.p, b.inst.data(), b.inst.size() * sizeof(Inst)) != hipSuccess ||
        h2d(t->tcoef64.p, b.coef.data(), b.coef.size() * sizeof(cpx<double>)) != hipSuccess) {
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

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// host staging of one upload: regions appended at 256-byte alignment
struct Arena {
    std::vector<unsigned char> h;
    size_t add(const void* p, size_t n)
    {
        const size_t o = align256(h.size());
        h.resize(o + n);
        if (n) std::memcpy(h.data() + o, p, n);
        return o;
    }
    template <typename T>
    size_t add(const std::vector<T>& v)
    {
        return add(v.data(), v.size() * sizeof(T));
    }
};

static int pinned_ensure(void** p, size_t* n, size_t bytes)
{
    if (*p && *n >= bytes) return 0;
    std::lock_guard<std::recursive_mutex> g(capture_mutex());
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *n = 0;
    if (hipHostMalloc(p, bytes ? bytes : 256, hipHostMallocDefault) != hipSuccess) {
        *p = nullptr;
        return -1;
    }
    *n = bytes;
    return 0;
}

// states per column of the tiled kernel: psi (f), + lam (gradient), + dpsi, dlam (HVP)
static int nstate_tile(int mode) { return mode == 2 ? 4 : mode == 1 ? 2 : 1; }

// window blocks of an instance list, reordered instances (lo = window-local pair
// index) and coefficients, block offsets for LDS (swizzled) and global states
struct Windowed {
    std::vector<Inst> inst;
    std::vector<cpx<double>> coef;
    std::vector<WBlock> blk_lds, blk_glob;
};

template <typename R>
static void make_windowed(const std::vector<Inst>& inst, const std::vector<cpx<double>>& coef, int L, int W,
                          Windowed* w, int kmax = kWinMaxGates, std::vector<int>* order_out = nullptr)
{
    std::vector<int> order;
    std::vector<WBlock> blocks;
    window_blocks(inst, L, W, kmax, &order, &blocks);
    const size_t G = inst.size();
    w->inst.resize(G);
    w->coef.resize(G * kCoefPerInst);
    for (const WBlock& bk : blocks)
        for (int k = bk.g0; k < bk.g1; k++) {
            Inst in = inst[order[k]];
            in.lo = window_case(bk, W, in.lo, in.hi);
            in.hi = 0;
            w->inst[k] = in;
            std::copy(coef.begin() + (size_t)order[k] * kCoefPerInst, coef.begin() + (size_t)(order[k] + 1) * kCoefPerInst,
                      w->coef.begin() + (size_t)k * kCoefPerInst);
        }
    w->blk_lds = blocks;
    w->blk_glob = blocks;
    for (size_t k = 0; k < blocks.size(); k++)
        for (int j = 0; j < (1 << W); j++) {
            uint32_t o = 0;
            for (int i = 0; i < W; i++) o |= ((uint32_t)(j >> i) & 1u) << blocks[k].pos[i];
            w->blk_glob[k].off[j] = o;
            w->blk_lds[k].off[j] = win_swz<batch_bank_bits<R>()>(o);
        }
    if (order_out) order_out->swap(order);
}

// tile blocks (12-qubit windows) and, per tile block, its 3-qubit window blocks, for the tiled
// register-window kernel: instances in the final order (lo = window-local pair index),
// coefficients, tile blocks, window blocks with LDS (swizzled tile-local) and global
// (amplitude-index) offsets
// gates of the largest window block: the kernels' staging slots
static int max_block_gates(const std::vector<WBlock>& blocks)
{
    int m = 1;
    for (const WBlock& bk : blocks) m = std::max(m, bk.g1 - bk.g0);
    return m;
}

struct WindowedTiles {
    std::vector<Inst> inst;
    std::vector<cpx<double>> coef;
    std::vector<WTBlock> tblocks;
    std::vector<WBlock> blk_lds, blk_glob;
};

static bool make_windowed_tiles(const std::vector<Inst>& inst, const std::vector<cpx<double>>& coef, int L,
                                WindowedTiles* w)
{
    constexpr int TB = 12, W = 3;
    std::vector<int> order1;
    std::vector<TBlock> tbs;
    tile_blocks(inst, L, TB, &order1, &tbs);
    if (order1.size() != inst.size()) return false;
    w->inst.clear();
    w->coef.clear();
    w->tblocks.clear();
    w->blk_lds.clear();
    w->blk_glob.clear();
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
        window_blocks(li, TB, W, kWinRegGates, &order2, &wbs);
        if (order2.size() != li.size()) return false;
        WTBlock t{};
        t.k0 = (int)w->blk_lds.size();
        const int base = (int)w->inst.size();
        for (int i = 0; i < TB; i++) t.tpos[i] = tb.pos[i];
        for (int i = 0; i < L - TB; i++) t.other[i] = tb.other[i];
        for (const WBlock& wb : wbs) {
            for (int k = wb.g0; k < wb.g1; k++) {
                const Inst& l = li[order2[k]];
                Inst f = inst[order1[tb.g0 + order2[k]]];
                f.lo = window_case(wb, W, l.lo, l.hi);
                f.hi = 0;
                w->inst.push_back(f);
                const size_t src = (size_t)order1[tb.g0 + order2[k]] * kCoefPerInst;
                w->coef.insert(w->coef.end(), coef.begin() + src, coef.begin() + src + kCoefPerInst);
            }
            WBlock bl = wb, bg = wb;
            bl.g0 = bg.g0 = base + wb.g0;
            bl.g1 = bg.g1 = base + wb.g1;
            for (int j = 0; j < (1 << W); j++) {
                uint32_t o = 0;
                for (int i = 0; i < W; i++) o |= ((uint32_t)(j >> i) & 1u) << wb.pos[i];
                bl.off[j] = win_swz<batch_bank_bits<double>()>(o);
                uint32_t gl = 0;
                for (int i = 0; i < TB; i++) gl |= ((o >> i) & 1u) << tb.pos[i];
                bg.off[j] = gl;
            }
            w->blk_lds.push_back(bl);
            w->blk_glob.push_back(bg);
        }
        t.k1 = (int)w->blk_lds.size();
        w->tblocks.push_back(t);
    }
    return true;
}

// the target program's 12-qubit tile form (window blocks inside tile blocks), uploaded once
static int prepare_tile_form(Context* c, Target& t)
{
    Target::TileForm& f = t.tf;
    if (f.ready) return 0;
    WindowedTiles wt;
    if (!make_windowed_tiles(t.tinst_h, t.tcoef_h, t.L, &wt)) {
        set_error("target materialisation: gate blocking failed");
        return -1;
    }
    MG_HIP(f.inst.ensure(wt.inst.size() * sizeof(Inst)));
    MG_HIP(f.coef.ensure(wt.coef.size() * sizeof(cpx<double>)));
    const size_t nb = wt.blk_lds.size() * sizeof(WBlock);
    MG_HIP(f.blk.ensure(2 * nb + wt.tblocks.size() * sizeof(WTBlock)));
    unsigned char* blk = (unsigned char*)f.blk.p;
    MG_HIP(h2d(f.inst.p, wt.inst.data(), wt.inst.size() * sizeof(Inst)));
    MG_HIP(h2d(f.coef.p, wt.coef.data(), wt.coef.size() * sizeof(cpx<double>)));
    MG_HIP(h2d(blk, wt.blk_lds.data(), nb));
    MG_HIP(h2d(blk + nb, wt.blk_glob.data(), nb));
    MG_HIP(h2d(blk + 2 * nb, wt.tblocks.data(), wt.tblocks.size() * sizeof(WTBlock)));
    f.G = (int)wt.inst.size();
    f.ntb = (int)wt.tblocks.size();
    f.nblk = (int)wt.blk_lds.size();
    f.slots = max_block_gates(wt.blk_lds);
    int rs = 0;
    win_sweep_lds_bytes(12, 0, 0, c->lds_limit, &f.acc_off, &f.buf_off, &rs, 2, f.slots);
    f.lds = f.buf_off + ((size_t)1 << 12) * sizeof(cpx<double>);
    f.ready = true;
    return 0;
}

// arguments of the materialisation of rows [b0, b1) into `rows` (row b at rows + (b - row0) S)
static WinTileArgs tile_form_args(const Target& t, long b0, long b1, void* rows, long row0)
{
    const Target::TileForm& f = t.tf;
    const size_t nb = (size_t)f.nblk * sizeof(WBlock);
    unsigned char* blk = (unsigned char*)f.blk.p;
    WinTileArgs a{};
    a.L = t.L;
    a.G = f.G;
    a.P = 0;
    a.ntb = f.ntb;
    a.inst = (const Inst*)f.inst.p;
    a.coef = f.coef.p;
    a.tblocks = (const WTBlock*)(blk + 2 * nb);
    a.blk_lds = (const WBlock*)blk;
    a.blk_glob = (const WBlock*)(blk + nb);
    a.b0 = b0;
    a.b1 = b1;
    a.nstate = 1;
    a.acc_off = (int)f.acc_off;
    a.buf_off = (int)f.buf_off;
    a.rs = 1;
    a.slots = f.slots;
    a.rows = rows;
    a.row0 = row0;
    return a;
}

static int tile_form_grid(const Context* c, const Target& t, long ncols)
{
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(2, c->lds_limit / t.tf.lds));
    return (int)std::min<long>(ncols, (long)c->ncu * per_cu);
}

// the target program's 4-qubit window form for the matrix-core materialisation, uploaded once
static int prepare_mat_form(Target& t)
{
    Target::MatForm& f = t.mf;
    if (f.ready) return 0;
    DenseWindows d;
    if (!make_dense_windows(t.tinst_h, t.tcoef_h, t.L, &d)) {
        set_error("target materialisation: 4-qubit window blocking failed");
        return -1;
    }
    MG_HIP(f.tb.ensure(d.tblocks.size() * sizeof(WTBlock)));
    MG_HIP(f.win.ensure(d.wins.size() * sizeof(DWin)));
    MG_HIP(f.af.ensure(d.afrag.size() * sizeof(double)));
    MG_HIP(f.af32.ensure(d.afrag32.size() * sizeof(float)));
    MG_HIP(h2d(f.tb.p, d.tblocks.data(), d.tblocks.size() * sizeof(WTBlock)));
    MG_HIP(h2d(f.win.p, d.wins.data(), d.wins.size() * sizeof(DWin)));
    MG_HIP(h2d(f.af.p, d.afrag.data(), d.afrag.size() * sizeof(double)));
    MG_HIP(h2d(f.af32.p, d.afrag32.data(), d.afrag32.size() * sizeof(float)));
    f.ntb = (int)d.tblocks.size();
    f.ready = true;
    return 0;
}

// rows [b0, b1) of a target program beyond 12 qubits, -conj(W e_b), into `rows` (row b at
// rows + (b - row0) S): on the matrix cores (4-qubit windows; FP64 MFMA, or FP32 MFMA into FP32
// rows for FP32 sweeps) or in 12-qubit register tiles (the tiled window kernel's forward sweep,
// FP64 rows only)
static bool rows_on_matrix_cores() { return knobs().mat_kernel != 0; }

static int launch_target_rows(Context* c, Target& t, long b0, long b1, void* rows, long row0, hipStream_t s,
                              int fp32 = 0)
{
    if (rows_on_matrix_cores()) {
        if (prepare_mat_form(t) < 0) return -1;
        MatArgs m{};
        m.L = t.L;
        m.ntb = t.mf.ntb;
        m.tblocks = (const WTBlock*)t.mf.tb.p;
        m.wins = (const DWin*)t.mf.win.p;
        m.afrag = fp32 ? t.mf.af32.p : t.mf.af.p;
        m.b0 = b0;
        m.b1 = b1;
        m.rows = rows;
        m.row0 = row0;
        const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(2, c->lds_limit / mat_mfma_lds(fp32)));
        const int grid = (int)std::min<long>(b1 - b0, (long)c->ncu * per_cu);
        MG_HIP(launch_materialize_mfma(m, fp32, grid, s));
        return 0;
    }
    if (fp32) {
        set_error("target materialisation: FP32 rows need the matrix-core kernel");
        return -1;
    }
    if (prepare_tile_form(c, t) < 0) return -1;
    const WinTileArgs a = tile_form_args(t, b0, b1, rows, row0);
    MG_HIP(launch_materialize_win_tile(a, tile_form_grid(c, t, b1 - b0), t.tf.lds, s));
    return 0;
}

// ---- evaluation -------------------------------------------------------------------------------
// Device buffers of one evaluation: the context's grow-only buffers, or (allocation-free
// evaluation) consecutive 256-byte aligned pieces of the static scratch array.
struct SweepMem {
    unsigned char* base = nullptr;  // null: the context's buffers
    size_t cap = 0, used = 0;
    int get(DevBuf& b, size_t bytes, void** p)
    {
        if (!base) {
            const hipError_t e = b.ensure(bytes);
            if (e != hipSuccess) {
                set_error(std::string("device allocation failed: ") + hipGetErrorString(e));
                return -1;
            }
            *p = b.p;
            return 0;
        }
        const size_t off = (used + 255) & ~(size_t)255;
        if (off + bytes > cap) {
            set_error("allocation-free evaluation needs more than the " + std::to_string(cap >> 20) +
                      " MiB static scratch (the states of this sweep fit neither the LDS nor registers, the gates need "
                      "checkpoints, or the partial sums are too large)");
            return -1;
        }
        used = off + bytes;
        *p = base + off;
        return 0;
    }
};

// Sends an evaluation's staged descriptors to the device in one copy and returns the device
// image.  From pinned staging the copy is asynchronous (a pageable copy stages through the driver
// and blocks the host, per region); the allocation-free evaluation, which allocates no pinned
// memory, copies from the host image and waits for it.
static int upload_flush(Context* c, SweepMem& ar, const Arena& up, hipStream_t s, unsigned char** dev)
{
    void* d = nullptr;
    const size_t n = std::max<size_t>(up.h.size(), 256);
    if (ar.get(c->e_up, n, &d) < 0) return -1;
    if (!up.h.empty()) {
        if (ar.base) {
            MG_HIP(hipMemcpyAsync(d, up.h.data(), up.h.size(), hipMemcpyHostToDevice, s));
            MG_HIP(hipStreamSynchronize(s));
        } else {
            // the staging is reused: the previous evaluation's copy out of it has completed (it
            // synchronised at its end, unless it failed in between)
            if (c->e_up_pending) MG_HIP(hipEventSynchronize(c->e_up_ev));
            c->e_up_pending = false;
            if (pinned_ensure(&c->e_up_h, &c->e_up_hn, n) < 0) {
                set_error("pinned staging allocation failed");
                return -1;
            }
            std::memcpy(c->e_up_h, up.h.data(), up.h.size());
            MG_HIP(hipMemcpyAsync(d, c->e_up_h, up.h.size(), hipMemcpyHostToDevice, s));
            MG_HIP(hipEventRecord(c->e_up_ev, s));
            c->e_up_pending = true;
        }
    }
    *dev = (unsigned char*)d;
    return 0;
}

// plan non-null: lay the sweep out in plan->mem (as ar) and record it instead of launching
template <typename R>
static int run_sweep(Context* c, Target& t, const Program& p, const Built& b, int mode, long b0, long b1, EvalOut* out,
                     EvalStats* st, SweepMem& ar, SweepPlan* plan = nullptr)
{
    hipStream_t s = c->s;
    const int L = p.L;
    const size_t S = (size_t)1 << L;
    const Knobs& kn = knobs();
    // threads of the LDS/global kernel: more waves per CU when one workgroup fills the LDS
    int lds_T = kn.lds_threads;
    {
        // 512 threads when the states fill the LDS (one workgroup per CU), else 256: two
        // 256-thread workgroups per CU when their states fit twice (measured at L=12 FP32:
        // f 0.66 -> 0.51 ms, f + gradient 1.94 -> 1.41 ms; each hides the other's barriers)
        const int nst0 = (mode == 2 ? 4 : 2);
        const size_t st0 = (size_t)nst0 * S * sizeof(cpx<R>);
        const bool fits = sweep_lds_head_bytes(p.P, 8, mode) + st0 <= c->lds_limit;
        const bool two = 2 * (sweep_lds_head_bytes(p.P, 4, mode) + st0) <= c->lds_limit;
        if (lds_T == 0) lds_T = fits && S >= 2048 && !two ? 512 : 256;
    }
    const int NW = lds_T / 64;

    // coefficients in precision R
    std::vector<cpx<R>> coefR(b.coef.size());
    for (size_t k = 0; k < coefR.size(); k++) coefR[k] = {(R)b.coef[k].re, (R)b.coef[k].im};
    // every descriptor of the evaluation is staged here (u_* offsets) and sent in one copy before
    // the first kernel
    Arena up;
    const size_t u_inst = up.add(b.inst), u_coef = up.add(coefR);

    // checkpointing: needed when psi_g cannot be recovered by V^{-1} (non-unitary gates)
    bool ckpt = false;
    if (mode >= 1) {
        if (kn.force_ckpt == 1) ckpt = true;
        else if (kn.force_ckpt == 0) ckpt = !b.invertible;
        else ckpt = !b.invertible || b.defect > 1e-10;
    }
    const int nstate = ckpt ? (mode == 2 ? 2 : 1) : (mode == 2 ? 4 : 2);
    const size_t head = sweep_lds_head_bytes(p.P, NW, mode);
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
        if (reg_T && sweep_lds_head_bytes(p.P, NW) + S * sizeof(cpx<R>) > c->lds_limit) reg_T = 0;
    }
    // register-resident gate-window kernel (FP64, 512 threads x 8 amplitudes, dense targets)
    size_t win_acc = 0, win_buf = 0, win_lds = 0;
    int win_rs = 0, win_slots = kWinRegGates;
    bool use_win = false;
    Windowed w_win;
    std::vector<int> win_order;  // the window form's instance order (program instance of each position)
    constexpr int fp32 = std::is_same<R, float>::value ? 1 : 0;
    if (!ckpt && b.defect <= 1e-12 && t.kind == 0 && win_sweep_fits(L, fp32) && kn.force_win != 0 && kn.force_reg != 0 &&
        kn.force_lds != 0 && (!fp32 || kn.force_win == 1)) {
        // the window form first: its largest block sizes the staging area
        make_windowed<R>(b.inst, b.coef, L, win_sweep_window_bits(fp32), &w_win, kWinRegGates, &win_order);
        win_slots = max_block_gates(w_win.blk_lds);
        win_lds = win_sweep_lds_bytes(L, p.P, fp32, c->lds_limit, &win_acc, &win_buf, &win_rs, mode, win_slots);
        use_win = win_lds <= c->lds_limit;
    }
    if (use_win) reg_T = 0;
    // tiled kernel: dense targets whose states do not fit the LDS (global state slices, one LDS
    // tile of 2^kw amplitudes per state and pass)
    const int tile_kw = tile_window_bits(fp32);
    // program targets beyond 12 qubits are read as dense rows too: the rows of one column chunk
    // at a time, materialised (12-qubit register tiles) into a chunk buffer just before the tiled
    // sweep of that chunk.  Applying the target's gates per column in the sweep instead streams
    // the states through memory once per gate window (HBM-bound at these sizes).
    const bool chunkable = t.kind == 1 && L > 12 && L - 12 <= 20 && kn.force_win != 0;
    size_t tile_off = 0, tile_lds = 0;
    bool use_tile = false;
    // the tiled kernel's accumulator: one slot per parameter gate, or, when those do not fit next
    // to the tile, one per parameter of the current tile block, added to the workgroup's partial
    // sums after each block of each column
    std::vector<int> tb_order, tb_slot, tb_param;
    std::vector<TBlock> tb_blocks;
    bool tile_slots = false;
    if (!use_win && !ckpt && !lds_states && (t.kind == 0 || chunkable) && L > tile_kw && kn.force_tile != 0 &&
        kn.force_lds == -1 && kn.force_reg != 1) {
        tile_lds = tile_sweep_lds_bytes(p.P, tile_kw, fp32, nstate_tile(mode), &tile_off);
        use_tile = tile_lds <= c->lds_limit;
        // f + gradient (two states, 64 KB): per-block slots also when they let two workgroups
        // share a CU and the full accumulator does not
        const bool want_two = mode == 1 && 2 * tile_lds > c->lds_limit;
        if ((!use_tile || want_two) && mode >= 1) {
            tile_blocks(b.inst, L, tile_kw, &tb_order, &tb_blocks);
            if (tb_order.size() == b.inst.size()) {
                tb_slot.resize(tb_order.size());
                int most = 0;
                for (TBlock& bk : tb_blocks) {
                    bk.s0 = (int)tb_param.size();
                    for (int k = bk.g0; k < bk.g1; k++) {
                        const int par = b.inst[tb_order[k]].param;
                        const auto it = std::find(tb_param.begin() + bk.s0, tb_param.end(), par);
                        tb_slot[k] = (int)(it - (tb_param.begin() + bk.s0));
                        if (it == tb_param.end()) tb_param.push_back(par);
                    }
                    bk.ns = (int)tb_param.size() - bk.s0;
                    most = std::max(most, bk.ns);
                }
                size_t slot_off = 0;
                const size_t slot_lds = tile_sweep_lds_bytes(most, tile_kw, fp32, nstate_tile(mode), &slot_off);
                if (slot_lds <= c->lds_limit && (!use_tile || 2 * slot_lds <= c->lds_limit)) {
                    use_tile = tile_slots = true;
                    tile_lds = slot_lds;
                    tile_off = slot_off;
                }
            }
        }
    }
    const bool chunk_rows = chunkable && use_tile;
    // FP32 tiled sweeps of a chunked target read FP32 rows made on the matrix cores (the FP64
    // register-window tiles need FP64 rows)
    const bool rows32 = chunk_rows && fp32 && rows_on_matrix_cores();
    auto row_bytes_of = [](bool r32, size_t n) { return n * (r32 ? sizeof(cpx<float>) : sizeof(cpx<double>)); };
    if (use_tile) reg_T = 0;
    // tiled register-window kernel (FP64, unitary gates): preferred over the LDS tiles
    bool use_wtile = false;
    size_t wt_acc = 0, wt_buf = 0, wt_lds = 0;
    int wt_rs = 0, wt_slots = kWinRegGates;
    WindowedTiles w_wt;
    if (use_tile && !fp32 && b.defect <= 1e-12 && kn.force_win != 0 && L > 12 && L - 12 <= 20 &&
        make_windowed_tiles(b.inst, b.coef, L, &w_wt)) {
        wt_slots = max_block_gates(w_wt.blk_lds);
        wt_lds = win_sweep_lds_bytes(12, p.P, 0, c->lds_limit, &wt_acc, &wt_buf, &wt_rs, mode, wt_slots);
        use_wtile = wt_lds <= c->lds_limit;
    }
    const int nw_reg = reg_T / 64;
    const size_t head_reg = reg_T ? sweep_lds_head_bytes(p.P, nw_reg) : 0;
    int reg_nst = 0;
    if (reg_T) {
        reg_nst = (int)std::min<size_t>(mode == 2 ? 4 : 2, (c->lds_limit - head_reg) / (S * sizeof(cpx<R>)));
        if (reg_nst == 3) reg_nst = 2;  // instantiated: 1, 2, 4 states per round
        if (reg_nst < 1) reg_T = 0;
    }
    const size_t lds = use_win ? win_lds
                     : use_wtile ? wt_lds
                     : use_tile ? tile_lds
                     : reg_T ? head_reg + (size_t)reg_nst * S * sizeof(cpx<R>) : (lds_states ? head + state_bytes : head);

    const long ncols = b1 - b0;
    int per_cu = kn.grid_per_cu > 0 ? kn.grid_per_cu : std::max<int>(1, std::min<size_t>(4, (160 * 1024) / std::max<size_t>(lds, 1)));
    if ((reg_T || use_win || use_tile) && kn.grid_per_cu <= 0) per_cu = std::max<int>(1, std::min<size_t>(2, (160 * 1024) / lds));
    long grid = std::min<long>(ncols, (long)c->ncu * per_cu);
    // bound the global scratch and checkpoint footprint (2 GiB each)
    const size_t lim = (size_t)2 << 30;
    const size_t tile_state_bytes = (size_t)nstate_tile(mode) * S * sizeof(cpx<R>);
    if (use_tile) grid = std::min<long>(grid, std::max<long>(1, lim / tile_state_bytes));
    else if (!lds_states && !reg_T && !use_win) grid = std::min<long>(grid, std::max<long>(1, lim / state_bytes));
    const size_t ck_per = ckpt ? (size_t)2 * (p.G + 1) * S * sizeof(cpx<R>) : 0;
    if (ckpt) grid = std::min<long>(grid, std::max<long>(1, lim / ck_per));
    grid = std::max<long>(grid, 1);
    // column chunks of materialised target rows: at most 4 GiB of rows, a multiple of the grid;
    // each chunk's workgroup partial sums are kept and all are reduced together
    long chunk = ncols, nchunks = 1;
    if (chunk_rows) {
        const long row_bytes = (long)(S * (rows32 ? sizeof(cpx<float>) : sizeof(cpx<double>)));
        const size_t cb = kn.row_chunk_bytes ? kn.row_chunk_bytes : ((size_t)4 << 30);
        chunk = std::max<long>(grid, (long)(cb / (size_t)row_bytes) / grid * grid);
        chunk = std::min<long>(chunk, ncols);
        nchunks = (ncols + chunk - 1) / chunk;
    }

    const int stride = 2 + 32 * p.P;
    void *d_partial = nullptr, *d_out = nullptr, *d_scratch = nullptr, *d_ckpt = nullptr, *d_rows = nullptr;
    if (ar.get(c->partial, (size_t)grid * nchunks * stride * sizeof(double), &d_partial) < 0 ||
        ar.get(c->out, (size_t)stride * sizeof(double), &d_out) < 0)
        return -1;
    if (chunk_rows) {
        if (ar.get(c->rows, (size_t)chunk * (size_t)row_bytes_of(rows32, S), &d_rows) < 0) return -1;
        if ((kn.mat_kernel != 0 ? prepare_mat_form(t) : prepare_tile_form(c, t)) < 0) return -1;
    }
    if (use_tile) {
        if (ar.get(c->scratch, (size_t)grid * tile_state_bytes, &d_scratch) < 0) return -1;
    } else if (!lds_states && !reg_T && !use_win) {
        if (ar.get(c->scratch, (size_t)grid * state_bytes, &d_scratch) < 0) return -1;
    }
    if (ckpt && ar.get(c->ckpt, (size_t)grid * ck_per, &d_ckpt) < 0) return -1;

    // the FP32 copy of dense rows: made for the kernels that read it (the tiled kernel rounds
    // the FP64 rows on load, so large targets keep one copy)
    if (fp32 && t.kind == 0 && !use_tile && !t.seed32_ready) {
        if (ar.base) {
            set_error("allocation-free FP32 evaluation: the target has no FP32 copy yet (one FP32 mg_evaluate makes it)");
            return -1;
        }
        if (ensure_fp32_target(t, s) < 0) return -1;
    }

    SweepArgs<R> a{};
    a.L = L;
    a.G = p.G;
    a.P = p.P;
    size_t ua_inst = u_inst, ua_coef = u_coef;  // a.inst, a.coef in the upload
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
    a.partial = (double*)d_partial;
    a.scratch = (cpx<R>*)d_scratch;
    a.ckpt = (cpx<R>*)d_ckpt;
    a.nstate = reg_T ? reg_nst : nstate;
    a.lds_state_offset = (int)(reg_T ? head_reg : head);
    // the L=12 forward-pass windows' matrices, formed on the device from this evaluation's
    // coefficients (launch_window_frags) once the upload is there: window gates at u_start and
    // u_gate, FP64 coefficients in program order at u_c64
    struct {
        int nwin = 0;
        size_t u_start = 0, u_gate = 0, u_c64 = 0;
        double* out64 = nullptr;
        float* out32 = nullptr;
    } frag;
    // FP32 sweeps with the states in LDS at L = 12: the forward pass (psi, and dpsi for the HVP) on
    // the matrix cores over 4-qubit windows; the instances go to the kernel in the windows' order (a
    // topological order of the program: the same product, and the reverse sweep runs it backwards)
    size_t ua_fwin = 0;
    if (fp32 && lds_states && !ckpt && !reg_T && !use_win && !use_tile && L == 12 && kn.fwd_mfma != 0 && NW <= 16) {
        std::vector<int> o2;
        const int nwin = forward_window_form(b.inst, false, &c->h_dwin, &c->h_wstart, &c->h_wgate, &o2);
        if (nwin < 1 || o2.size() != b.inst.size()) {
            set_error("LDS sweep: forward window blocking failed");
            return -1;
        }
        std::vector<Inst> wi(o2.size());
        std::vector<cpx<R>> wc(o2.size() * kCoefPerInst);
        for (size_t k = 0; k < o2.size(); k++) {
            wi[k] = b.inst[o2[k]];
            std::copy(coefR.begin() + (size_t)o2[k] * kCoefPerInst, coefR.begin() + (size_t)(o2[k] + 1) * kCoefPerInst,
                      wc.begin() + k * kCoefPerInst);
        }
        void* d_fa = nullptr;
        if (ar.get(c->w_fwaf, (size_t)nwin * (mode == 2 ? 16 : 8) * 64 * sizeof(float), &d_fa) < 0) return -1;
        ua_inst = up.add(wi);
        ua_coef = up.add(wc);
        ua_fwin = up.add(c->h_dwin.data(), nwin * sizeof(FWin));
        frag.nwin = nwin;
        frag.u_start = up.add(c->h_wstart);
        frag.u_gate = up.add(c->h_wgate);
        frag.u_c64 = up.add(b.coef);
        frag.out32 = (float*)d_fa;
        a.fwaf = (const float*)d_fa;
        a.nwin = nwin;
        if (plan) plan->korder = o2;
    }

    WinSweepArgs wa{};
    size_t uw_inst = 0, uw_coef = 0, uw_blk = 0, uw_glob = 0, uw_fwin = 0;  // wa's descriptors in the upload
    if (use_win) {
        const Windowed& w = w_win;
        std::vector<cpx<R>> wc(w.coef.size());
        for (size_t k = 0; k < wc.size(); k++) wc[k] = {(R)w.coef[k].re, (R)w.coef[k].im};
        uw_inst = up.add(w.inst);
        uw_coef = up.add(wc);
        uw_blk = up.add(w.blk_lds);
        uw_glob = up.add(w.blk_glob);
        wa.L = L;
        wa.G = p.G;
        wa.P = p.P;
        wa.nblocks = (int)w.blk_lds.size();
        wa.seed = fp32 ? t.seed32.p : t.seed64.p;
        wa.b0 = b0;
        wa.b1 = b1;
        wa.partial = (double*)d_partial;
        wa.acc_off = (int)win_acc;
        wa.buf_off = (int)win_buf;
        wa.rs = win_rs;
        wa.slots = win_slots;
        // FP64: the forward pass (psi, and dpsi for the HVP) on the matrix cores over 4-qubit
        // windows of the column held in the re-layout buffer (the HVP needs two states per round);
        // the windows' matrices are formed on the device from the coefficients in program order
        if (!fp32 && L == 12 && kn.fwd_mfma != 0 && (mode < 2 || win_rs >= 2)) {
            std::vector<int> o2;
            const int nwin = forward_window_form(b.inst, true, &c->h_dwin, &c->h_wstart, &c->h_wgate, &o2);
            if (nwin < 1) {
                set_error("window sweep: forward window blocking failed");
                return -1;
            }
            void* d_fa = nullptr;
            if (ar.get(c->w_fwaf, (size_t)nwin * (mode == 2 ? 16 : 8) * 64 * sizeof(double), &d_fa) < 0) return -1;
            uw_fwin = up.add(c->h_dwin.data(), nwin * sizeof(DWin));
            frag.nwin = nwin;
            frag.u_start = up.add(c->h_wstart);
            frag.u_gate = up.add(c->h_wgate);
            frag.u_c64 = u_coef;  // the sweep's coefficients in program order (R = double here)
            frag.out64 = (double*)d_fa;
            wa.fwaf = (const double*)d_fa;
            wa.nfwin = nwin;
        }
    }
    WinTileArgs wta{};
    size_t ut_inst = 0, ut_coef = 0, ut_lds = 0, ut_glob = 0, ut_tb = 0;  // wta's descriptors in the upload
    if (use_wtile) {
        const WindowedTiles& w = w_wt;
        ut_inst = up.add(w.inst);
        ut_coef = up.add(w.coef);
        ut_lds = up.add(w.blk_lds);
        ut_glob = up.add(w.blk_glob);
        ut_tb = up.add(w.tblocks);
        wta.L = L;
        wta.G = p.G;
        wta.P = p.P;
        wta.ntb = (int)w.tblocks.size();
        wta.seed = t.seed64.p;
        wta.b0 = b0;
        wta.b1 = b1;
        wta.partial = (double*)d_partial;
        wta.scratch = d_scratch;
        wta.nstate = nstate_tile(mode);
        wta.acc_off = (int)wt_acc;
        wta.buf_off = (int)wt_buf;
        wta.rs = wt_rs;
        wta.slots = wt_slots;
    }
    TileArgs<R> ta{};
    size_t ut_fwaf = 0, ut_slot = 0, ut_bpar = 0;  // ta's descriptors in the upload (with ut_inst ...)
    bool ut_fw = false;
    if (use_tile && !use_wtile) {
        std::vector<int> order;
        std::vector<TBlock> blocks;
        if (tile_slots) {
            order = tb_order;
            blocks = tb_blocks;
        } else
            tile_blocks(b.inst, L, tile_kw, &order, &blocks);
        if (order.size() != b.inst.size()) {
            set_error("tiled sweep: gate blocking failed");
            return -1;
        }
        // FP32: the forward pass of each tile block on the matrix cores, over 4-qubit windows (the
        // block's instances reordered into window order; the accumulator slots follow them)
        std::vector<FWin> fwins;
        std::vector<float> fwaf;
        if (fp32 && kn.fwd_mfma != 0) {
            std::vector<int> order_w(order.size()), slot_w(tb_slot.size());
            for (TBlock& bk : blocks) {
                std::vector<Inst> li;
                std::vector<cpx<double>> lc;
                for (int k = bk.g0; k < bk.g1; k++) {
                    Inst in = b.inst[order[k]];
                    in.lo = (int)(std::find(bk.pos, bk.pos + tile_kw, in.lo) - bk.pos);
                    in.hi = (int)(std::find(bk.pos, bk.pos + tile_kw, in.hi) - bk.pos);
                    if (in.lo >= tile_kw || in.hi >= tile_kw) {
                        set_error("tiled sweep: gate outside its window");
                        return -1;
                    }
                    li.push_back(in);
                    lc.insert(lc.end(), b.coef.begin() + (size_t)order[k] * kCoefPerInst,
                              b.coef.begin() + (size_t)(order[k] + 1) * kCoefPerInst);
                }
                std::vector<int> o2;
                bk.w0 = (int)fwins.size();
                if (forward_windows(li, lc, &o2, &fwins, &fwaf, mode == 2) < 0) {
                    set_error("tiled sweep: forward window blocking failed");
                    return -1;
                }
                bk.w1 = (int)fwins.size();
                for (int i = 0; i < bk.g1 - bk.g0; i++) {
                    order_w[bk.g0 + i] = order[bk.g0 + o2[i]];
                    if (tile_slots) slot_w[bk.g0 + i] = tb_slot[bk.g0 + o2[i]];
                }
            }
            order.swap(order_w);
            if (tile_slots) tb_slot.swap(slot_w);
        }
        std::vector<Inst> ti(order.size());
        std::vector<cpx<R>> tc(order.size() * kCoefPerInst);
        for (const TBlock& bk : blocks)
            for (int k = bk.g0; k < bk.g1; k++) {
                Inst in = b.inst[order[k]];
                // window-local bit positions (the window's positions are ascending)
                const int lo = (int)(std::find(bk.pos, bk.pos + tile_kw, in.lo) - bk.pos);
                const int hi = (int)(std::find(bk.pos, bk.pos + tile_kw, in.hi) - bk.pos);
                if (lo >= tile_kw || hi >= tile_kw) {
                    set_error("tiled sweep: gate outside its window");
                    return -1;
                }
                in.lo = lo;
                in.hi = hi;
                ti[k] = in;
                for (int e = 0; e < kCoefPerInst; e++) {
                    const cpx<double> v = b.coef[(size_t)order[k] * kCoefPerInst + e];
                    tc[(size_t)k * kCoefPerInst + e] = {(R)v.re, (R)v.im};
                }
            }
        ut_inst = up.add(ti);
        ut_coef = up.add(tc);
        ut_tb = up.add(blocks);
        if (!fwins.empty()) {
            uw_fwin = up.add(fwins);
            ut_fwaf = up.add(fwaf);
            ut_fw = true;
        }
        if (tile_slots) {
            ut_slot = up.add(tb_slot);
            ut_bpar = up.add(tb_param);
            int most = 0;
            for (const TBlock& bk : blocks) most = std::max(most, bk.ns);
            ta.nacc = most;
        }
        ta.L = L;
        ta.G = p.G;
        ta.P = p.P;
        ta.kw = tile_kw;
        ta.nblocks = (int)blocks.size();
        ta.seed = fp32 ? nullptr : (const cpx<R>*)t.seed64.p;
        ta.seed64 = fp32 ? (const cpx<double>*)t.seed64.p : nullptr;
        ta.b0 = b0;
        ta.b1 = b1;
        ta.partial = (double*)d_partial;
        ta.scratch = (cpx<R>*)d_scratch;
        ta.nstate = nstate_tile(mode);
        ta.tile_off = (int)tile_off;
    }

    // the upload, then the device pointers into it, and the forward windows' matrices
    unsigned char* dev = nullptr;
    if (upload_flush(c, ar, up, s, &dev) < 0) return -1;
    a.inst = (const Inst*)(dev + ua_inst);
    a.coef = (const cpx<R>*)(dev + ua_coef);
    if (frag.out32) a.fwin = dev + ua_fwin;
    if (use_win) {
        wa.inst = (const Inst*)(dev + uw_inst);
        wa.coef = dev + uw_coef;
        wa.blk_lds = (const WBlock*)(dev + uw_blk);
        wa.blk_glob = (const WBlock*)(dev + uw_glob);
        if (frag.out64) wa.fwin = dev + uw_fwin;
    }
    if (use_wtile) {
        wta.inst = (const Inst*)(dev + ut_inst);
        wta.coef = dev + ut_coef;
        wta.blk_lds = (const WBlock*)(dev + ut_lds);
        wta.blk_glob = (const WBlock*)(dev + ut_glob);
        wta.tblocks = (const WTBlock*)(dev + ut_tb);
    } else if (use_tile) {
        ta.inst = (const Inst*)(dev + ut_inst);
        ta.coef = (const cpx<R>*)(dev + ut_coef);
        ta.blocks = (const TBlock*)(dev + ut_tb);
        if (ut_fw) {
            ta.fwin = dev + uw_fwin;
            ta.fwaf = (const float*)(dev + ut_fwaf);
        }
        if (tile_slots) {
            ta.slot = (const int*)(dev + ut_slot);
            ta.bparam = (const int*)(dev + ut_bpar);
        }
    }
    if (plan) {
        // the sweeps with a plan: LDS-resident states (no checkpoints, registers or tiles), or the
        // register gate-window kernel
        const bool lds_kernel = lds_states && !ckpt && !reg_T && !use_win && !use_tile;
        plan->kind = use_win ? 2 : lds_kernel ? 1 : 0;
        if (plan->kind == 0) return 0;
        plan->mode = mode;
        plan->fp32 = fp32;
        plan->grid = (int)grid;
        plan->threads = use_win ? 512 : lds_T;
        plan->lds = lds;
        plan->stride = 2 + 32 * p.P;
        plan->partial = (double*)d_partial;
        plan->out = (double*)d_out;
        plan->nwin = frag.nwin;
        plan->wstart = frag.nwin ? (const int*)(dev + frag.u_start) : nullptr;
        plan->wgate = frag.nwin ? (const WinGate*)(dev + frag.u_gate) : nullptr;
        plan->frag64 = frag.out64;
        plan->frag32 = frag.out32;
        plan->prog64 = frag.nwin ? (cpx<double>*)(dev + frag.u_c64) : nullptr;
        if (use_win) {
            plan->wa = wa;
            plan->kcoef = const_cast<void*>(wa.coef);
            plan->korder = win_order;
        } else {
            if constexpr (std::is_same<R, double>::value) plan->a64 = a;
            else plan->a32 = a;
            plan->kcoef = const_cast<cpx<R>*>(a.coef);
            if (!frag.out32) {
                plan->korder.resize(p.G);
                for (int g = 0; g < p.G; g++) plan->korder[g] = g;
            }
        }
        if ((int)plan->korder.size() != p.G) {
            set_error("sweep plan: instance order incomplete");
            return -1;
        }
        return 0;
    }
    if (frag.nwin)
        MG_HIP(launch_window_frags(frag.nwin, (const int*)(dev + frag.u_start), (const WinGate*)(dev + frag.u_gate),
                                   (const cpx<double>*)(dev + frag.u_c64), mode == 2, frag.out64, frag.out32, s));
    if (kn.time_kernels) MG_HIP(hipEventRecord(c->ev0, s));
    if (chunk_rows) {
        for (long ci = 0; ci < nchunks; ci++) {
            const long c0 = b0 + ci * chunk, c1 = std::min<long>(b1, c0 + chunk);
            if (launch_target_rows(c, t, c0, c1, d_rows, c0, s, rows32 ? 1 : 0) < 0) return -1;
            double* part = (double*)d_partial + (size_t)ci * grid * stride;
            if (use_wtile) {
                WinTileArgs w2 = wta;
                w2.b0 = c0;
                w2.b1 = c1;
                w2.seed = d_rows;
                w2.row0 = c0;
                w2.partial = part;
                MG_HIP(launch_sweep_win_tile(w2, mode, (int)grid, lds, s));
            } else {
                TileArgs<R> t2 = ta;
                t2.b0 = c0;
                t2.b1 = c1;
                t2.row0 = c0;
                // FP32 sweeps read FP32 rows from the matrix cores, else the FP64 rows (rounded on load)
                t2.seed = (!fp32 || rows32) ? (const cpx<R>*)d_rows : nullptr;
                t2.seed64 = (fp32 && !rows32) ? (const cpx<double>*)d_rows : nullptr;
                t2.partial = part;
                MG_HIP(launch_sweep_tile<R>(t2, mode, (int)grid, lds, s));
            }
        }
    } else if (use_wtile)
        MG_HIP(launch_sweep_win_tile(wta, mode, (int)grid, lds, s));
    else if (use_tile)
        MG_HIP(launch_sweep_tile<R>(ta, mode, (int)grid, lds, s));
    else if (use_win)
        MG_HIP(launch_sweep_win(wa, mode, fp32, (int)grid, lds, s));
    else if (reg_T)
        MG_HIP(launch_sweep_reg<R>(a, mode, reg_T, reg_qpt, (int)grid, lds, s));
    else
        MG_HIP(launch_sweep<R>(a, mode, lds_states, ckpt, lds_T, (int)grid, lds, s));
    if (kn.time_kernels) MG_HIP(hipEventRecord(c->ev1, s));
    MG_HIP(launch_reduce((const double*)d_partial, (int)(grid * nchunks), stride, (double*)d_out, s));
    // the result into pinned memory (the allocation-free evaluation: pageable)
    double* oh = nullptr;
    if (!ar.base && pinned_ensure(&c->e_out_h, &c->e_out_hn, stride * sizeof(double)) == 0) {
        oh = (double*)c->e_out_h;
    } else {
        c->out_h.resize(stride);
        oh = c->out_h.data();
    }
    MG_HIP(hipMemcpyAsync(oh, d_out, stride * sizeof(double), hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
    c->e_up_pending = false;

    const double* o = oh;
    out->f = zd(o[0], o[1]);
    out->grad.assign((size_t)8 * p.P, zd(0.0));
    out->dhz.assign((size_t)8 * p.P, zd(0.0));
    for (int q = 0; q < p.P; q++)
        for (int e = 0; e < 8; e++) {
            out->grad[8 * q + e] = zd(o[2 + 32 * q + 2 * e], o[2 + 32 * q + 2 * e + 1]);
            out->dhz[8 * q + e] = zd(o[2 + 32 * q + 16 + 2 * e], o[2 + 32 * q + 16 + 2 * e + 1]);
        }
    if (st) {
        st->lds_states = lds_states && !reg_T && !use_win && !use_tile;
        st->ckpt = ckpt;
        st->reg_threads = use_win ? 512 : reg_T;
        st->threads = use_win || use_tile ? 512 : reg_T ? reg_T : lds_T;  // the tiled kernel: 512 in both precisions
        st->kernel = use_win     ? "sweep_win_kernel"
                     : use_wtile ? "sweep_win_tile_kernel"
                     : use_tile  ? "sweep_tile_kernel"
                     : reg_T     ? "sweep_reg_kernel"
                                 : "sweep_kernel";
        st->grid = (int)grid;
        st->lds_bytes = lds;
        st->kernel_ms = 0.f;
        if (kn.time_kernels) MG_HIP(hipEventElapsedTime(&st->kernel_ms, c->ev0, c->ev1));
    }
    return 0;
}

int evaluate(const Target& tc, const Program& p, const zd* gates, const zd* Z, int mode, int precision, EvalOut* out,
             long b0, long b1, EvalStats* stats, bool noalloc)
{
    Context* c = ctx();
    if (!c) { set_error("no HIP device available"); return -1; }
    SweepMem ar;
    if (noalloc) {
        ar.base = c->noalloc;
        ar.cap = kNoallocScratchBytes;
    }
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
        // FP32 coefficients of a target program (dense rows: see run_sweep)
        if (t.kind == 1 && !t.tcoef32_ready) {
            if (noalloc) {
                set_error("allocation-free FP32 evaluation: the target has no FP32 copy yet (one FP32 mg_evaluate makes it)");
                return -1;
            }
            if (ensure_fp32_target(t, c->s) < 0) return -1;
        }
        return run_sweep<float>(c, t, p, b, mode, b0, b1, out, stats, ar);
    }
    return run_sweep<double>(c, t, p, b, mode, b0, b1, out, stats, ar);
}

int plan_sweep(const Target& tc, const Program& p, const zd* gates, int mode, int precision, SweepPlan* plan)
{
    Context* c = ctx();
    if (!c) { set_error("no HIP device available"); return -1; }
    Target& t = const_cast<Target&>(tc);
    std::string why;
    if (!p.validate(&why)) { set_error("program: " + why); return -1; }
    if (t.L != p.L || t.kind != 0) { set_error("sweep plan: needs a dense target of the program's qubit count"); return -1; }
    *plan = SweepPlan{};
    // the FP32 copy of the rows now: the plan's layout takes no allocation of its own
    if (precision == kFP32 && !t.seed32_ready && ensure_fp32_target(t, c->s) < 0) return -1;
    // directions of zero in the layout; the plan's users rewrite all coefficients before a launch
    std::vector<zd> zero((size_t)8 * p.P, zd(0.0));
    Built b;
    build_instances(p, gates, zero.data(), &b);
    // everything in one allocation of the plan's own, laid out as the allocation-free evaluation
    // does: the workgroup partial sums (at most 4 workgroups per CU) and 8 MiB for the descriptors
    const size_t bytes = (size_t)c->ncu * 4 * (2 + 32 * (size_t)p.P) * sizeof(double) + ((size_t)8 << 20);
    MG_HIP(plan->mem.ensure(bytes));
    SweepMem ar;
    ar.base = (unsigned char*)plan->mem.p;
    ar.cap = bytes;
    const long S = 1L << p.L;
    if (precision == kFP32) return run_sweep<float>(c, t, p, b, mode, 0, S, nullptr, nullptr, ar, plan);
    return run_sweep<double>(c, t, p, b, mode, 0, S, nullptr, nullptr, ar, plan);
}

hipError_t launch_plan(const SweepPlan& plan, const int* skip, hipStream_t s)
{
    hipError_t e = hipSuccess;
    if (plan.nwin)
        e = launch_window_frags(plan.nwin, plan.wstart, plan.wgate, plan.prog64, plan.mode == 2, plan.frag64, plan.frag32,
                                s, skip);
    if (e != hipSuccess) return e;
    if (plan.kind == 2) {
        WinSweepArgs wa = plan.wa;
        wa.skip = skip;
        e = launch_sweep_win(wa, plan.mode, plan.fp32, plan.grid, plan.lds, s);
    } else if (plan.kind == 1 && plan.fp32) {
        SweepArgs<float> a = plan.a32;
        a.skip = skip;
        e = launch_sweep<float>(a, plan.mode, true, false, plan.threads, plan.grid, plan.lds, s);
    } else if (plan.kind == 1) {
        SweepArgs<double> a = plan.a64;
        a.skip = skip;
        e = launch_sweep<double>(a, plan.mode, true, false, plan.threads, plan.grid, plan.lds, s);
    } else {
        return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    return launch_reduce(plan.partial, plan.grid, plan.stride, plan.out, s, skip);
}

}  // namespace mg

// ---- batched evaluation ------------------------------------------------------------------------
namespace mg {

BatchTuning& batch_tuning()
{
    static BatchTuning t;
    return t;
}

bool batch_single_request(const Target& t, const Program& p)
{
    const BatchTuning& bt = batch_tuning();
    // dense targets, and program targets beyond 12 qubits (read in materialised column chunks)
    return (t.kind == 0 || p.L > 12) && p.L >= bt.single_min && p.L <= bt.single_max;
}

BatchStats& last_batch_stats()
{
    static thread_local BatchStats s;
    return s;
}

void batch_shape(int L, int precision, int* cgw, int* cpi)
{
    const BatchTuning& bt = batch_tuning();
    const int pi = precision == kFP32 ? 1 : 0;
    const int W = bt.wbits[pi], threads = bt.threads[pi];
    const long S = 1L << L, ntup = std::max(1L, S >> W);
    const long wmax = std::min<long>(threads, ntup);
    long w = (L < 31 && bt.cgw[pi][L]) ? bt.cgw[pi][L] : wmax;
    w = std::max(1L, std::min(wmax, w));
    while (w & (w - 1)) w &= w - 1;  // a power of two
    const int ncc = (int)(threads / w);
    long c = (L < 31 && bt.cpi[pi][L]) ? bt.cpi[pi][L] : std::max<long>(ncc, S / 256);
    c = (c + ncc - 1) / ncc * ncc;
    *cgw = (int)w;
    *cpi = (int)std::min<long>(c, S);
}

// device copy of a target program in window form (per precision, built on first use)
template <typename R>
static int ensure_windowed_target(Target& t)
{
    const int pi = std::is_same<R, float>::value ? 1 : 0;
    Target::WinCache& wc = t.wc[pi];
    const int W = batch_tuning().wbits[pi];
    if (wc.wbits == W) return 0;
    Windowed w;
    make_windowed<R>(t.tinst_h, t.tcoef_h, t.L, W, &w);
    std::vector<cpx<R>> cR(w.coef.size());
    for (size_t k = 0; k < cR.size(); k++) cR[k] = {(R)w.coef[k].re, (R)w.coef[k].im};
    MG_HIP(wc.inst.ensure(w.inst.size() * sizeof(Inst)));
    MG_HIP(wc.coef.ensure(cR.size() * sizeof(cpx<R>)));
    MG_HIP(wc.blk_lds.ensure(w.blk_lds.size() * sizeof(WBlock)));
    MG_HIP(wc.blk_glob.ensure(w.blk_glob.size() * sizeof(WBlock)));
    MG_HIP(h2d(wc.inst.p, w.inst.data(), w.inst.size() * sizeof(Inst)));
    MG_HIP(h2d(wc.coef.p, cR.data(), cR.size() * sizeof(cpx<R>)));
    MG_HIP(h2d(wc.blk_lds.p, w.blk_lds.data(), w.blk_lds.size() * sizeof(WBlock)));
    MG_HIP(h2d(wc.blk_glob.p, w.blk_glob.data(), w.blk_glob.size() * sizeof(WBlock)));
    wc.nblocks = (int)w.blk_lds.size();
    wc.wbits = W;
    return 0;
}

int target_densify(Target& t)
{
    if (t.kind != 1 || t.L < 3) return 0;
    Context* c = ctx();
    if (!c) { set_error("no HIP device"); return -1; }
    const int L = t.L;
    const size_t S = (size_t)1 << L;
    // the FP64 window form (3 bits) of the target program, LDS offsets
    Windowed w;
    make_windowed<double>(t.tinst_h, t.tcoef_h, L, 3, &w);
    DevBuf di, dc, db;
    MG_HIP(di.ensure(w.inst.size() * sizeof(Inst)));
    MG_HIP(dc.ensure(w.coef.size() * sizeof(cpx<double>)));
    MG_HIP(db.ensure(w.blk_lds.size() * sizeof(WBlock)));
    MG_HIP(h2d(di.p, w.inst.data(), w.inst.size() * sizeof(Inst)));
    MG_HIP(h2d(dc.p, w.coef.data(), w.coef.size() * sizeof(cpx<double>)));
    const long cgw = std::min<long>(256, std::max<long>(1, (long)(S >> 3)));
    const size_t lds = (size_t)(256 / cgw) * S * sizeof(cpx<double>);
    const bool in_lds = lds <= c->lds_limit - 256;
    // LDS-resident columns use the swizzled block offsets, in-place rows the plain ones
    const std::vector<WBlock>& bl = in_lds ? w.blk_lds : w.blk_glob;
    MG_HIP(h2d(db.p, bl.data(), bl.size() * sizeof(WBlock)));
    if (t.seed64.ensure(S * S * sizeof(cpx<double>)) != hipSuccess) { set_error("out of device memory for the target"); return -1; }
    // beyond the LDS: 12-qubit tiles of each row in registers (the tiled window kernel's
    // forward sweep), unless the knob asks for the windowed in-place rows
    if (!in_lds && L > 12 && L - 12 <= 20 && knobs().force_win != 0) {
        if (launch_target_rows(c, t, 0, (long)S, t.seed64.p, 0, c->s) < 0) return -1;
        MG_HIP(hipStreamSynchronize(c->s));
        t.kind = 0;
        t.seed32_ready = false;
        return 0;
    }
    if (in_lds) {
        const int grid = (int)std::max<long>(1, std::min<long>((long)(S / (256 / cgw)), (long)c->ncu * 2));
        MG_HIP(launch_materialize(L, (const Inst*)di.p, (const cpx<double>*)dc.p, (const WBlock*)db.p, (int)bl.size(),
                                  (cpx<double>*)t.seed64.p, grid, lds, c->s));
    } else {
        const int grid = (int)std::min<long>((long)S, (long)c->ncu * 4);
        MG_HIP(launch_materialize_global(L, (const Inst*)di.p, (const cpx<double>*)dc.p, (const WBlock*)db.p,
                                         (int)bl.size(), (cpx<double>*)t.seed64.p, grid, c->s));
    }
    MG_HIP(hipStreamSynchronize(c->s));
    t.kind = 0;
    t.seed32_ready = false;
    return 0;
}

template <typename R>
static int run_batch(Context* c, const std::vector<int>& idx, BatchReq* reqs, BatchStats* st)
{
    const auto t_start = std::chrono::steady_clock::now();
    hipStream_t s = c->s;
    const int prec = std::is_same<R, float>::value ? kFP32 : kFP64;
    const int pi = prec == kFP32 ? 1 : 0;
    const int T = batch_tuning().threads[pi], W = batch_tuning().wbits[pi];
    if (!batch_variant_exists(pi, T, W)) {
        set_error("batched evaluation: no kernel for " + std::to_string(T) + " threads and " + std::to_string(W) +
                  " window bits");
        return -1;
    }
    const int n = (int)idx.size();
    const size_t fixed = batch_lds_fixed_bytes(T);
    const size_t lds_cap = c->lds_limit - 256;  // the kernel also holds 16 bytes of static LDS

    std::vector<BatchInst<R>> insts(n);
    std::vector<int> cpis(n);
    size_t lds = fixed, scratch_stride = 0;
    Arena ar;
    // per instance: offsets of its gate instances, coefficients and blocks in the arena
    std::vector<size_t> o_inst(n), o_coef(n), o_blk(n);
    for (int i = 0; i < n; i++) {
        const BatchReq& r = reqs[idx[i]];
        Built built;
        build_instances(*r.p, r.gates, r.Z, &built);
        if (r.mode >= 1 && !built.invertible) {
            set_error("batched evaluation: singular gate block (cannot uncompute)");
            return -1;
        }
        if (prec == kFP32 && ensure_fp32_target(*r.t, s) < 0) return -1;
        if (r.t->kind == 1 && ensure_windowed_target<R>(*r.t) < 0) return -1;
        const int L = r.p->L;
        const size_t S = (size_t)1 << L;
        const int nsl = r.mode == 2 ? 4 : 2;
        int cgw, cpi;
        batch_shape(L, prec, &cgw, &cpi);
        BatchInst<R>& in = insts[i];
        in = BatchInst<R>{};
        in.L = L;
        in.G = r.p->G;
        in.P = r.p->P;
        in.mode = r.mode;
        in.target_kind = r.t->kind;
        in.Gt = r.t->kind == 1 ? r.t->Gt : 0;
        in.acc_off = (int)fixed;
        in.state_off = (int)align256(fixed + sizeof(double) * 32 * (size_t)in.P);
        if ((size_t)in.state_off > lds_cap) {
            set_error("batched evaluation: too many parameter gates for the LDS accumulator");
            return -1;
        }
        in.cgw = cgw;
        const size_t col_states = (size_t)(T / cgw) * nsl * S;
        size_t need = in.state_off + col_states * sizeof(cpx<R>);
        in.lds = need <= lds_cap ? 1 : 0;
        if (!in.lds) {
            need = in.state_off;
            scratch_stride = std::max(scratch_stride, col_states);
        }
        cpis[i] = cpi;
        lds = std::max(lds, need);

        Windowed w;
        make_windowed<R>(built.inst, built.coef, L, W, &w);
        std::vector<cpx<R>> cR(w.coef.size());
        for (size_t k = 0; k < cR.size(); k++) cR[k] = {(R)w.coef[k].re, (R)w.coef[k].im};
        o_inst[i] = ar.add(w.inst.data(), w.inst.size() * sizeof(Inst));
        o_coef[i] = ar.add(cR.data(), cR.size() * sizeof(cpx<R>));
        const std::vector<WBlock>& bl = in.lds ? w.blk_lds : w.blk_glob;
        o_blk[i] = ar.add(bl.data(), bl.size() * sizeof(WBlock));
        in.nblocks = (int)bl.size();
        if (r.t->kind == 0) {
            in.seed = prec == kFP32 ? (const cpx<R>*)r.t->seed32.p : (const cpx<R>*)r.t->seed64.p;
        } else {
            const Target::WinCache& wc = r.t->wc[pi];
            in.tinst = (const Inst*)wc.inst.p;
            in.tcoef = (const cpx<R>*)wc.coef.p;
            in.tblocks = (const WBlock*)(in.lds ? wc.blk_lds.p : wc.blk_glob.p);
            in.ntblocks = wc.nblocks;
        }
    }

    // work items and reduction segments
    struct ItemCost {
        BatchItem it;
        double cost;
    };
    std::vector<ItemCost> items;
    std::vector<ReduceSeg> segs(n);
    std::vector<int2> blocks;
    long long poff = 0, ooff = 0;
    std::vector<long long> out_off(n);
    for (int i = 0; i < n; i++) {
        const BatchInst<R>& in = insts[i];
        const BatchReq& rq = reqs[idx[i]];
        const long S = 1L << in.L;
        const long c0 = rq.b1 > 0 ? rq.b0 : 0, c1 = rq.b1 > 0 ? rq.b1 : S;
        const int stride = 2 + 32 * in.P;
        const double per_col = (double)S * ((double)in.G * (in.mode == 2 ? 18 : in.mode == 1 ? 6 : 2) + 2.0 * in.Gt + 1.0);
        segs[i].src = poff;
        segs[i].dst = ooff;
        segs[i].stride = stride;
        int cnt = 0;
        for (long b = c0; b < c1; b += cpis[i]) {
            const long e = std::min<long>(c1, b + cpis[i]);
            items.push_back({BatchItem{i, (int)b, (int)e, 0, poff}, per_col * (double)(e - b)});
            poff += stride;
            cnt++;
        }
        segs[i].n = cnt;
        out_off[i] = ooff;
        for (int k0 = 0; k0 < stride; k0 += 64) blocks.push_back(int2{i, k0});
        ooff += stride;
    }
    std::stable_sort(items.begin(), items.end(), [](const ItemCost& a, const ItemCost& b) { return a.cost > b.cost; });
    const int nitems = (int)items.size();
    std::vector<BatchItem> items_h(nitems);
    for (int k = 0; k < nitems; k++) items_h[k] = items[k].it;

    // packed upload: per-instance arrays (above), items, segments, reduce blocks, descriptors
    const size_t o_items = ar.add(items_h.data(), items_h.size() * sizeof(BatchItem));
    const size_t o_segs = ar.add(segs.data(), segs.size() * sizeof(ReduceSeg));
    const size_t o_rblk = ar.add(blocks.data(), blocks.size() * sizeof(int2));
    const size_t o_insts = align256(ar.h.size());
    const size_t up_bytes = align256(o_insts + n * sizeof(BatchInst<R>));
    MG_HIP(c->b_up.ensure(up_bytes));
    unsigned char* d = (unsigned char*)c->b_up.p;
    for (int i = 0; i < n; i++) {
        insts[i].inst = (const Inst*)(d + o_inst[i]);
        insts[i].coef = (const cpx<R>*)(d + o_coef[i]);
        insts[i].blocks = (const WBlock*)(d + o_blk[i]);
    }
    if (pinned_ensure(&c->b_up_h, &c->b_up_hn, up_bytes) < 0) { set_error("pinned staging allocation failed"); return -1; }
    unsigned char* h = (unsigned char*)c->b_up_h;
    std::memcpy(h, ar.h.data(), ar.h.size());
    std::memcpy(h + o_insts, insts.data(), n * sizeof(BatchInst<R>));

    const Knobs& kn = knobs();
    int per_cu = batch_tuning().grid_per_cu;
    if (per_cu <= 0) per_cu = (int)std::max<size_t>(1, std::min<size_t>(T == 256 ? 4 : 2, c->lds_limit / lds));
    const int grid = (int)std::max<long>(1, std::min<long>(nitems, (long)c->ncu * per_cu));
    MG_HIP(c->b_partial.ensure((size_t)poff * sizeof(double)));
    MG_HIP(c->b_out.ensure((size_t)ooff * sizeof(double)));
    MG_HIP(c->b_counter.ensure(sizeof(int)));
    if (scratch_stride) MG_HIP(c->b_scratch.ensure((size_t)grid * scratch_stride * sizeof(cpx<R>)));
    if (pinned_ensure(&c->b_out_h, &c->b_out_hn, (size_t)ooff * sizeof(double)) < 0) {
        set_error("pinned staging allocation failed");
        return -1;
    }

    const double prep_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    // upload, counter reset + persistent kernel, reduction, download; `events` brackets the kernel
    auto enqueue = [&](bool events) -> hipError_t {
        hipError_t e = hipMemcpyAsync(c->b_up.p, c->b_up_h, up_bytes, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && events) e = hipEventRecord(c->ev0, s);
        if (e == hipSuccess)
            e = launch_batch<R>((const BatchInst<R>*)(d + o_insts), (const BatchItem*)(d + o_items), nitems,
                                (int*)c->b_counter.p, (double*)c->b_partial.p,
                                scratch_stride ? (cpx<R>*)c->b_scratch.p : nullptr, (long long)scratch_stride, T, W, grid,
                                lds, s);
        if (e == hipSuccess && events) e = hipEventRecord(c->ev1, s);
        if (e == hipSuccess)
            e = launch_batch_reduce((const ReduceSeg*)(d + o_segs), (const int2*)(d + o_rblk), (int)blocks.size(),
                                    (const double*)c->b_partial.p, (double*)c->b_out.p, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->b_out_h, c->b_out.p, (size_t)ooff * sizeof(double), hipMemcpyDeviceToHost, s);
        return e;
    };
    // Every node argument of the sequence (sizes, offsets, buffer and staging addresses) is in the
    // key; the request descriptors travel in the upload.  A key launched a second time is captured
    // once and replayed from then on.
    hipGraphExec_t graph = nullptr;
    bool captured = false;
    if (batch_tuning().graphs) {
        const Context::GraphKey key{prec, T, W, nitems, grid, (long long)lds, (long long)up_bytes, ooff,
                                    (long long)blocks.size(), (long long)scratch_stride, (long long)o_insts,
                                    (long long)o_items, (long long)o_segs, (long long)o_rblk, (long long)(uintptr_t)c->b_up.p,
                                    (long long)(uintptr_t)c->b_up_h, (long long)(uintptr_t)c->b_partial.p,
                                    (long long)(uintptr_t)c->b_out.p, (long long)(uintptr_t)c->b_out_h,
                                    (long long)(uintptr_t)c->b_counter.p, (long long)(uintptr_t)c->b_scratch.p};
        graph = c->graph_find(key);
        if (!graph && c->graph_seen(key)) {
            // the dynamic-LDS attribute is a function attribute, set outside the capture
            MG_HIP(launch_batch<R>(nullptr, nullptr, 0, nullptr, nullptr, nullptr, 0, T, W, -1, lds, s));
            std::unique_lock<std::recursive_mutex> cap(capture_mutex());
            MG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const hipError_t e = enqueue(false);
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(s, &g);
            cap.unlock();
            if (e != hipSuccess || ee != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                set_error(std::string("batched evaluation: graph capture failed: ") +
                          hipGetErrorString(e != hipSuccess ? e : ee));
                return -1;
            }
            const hipError_t ei = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                set_error(std::string("batched evaluation: graph instantiation failed: ") + hipGetErrorString(ei));
                return -1;
            }
            c->graph_insert(key, graph);
            captured = true;
        }
    }
    if (graph) {
        // the kernel's dynamic-LDS limit as at capture (another launch may have changed it)
        if (!captured) MG_HIP(launch_batch<R>(nullptr, nullptr, 0, nullptr, nullptr, nullptr, 0, T, W, -1, lds, s));
        // timing brackets the whole graph (upload and download included)
        if (kn.time_kernels) MG_HIP(hipEventRecord(c->ev0, s));
        MG_HIP(hipGraphLaunch(graph, s));
        if (kn.time_kernels) MG_HIP(hipEventRecord(c->ev1, s));
    } else {
        MG_HIP(enqueue(kn.time_kernels));
    }
    MG_HIP(hipStreamSynchronize(s));

    const double* o = (const double*)c->b_out_h;
    for (int i = 0; i < n; i++) {
        const BatchReq& r = reqs[idx[i]];
        const double* oi = o + out_off[i];
        const int P = r.p->P;
        EvalOut* out = r.out;
        out->f = zd(oi[0], oi[1]);
        out->grad.assign((size_t)8 * P, zd(0.0));
        out->dhz.assign((size_t)8 * P, zd(0.0));
        for (int q = 0; q < P; q++)
            for (int e = 0; e < 8; e++) {
                out->grad[8 * q + e] = zd(oi[2 + 32 * q + 2 * e], oi[2 + 32 * q + 2 * e + 1]);
                out->dhz[8 * q + e] = zd(oi[2 + 32 * q + 16 + 2 * e], oi[2 + 32 * q + 16 + 2 * e + 1]);
            }
    }
    if (st) {
        st->launches++;
        st->items += nitems;
        st->grid = grid;
        st->lds_bytes = std::max(st->lds_bytes, lds);
        st->prep_ms += prep_ms;
        st->graph_replays += graph ? 1 : 0;
        st->graph_captures += captured ? 1 : 0;
        float ms = 0.f;
        if (kn.time_kernels) MG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        st->kernel_ms += ms;
    }
    return 0;
}

int evaluate_batch(BatchReq* reqs, int n, BatchStats* st)
{
    Context* c = ctx();
    if (!c) { set_error("no HIP device available"); return -1; }
    if (st) *st = BatchStats{};
    std::vector<int> i32, i64, single;
    for (int i = 0; i < n; i++) {
        BatchReq& r = reqs[i];
        std::string why;
        if (!r.t || !r.p || !r.gates || !r.out) { set_error("batched evaluation: null request field"); return -1; }
        if (!r.p->validate(&why)) { set_error("batched evaluation: program: " + why); return -1; }
        if (r.t->L != r.p->L) { set_error("batched evaluation: target and program have different qubit counts"); return -1; }
        if (r.p->L > 16) { set_error("batched evaluation: at most 16 qubits"); return -1; }
        if (r.p->L < batch_tuning().wbits[r.precision == kFP32 ? 1 : 0]) {
            set_error("batched evaluation: fewer qubits than the gate-window bits");
            return -1;
        }
        if (r.mode < 0 || r.mode > 2) { set_error("batched evaluation: invalid mode"); return -1; }
        if (r.mode == 2 && !r.Z) { set_error("batched evaluation: Hessian-vector product needs a direction"); return -1; }
        if (r.b1 > 0 && (r.b0 < 0 || r.b0 >= r.b1 || r.b1 > (1L << r.p->L))) {
            set_error("batched evaluation: invalid column range");
            return -1;
        }
        if (batch_single_request(*r.t, *r.p)) {
            single.push_back(i);
            continue;
        }
        (r.precision == kFP32 ? i32 : i64).push_back(i);
    }
    if (!i32.empty() && run_batch<float>(c, i32, reqs, st) < 0) return -1;
    if (!i64.empty() && run_batch<double>(c, i64, reqs, st) < 0) return -1;
    for (int i : single) {
        BatchReq& r = reqs[i];
        EvalStats es;
        if (evaluate(*r.t, *r.p, r.gates, r.Z, r.mode, r.precision, r.out, r.b1 > 0 ? r.b0 : 0, r.b1 > 0 ? r.b1 : -1,
                     &es) < 0)
            return -1;
        if (st) {
            st->launches++;
            st->kernel_ms += es.kernel_ms;
        }
    }
    return 0;
}

}  // namespace mg
