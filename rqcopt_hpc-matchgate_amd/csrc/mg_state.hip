// Single-column kernels on caller-owned host statevectors, executed on the GPU
// (C-ABI of include/gate.h, include/quantum_circuit.h, include/matchgate_brickwall.h).
//
// These keep the reference's per-column contracts (src/gate.c:575-1143,
// src/quantum_circuit.c:13-140, src/matchgate_brickwall.c:14-435): inputs are
// copied to the device, the sweep runs there in FP64, outputs are copied back.
// The batched column path (mg_sweep.hip) is the fast path; these entry points
// exist so that per-column callers (tests, micro-benchmarks, custom drivers) run
// unchanged.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/gate.h"
#include "../../include/matchgate_brickwall.h"
#include "../../include/quantum_circuit.h"
#include "mg_device.hpp"
#include "mg_runtime.hpp"

namespace mg {
namespace {

using C = cpx<double>;

struct SortedGate {
    C g[8];
    int lo, hi, flip;
};

SortedGate sort_gate(const struct matchgate* gate, int i, int j, int L)
{
    SortedGate s;
    const numeric* c = gate->center_block.data;
    const numeric* k = gate->corner_block.data;
    s.flip = j < i;
    if (s.flip) std::swap(i, j);
    s.hi = L - 1 - i;
    s.lo = L - 1 - j;
    for (int e = 0; e < 4; e++) {
        const numeric v = s.flip ? c[3 - e] : c[e];
        s.g[e] = {v.real(), v.imag()};
        s.g[4 + e] = {k[e].real(), k[e].imag()};
    }
    return s;
}

// out = V in for ns interleaved states (reference src/gate.c:581-652, 779-851)
__global__ void k_apply(SortedGate s, const C* __restrict__ in, C* __restrict__ out, uint32_t Q, int ns, int transpose)
{
    const uint32_t n = Q * ns;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t q = t / ns, k = t % ns;
        const uint32_t b = quad_base(q, s.lo, s.hi);
        const uint32_t ix = b * ns + k, iy = (b | (1u << s.lo)) * ns + k, iz = (b | (1u << s.hi)) * ns + k,
                       iw = (b | (1u << s.lo) | (1u << s.hi)) * ns + k;
        const C x = in[ix], y = in[iy], z = in[iz], w = in[iw];
        const C* g = s.g;
        if (!transpose) {
            out[ix] = cfma(g[4], x, cmul(g[5], w));
            out[iy] = cfma(g[0], y, cmul(g[1], z));
            out[iz] = cfma(g[2], y, cmul(g[3], z));
            out[iw] = cfma(g[6], x, cmul(g[7], w));
        } else {
            out[ix] = cfma(g[4], x, cmul(g[6], w));
            out[iy] = cfma(g[0], y, cmul(g[2], z));
            out[iz] = cfma(g[1], y, cmul(g[3], z));
            out[iw] = cfma(g[5], x, cmul(g[7], w));
        }
    }
}

// dg[k] = dout (x) psi_k for ns interleaved states psi; one workgroup per state,
// fixed-order tree reduction (reference src/gate.c:753-757, 858-973)
__global__ void k_outer(SortedGate s, const C* __restrict__ psi, const C* __restrict__ dout, uint32_t Q, int ns,
                        C* __restrict__ dg)
{
    __shared__ double red[256][16];
    const int k = blockIdx.x;
    C acc[8];
    for (int e = 0; e < 8; e++) acc[e] = {0.0, 0.0};
    for (uint32_t q = threadIdx.x; q < Q; q += blockDim.x) {
        const uint32_t b = quad_base(q, s.lo, s.hi);
        const uint32_t a[4] = {b, b | (1u << s.lo), b | (1u << s.hi), b | (1u << s.lo) | (1u << s.hi)};
        const C x = psi[a[0] * ns + k], y = psi[a[1] * ns + k], z = psi[a[2] * ns + k], w = psi[a[3] * ns + k];
        const C dx = dout[a[0]], dy = dout[a[1]], dz = dout[a[2]], dw = dout[a[3]];
        acc[0] = cfma(dy, y, acc[0]);
        acc[1] = cfma(dy, z, acc[1]);
        acc[2] = cfma(dz, y, acc[2]);
        acc[3] = cfma(dz, z, acc[3]);
        acc[4] = cfma(dx, x, acc[4]);
        acc[5] = cfma(dx, w, acc[5]);
        acc[6] = cfma(dw, x, acc[6]);
        acc[7] = cfma(dw, w, acc[7]);
    }
    for (int e = 0; e < 8; e++) {
        red[threadIdx.x][2 * e] = acc[e].re;
        red[threadIdx.x][2 * e + 1] = acc[e].im;
    }
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int v = 0; v < 16; v++) red[threadIdx.x][v] += red[threadIdx.x + h][v];
        __syncthreads();
    }
    if (threadIdx.x < 8) {
        int e = threadIdx.x;
        const C v = {red[0][2 * e], red[0][2 * e + 1]};
        if (s.flip && e < 4) e = 3 - e;  // back to the caller's wire order
        dg[8 * k + e] = v;
    }
}

// the 8 states d(V psi)/dV[e] (reference src/gate.c:982-1057)
__global__ void k_placeholder(SortedGate s, const C* __restrict__ psi, C* __restrict__ out, uint32_t Q)
{
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < Q; q += gridDim.x * blockDim.x) {
        const uint32_t b = quad_base(q, s.lo, s.hi);
        const uint32_t ix = b, iy = b | (1u << s.lo), iz = b | (1u << s.hi), iw = b | (1u << s.lo) | (1u << s.hi);
        const C x = psi[ix], y = psi[iy], z = psi[iz], w = psi[iw];
        const C zero = {0.0, 0.0};
        C* ox = out + 8 * (size_t)ix; C* oy = out + 8 * (size_t)iy; C* oz = out + 8 * (size_t)iz; C* ow = out + 8 * (size_t)iw;
        for (int e = 0; e < 8; e++) { ox[e] = zero; oy[e] = zero; oz[e] = zero; ow[e] = zero; }
        const int f = s.flip;
        oy[f ? 3 : 0] = y;
        oy[f ? 2 : 1] = z;
        oz[f ? 1 : 2] = y;
        oz[f ? 0 : 3] = z;
        ox[4] = x;
        ox[5] = w;
        ow[6] = x;
        ow[7] = w;
    }
}

// the 16 states d(G psi)/dG[r][c] of a general two-qubit gate G (reference src/gate.c:1066-1143):
// with r, c pair values in the caller's wire order (bit of wire i high, of wire j low), state
// 4 r + c holds psi's amplitude of pair value c at the amplitude of pair value r, zero elsewhere
__global__ void k_gate_placeholder(int lo, int hi, int flip, const C* __restrict__ psi, C* __restrict__ out, uint32_t Q)
{
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < Q; q += gridDim.x * blockDim.x) {
        const uint32_t b = quad_base(q, lo, hi);
        // amplitude index of pair value p = 2 k + l (k: bit of wire i, l: bit of wire j)
        uint32_t at[4];
        for (int p = 0; p < 4; p++) {
            const uint32_t k = p >> 1, l = p & 1;
            at[p] = flip ? (b | (k << lo) | (l << hi)) : (b | (k << hi) | (l << lo));
        }
        C v[4];
        for (int p = 0; p < 4; p++) v[p] = psi[at[p]];
        const C zero = {0.0, 0.0};
        for (int r = 0; r < 4; r++) {
            C* o = out + 16 * (size_t)at[r];
            for (int e = 0; e < 16; e++) o[e] = (e >> 2) == r ? v[e & 3] : zero;
        }
    }
}

struct Dev {
    void* p = nullptr;
    ~Dev()
    {
        std::lock_guard<std::recursive_mutex> g(capture_mutex());
        if (p) (void)hipFree(p);
    }
};

#define ST_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            set_error(std::string(#call) + " failed: " + hipGetErrorString(e_));       \
            return -1;                                                                 \
        }                                                                              \
    } while (0)

int dalloc(Dev& d, size_t bytes)
{
    std::lock_guard<std::recursive_mutex> g(capture_mutex());
    return hipMalloc(&d.p, std::max<size_t>(bytes, 256)) == hipSuccess ? 0 : -1;
}

int grid_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 4096); }

// ---- building blocks on device buffers ----
int d_apply(const SortedGate& s, const C* in, C* out, int L, int ns, int transpose, hipStream_t st)
{
    const uint32_t Q = 1u << (L - 2);
    hipLaunchKernelGGL(k_apply, dim3(grid_for((size_t)Q * ns)), dim3(256), 0, st, s, in, out, Q, ns, transpose);
    ST_HIP(hipGetLastError());
    return 0;
}

int d_outer(const SortedGate& s, const C* psi, const C* dout, int L, int ns, C* dg, hipStream_t st)
{
    const uint32_t Q = 1u << (L - 2);
    hipLaunchKernelGGL(k_outer, dim3(ns), dim3(256), 0, st, s, psi, dout, Q, ns, dg);
    ST_HIP(hipGetLastError());
    return 0;
}

int d_placeholder(const SortedGate& s, const C* psi, C* out, int L, hipStream_t st)
{
    const uint32_t Q = 1u << (L - 2);
    hipLaunchKernelGGL(k_placeholder, dim3(grid_for(Q)), dim3(256), 0, st, s, psi, out, Q);
    ST_HIP(hipGetLastError());
    return 0;
}

int up(void* d, const void* h, size_t n, hipStream_t st) { ST_HIP(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, st)); return 0; }
int down(void* h, const void* d, size_t n, hipStream_t st)
{
    ST_HIP(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, st));
    ST_HIP(hipStreamSynchronize(st));
    return 0;
}

// gates of a sequence: either the reference brickwall (vlist, perms) or a sequential circuit
struct Seq {
    std::vector<SortedGate> g;
    std::vector<int> param;
};

Seq brickwall_seq(const struct matchgate* vlist, int nlayers, const int* const* perms, int L)
{
    Seq q;
    for (int i = 0; i < nlayers; i++)
        for (int j = 0; j < L; j += 2) {
            q.g.push_back(sort_gate(&vlist[i], perms[i][j], perms[i][j + 1], L));
            q.param.push_back(i);
        }
    return q;
}

Seq circuit_seq(const struct matchgate* gates, int ngates, const int* wires, int L)
{
    Seq q;
    for (int k = 0; k < ngates; k++) {
        q.g.push_back(sort_gate(&gates[k], wires[2 * k], wires[2 * k + 1], L));
        q.param.push_back(k);
    }
    return q;
}

// forward: device buffer states[(G+1) * S]: states[k] = state before gate k, states[G] = output
int d_forward(const Seq& q, const numeric* psi, int L, Dev& states, hipStream_t st)
{
    const size_t S = (size_t)1 << L, G = q.g.size();
    if (dalloc(states, (G + 1) * S * sizeof(C)) < 0) { set_error("device allocation failed"); return -1; }
    C* d = (C*)states.p;
    if (up(d, psi, S * sizeof(C), st) < 0) return -1;
    for (size_t k = 0; k < G; k++)
        if (d_apply(q.g[k], d + k * S, d + (k + 1) * S, L, 1, 0, st) < 0) return -1;
    return 0;
}

// backward with the gradient states kept: lam[k] = dpsi_out of gate k; lam[G] unused
int d_backward(const Seq& q, const C* fwd, const numeric* dpsi_out, int L, Dev& lam, C* dg_dev, hipStream_t st)
{
    const size_t S = (size_t)1 << L, G = q.g.size();
    if (dalloc(lam, (G + 1) * S * sizeof(C)) < 0) { set_error("device allocation failed"); return -1; }
    C* l = (C*)lam.p;
    // l[k+1] holds the output gradient of gate k; l[0] the input gradient
    if (up(l + G * S, dpsi_out, S * sizeof(C), st) < 0) return -1;
    for (size_t k = G; k-- > 0;) {
        if (d_outer(q.g[k], fwd + k * S, l + (k + 1) * S, L, 1, dg_dev + 8 * k, st) < 0) return -1;
        if (d_apply(q.g[k], l + (k + 1) * S, l + k * S, L, 1, 1, st) < 0) return -1;
    }
    return 0;
}

void sum_per_param(const std::vector<C>& dgi, const Seq& q, int P, struct matchgate* out)
{
    for (int p = 0; p < P; p++)
        for (int e = 0; e < 8; e++) (&out[p].center_block.data[0])[e] = 0.0;
    for (size_t k = 0; k < q.g.size(); k++)
        for (int e = 0; e < 8; e++) (&out[q.param[k]].center_block.data[0])[e] += numeric(dgi[8 * k + e].re, dgi[8 * k + e].im);
}

}  // namespace
}  // namespace mg

using namespace mg;

extern "C" {

// ---------------- gate.h ----------------

void apply_matchgate(const struct matchgate* gate, const int i, const int j, const struct statevector* psi,
                     struct statevector* psi_out)
{
    const int L = psi->nqubits;
    const size_t S = (size_t)1 << L;
    hipStream_t st = stream();
    Dev a, b;
    if (dalloc(a, S * sizeof(C)) < 0 || dalloc(b, S * sizeof(C)) < 0) { set_error("device allocation failed"); return; }
    if (up(a.p, psi->data, S * sizeof(C), st) < 0) return;
    if (d_apply(sort_gate(gate, i, j, L), (C*)a.p, (C*)b.p, L, 1, 0, st) < 0) return;
    down(psi_out->data, b.p, S * sizeof(C), st);
}

void apply_matchgate_backward(const struct matchgate* gate, const int i, const int j, const struct statevector* psi,
                              const struct statevector* dpsi_out, struct statevector* dpsi, struct matchgate* dgate)
{
    const int L = psi->nqubits;
    const size_t S = (size_t)1 << L;
    hipStream_t st = stream();
    Dev a, d, o, g;
    if (dalloc(a, S * sizeof(C)) < 0 || dalloc(d, S * sizeof(C)) < 0 || dalloc(o, S * sizeof(C)) < 0 || dalloc(g, 8 * sizeof(C)) < 0) {
        set_error("device allocation failed");
        return;
    }
    if (up(a.p, psi->data, S * sizeof(C), st) < 0 || up(d.p, dpsi_out->data, S * sizeof(C), st) < 0) return;
    const SortedGate s = sort_gate(gate, i, j, L);
    if (d_outer(s, (C*)a.p, (C*)d.p, L, 1, (C*)g.p, st) < 0) return;
    if (d_apply(s, (C*)d.p, (C*)o.p, L, 1, 1, st) < 0) return;
    if (down(dgate, g.p, 8 * sizeof(C), st) < 0) return;
    down(dpsi->data, o.p, S * sizeof(C), st);
}

void apply_matchgate_to_array(const struct matchgate* gate, const int i, const int j,
                              const struct statevector_array* psi, struct statevector_array* psi_out)
{
    const int L = psi->nqubits, ns = psi->nstates;
    const size_t n = ((size_t)1 << L) * ns;
    hipStream_t st = stream();
    Dev a, b;
    if (dalloc(a, n * sizeof(C)) < 0 || dalloc(b, n * sizeof(C)) < 0) { set_error("device allocation failed"); return; }
    if (up(a.p, psi->data, n * sizeof(C), st) < 0) return;
    if (d_apply(sort_gate(gate, i, j, L), (C*)a.p, (C*)b.p, L, ns, 0, st) < 0) return;
    down(psi_out->data, b.p, n * sizeof(C), st);
}

void apply_matchgate_backward_array(const struct matchgate* gate, const int i, const int j,
                                    const struct statevector_array* psi, const struct statevector* dpsi_out,
                                    struct matchgate* dgates)
{
    const int L = psi->nqubits, ns = psi->nstates;
    const size_t S = (size_t)1 << L;
    hipStream_t st = stream();
    Dev a, d, g;
    if (dalloc(a, S * ns * sizeof(C)) < 0 || dalloc(d, S * sizeof(C)) < 0 || dalloc(g, 8 * ns * sizeof(C)) < 0) {
        set_error("device allocation failed");
        return;
    }
    if (up(a.p, psi->data, S * ns * sizeof(C), st) < 0 || up(d.p, dpsi_out->data, S * sizeof(C), st) < 0) return;
    if (d_outer(sort_gate(gate, i, j, L), (C*)a.p, (C*)d.p, L, ns, (C*)g.p, st) < 0) return;
    down(dgates, g.p, 8 * ns * sizeof(C), st);
}

void apply_matchgate_placeholder(const int i, const int j, const struct statevector* psi, struct statevector_array* psi_out)
{
    const int L = psi->nqubits;
    const size_t S = (size_t)1 << L;
    hipStream_t st = stream();
    Dev a, b;
    if (dalloc(a, S * sizeof(C)) < 0 || dalloc(b, 8 * S * sizeof(C)) < 0) { set_error("device allocation failed"); return; }
    if (up(a.p, psi->data, S * sizeof(C), st) < 0) return;
    struct matchgate unit;
    identity_matchgate(&unit);
    if (d_placeholder(sort_gate(&unit, i, j, L), (C*)a.p, (C*)b.p, L, st) < 0) return;
    down(psi_out->data, b.p, 8 * S * sizeof(C), st);
}

void apply_gate_placeholder(const int i, const int j, const struct statevector* psi, struct statevector_array* psi_out)
{
    const int L = psi->nqubits;
    if (psi_out->nqubits != L || psi_out->nstates != 16 || i < 0 || j < 0 || i >= L || j >= L || i == j) {
        set_error("apply_gate_placeholder: invalid wires or output array (16 states of the same qubit count)");
        return;
    }
    const size_t S = (size_t)1 << L;
    hipStream_t st = stream();
    Dev a, b;
    if (dalloc(a, S * sizeof(C)) < 0 || dalloc(b, 16 * S * sizeof(C)) < 0) { set_error("device allocation failed"); return; }
    if (up(a.p, psi->data, S * sizeof(C), st) < 0) return;
    const int flip = j < i;
    const int hi = L - 1 - std::min(i, j), lo = L - 1 - std::max(i, j);
    const uint32_t Q = 1u << (L - 2);
    hipLaunchKernelGGL(k_gate_placeholder, dim3(grid_for(Q)), dim3(256), 0, st, lo, hi, flip, (const C*)a.p, (C*)b.p, Q);
    if (hipGetLastError() != hipSuccess) { set_error("apply_gate_placeholder: launch failed"); return; }
    down(psi_out->data, b.p, 16 * S * sizeof(C), st);
}

// ---------------- quantum_circuit.h ----------------

int allocate_quantum_circuit_cache(const int nqubits, const int ngates, struct quantum_circuit_cache* cache)
{
    cache->nqubits = nqubits;
    cache->ngates = ngates;
    cache->psi_list = (struct statevector*)calloc(std::max(ngates, 1), sizeof(struct statevector));
    if (!cache->psi_list) return -1;
    for (int k = 0; k < ngates; k++)
        if (allocate_statevector(nqubits, &cache->psi_list[k]) < 0) return -1;
    return 0;
}

void free_quantum_circuit_cache(struct quantum_circuit_cache* cache)
{
    for (int k = 0; k < cache->ngates; k++) free_statevector(&cache->psi_list[k]);
    free(cache->psi_list);
    cache->psi_list = nullptr;
    cache->nqubits = 0;
    cache->ngates = 0;
}

static int run_forward(const Seq& q, const struct statevector* psi, struct quantum_circuit_cache* cache,
                       struct statevector* psi_out)
{
    const int L = psi->nqubits;
    const size_t S = (size_t)1 << L, G = q.g.size();
    hipStream_t st = stream();
    Dev states;
    if (d_forward(q, psi->data, L, states, st) < 0) return -1;
    const C* d = (const C*)states.p;
    if (cache)
        for (size_t k = 0; k < G; k++)
            if (hipMemcpyAsync(cache->psi_list[k].data, d + k * S, S * sizeof(C), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    return down(psi_out->data, d + G * S, S * sizeof(C), st);
}

// backward from a host cache of forward states
static int run_backward(const Seq& q, int P, const struct quantum_circuit_cache* cache,
                        const struct statevector* dpsi_out, struct statevector* dpsi, struct matchgate* dout, bool per_instance)
{
    const int L = dpsi_out->nqubits;
    const size_t S = (size_t)1 << L, G = q.g.size();
    hipStream_t st = stream();
    Dev fwd, lam, dg;
    if (dalloc(fwd, G * S * sizeof(C)) < 0 || dalloc(dg, 8 * G * sizeof(C)) < 0) { set_error("device allocation failed"); return -1; }
    for (size_t k = 0; k < G; k++)
        if (up((C*)fwd.p + k * S, cache->psi_list[k].data, S * sizeof(C), st) < 0) return -1;
    if (d_backward(q, (const C*)fwd.p, dpsi_out->data, L, lam, (C*)dg.p, st) < 0) return -1;
    std::vector<C> dgi(8 * G);
    if (down(dgi.data(), dg.p, 8 * G * sizeof(C), st) < 0) return -1;
    if (down(dpsi->data, lam.p, S * sizeof(C), st) < 0) return -1;
    if (per_instance) {
        for (size_t k = 0; k < G; k++)
            for (int e = 0; e < 8; e++) (&dout[k].center_block.data[0])[e] = numeric(dgi[8 * k + e].re, dgi[8 * k + e].im);
    } else {
        sum_per_param(dgi, q, P, dout);
    }
    return 0;
}

int apply_quantum_matchgate_circuit(const struct matchgate gates[], const int ngates, const int wires[],
                                    const struct statevector* psi, struct statevector* psi_out)
{
    return run_forward(circuit_seq(gates, ngates, wires, psi->nqubits), psi, nullptr, psi_out);
}

int quantum_matchgate_circuit_forward(const struct matchgate gates[], const int ngates, const int wires[],
                                      const struct statevector* psi, struct quantum_circuit_cache* cache,
                                      struct statevector* psi_out)
{
    return run_forward(circuit_seq(gates, ngates, wires, psi->nqubits), psi, cache, psi_out);
}

int quantum_matchgate_circuit_backward(const struct matchgate gates[], const int ngates, const int wires[],
                                       const struct quantum_circuit_cache* cache, const struct statevector* dpsi_out,
                                       struct statevector* dpsi, struct matchgate dgates[])
{
    return run_backward(circuit_seq(gates, ngates, wires, dpsi_out->nqubits), ngates, cache, dpsi_out, dpsi, dgates, true);
}

// ---------------- matchgate_brickwall.h ----------------

int apply_matchgate_brickwall_unitary(const struct matchgate vlist[], const int nlayers, const int* perms[],
                                      const struct statevector* psi, struct statevector* psi_out)
{
    return run_forward(brickwall_seq(vlist, nlayers, perms, psi->nqubits), psi, nullptr, psi_out);
}

int matchgate_brickwall_unitary_forward(const struct matchgate vlist[], int nlayers, const int* perms[],
                                        const struct statevector* psi, struct quantum_circuit_cache* cache,
                                        struct statevector* psi_out)
{
    return run_forward(brickwall_seq(vlist, nlayers, perms, psi->nqubits), psi, cache, psi_out);
}

int matchgate_brickwall_unitary_backward(const struct matchgate vlist[], int nlayers, const int* perms[],
                                         const struct quantum_circuit_cache* cache, const struct statevector* dpsi_out,
                                         struct statevector* dpsi, struct matchgate dvlist[])
{
    return run_backward(brickwall_seq(vlist, nlayers, perms, dpsi_out->nqubits), nlayers, cache, dpsi_out, dpsi, dvlist, false);
}

// gradient + complex Hessian of one column: the reference's placeholder propagation
// (src/matchgate_brickwall.c:137-260) on the device; hess = A + A^T with A the
// upper (earlier instance, later instance) blocks.  ti != 0: the translation-invariant
// shortcut (:308-435).
static int backward_hessian(const struct matchgate* vlist, int nlayers, const int* const* perms,
                            const struct quantum_circuit_cache* cache, const struct statevector* dpsi_out,
                            struct statevector* dpsi, struct matchgate* dvlist, numeric* hess, bool ti)
{
    const int L = dpsi_out->nqubits, per = L / 2, m = 8 * nlayers;
    const Seq q = brickwall_seq(vlist, nlayers, perms, L);
    const size_t S = (size_t)1 << L, G = q.g.size();
    hipStream_t st = stream();
    Dev fwd, lam, dg, t0, t1, h;
    if (dalloc(fwd, G * S * sizeof(C)) < 0 || dalloc(dg, 8 * G * sizeof(C)) < 0 || dalloc(t0, 8 * S * sizeof(C)) < 0 ||
        dalloc(t1, 8 * S * sizeof(C)) < 0 || dalloc(h, 64 * G * sizeof(C)) < 0) {
        set_error("device allocation failed");
        return -1;
    }
    for (size_t k = 0; k < G; k++)
        if (up((C*)fwd.p + k * S, cache->psi_list[k].data, S * sizeof(C), st) < 0) return -1;
    if (d_backward(q, (const C*)fwd.p, dpsi_out->data, L, lam, (C*)dg.p, st) < 0) return -1;
    std::vector<C> dgi(8 * G);
    if (down(dgi.data(), dg.p, 8 * G * sizeof(C), st) < 0) return -1;
    if (down(dpsi->data, lam.p, S * sizeof(C), st) < 0) return -1;
    sum_per_param(dgi, q, nlayers, dvlist);

    std::vector<numeric> A((size_t)m * m, numeric(0.0));
    std::vector<C> hh(64 * G);
    const C* l = (const C*)lam.p;
    for (size_t p = 0; p < G; p++) {
        if (ti && p % per != 0) continue;
        const int i = q.param[p];
        if (d_placeholder(q.g[p], (const C*)fwd.p + p * S, (C*)t0.p, L, st) < 0) return -1;
        C* cur = (C*)t0.p;
        C* nxt = (C*)t1.p;
        for (size_t r = p + 1; r < G; r++) {
            if (d_outer(q.g[r], cur, l + (r + 1) * S, L, 8, (C*)h.p + 64 * r, st) < 0) return -1;
            if (r + 1 < G) {
                if (d_apply(q.g[r], cur, nxt, L, 8, 0, st) < 0) return -1;
                std::swap(cur, nxt);
            }
        }
        if (down(hh.data(), h.p, 64 * G * sizeof(C), st) < 0) return -1;
        for (size_t r = p + 1; r < G; r++) {
            const int j = q.param[r];
            double mult = 1.0;
            if (ti) {
                const int s = (int)(r - (size_t)j * per);
                mult = (i == j) ? (double)(per - s) : (double)per;
            }
            for (int x = 0; x < 8; x++)
                for (int y = 0; y < 8; y++)
                    A[(size_t)(8 * i + x) * m + 8 * j + y] += mult * numeric(hh[64 * r + 8 * x + y].re, hh[64 * r + 8 * x + y].im);
        }
    }
    for (int a = 0; a < m; a++)
        for (int b = 0; b < m; b++) hess[(size_t)a * m + b] = A[(size_t)a * m + b] + A[(size_t)b * m + a];
    return 0;
}

int matchgate_brickwall_unitary_backward_hessian(const struct matchgate vlist[], int nlayers, const int* perms[],
                                                 const struct quantum_circuit_cache* cache,
                                                 const struct statevector* dpsi_out, struct statevector* dpsi,
                                                 struct matchgate dvlist[], numeric* hess)
{
    return backward_hessian(vlist, nlayers, perms, cache, dpsi_out, dpsi, dvlist, hess, false);
}

int matchgate_ti_brickwall_unitary_backward_hessian(const struct matchgate vlist[], int nlayers, const int* perms[],
                                                    const struct quantum_circuit_cache* cache,
                                                    const struct statevector* dpsi_out, struct statevector* dpsi,
                                                    struct matchgate dvlist[], numeric* hess)
{
    return backward_hessian(vlist, nlayers, perms, cache, dpsi_out, dpsi, dvlist, hess, true);
}

// translation-invariant gradient (reference src/matchgate_brickwall.c:266-306): the
// last pair of each layer only, scaled by L/2; the backpropagated state skips the
// other pairs of the layer exactly as the reference does
int matchgate_ti_brickwall_unitary_backward(const struct matchgate vlist[], int nlayers, const int* perms[],
                                            const struct quantum_circuit_cache* cache, const struct statevector* dpsi_out,
                                            struct statevector* dpsi, struct matchgate dvlist[])
{
    const int L = dpsi_out->nqubits, per = L / 2;
    const size_t S = (size_t)1 << L;
    hipStream_t st = stream();
    Dev psi, a, b, g;
    if (dalloc(psi, S * sizeof(C)) < 0 || dalloc(a, S * sizeof(C)) < 0 || dalloc(b, S * sizeof(C)) < 0 ||
        dalloc(g, 8 * nlayers * sizeof(C)) < 0) {
        set_error("device allocation failed");
        return -1;
    }
    if (up(a.p, dpsi_out->data, S * sizeof(C), st) < 0) return -1;
    C* cur = (C*)a.p;
    C* nxt = (C*)b.p;
    int k = nlayers * per - 1;
    for (int i = nlayers - 1; i >= 0; i--) {
        const SortedGate s = sort_gate(&vlist[i], perms[i][L - 2], perms[i][L - 1], L);
        if (up(psi.p, cache->psi_list[k].data, S * sizeof(C), st) < 0) return -1;
        if (d_outer(s, (const C*)psi.p, cur, L, 1, (C*)g.p + 8 * i, st) < 0) return -1;
        if (d_apply(s, cur, nxt, L, 1, 1, st) < 0) return -1;
        std::swap(cur, nxt);
        k -= per;
    }
    std::vector<C> dg(8 * nlayers);
    if (down(dg.data(), g.p, 8 * nlayers * sizeof(C), st) < 0) return -1;
    if (down(dpsi->data, cur, S * sizeof(C), st) < 0) return -1;
    for (int i = 0; i < nlayers; i++)
        for (int e = 0; e < 8; e++) (&dvlist[i].center_block.data[0])[e] = (double)per * numeric(dg[8 * i + e].re, dg[8 * i + e].im);
    return 0;
}

}  // extern "C"
