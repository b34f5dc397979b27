// C-ABI entry points of the GPU evaluation path: the reference's target functions
// (include/matchgate_target.h), the GPU-native interface (include/mg_gpu.h) and the
// brickwall optimisation drivers (include/mg_brickwall_opt.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include <thread>

#include "../../include/matchgate_brickwall.h"
#include "../../include/matchgate_target.h"
#include "../../include/mg_brickwall_opt.h"
#include "../../include/mg_gpu.h"
#include "../../include/trust_region.h"
#include "mg_device_tr.hpp"
#include "mg_host.hpp"
#include "mg_runtime.hpp"

using mg::zd;

namespace mg {

// ---- Riemannian projection (reference src/matchgate_target.c:903-967) ----

void riemannian_gradient(const struct matchgate* V, int P, const zd* dgrad, double* grad_vec)
{
    for (int i = 0; i < P; i++) {
        struct matchgate g;
        std::memcpy(&g, dgrad + 8 * i, sizeof(g));
        conjugate_matchgate(&g);
        tangent_matchgate_to_real(&V[i], &g, grad_vec + 8 * i);
    }
}

// G = P_V(conj(hz)) - (Z g^T V + V g^T Z)/2 for one parameter gate
void weingarten_block(const struct matchgate* V, const zd* g, const struct matchgate* Z, const zd* hz, double* r)
{
    struct matchgate G, Gp, gc, gh, T;
    std::memcpy(&G, hz, sizeof(G));
    conjugate_matchgate(&G);
    project_tangent_matchgate(V, &G, &Gp);
    std::memcpy(&gc, g, sizeof(gc));
    conjugate_matchgate(&gc);
    adjoint_matchgate(&gc, &gh);
    symmetric_triple_matchgate_product(Z, &gh, V, &T);
    sub_matchgates(&Gp, &T, &G);
    tangent_matchgate_to_real(V, &G, r);
}

}  // namespace mg

using mg::riemannian_gradient;
using mg::weingarten_block;

namespace {

const zd* Zp(const struct matchgate* g) { return reinterpret_cast<const zd*>(g); }
zd* Zp(struct matchgate* g) { return reinterpret_cast<zd*>(g); }

mg::EvalStats g_last_stats;

// ---- targets of linear_func callbacks ----
// The reference calls ufunc for every basis column on every evaluation
// (src/matchgate_target.c:795), so a map whose udata changed between two calls takes
// effect at once.  The entry points do the same by default: every call materialises the
// callback's map again (2^n host calls) into a device target that lives until the next
// call.  With mg_target_cache_reuse(1) a caller that promises an unchanged map keeps the
// materialised targets per (ufunc, udata, nqubits) instead, until mg_target_cache_clear()
// or until the least recently used entry is evicted (at most kCacheEntries targets and
// kCacheBytes of device rows).
constexpr size_t kCacheEntries = 8;
constexpr double kCacheBytes = 64.0 * (1ull << 30);

using CacheKey = std::tuple<void*, void*, int>;

struct TargetCache {
    bool reuse = false;
    std::list<std::pair<CacheKey, std::unique_ptr<mg::Target>>> lru;  // most recently used first
    std::unique_ptr<mg::Target> last;                                  // the per-call target (reuse off)
};

TargetCache& target_cache()
{
    static TargetCache c;
    return c;
}

// device bytes of a materialised target: FP64 rows and the FP32 copy made on first FP32 use
double target_bytes(int L) { return 24.0 * (double)((size_t)1 << L) * (double)((size_t)1 << L); }

// Calls ufunc on every basis vector (several host threads, as the reference's
// OpenMP column loop does, src/matchgate_target.c:788-805) and stores the seed rows
// -conj(U e_b).  Returns nullptr and *err = -2 if ufunc fails.
mg::Target* materialise(linear_func ufunc, void* udata, int L, int* err)
{
    const size_t S = (size_t)1 << L;
    std::vector<mg::cpx<double>> seed(S * S);
    std::atomic<int> failed{0};
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](int tid) {
        statevector psi{}, out{};
        if (allocate_statevector(L, &psi) < 0 || allocate_statevector(L, &out) < 0) {
            failed = -1;
        } else {
            for (size_t b = tid; b < S && failed.load() == 0; b += nt) {
                std::memset((void*)psi.data, 0, S * sizeof(numeric));
                psi.data[b] = 1.0;
                const int ret = ufunc(&psi, udata, &out);
                if (ret < 0) {
                    std::fprintf(stderr, "call of 'ufunc' failed, return value: %i\n", ret);
                    failed = -2;
                    break;
                }
                mg::cpx<double>* row = &seed[b * S];
                for (size_t a = 0; a < S; a++) row[a] = {-out.data[a].real(), out.data[a].imag()};
            }
        }
        if (psi.data) free_statevector(&psi);
        if (out.data) free_statevector(&out);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    if (failed.load() != 0) {
        *err = failed.load();
        return nullptr;
    }
    mg::Target* t = mg::target_dense_seed(std::move(seed), L);
    if (!t) *err = -1;
    return t;
}

mg::Target* cached_target(linear_func ufunc, void* udata, int L, int* err)
{
    *err = 0;
    TargetCache& c = target_cache();
    if (!c.reuse) {
        c.last.reset();
        mg::Target* t = materialise(ufunc, udata, L, err);
        c.last.reset(t);
        return t;
    }
    const CacheKey key = std::make_tuple((void*)ufunc, udata, L);
    for (auto it = c.lru.begin(); it != c.lru.end(); ++it)
        if (it->first == key) {
            c.lru.splice(c.lru.begin(), c.lru, it);
            return c.lru.front().second.get();
        }
    mg::Target* t = materialise(ufunc, udata, L, err);
    if (!t) return nullptr;
    c.lru.emplace_front(key, std::unique_ptr<mg::Target>(t));
    double bytes = 0;
    for (const auto& e : c.lru) bytes += target_bytes(std::get<2>(e.first));
    while (c.lru.size() > 1 && (c.lru.size() > kCacheEntries || bytes > kCacheBytes)) {
        bytes -= target_bytes(std::get<2>(c.lru.back().first));
        c.lru.pop_back();
    }
    return t;
}

bool check_brickwall(int nlayers, int nqubits, const int* const* perms)
{
    if (nlayers < 1 || nqubits < 2 || nqubits % 2 != 0 || !perms) {
        mg::set_error("brickwall needs nlayers >= 1 and an even number of qubits");
        return false;
    }
    return true;
}

int full_hessian(const mg::Target& t, const mg::Program& p, const zd* gates, int precision, zd* f, zd* dgrad, zd* hess)
{
    const int m = 8 * p.P;
    std::vector<zd> Z((size_t)m, zd(0.0));
    mg::EvalOut out;
    for (int k = 0; k < m; k++) {
        std::fill(Z.begin(), Z.end(), zd(0.0));
        Z[k] = 1.0;
        if (mg::evaluate(t, p, gates, Z.data(), mg::kFGradHvp, precision, &out, 0, -1, &g_last_stats) < 0) return -1;
        for (int r = 0; r < m; r++) hess[(size_t)r * m + k] = out.dhz[r];
    }
    // D^2 f is symmetric: remove the rounding asymmetry of the column-wise evaluation
    for (int a = 0; a < m; a++)
        for (int b = 0; b < a; b++) {
            const zd s = 0.5 * (hess[(size_t)a * m + b] + hess[(size_t)b * m + a]);
            hess[(size_t)a * m + b] = s;
            hess[(size_t)b * m + a] = s;
        }
    if (f) *f = out.f;
    if (dgrad) std::copy(out.grad.begin(), out.grad.end(), dgrad);
    return 0;
}

// real Riemannian Hessian matrix from the complex one (reference src/matchgate_target.c:910-967)
void project_hessian(const struct matchgate* V, int P, const zd* dgrad, const zd* hess, double* H)
{
    const int m = 8 * P;
    for (int i = 0; i < P; i++)
        for (int j = i; j < P; j++)
            for (int k = 0; k < 8; k++) {
                double r[8] = {0};
                r[k] = 1;
                struct matchgate Z;
                real_to_tangent_matchgate(r, &V[j], &Z);
                const zd* Zz = Zp(&Z);
                zd G[8];
                for (int x = 0; x < 8; x++) {
                    G[x] = 0.0;
                    for (int y = 0; y < 8; y++) G[x] += hess[(size_t)(8 * i + x) * m + 8 * j + y] * Zz[y];
                }
                if (i == j) {
                    weingarten_block(&V[i], dgrad + 8 * i, &Z, G, r);
                } else {
                    struct matchgate Gm;
                    std::memcpy(&Gm, G, sizeof(Gm));
                    conjugate_matchgate(&Gm);
                    tangent_matchgate_to_real(&V[i], &Gm, r);
                }
                for (int x = 0; x < 8; x++) H[(size_t)(8 * i + x) * m + 8 * j + k] = r[x];
            }
    for (int a = 0; a < m; a++)
        for (int b = 0; b < a; b++) H[(size_t)a * m + b] = H[(size_t)b * m + a];
}

int eval_plain(const mg::Target& t, const mg::Program& p, const zd* gates, int mode, zd* f, zd* grad)
{
    mg::EvalOut out;
    if (mg::evaluate(t, p, gates, nullptr, mode, mg::kFP64, &out, 0, -1, &g_last_stats) < 0) return -1;
    *f = out.f;
    if (grad) std::copy(out.grad.begin(), out.grad.end(), grad);
    return 0;
}

}  // namespace

extern "C" {

// =================== reference target functions (include/matchgate_target.h) ===================

void mg_target_cache_clear(void)
{
    target_cache().lru.clear();
    target_cache().last.reset();
}

void mg_target_cache_reuse(int on)
{
    target_cache().reuse = on != 0;
    if (!on) target_cache().lru.clear();
}

int matchgate_circuit_unitary_target(linear_func ufunc, void* udata, const struct matchgate gates[], const int ngates,
                                     const int wires[], const int nqubits, numeric* fval)
{
    int err;
    mg::Target* t = cached_target(ufunc, udata, nqubits, &err);
    if (!t) return err;
    mg::Program p;
    p.L = nqubits;
    p.P = p.G = ngates;
    p.param.resize(ngates);
    for (int g = 0; g < ngates; g++) p.param[g] = g;
    p.wires.assign(wires, wires + 2 * ngates);
    zd f;
    if (eval_plain(*t, p, Zp(gates), mg::kF, &f, nullptr) < 0) return -1;
    *fval = f;
    return 0;
}

int parallel_matchgate_circuit_unitary_target(linear_func ufunc, void* udata, const struct matchgate gates[],
                                              const int ngates, const int wires[], const int nqubits, numeric* fval)
{
    return matchgate_circuit_unitary_target(ufunc, udata, gates, ngates, wires, nqubits, fval);
}

int matchgate_circuit_unitary_target_and_gradient(linear_func ufunc, void* udata, const struct matchgate gates[],
                                                  const int ngates, const int wires[], const int nqubits,
                                                  numeric* fval, struct matchgate dgates[])
{
    int err;
    mg::Target* t = cached_target(ufunc, udata, nqubits, &err);
    if (!t) return err;
    mg::Program p;
    p.L = nqubits;
    p.P = p.G = ngates;
    p.param.resize(ngates);
    for (int g = 0; g < ngates; g++) p.param[g] = g;
    p.wires.assign(wires, wires + 2 * ngates);
    zd f;
    if (eval_plain(*t, p, Zp(gates), mg::kFGrad, &f, Zp(dgates)) < 0) return -1;
    *fval = f;
    return 0;
}

int matchgate_brickwall_unitary_target(linear_func ufunc, void* udata, const struct matchgate vlist[],
                                       const int nlayers, const int nqubits, const int* perms[], numeric* fval)
{
    if (!check_brickwall(nlayers, nqubits, perms)) return -1;
    int err;
    mg::Target* t = cached_target(ufunc, udata, nqubits, &err);
    if (!t) return err;
    const mg::Program p = mg::brickwall_program(nqubits, nlayers, perms);
    zd f;
    if (eval_plain(*t, p, Zp(vlist), mg::kF, &f, nullptr) < 0) return -1;
    *fval = f;
    return 0;
}

int matchgate_brickwall_unitary_target_and_gradient(linear_func ufunc, void* udata, const struct matchgate vlist[],
                                                    const int nlayers, const int nqubits, const int* perms[],
                                                    numeric* fval, struct matchgate dvlist[])
{
    if (!check_brickwall(nlayers, nqubits, perms)) return -1;
    int err;
    mg::Target* t = cached_target(ufunc, udata, nqubits, &err);
    if (!t) return err;
    const mg::Program p = mg::brickwall_program(nqubits, nlayers, perms);
    zd f;
    if (eval_plain(*t, p, Zp(vlist), mg::kFGrad, &f, Zp(dvlist)) < 0) return -1;
    *fval = f;
    return 0;
}

int matchgate_brickwall_unitary_target_gradient_hessian(linear_func ufunc, void* udata, const struct matchgate vlist[],
                                                        const int nlayers, const int nqubits, const int* perms[],
                                                        numeric* fval, struct matchgate dvlist[], numeric* hess)
{
    if (!check_brickwall(nlayers, nqubits, perms)) return -1;
    int err;
    mg::Target* t = cached_target(ufunc, udata, nqubits, &err);
    if (!t) return err;
    const mg::Program p = mg::brickwall_program(nqubits, nlayers, perms);
    zd f;
    if (full_hessian(*t, p, Zp(vlist), mg::kFP64, &f, Zp(dvlist), hess) < 0) return -1;
    *fval = f;
    return 0;
}

int parallel_matchgate_brickwall_unitary_target_gradient_hessian(linear_func ufunc, void* udata,
                                                                 const struct matchgate vlist[], const int nlayers,
                                                                 const int nqubits, const int* perms[], numeric* fval,
                                                                 struct matchgate dvlist[], numeric* hess)
{
    return matchgate_brickwall_unitary_target_gradient_hessian(ufunc, udata, vlist, nlayers, nqubits, perms, fval,
                                                               dvlist, hess);
}

// Translation-invariant shortcut (reference src/matchgate_brickwall.c:308-435):
// second derivatives only against the first pair (r = 0) of each layer, weighted by
// the multiplicity L/2 (off-diagonal layer blocks) or L/2 - s (same layer, pair s),
// then hess = A + A^T.  Computed from per-instance directional sweeps: the
// direction sits on instance (i, 0) alone, the outputs are kept per instance.
int matchgate_ti_brickwall_unitary_target_gradient_hessian(linear_func ufunc, void* udata,
                                                           const struct matchgate vlist[], const int nlayers,
                                                           const int nqubits, const int* perms[], numeric* fval,
                                                           struct matchgate dvlist[], numeric* hess)
{
    if (!check_brickwall(nlayers, nqubits, perms)) return -1;
    int err;
    mg::Target* t = cached_target(ufunc, udata, nqubits, &err);
    if (!t) return err;
    const mg::Program tied = mg::brickwall_program(nqubits, nlayers, perms);
    mg::Program untied = tied;
    untied.P = tied.G;
    for (int g = 0; g < tied.G; g++) untied.param[g] = g;
    std::vector<zd> gi((size_t)8 * tied.G);
    for (int g = 0; g < tied.G; g++) std::copy(Zp(&vlist[tied.param[g]]), Zp(&vlist[tied.param[g]]) + 8, &gi[8 * g]);
    const int per = nqubits / 2, m = 8 * nlayers;
    std::vector<zd> A((size_t)m * m, zd(0.0)), Z(gi.size());
    mg::EvalOut out;
    for (int i = 0; i < nlayers; i++) {
        const int p0 = i * per;
        for (int x = 0; x < 8; x++) {
            std::fill(Z.begin(), Z.end(), zd(0.0));
            Z[8 * p0 + x] = 1.0;
            if (mg::evaluate(*t, untied, gi.data(), Z.data(), mg::kFGradHvp, mg::kFP64, &out, 0, -1, &g_last_stats) < 0)
                return -1;
            for (int q = p0 + 1; q < tied.G; q++) {
                const int j = q / per, s = q - j * per;
                const double mult = (i == j) ? (double)(per - s) : (double)per;
                for (int y = 0; y < 8; y++) A[(size_t)(8 * i + x) * m + 8 * j + y] += mult * out.dhz[8 * q + y];
            }
        }
    }
    for (int a = 0; a < m; a++)
        for (int b = 0; b < m; b++) hess[(size_t)a * m + b] = A[(size_t)a * m + b] + A[(size_t)b * m + a];
    // full per-layer gradient (the TI variant computes the gradient without the shortcut)
    for (int l = 0; l < nlayers; l++)
        for (int e = 0; e < 8; e++) {
            zd s = 0.0;
            for (int k = 0; k < per; k++) s += out.grad[8 * (l * per + k) + e];
            Zp(&dvlist[l])[e] = s;
        }
    *fval = out.f;
    return 0;
}

int matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix(linear_func ufunc, void* udata,
                                                                      const struct matchgate vlist[], const int nlayers,
                                                                      const int nqubits, const int* perms[],
                                                                      numeric* fval, double* grad_vec, double* H)
{
    if (!check_brickwall(nlayers, nqubits, perms)) return -1;
    int err;
    mg::Target* t = cached_target(ufunc, udata, nqubits, &err);
    if (!t) return err;
    const mg::Program p = mg::brickwall_program(nqubits, nlayers, perms);
    const int m = 8 * nlayers;
    std::vector<zd> dgrad(m), hess((size_t)m * m);
    zd f;
    if (full_hessian(*t, p, Zp(vlist), mg::kFP64, &f, dgrad.data(), hess.data()) < 0) return -1;
    riemannian_gradient(vlist, nlayers, dgrad.data(), grad_vec);
    project_hessian(vlist, nlayers, dgrad.data(), hess.data(), H);
    *fval = f;
    return 0;
}

// =================== GPU-native interface (include/mg_gpu.h) ===================

const char* mg_last_error(void) { return mg::last_error(); }

mg_target* mg_target_dense(const numeric* U, int nqubits)
{
    if (!U || nqubits < 1 || nqubits > 15) {
        mg::set_error("dense target: invalid arguments (2^n x 2^n doubles complex, n <= 15)");
        return nullptr;
    }
    mg::Target* t = mg::target_dense(reinterpret_cast<const zd*>(U), nqubits);
    if (!t) return nullptr;
    auto* h = new mg_target;
    h->t.reset(t);
    return h;
}

mg_target* mg_target_circuit(const struct matchgate gates[], int ngates, const int wires[], int nqubits)
{
    mg::Program p;
    p.L = nqubits;
    p.P = p.G = ngates;
    p.param.resize(ngates);
    for (int g = 0; g < ngates; g++) p.param[g] = g;
    p.wires.assign(wires, wires + 2 * ngates);
    mg::Target* t = mg::target_program(p, Zp(gates));
    if (!t) return nullptr;
    auto* h = new mg_target;
    h->t.reset(t);
    return h;
}

mg_target* mg_target_brickwall(const struct matchgate ulist[], int ulayers, const int* uperms[], int nqubits)
{
    if (!check_brickwall(ulayers, nqubits, uperms)) return nullptr;
    const mg::Program p = mg::brickwall_program(nqubits, ulayers, uperms);
    mg::Target* t = mg::target_program(p, Zp(ulist));
    if (!t) return nullptr;
    auto* h = new mg_target;
    h->t.reset(t);
    return h;
}

mg_target* mg_target_from_linear_func(linear_func ufunc, void* udata, int nqubits)
{
    int err = 0;
    mg::Target* t = materialise(ufunc, udata, nqubits, &err);
    if (!t) return nullptr;
    auto* h = new mg_target;
    h->t.reset(t);
    return h;
}

void mg_target_free(mg_target* t) { delete t; }

int mg_target_densify(mg_target* t)
{
    if (!t || !t->t) {
        mg::set_error("mg_target_densify: null target");
        return -1;
    }
    return mg::target_densify(*t->t);
}

mg_program* mg_program_create(int nqubits, int nparams, int ninst, const int* param, const int* wires)
{
    auto* h = new mg_program;
    h->p.L = nqubits;
    h->p.P = nparams;
    h->p.G = ninst;
    if (ninst > 0 && param && wires) {
        h->p.param.assign(param, param + ninst);
        h->p.wires.assign(wires, wires + 2 * ninst);
    }
    std::string why;
    if (!h->p.validate(&why)) {
        mg::set_error("program: " + why);
        delete h;
        return nullptr;
    }
    return h;
}

mg_program* mg_program_brickwall(int nqubits, int nlayers, const int* perms[])
{
    if (!check_brickwall(nlayers, nqubits, perms)) return nullptr;
    auto* h = new mg_program;
    h->p = mg::brickwall_program(nqubits, nlayers, perms);
    return h;
}

void mg_program_free(mg_program* p) { delete p; }
int mg_program_nparams(const mg_program* p) { return p ? p->p.P : -1; }
int mg_program_ninst(const mg_program* p) { return p ? p->p.G : -1; }

int mg_evaluate(const mg_target* t, const mg_program* p, const struct matchgate gates[], const struct matchgate Z[],
                int what, int precision, numeric* f, struct matchgate dgrad[], struct matchgate dhz[])
{
    if (!t || !p || !gates) {
        mg::set_error("mg_evaluate: null argument");
        return -1;
    }
    mg::EvalOut out;
    if (mg::evaluate(*t->t, p->p, Zp(gates), Z ? Zp(Z) : nullptr, what, precision, &out, 0, -1, &g_last_stats) < 0)
        return -1;
    if (f) *f = out.f;
    if (dgrad && what >= 1) std::copy(out.grad.begin(), out.grad.end(), Zp(dgrad));
    if (dhz && what >= 2) std::copy(out.dhz.begin(), out.dhz.end(), Zp(dhz));
    return 0;
}

int mg_evaluate_noalloc(const mg_target* t, const mg_program* p, const struct matchgate gates[],
                        const struct matchgate Z[], int what, int precision, numeric* f, struct matchgate dgrad[],
                        struct matchgate dhz[])
{
    if (!t || !p || !gates) {
        mg::set_error("mg_evaluate_noalloc: null argument");
        return -1;
    }
    mg::EvalOut out;
    if (mg::evaluate(*t->t, p->p, Zp(gates), Z ? Zp(Z) : nullptr, what, precision, &out, 0, -1, &g_last_stats, true) < 0)
        return -1;
    if (f) *f = out.f;
    if (dgrad && what >= 1) std::copy(out.grad.begin(), out.grad.end(), Zp(dgrad));
    if (dhz && what >= 2) std::copy(out.dhz.begin(), out.dhz.end(), Zp(dhz));
    return 0;
}

int mg_evaluate_columns(const mg_target* t, const mg_program* p, const struct matchgate gates[],
                        const struct matchgate Z[], int what, int precision, long b0, long b1, numeric* f,
                        struct matchgate dgrad[], struct matchgate dhz[])
{
    if (!t || !p || !gates) {
        mg::set_error("mg_evaluate_columns: null argument");
        return -1;
    }
    mg::EvalOut out;
    if (mg::evaluate(*t->t, p->p, Zp(gates), Z ? Zp(Z) : nullptr, what, precision, &out, b0, b1, &g_last_stats) < 0)
        return -1;
    if (f) *f = out.f;
    if (dgrad && what >= 1) std::copy(out.grad.begin(), out.grad.end(), Zp(dgrad));
    if (dhz && what >= 2) std::copy(out.dhz.begin(), out.dhz.end(), Zp(dhz));
    return 0;
}

int mg_riemannian_gradient_hvp(const mg_target* t, const mg_program* p, const struct matchgate gates[], const double* v,
                               int precision, numeric* f, double* grad_vec, double* hvp)
{
    if (!t || !p || !gates) {
        mg::set_error("mg_riemannian_gradient_hvp: null argument");
        return -1;
    }
    const int P = p->p.P;
    std::vector<struct matchgate> Z(v ? P : 0);
    for (int i = 0; v && i < P; i++) real_to_tangent_matchgate(v + 8 * i, &gates[i], &Z[i]);
    mg::EvalOut out;
    if (mg::evaluate(*t->t, p->p, Zp(gates), v ? Zp(Z.data()) : nullptr, v ? mg::kFGradHvp : mg::kFGrad, precision, &out,
                     0, -1, &g_last_stats) < 0)
        return -1;
    if (f) *f = out.f;
    if (grad_vec) riemannian_gradient(gates, P, out.grad.data(), grad_vec);
    if (v && hvp)
        for (int i = 0; i < P; i++) weingarten_block(&gates[i], &out.grad[8 * i], &Z[i], &out.dhz[8 * i], hvp + 8 * i);
    return 0;
}

int mg_riemannian_project(const struct matchgate gates[], int nparams, const struct matchgate dgrad[], const double* v,
                          const struct matchgate dhz[], double* grad_vec, double* hvp)
{
    if (!gates || nparams < 1 || !dgrad || (v && !dhz)) {
        mg::set_error("mg_riemannian_project: null argument");
        return -1;
    }
    const zd* g = Zp(dgrad);
    if (grad_vec) riemannian_gradient(gates, nparams, g, grad_vec);
    if (v && hvp)
        for (int i = 0; i < nparams; i++) {
            struct matchgate Z;
            real_to_tangent_matchgate(v + 8 * i, &gates[i], &Z);
            weingarten_block(&gates[i], g + 8 * i, &Z, Zp(dhz) + 8 * i, hvp + 8 * i);
        }
    return 0;
}

int mg_hessian(const mg_target* t, const mg_program* p, const struct matchgate gates[], int precision, numeric* f,
               struct matchgate dgrad[], numeric* hess)
{
    if (!t || !p || !gates || !hess) {
        mg::set_error("mg_hessian: null argument");
        return -1;
    }
    zd fz;
    std::vector<zd> g((size_t)8 * p->p.P);
    if (full_hessian(*t->t, p->p, Zp(gates), precision, &fz, g.data(), hess) < 0) return -1;
    if (f) *f = fz;
    if (dgrad) std::copy(g.begin(), g.end(), Zp(dgrad));
    return 0;
}

int mg_gradient_vector_hessian_matrix(const mg_target* t, const mg_program* p, const struct matchgate gates[],
                                      int precision, numeric* f, double* grad_vec, double* H)
{
    if (!t || !p || !gates) {
        mg::set_error("mg_gradient_vector_hessian_matrix: null argument");
        return -1;
    }
    const int m = 8 * p->p.P;
    std::vector<zd> g(m), hess((size_t)m * m);
    zd fz;
    if (full_hessian(*t->t, p->p, Zp(gates), precision, &fz, g.data(), hess.data()) < 0) return -1;
    if (grad_vec) riemannian_gradient(gates, p->p.P, g.data(), grad_vec);
    if (H) project_hessian(gates, p->p.P, g.data(), hess.data(), H);
    if (f) *f = fz;
    return 0;
}

void mg_set_checkpointing(int mode) { mg::knobs().force_ckpt = mode; }
void mg_set_lds_states(int mode) { mg::knobs().force_lds = mode; }
void mg_set_kernel_timing(int on) { mg::knobs().time_kernels = on != 0; }
void mg_set_register_kernel(int mode, int threads)
{
    mg::knobs().force_reg = mode;
    mg::knobs().reg_threads = threads;
}
void mg_set_grid_per_cu(int n) { mg::knobs().grid_per_cu = n; }
void mg_set_sweep_threads(int threads) { mg::knobs().lds_threads = threads; }
int mg_last_sweep_threads(void) { return g_last_stats.reg_threads; }
const char* mg_last_sweep_kernel(void) { return g_last_stats.kernel ? g_last_stats.kernel : ""; }
void mg_set_window_kernel(int mode) { mg::knobs().force_win = mode; }

void mg_set_tile_kernel(int mode) { mg::knobs().force_tile = mode != 0 ? -1 : 0; }
void mg_set_row_chunk_bytes(long bytes) { mg::knobs().row_chunk_bytes = bytes > 0 ? (size_t)bytes : 0; }
void mg_set_materialize_kernel(int mode) { mg::knobs().mat_kernel = mode; }
void mg_set_forward_mfma(int mode) { mg::knobs().fwd_mfma = mode; }
void mg_set_device_trust_region(int mode) { mg::knobs().device_tr = mode; }
static thread_local long g_tr_graphs = 0;
long mg_last_trust_region_graphs(void) { return g_tr_graphs; }
float mg_last_kernel_ms(void) { return g_last_stats.kernel_ms; }
void mg_last_sweep_info(int* lds_states, int* ckpt, int* grid, long* lds_bytes)
{
    if (lds_states) *lds_states = g_last_stats.lds_states;
    if (ckpt) *ckpt = g_last_stats.ckpt;
    if (grid) *grid = g_last_stats.grid;
    if (lds_bytes) *lds_bytes = (long)g_last_stats.lds_bytes;
}

// =================== optimisation ===================

struct opt_ctx {
    const mg::Target* t;
    const mg::Program* p;
    int precision;      // requested sweep precision
    int switch_last;    // iterations (from the end) forced to FP64
    int niter;
    int iter;           // iteration index, advanced by every gradient evaluation
    int nhvp;
    int failed;
    numeric* f_citer;           // complex f per iteration (nullable)
};

static int cur_precision(const opt_ctx* c)
{
    if (c->precision == MG_FP64) return MG_FP64;
    return (c->iter >= c->niter - c->switch_last) ? MG_FP64 : MG_FP32;
}

static double opt_f(const double* x, void* fdata, numeric* f_complex)
{
    opt_ctx* c = (opt_ctx*)fdata;
    mg::EvalOut out;
    if (mg::evaluate(*c->t, *c->p, reinterpret_cast<const zd*>(x), nullptr, mg::kF, cur_precision(c), &out, 0, -1,
                     &g_last_stats) < 0) {
        c->failed = 1;
        return NAN;
    }
    *f_complex = out.f;
    return out.f.real();
}

static double opt_f_deriv(const double* x, void* fdata, double* grad)
{
    opt_ctx* c = (opt_ctx*)fdata;
    c->iter++;
    mg::EvalOut out;
    if (mg::evaluate(*c->t, *c->p, reinterpret_cast<const zd*>(x), nullptr, mg::kFGrad, cur_precision(c), &out, 0, -1,
                     &g_last_stats) < 0) {
        c->failed = 1;
        return NAN;
    }
    riemannian_gradient(reinterpret_cast<const struct matchgate*>(x), c->p->P, out.grad.data(), grad);
    if (c->f_citer && c->iter >= 0 && c->iter < c->niter) c->f_citer[c->iter] = out.f;
    return out.f.real();
}

static void opt_hvp(const double* x, void* fdata, double* vec, double* hvp)
{
    opt_ctx* c = (opt_ctx*)fdata;
    const int P = c->p->P;
    const struct matchgate* V = reinterpret_cast<const struct matchgate*>(x);
    std::vector<struct matchgate> Z(P);
    for (int i = 0; i < P; i++) real_to_tangent_matchgate(vec + 8 * i, &V[i], &Z[i]);
    mg::EvalOut out;
    if (mg::evaluate(*c->t, *c->p, reinterpret_cast<const zd*>(x), Zp(Z.data()), mg::kFGradHvp, cur_precision(c), &out, 0,
                     -1, &g_last_stats) < 0) {
        c->failed = 1;
        std::fill(hvp, hvp + 8 * P, NAN);
        return;
    }
    c->nhvp++;
    for (int i = 0; i < P; i++) weingarten_block(&V[i], &out.grad[8 * i], &Z[i], &out.dhz[8 * i], hvp + 8 * i);
}

static double opt_f_deriv_hess(const double* x, void* fdata, numeric* f_complex, double* grad, double* H)
{
    opt_ctx* c = (opt_ctx*)fdata;
    c->iter++;
    const int P = c->p->P, m = 8 * P;
    std::vector<zd> g(m), hess((size_t)m * m);
    zd f;
    if (full_hessian(*c->t, *c->p, reinterpret_cast<const zd*>(x), cur_precision(c), &f, g.data(), hess.data()) < 0) {
        c->failed = 1;
        return NAN;
    }
    const struct matchgate* V = reinterpret_cast<const struct matchgate*>(x);
    riemannian_gradient(V, P, g.data(), grad);
    project_hessian(V, P, g.data(), hess.data(), H);
    *f_complex = f;
    return f.real();
}

// per-gate retraction (reference src/mg_brickwall_opt.c:79-94)
static void opt_retract(const double* x, const double* eta, void* rdata, double* xs)
{
    const int P = *(const int*)rdata;
    const struct matchgate* V = reinterpret_cast<const struct matchgate*>(x);
    struct matchgate* W = reinterpret_cast<struct matchgate*>(xs);
    for (int i = 0; i < P; i++) retract_matchgate(&V[i], eta + 8 * i, &W[i]);
}

int mg_optimize(const mg_target* t, const mg_program* p, const struct matchgate gates_start[],
                const struct rtr_params* params, int niter, int hvp, int precision, int precision_switch,
                double* f_iter, numeric* f_citer, struct matchgate gates_opt[], int* nhvp)
{
    if (!t || !p || !gates_start || !params || !f_iter || !gates_opt || niter < 0) {
        mg::set_error("mg_optimize: invalid argument");
        return -1;
    }
    opt_ctx c{t->t.get(), &p->p, precision, std::max(0, precision_switch), niter, -1, 0, 0, hvp ? f_citer : nullptr};
    int P = p->p.P;
    const int n = 8 * P, m = 16 * P;
    g_tr_graphs = 0;
    if (hvp && mg::knobs().device_tr != 0) {
        // the trust region on the device, one graph per iteration, where its sweeps have plans
        mg::DevTRStats ds;
        const int r = mg::optimize_device(*t->t, p->p, reinterpret_cast<const zd*>(gates_start), params, niter, precision,
                                          precision_switch, mg::knobs().device_tr == 2 ? 1 : 0, f_iter,
                                          reinterpret_cast<zd*>(f_citer),
                                          reinterpret_cast<zd*>(gates_opt), nhvp, &ds);
        if (r < 0) return -1;
        if (r > 0) {
            g_tr_graphs = ds.graph_launches;
            return 0;
        }
    }
    if (hvp) {
        riemannian_trust_region_optimize_hvp(opt_f, opt_f_deriv, opt_hvp, &c, opt_retract, &P, n,
                                             reinterpret_cast<const double*>(gates_start), m, params, niter, f_iter,
                                             reinterpret_cast<double*>(gates_opt));
        if (f_citer) {
            numeric fc;
            c.iter = niter;
            opt_f(reinterpret_cast<const double*>(gates_opt), &c, &fc);
            f_citer[niter] = fc;
        }
    } else {
        riemannian_trust_region_optimize_hmat(opt_f, opt_f_deriv_hess, &c, opt_retract, &P, n,
                                              reinterpret_cast<const double*>(gates_start), m, params, niter, f_iter,
                                              f_citer, reinterpret_cast<double*>(gates_opt));
    }
    if (nhvp) *nhvp = c.nhvp;
    if (c.failed) return -1;
    return 0;
}

// =================== brickwall drivers (include/mg_brickwall_opt.h) ===================

static void brickwall_opt(linear_func ufunc, void* udata, const struct matchgate vlist_start[], const int nlayers,
                          const int nqubits, const int* perms[], struct rtr_params* params, const int niter,
                          double* f_iter, numeric* f_citer, struct matchgate vlist_opt[], int hvp)
{
    // the reference resets the per-iteration callback (src/mg_brickwall_opt.c:114-116)
    params->g_func = NULL;
    params->g_data = NULL;
    params->g_iter = NULL;
    // on any failure the caller gets NaN objective values and the start gates, never
    // uninitialised output, and the reason on stderr (the reference reports its failures there)
    auto fail = [&](const char* what) {
        std::fprintf(stderr, "%s: %s\n", what, mg::last_error());
        for (int k = 0; k <= niter; k++) {
            if (f_iter) f_iter[k] = NAN;
            if (f_citer) f_citer[k] = numeric(NAN, NAN);
        }
        if (vlist_opt && vlist_start && nlayers > 0) std::memcpy(vlist_opt, vlist_start, sizeof(struct matchgate) * nlayers);
    };
    const char* name = hvp ? "optimize_matchgate_brickwall_circuit_hvp" : "optimize_matchgate_brickwall_circuit_hmat";
    if (!check_brickwall(nlayers, nqubits, perms)) return fail(name);
    int err;
    mg::Target* t = cached_target(ufunc, udata, nqubits, &err);
    if (!t) {
        if (err == -2) mg::set_error("call of 'ufunc' failed");
        return fail(name);
    }
    mg_program prog;
    prog.p = mg::brickwall_program(nqubits, nlayers, perms);
    mg_target th;
    th.t.reset(t);
    const int ret = mg_optimize(&th, &prog, vlist_start, params, niter, hvp, MG_FP64, 0, f_iter, f_citer, vlist_opt, nullptr);
    (void)th.t.release();  // owned by the target cache
    if (ret < 0) fail(name);
}

void optimize_matchgate_brickwall_circuit_hmat(linear_func ufunc, void* udata, const struct matchgate vlist_start[],
                                               const int nlayers, const int nqubits, const int* perms[],
                                               struct rtr_params* params, const int niter, double* f_iter,
                                               numeric* f_citer, struct matchgate vlist_opt[])
{
    brickwall_opt(ufunc, udata, vlist_start, nlayers, nqubits, perms, params, niter, f_iter, f_citer, vlist_opt, 0);
}

void optimize_matchgate_brickwall_circuit_hvp(linear_func ufunc, void* udata, const struct matchgate vlist_start[],
                                              const int nlayers, const int nqubits, const int* perms[],
                                              struct rtr_params* params, const int niter, double* f_iter,
                                              numeric* f_citer, struct matchgate vlist_opt[])
{
    brickwall_opt(ufunc, udata, vlist_start, nlayers, nqubits, perms, params, niter, f_iter, f_citer, vlist_opt, 1);
}

}  // extern "C"
