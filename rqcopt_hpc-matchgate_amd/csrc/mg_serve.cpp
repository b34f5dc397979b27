// Batched evaluation and the optimisation-job server (include/mg_gpu.h).
//
// mg_evaluate_batch merges requests of any targets, programs, sizes and modes into
// one persistent launch per precision (mg_batch.hip) and projects the results like
// mg_riemannian_gradient_hvp.
//
// mg_serve runs Riemannian trust-region optimisations (the reference's
// riemannian_trust_region_optimize_* loop, src/trust_region.c:235-306, in its HVP
// form) for a stream of jobs.  With batching on, every admitted job runs its trust
// region in a host thread of its own; each f / gradient / HVP evaluation the job
// needs is posted to the server, which launches one batched evaluation as soon as
// every admitted job has posted one (continuous batching: jobs that arrive while a
// launch is on the GPU are admitted before the next one, finished jobs leave at once).
// Evaluations that fill the GPU on their own (dense targets of 12 or more qubits, which run
// on the single-request kernels) do not wait for these rounds: a second lane, a worker thread
// with its own stream and buffers, runs them as they are posted, so the large jobs' sweeps
// and the small jobs' batched rounds overlap on the device and neither waits for the other.
// Targets are created on the device at admission and freed at completion; admission
// is limited by a job count and by the device memory the admitted jobs hold.
// Without batching the jobs run one after another, each evaluation its own launch
// (the plain single-job path).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/mg_gpu.h"
#include "../../include/trust_region.h"
#include "mg_host.hpp"

using mg::zd;

using mg::riemannian_gradient;

namespace {

const zd* Zp(const struct matchgate* g) { return reinterpret_cast<const zd*>(g); }

// Riemannian outputs of one request from the sweep outputs
void project_request(const struct matchgate* V, int P, int mode, const struct matchgate* Z, const mg::EvalOut& o,
                     numeric* f, double* grad_vec, double* hvp)
{
    if (f) *f = o.f;
    if (mode >= 1 && grad_vec) riemannian_gradient(V, P, o.grad.data(), grad_vec);
    if (mode == 2 && hvp)
        for (int i = 0; i < P; i++) mg::weingarten_block(&V[i], &o.grad[8 * i], &Z[i], &o.dhz[8 * i], hvp + 8 * i);
}

// ---- the job server ------------------------------------------------------------------------

struct Server;

struct Posted {
    mg::BatchReq req;
    mg::EvalOut out;
    int status = 0;
    bool done = false;
};

struct JobRt {
    Server* srv = nullptr;
    mg_job* spec = nullptr;
    mg::Target* target = nullptr;             // the job's target: its own, or the caller's resident one
    std::unique_ptr<mg::Target> own_target;
    size_t footprint = 0;
    int precision = MG_FP64, switch_last = 0, niter = 0, iter = -1, nhvp = 0, failed = 0;
    bool reuse = false;
    // the last two evaluations of f (+ Riemannian gradient): gates, precision, results
    struct Kept {
        std::vector<double> x, grad;
        int prec = -1;
        bool has_grad = false;
        mg::zd f{0.0, 0.0};
    };
    Kept kept[2];
    int kept_next = 0;
    std::thread th;
    bool finished = false;
    bool big = false;  // its evaluations fill the GPU alone: posted to the second lane
    bool cut = false;  // stopped by the server's time limit
};

struct Server {
    std::mutex mu;
    std::condition_variable cv_server, cv_jobs, cv_big;
    std::vector<Posted*> posted;      // batched rounds
    std::vector<Posted*> big_posted;  // the second lane
    bool stop_big = false;
    int running = 0, running_small = 0;
    std::chrono::steady_clock::time_point t0;
    std::function<void()> on_eval;  // after each evaluation outside the batched rounds (progress)
    double time_limit_s = 0.0;  // > 0: evaluations that would start later fail (jobs end with status -2)
    bool past_limit() const
    {
        return time_limit_s > 0.0 &&
               std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > time_limit_s;
    }
    bool batched = true;
    long rounds = 0, requests = 0;
    double kernel_ms = 0.0, eval_s = 0.0, prep_ms = 0.0;
    long graph_replays = 0;
};

int job_precision(const JobRt* j)
{
    if (j->precision == MG_FP64) return MG_FP64;
    return (j->iter >= j->niter - j->switch_last) ? MG_FP64 : MG_FP32;
}

// one evaluation for a job: posted to the server (batched) or launched directly
int job_eval(JobRt* j, const zd* gates, const zd* Z, int mode, mg::EvalOut* out)
{
    const mg::Program& prog = j->spec->program->p;
    if (j->srv->past_limit()) {
        j->cut = true;
        return -1;
    }
    if (!j->srv->batched) {
        mg::EvalStats st;
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = mg::evaluate(*j->target, prog, gates, Z, mode, job_precision(j), out, 0, -1, &st);
        j->srv->eval_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        j->srv->rounds++;
        j->srv->requests++;
        j->srv->kernel_ms += st.kernel_ms;
        if (j->srv->on_eval) j->srv->on_eval();
        return rc;
    }
    Posted p;
    p.req.t = j->target;
    p.req.p = &prog;
    p.req.gates = gates;
    p.req.Z = Z;
    p.req.mode = mode;
    p.req.precision = job_precision(j);
    p.req.out = &p.out;
    std::unique_lock<std::mutex> lk(j->srv->mu);
    if (j->big) {
        j->srv->big_posted.push_back(&p);
        j->srv->cv_big.notify_one();
    } else {
        j->srv->posted.push_back(&p);
        j->srv->cv_server.notify_one();
    }
    j->srv->cv_jobs.wait(lk, [&] { return p.done; });
    if (p.status == 0) *out = std::move(p.out);
    return p.status;
}

// with reuse: f (+ Riemannian gradient) at x from one of the last two evaluations when it had
// the same gates and precision (bitwise: evaluations are deterministic)
const JobRt::Kept* job_kept(const JobRt* j, const double* x, bool need_grad)
{
    if (!j->reuse) return nullptr;
    const size_t m = 16 * (size_t)j->spec->program->p.P;
    for (const JobRt::Kept& k : j->kept)
        if ((k.has_grad || !need_grad) && k.prec == job_precision(j) && k.x.size() == m &&
            std::memcmp(k.x.data(), x, m * sizeof(double)) == 0)
            return &k;
    return nullptr;
}

JobRt::Kept& job_keep(JobRt* j, const double* x, const mg::zd& f)
{
    JobRt::Kept& k = j->kept[j->kept_next];
    j->kept_next ^= 1;
    k.x.assign(x, x + 16 * (size_t)j->spec->program->p.P);
    k.prec = job_precision(j);
    k.f = f;
    k.has_grad = false;
    return k;
}

// f + Riemannian gradient at x (one evaluation), kept
const JobRt::Kept* job_f_grad(JobRt* j, const double* x)
{
    const int P = j->spec->program->p.P;
    mg::EvalOut out;
    if (job_eval(j, reinterpret_cast<const zd*>(x), nullptr, mg::kFGrad, &out) < 0) return nullptr;
    JobRt::Kept& k = job_keep(j, x, out.f);
    k.grad.resize(8 * (size_t)P);
    riemannian_gradient(reinterpret_cast<const struct matchgate*>(x), P, out.grad.data(), k.grad.data());
    k.has_grad = true;
    return &k;
}

double job_f(const double* x, void* fdata, numeric* f_complex)
{
    JobRt* j = (JobRt*)fdata;
    if (const JobRt::Kept* k = job_kept(j, x, false)) {
        *f_complex = k->f;
        return k->f.real();
    }
    // with reuse, f at a trial point comes with its gradient while iterations remain: the
    // next iteration starts there when the step is accepted
    if (j->reuse && j->iter + 1 < j->niter) {
        const JobRt::Kept* k = job_f_grad(j, x);
        if (!k) {
            j->failed = 1;
            return NAN;
        }
        *f_complex = k->f;
        return k->f.real();
    }
    mg::EvalOut out;
    if (job_eval(j, reinterpret_cast<const zd*>(x), nullptr, mg::kF, &out) < 0) {
        j->failed = 1;
        return NAN;
    }
    // the final f after the last iteration is asked at these gates when the step is accepted
    if (j->reuse) job_keep(j, x, out.f);
    *f_complex = out.f;
    return out.f.real();
}

double job_f_deriv(const double* x, void* fdata, double* grad)
{
    JobRt* j = (JobRt*)fdata;
    j->iter++;
    const int P = j->spec->program->p.P;
    // after a rejected step the trust region asks again at the same gates; after an
    // accepted one, at the trial point whose f came with its gradient
    if (j->reuse) {
        const JobRt::Kept* k = job_kept(j, x, true);
        if (!k) k = job_f_grad(j, x);
        if (!k) {
            j->failed = 1;
            return NAN;
        }
        std::memcpy(grad, k->grad.data(), 8 * (size_t)P * sizeof(double));
        return k->f.real();
    }
    mg::EvalOut out;
    if (job_eval(j, reinterpret_cast<const zd*>(x), nullptr, mg::kFGrad, &out) < 0) {
        j->failed = 1;
        return NAN;
    }
    riemannian_gradient(reinterpret_cast<const struct matchgate*>(x), P, out.grad.data(), grad);
    return out.f.real();
}

void job_hvp(const double* x, void* fdata, double* vec, double* hvp)
{
    JobRt* j = (JobRt*)fdata;
    const int P = j->spec->program->p.P;
    const struct matchgate* V = reinterpret_cast<const struct matchgate*>(x);
    std::vector<struct matchgate> Z(P);
    for (int i = 0; i < P; i++) real_to_tangent_matchgate(vec + 8 * i, &V[i], &Z[i]);
    mg::EvalOut out;
    if (job_eval(j, reinterpret_cast<const zd*>(x), Zp(Z.data()), mg::kFGradHvp, &out) < 0) {
        j->failed = 1;
        std::fill(hvp, hvp + 8 * P, NAN);
        return;
    }
    j->nhvp++;
    project_request(V, P, 2, Z.data(), out, nullptr, nullptr, hvp);
}

void job_retract(const double* x, const double* eta, void* rdata, double* xs)
{
    const int P = *(const int*)rdata;
    const struct matchgate* V = reinterpret_cast<const struct matchgate*>(x);
    struct matchgate* W = reinterpret_cast<struct matchgate*>(xs);
    for (int i = 0; i < P; i++) retract_matchgate(&V[i], eta + 8 * i, &W[i]);
}

// the job's trust region (reference src/mg_brickwall_opt.c:99-122 with the HVP solver)
void run_job(JobRt* j)
{
    mg_job* s = j->spec;
    int P = s->program->p.P;
    struct rtr_params prm;
    set_rtr_default_params(16 * P, &prm);
    // a failed (or time-limited) evaluation ends the job's trust region at once
    const std::function<bool()> stop = [j] { return j->failed != 0; };
    mg::rtr_optimize_hvp(job_f, job_f_deriv, job_hvp, j, job_retract, &P, 8 * P,
                         reinterpret_cast<const double*>(s->gates_start), 16 * P, &prm, s->niter, s->f_iter,
                         reinterpret_cast<double*>(s->gates_opt), j->reuse, &stop);
    s->nhvp = j->nhvp;
    s->status = j->cut ? -2 : j->failed ? -1 : 0;
}

// device memory a job holds while admitted: its target, its coefficients and its
// share of the batched partial sums (two precisions)
size_t job_footprint(const mg_job& s, int densify_qubits)
{
    const int L = s.nqubits;
    const size_t S = (size_t)1 << L;
    const mg::Program& p = s.program->p;
    size_t b = 0;
    // dense rows: FP64, plus an FP32 copy up to 12 qubits (the tiled FP32 kernel of larger
    // targets rounds the FP64 rows on load)
    if (s.target) {
        // the caller's resident target: not the job's to hold
    } else if (s.U || L <= densify_qubits) {
        b += S * S * (sizeof(double) * 2 + (L <= 12 ? sizeof(float) * 2 : 0));
    } else {
        b += (size_t)s.target_ngates * (mg::kCoefPerInst * 24 + sizeof(mg::Inst));
    }
    b += (size_t)p.G * (mg::kCoefPerInst * 16 + sizeof(mg::Inst));
    int cgw, cpi;
    mg::batch_shape(L, mg::kFP32, &cgw, &cpi);
    b += 2 * (S / cpi + 1) * (2 + 32 * (size_t)p.P) * sizeof(double);
    return b;
}

mg::Target* make_target(const mg_job& s, int densify_qubits)
{
    if (s.U) return mg::target_dense(reinterpret_cast<const zd*>(s.U), s.nqubits);
    mg::Program tp;
    tp.L = s.nqubits;
    tp.G = s.target_ngates;
    tp.P = s.target_ngates;
    tp.param.resize(tp.G);
    tp.wires.assign(s.target_wires, s.target_wires + 2 * tp.G);
    for (int g = 0; g < tp.G; g++) tp.param[g] = g;
    std::unique_ptr<mg::Target> t(mg::target_program(tp, Zp(s.target_gates)));
    if (t && s.nqubits <= densify_qubits && mg::target_densify(*t) < 0) return nullptr;
    return t.release();
}

bool valid_job(const mg_job& s)
{
    if (!s.program || !s.gates_start || !s.f_iter || !s.gates_opt || s.niter < 0) return false;
    if (s.nqubits != s.program->p.L || s.nqubits > 16) return false;
    if (s.target) return s.target->t && s.target->t->L == s.nqubits;
    if (!s.U && (s.target_ngates < 1 || !s.target_gates || !s.target_wires)) return false;
    return true;
}

}  // namespace

extern "C" {

int mg_evaluate_batch(mg_request* reqs, int n)
{
    if (n < 0 || (n > 0 && !reqs)) {
        mg::set_error("mg_evaluate_batch: invalid argument");
        return -1;
    }
    std::vector<mg::BatchReq> br(n);
    std::vector<mg::EvalOut> outs(n);
    std::vector<std::vector<struct matchgate>> Z(n);
    for (int i = 0; i < n; i++) {
        const mg_request& r = reqs[i];
        if (!r.target || !r.program || !r.gates || r.what < MG_EVAL_F || r.what > MG_EVAL_F_GRAD_HVP ||
            (r.what == MG_EVAL_F_GRAD_HVP && !r.v)) {
            mg::set_error("mg_evaluate_batch: invalid request " + std::to_string(i));
            return -1;
        }
        const int P = r.program->p.P;
        if (r.what == MG_EVAL_F_GRAD_HVP) {
            Z[i].resize(P);
            for (int k = 0; k < P; k++) real_to_tangent_matchgate(r.v + 8 * k, &r.gates[k], &Z[i][k]);
        }
        br[i].t = r.target->t.get();
        br[i].p = &r.program->p;
        br[i].gates = Zp(r.gates);
        br[i].Z = Z[i].empty() ? nullptr : Zp(Z[i].data());
        br[i].mode = r.what;
        br[i].precision = r.precision == MG_FP32 ? mg::kFP32 : mg::kFP64;
        br[i].out = &outs[i];
        br[i].b0 = r.b0;
        br[i].b1 = r.b1;
    }
    mg::BatchStats st;
    if (mg::evaluate_batch(br.data(), n, &st) < 0) return -1;
    mg::last_batch_stats() = st;
    for (int i = 0; i < n; i++) {
        const mg_request& r = reqs[i];
        project_request(r.gates, r.program->p.P, r.what, Z[i].empty() ? nullptr : Z[i].data(), outs[i], r.f,
                        r.grad_vec, r.hvp);
    }
    return 0;
}

int mg_batch_set_kernel(int precision, int threads, int window_bits)
{
    const int pi = precision == MG_FP32 ? 1 : 0;
    if (!mg::batch_variant_exists(pi, threads, window_bits)) {
        mg::set_error("mg_batch_set_kernel: no kernel for this precision, thread count and window size");
        return -1;
    }
    mg::batch_tuning().threads[pi] = threads;
    mg::batch_tuning().wbits[pi] = window_bits;
    return 0;
}

void mg_batch_set_single_range(int min_qubits, int max_qubits)
{
    mg::batch_tuning().single_min = min_qubits;
    mg::batch_tuning().single_max = max_qubits;
}

void mg_batch_get_kernel(int precision, int* threads, int* window_bits)
{
    const int pi = precision == MG_FP32 ? 1 : 0;
    if (threads) *threads = mg::batch_tuning().threads[pi];
    if (window_bits) *window_bits = mg::batch_tuning().wbits[pi];
}

void mg_batch_set_grid_per_cu(int n) { mg::batch_tuning().grid_per_cu = n > 0 ? n : 0; }

void mg_batch_set_shape(int precision, int nqubits, int threads_per_column, int columns_per_item)
{
    if (nqubits < 0 || nqubits > 30) return;
    const int pi = precision == MG_FP32 ? 1 : 0;
    mg::batch_tuning().cgw[pi][nqubits] = threads_per_column;
    mg::batch_tuning().cpi[pi][nqubits] = columns_per_item;
}

void mg_batch_get_shape(int precision, int nqubits, int* threads_per_column, int* columns_per_item)
{
    int cgw = 0, cpi = 0;
    if (nqubits >= 2 && nqubits <= 30) mg::batch_shape(nqubits, precision == MG_FP32 ? mg::kFP32 : mg::kFP64, &cgw, &cpi);
    if (threads_per_column) *threads_per_column = cgw;
    if (columns_per_item) *columns_per_item = cpi;
}

float mg_batch_last_kernel_ms(void) { return mg::last_batch_stats().kernel_ms; }

void mg_batch_set_graphs(int on) { mg::batch_tuning().graphs = on != 0; }

void mg_batch_last_graph_stats(int* replays, int* captures)
{
    if (replays) *replays = mg::last_batch_stats().graph_replays;
    if (captures) *captures = mg::last_batch_stats().graph_captures;
}

int mg_batch_autotune(mg_request* reqs, int n, int precision, int reps, int* threads, int* window_bits, int* per_cu,
                      float* kernel_ms)
{
    if (n <= 0 || !reqs || reps < 1) {
        mg::set_error("mg_batch_autotune: invalid argument");
        return -1;
    }
    const int pi = precision == MG_FP32 ? 1 : 0;
    std::vector<mg::BatchReq> br(n);
    std::vector<mg::EvalOut> outs(n);
    std::vector<std::vector<struct matchgate>> Z(n);
    for (int i = 0; i < n; i++) {
        const mg_request& r = reqs[i];
        if (!r.target || !r.program || !r.gates || (r.what == MG_EVAL_F_GRAD_HVP && !r.v)) {
            mg::set_error("mg_batch_autotune: invalid request " + std::to_string(i));
            return -1;
        }
        if (r.what == MG_EVAL_F_GRAD_HVP) {
            Z[i].resize(r.program->p.P);
            for (int k = 0; k < r.program->p.P; k++) real_to_tangent_matchgate(r.v + 8 * k, &r.gates[k], &Z[i][k]);
        }
        br[i] = mg::BatchReq{r.target->t.get(), &r.program->p, Zp(r.gates), Z[i].empty() ? nullptr : Zp(Z[i].data()),
                             r.what, pi ? mg::kFP32 : mg::kFP64, &outs[i], r.b0, r.b1};
    }
    mg::BatchTuning& bt = mg::batch_tuning();
    const mg::BatchTuning saved = bt;
    const bool timing = mg::knobs().time_kernels;
    mg::knobs().time_kernels = true;
    bt.graphs = false;  // time the kernels themselves, not graph replays
    float best = 1e30f;
    int bT = saved.threads[pi], bW = saved.wbits[pi], bP = saved.grid_per_cu;
    for (int T : {256, 512})
        for (int W = 2; W <= 4; W++) {
            if (!mg::batch_variant_exists(pi, T, W)) continue;
            for (int pc : {1, 2}) {
                bt.threads[pi] = T;
                bt.wbits[pi] = W;
                bt.grid_per_cu = pc;
                mg::BatchStats st;
                if (mg::evaluate_batch(br.data(), n, &st) < 0) continue;  // warm-up (and shape check)
                float ms = 0.f;
                bool ok = true;
                for (int k = 0; k < reps && ok; k++) {
                    ok = mg::evaluate_batch(br.data(), n, &st) == 0;
                    ms += st.kernel_ms;
                }
                if (ok && ms < best) {
                    best = ms;
                    bT = T;
                    bW = W;
                    bP = pc;
                }
            }
        }
    mg::knobs().time_kernels = timing;
    bt = saved;
    if (best >= 1e30f) {
        mg::set_error("mg_batch_autotune: no variant ran");
        return -1;
    }
    bt.threads[pi] = bT;
    bt.wbits[pi] = bW;
    bt.grid_per_cu = bP;
    if (threads) *threads = bT;
    if (window_bits) *window_bits = bW;
    if (per_cu) *per_cu = bP;
    if (kernel_ms) *kernel_ms = best / reps;
    return 0;
}

int mg_serve(mg_job* jobs, int njobs, const mg_serve_options* opt, mg_serve_stats* stats)
{
    if (njobs < 0 || (njobs > 0 && !jobs) || !opt) {
        mg::set_error("mg_serve: invalid argument");
        return -1;
    }
    for (int i = 0; i < njobs; i++)
        if (!valid_job(jobs[i])) {
            mg::set_error("mg_serve: invalid job " + std::to_string(i));
            return -1;
        }
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto now_s = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
    std::vector<int> order(njobs);
    for (int i = 0; i < njobs; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return jobs[a].arrival < jobs[b].arrival; });
    auto arrival = [&](int k) { return opt->realtime ? jobs[order[k]].arrival : 0.0; };

    Server srv;
    srv.batched = opt->batched != 0;
    srv.t0 = t0;
    srv.time_limit_s = opt->time_limit_s;
    std::vector<std::unique_ptr<JobRt>> rt(njobs);
    int next = 0, done = 0, peak_active = 0, rc = 0;
    size_t held = 0, peak_held = 0;
    double last_report = 0.0;
    // progress line after a round or a finished job (callers with long runs ask for it)
    auto report = [&](long rounds, long requests) {
        if (opt->progress_s <= 0.0) return;
        const double t = now_s();
        if (t - last_report < opt->progress_s) return;
        last_report = t;
        std::fprintf(stderr, "[mg_serve] %.1f s: %d of %d jobs done, %ld launches, %ld evaluations\n", t, done, njobs,
                     rounds, requests);
        std::fflush(stderr);
    };

    srv.on_eval = [&] { report(srv.rounds, srv.requests); };

    auto start_job = [&](int i) -> bool {
        auto j = std::make_unique<JobRt>();
        j->srv = &srv;
        j->spec = &jobs[i];
        j->precision = opt->precision == MG_FP32 ? MG_FP32 : MG_FP64;
        j->switch_last = std::max(0, opt->precision_switch);
        j->niter = jobs[i].niter;
        j->reuse = opt->tr_reuse != 0;
        j->footprint = job_footprint(jobs[i], opt->densify_qubits);
        jobs[i].t_admit = now_s();
        if (jobs[i].target) {
            j->target = jobs[i].target->t.get();
        } else {
            j->own_target.reset(make_target(jobs[i], opt->densify_qubits));
            j->target = j->own_target.get();
        }
        if (!j->target) {
            jobs[i].status = -1;
            return false;
        }
        rt[i] = std::move(j);
        return true;
    };

    if (!srv.batched) {
        // one job at a time, one launch per evaluation
        for (int k = 0; k < njobs; k++) {
            const int i = order[k];
            const double wait = arrival(k) - now_s();
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            if (srv.past_limit()) {
                jobs[i].status = -2;
                jobs[i].t_admit = jobs[i].t_done = now_s();
                done++;
                continue;
            }
            if (!start_job(i)) { rc = -1; continue; }
            peak_active = 1;
            peak_held = std::max(peak_held, rt[i]->footprint);
            run_job(rt[i].get());
            rt[i]->own_target.reset();
            jobs[i].t_done = now_s();
            if (jobs[i].status == -1) rc = -1;
            done++;
            report(srv.rounds, srv.requests);
        }
    } else {
        // the second lane: evaluations that fill the GPU alone, run as they are posted on this
        // worker's own stream and buffers (they overlap the batched rounds on the device)
        std::thread big_worker([&srv] {
            std::unique_lock<std::mutex> g(srv.mu);
            for (;;) {
                srv.cv_big.wait(g, [&] { return srv.stop_big || !srv.big_posted.empty(); });
                if (srv.big_posted.empty()) break;
                std::vector<Posted*> batch;
                batch.swap(srv.big_posted);
                g.unlock();
                std::vector<mg::BatchReq> reqs(batch.size());
                for (size_t k = 0; k < batch.size(); k++) reqs[k] = batch[k]->req;
                mg::BatchStats st;
                const auto tb = clock::now();
                const int r = mg::evaluate_batch(reqs.data(), (int)reqs.size(), &st);
                const double dt = std::chrono::duration<double>(clock::now() - tb).count();
                g.lock();
                srv.rounds++;
                srv.requests += (long)batch.size();
                srv.kernel_ms += st.kernel_ms;
                srv.eval_s += dt;
                srv.prep_ms += st.prep_ms;
                srv.graph_replays += st.graph_replays;
                for (Posted* p : batch) {
                    p->status = r;
                    p->done = true;
                }
                if (srv.on_eval) srv.on_eval();
                srv.cv_jobs.notify_all();
            }
        });
        std::unique_lock<std::mutex> lk(srv.mu);
        while (done < njobs) {
            // admit arrived jobs while slots and device memory allow
            while (next < njobs && arrival(next) <= now_s()) {
                const int i = order[next];
                if (srv.past_limit()) {
                    jobs[i].status = -2;
                    jobs[i].t_admit = jobs[i].t_done = now_s();
                    next++;
                    done++;
                    continue;
                }
                const size_t fp = job_footprint(jobs[i], opt->densify_qubits);
                const bool slot = opt->max_active <= 0 || srv.running < opt->max_active;
                const bool mem = opt->hbm_budget <= 0 || srv.running == 0 || (double)(held + fp) <= opt->hbm_budget;
                if (!slot || !mem) break;
                lk.unlock();
                const bool ok = start_job(i);
                lk.lock();
                next++;
                if (!ok) {
                    jobs[i].t_done = now_s();
                    done++;
                    rc = -1;
                    continue;
                }
                held += rt[i]->footprint;
                peak_held = std::max(peak_held, held);
                JobRt* j = rt[i].get();
                j->big = mg::batch_single_request(*j->target, jobs[i].program->p);
                srv.running++;
                if (!j->big) srv.running_small++;
                peak_active = std::max(peak_active, srv.running);
                j->th = std::thread([j, &srv] {
                    run_job(j);
                    std::lock_guard<std::mutex> g(srv.mu);
                    j->finished = true;
                    srv.running--;
                    if (!j->big) srv.running_small--;
                    srv.cv_server.notify_one();
                });
            }
            // retire finished jobs
            for (int k = 0; k < next; k++) {
                JobRt* j = rt[order[k]].get();
                if (j && j->finished && j->th.joinable()) {
                    j->th.join();
                    j->own_target.reset();
                    held -= j->footprint;
                    j->spec->t_done = now_s();
                    if (j->spec->status == -1) rc = -1;
                    done++;
                }
            }
            if (done >= njobs) break;
            // launch when every admitted job of this lane waits on an evaluation
            if (srv.running_small > 0 && (int)srv.posted.size() == srv.running_small) {
                std::vector<Posted*> batch;
                batch.swap(srv.posted);
                lk.unlock();
                std::vector<mg::BatchReq> reqs(batch.size());
                for (size_t k = 0; k < batch.size(); k++) reqs[k] = batch[k]->req;
                mg::BatchStats st;
                const auto t0 = clock::now();
                const int r = mg::evaluate_batch(reqs.data(), (int)reqs.size(), &st);
                const double dt = std::chrono::duration<double>(clock::now() - t0).count();
                lk.lock();
                srv.rounds++;
                srv.requests += (long)batch.size();
                report(srv.rounds, srv.requests);
                srv.kernel_ms += st.kernel_ms;
                srv.eval_s += dt;
                srv.prep_ms += st.prep_ms;
                srv.graph_replays += st.graph_replays;
                for (Posted* p : batch) {
                    p->status = r;
                    p->done = true;
                }
                srv.cv_jobs.notify_all();
                continue;
            }
            // otherwise wait for a posting, a finished job or the next arrival
            const bool jobs_pending = srv.running > 0 || next < njobs;
            if (!jobs_pending) break;
            if (next < njobs && srv.running == 0) {
                const double wait = arrival(next) - now_s();
                if (wait > 0) srv.cv_server.wait_for(lk, std::chrono::duration<double>(wait));
            } else if (next < njobs && arrival(next) > now_s()) {
                srv.cv_server.wait_for(lk, std::chrono::duration<double>(arrival(next) - now_s()));
            } else {
                srv.cv_server.wait(lk);
            }
        }
        srv.stop_big = true;
        srv.cv_big.notify_all();
        lk.unlock();
        big_worker.join();
    }
    if (stats) {
        stats->wall_s = now_s();
        stats->rounds = srv.rounds;
        stats->requests = srv.requests;
        stats->kernel_ms = srv.kernel_ms;
        stats->max_active = peak_active;
        stats->peak_footprint = (double)peak_held;
        stats->eval_s = srv.eval_s;
        stats->prep_s = srv.prep_ms * 1e-3;
        stats->graph_replays = srv.graph_replays;
    }
    return rc;
}

}  // extern "C"
