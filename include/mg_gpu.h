/*
 * GPU-native interface of the matchgate target / gradient / Hessian-vector
 * product path (no reference counterpart: the reference evaluates targets through
 * a per-column host callback, src/matchgate_target.h:10, and has no HVP).
 *
 * Targets live on the device; gate programs generalise the reference's tied
 * brickwall (one gate per layer, src/matchgate_target.c:308-324) to any map from
 * gate instances to parameter gates, e.g. the open-boundary brickwall with an
 * independent gate per instance.  Sweeps run in FP64 or FP32 (MG_FP32); traces,
 * gradients and second derivatives are always accumulated in FP64 and returned as
 * FP64.  Results are deterministic run to run.
 *
 * All functions return 0 on success and -1 on error (message on stderr and in
 * mg_last_error()).  The library is used from one host thread at a time.
 */
#pragma once
#include "matchgate.h"
#include "matchgate_target.h"
#include "trust_region.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mg_target mg_target;
typedef struct mg_program mg_program;

enum { MG_FP64 = 0, MG_FP32 = 1 };
enum { MG_EVAL_F = 0, MG_EVAL_F_GRAD = 1, MG_EVAL_F_GRAD_HVP = 2 };

const char* mg_last_error(void);

/* ---- targets U (f = -Tr(U^dagger W)) ---- */
/* dense U, row-major 2^n x 2^n (the zgemv 'ufunc' of examples/hubbard1d/hubbard1d_opt.c:41-52) */
mg_target* mg_target_dense(const numeric* U, int nqubits);
/* matrix-free brickwall U (the 'ufunc_matfree' of examples/hubbard1d/hubbard1d_opt.c:23-39) */
mg_target* mg_target_brickwall(const struct matchgate ulist[], int ulayers, const int* uperms[], int nqubits);
/* sequential circuit U: gate k on wires (wires[2k], wires[2k+1]) */
mg_target* mg_target_circuit(const struct matchgate gates[], int ngates, const int wires[], int nqubits);
/* any linear map, materialised once by 2^n calls of ufunc (reentrant, called from several threads) */
mg_target* mg_target_from_linear_func(linear_func ufunc, void* udata, int nqubits);
void mg_target_free(mg_target* t);
/* materialise a circuit or brickwall target once on the GPU as dense rows (2^n x 2^n FP64);
 * later evaluations read the rows instead of applying the gates per column.  No-op for dense
 * targets. */
int mg_target_densify(mg_target* t);

/* ---- gate programs ---- */
/* ninst instances; instance g applies parameter gate param[g] to (wires[2g], wires[2g+1]) */
mg_program* mg_program_create(int nqubits, int nparams, int ninst, const int* param, const int* wires);
/* the reference's tied brickwall: nlayers parameter gates */
mg_program* mg_program_brickwall(int nqubits, int nlayers, const int* perms[]);
void mg_program_free(mg_program* p);
int mg_program_nparams(const mg_program* p);
int mg_program_ninst(const mg_program* p);

/* ---- evaluation over all 2^n basis columns ----
 * what = MG_EVAL_F: f;  MG_EVAL_F_GRAD: + complex gradient dgrad[P] (df/dV, no conjugation,
 * src/gate.c:718-722);  MG_EVAL_F_GRAD_HVP: + dhz[P] = D^2 f[Z] for the complex direction Z[P]. */
int mg_evaluate(const mg_target* t, const mg_program* p, const struct matchgate gates[], const struct matchgate Z[],
                int what, int precision, numeric* f, struct matchgate dgrad[], struct matchgate dhz[]);

/* mg_evaluate without device allocation (no reference counterpart: for callers that cannot
 * allocate device memory, e.g. code under HIP graph capture or tight serving loops).  Every
 * device buffer of the sweep (gate descriptors, per-workgroup partial sums, the reduced
 * result) comes from a static 32 MiB array that each device holds for the life of the
 * process.  Sweeps that need more -- states that fit neither the LDS nor registers (more
 * than 12 qubits), checkpointed sweeps of gates that are not unitary, or partial sums beyond
 * the array -- fail with an error, as does an FP32 evaluation of a target whose FP32 copy
 * one FP32 mg_evaluate has not yet made.  Results equal mg_evaluate's.  Calls share the
 * array: calls that overlap on one device (several host threads) race and must not be made. */
int mg_evaluate_noalloc(const mg_target* t, const mg_program* p, const struct matchgate gates[],
                        const struct matchgate Z[], int what, int precision, numeric* f, struct matchgate dgrad[],
                        struct matchgate dhz[]);

/* the same over the basis columns [b0, b1) only: partial sums that add up to the full
 * trace (one GPU's share when the columns are split across devices) */
int mg_evaluate_columns(const mg_target* t, const mg_program* p, const struct matchgate gates[],
                        const struct matchgate Z[], int what, int precision, long b0, long b1, numeric* f,
                        struct matchgate dgrad[], struct matchgate dhz[]);

/* the Riemannian projection alone (host, FP64): summed partial sums of mg_evaluate_columns
 * (after adding them across devices) -> Riemannian gradient (8P reals) and, if v != NULL, the
 * Hessian-vector product H v (8P reals) of the direction v whose tangent was passed as Z;
 * mg_riemannian_gradient_hvp applies the same projection to a whole-trace evaluation */
int mg_riemannian_project(const struct matchgate gates[], int nparams, const struct matchgate dgrad[], const double* v,
                          const struct matchgate dhz[], double* grad_vec, double* hvp);

/* Riemannian gradient (8P reals, src/matchgate_target.c:903-908) and, if v != NULL,
 * the Riemannian Hessian-vector product H v (8P reals), one sweep */
int mg_riemannian_gradient_hvp(const mg_target* t, const mg_program* p, const struct matchgate gates[], const double* v,
                               int precision, numeric* f, double* grad_vec, double* hvp);

/* complex Hessian (8P)^2 (layout of src/matchgate_brickwall.c:135) from 8P directional sweeps */
int mg_hessian(const mg_target* t, const mg_program* p, const struct matchgate gates[], int precision, numeric* f,
               struct matchgate dgrad[], numeric* hess);

/* real gradient and Riemannian Hessian matrix (src/matchgate_target.c:877-973) */
int mg_gradient_vector_hessian_matrix(const mg_target* t, const mg_program* p, const struct matchgate gates[],
                                      int precision, numeric* f, double* grad_vec, double* H);

/* ---- Riemannian trust-region optimisation on the GPU path ----
 * hvp != 0: truncated CG on GPU Hessian-vector products; hvp == 0: explicit Hessian matrix.
 * precision_switch: the last precision_switch iterations (and the final f) run in FP64
 * when precision == MG_FP32.  nhvp (nullable) receives the number of HVP sweeps. */
int mg_optimize(const mg_target* t, const mg_program* p, const struct matchgate gates_start[],
                const struct rtr_params* params, int niter, int hvp, int precision, int precision_switch,
                double* f_iter, numeric* f_citer, struct matchgate gates_opt[], int* nhvp);

/* ---- host trust region with reuse (no device work) ----
 * riemannian_trust_region_optimize_hvp (include/trust_region.h, same arguments); with
 * heta_from_cg != 0, H eta of each ratio test is the truncated CG's own sum of its steps times
 * the products H d it already evaluated: one f_hvp call per iteration fewer, equal to a separate
 * product up to rounding (the served jobs' trust region, mg_serve_options.tr_reuse). */
void mg_trust_region_optimize_hvp(const target_func f, const target_gradient_func f_deriv,
                                  const hessian_vector_product_func f_hvp, void* fdata, retract_func retract,
                                  void* rdata, const int n, const double* x_init, const int m,
                                  const struct rtr_params* params, const int niter, double* f_iter, double* x_final,
                                  int heta_from_cg);

/* ---- batched evaluation ----
 * Requests of any targets, programs, qubit counts (<= 16) and modes, merged into one
 * persistent launch per precision.  Outputs as mg_riemannian_gradient_hvp.  A request's
 * results are the same bits whether it is evaluated alone or with others (they depend
 * on the batch shape of its qubit count, mg_batch_set_shape, only). */
typedef struct mg_request {
    const mg_target* target;
    const mg_program* program;
    const struct matchgate* gates; /* [nparams] */
    const double* v;               /* real tangent direction [8 nparams] (what == MG_EVAL_F_GRAD_HVP) */
    int what;                      /* MG_EVAL_F, MG_EVAL_F_GRAD or MG_EVAL_F_GRAD_HVP */
    int precision;                 /* MG_FP64 or MG_FP32 sweeps */
    numeric* f;                    /* outputs, each nullable */
    double* grad_vec;              /* Riemannian gradient [8 nparams] (what >= MG_EVAL_F_GRAD) */
    double* hvp;                   /* Riemannian Hessian-vector product [8 nparams] */
    long b0, b1;                   /* basis columns [b0, b1) (partial sums, as mg_evaluate_columns);
                                      b1 <= 0: all 2^n columns */
} mg_request;
int mg_evaluate_batch(mg_request* reqs, int n);
/* kernel variant of the batched launch per precision: threads per workgroup and gate-window
 * bits (each thread applies gates to 2^window_bits amplitudes per state in registers).
 * Compiled: FP32 (256, 2..4), (512, 2..3); FP64 (256, 2..3), (512, 2).  Returns -1 for others. */
int mg_batch_set_kernel(int precision, int threads, int window_bits);
void mg_batch_get_kernel(int precision, int* threads, int* window_bits);
/* requests with dense targets of min_qubits..max_qubits qubits (default 12..30) fill the GPU on
 * their own: mg_evaluate_batch runs them on the single-request kernels (register gate-window,
 * LDS and tiled sweeps), one launch each, and the rest in the persistent batch; (0, 0) batches
 * everything */
void mg_batch_set_single_range(int min_qubits, int max_qubits);
/* workgroups per CU of the batched launch (0 automatic) */
void mg_batch_set_grid_per_cu(int n);
/* batch shape for nqubits in the given precision: threads per column (64 .. threads, a power
 * of two) and basis columns per work item; 0 restores the default */
void mg_batch_set_shape(int precision, int nqubits, int threads_per_column, int columns_per_item);
void mg_batch_get_shape(int precision, int nqubits, int* threads_per_column, int* columns_per_item);
/* duration of the last mg_evaluate_batch kernels in ms (needs mg_set_kernel_timing(1); a launch
 * replayed from a graph is timed whole, upload and download included) */
float mg_batch_last_kernel_ms(void);
/* HIP graphs for batched launches (default on): a launch whose arguments, buffers and sizes
 * repeat an earlier launch is captured once (upload, persistent kernel, reduction, download) and
 * replayed from then on; the request descriptors travel in the upload, so one graph serves every
 * trust-region step of a fixed set of requests.  Results are identical with and without graphs. */
void mg_batch_set_graphs(int on);
/* replayed and newly captured launches of the last mg_evaluate_batch call */
void mg_batch_last_graph_stats(int* replays, int* captures);
/* autotuner: times every compiled (threads, window bits) variant with 1 and 2 workgroups per
 * CU on the given calibration requests (their precision field is ignored; `precision` is
 * tuned), reps timed launches each, and keeps the fastest for that precision.  Outputs
 * (nullable) the choice and its mean kernel time.  Run it before serving: results depend on
 * the kernel variant, so it is fixed while requests are compared. */
int mg_batch_autotune(mg_request* reqs, int n, int precision, int reps, int* threads, int* window_bits, int* per_cu,
                      float* kernel_ms);

/* ---- serving: a stream of trust-region optimisation jobs ----
 * Each job runs riemannian_trust_region_optimize_hvp with set_rtr_default_params(16 nparams)
 * on its circuit and target (with tr_reuse, the variant described at mg_serve_options).  batched = 1: jobs are admitted as they arrive (and as slots and
 * device memory allow) and all admitted jobs' evaluations are merged into shared launches;
 * batched = 0: jobs run one at a time, one launch per evaluation.  The target is created on
 * the device at admission and freed when the job ends, unless the job names a resident target
 * (`target`), which the caller keeps: several jobs may share one, and it costs no admission time
 * or job footprint. */
typedef struct mg_job {
    int nqubits;
    const mg_program* program;           /* circuit to optimise */
    const struct matchgate* gates_start; /* [nparams] */
    const numeric* U;                    /* dense target (row-major 2^n x 2^n), or NULL for a circuit target: */
    int target_ngates;                   /*   gate k on (target_wires[2k], target_wires[2k+1]) */
    const struct matchgate* target_gates;
    const int* target_wires;
    int niter;
    double arrival;                      /* seconds after the start of mg_serve */
    double* f_iter;                      /* out: [niter + 1] as riemannian_trust_region_optimize_hvp */
    struct matchgate* gates_opt;         /* out: [nparams] */
    double t_admit, t_done;              /* out: seconds after the start */
    int nhvp;                            /* out: HVP evaluations */
    int status;                          /* out: 0 ok, -1 failed, -2 stopped or not run (time_limit_s) */
    const mg_target* target;             /* a resident target of nqubits (mg_target_*), used instead of U or
                                            the circuit and left to the caller; NULL: U or the circuit */
} mg_job;

typedef struct mg_serve_options {
    int batched;          /* 1: continuous batching, 0: one job at a time */
    int precision;        /* sweep precision (MG_FP32 or MG_FP64) */
    int precision_switch; /* with MG_FP32: the last iterations run in FP64 */
    int max_active;       /* admitted jobs at most (0: no limit) */
    double hbm_budget;    /* device bytes the admitted jobs may hold (0: no limit) */
    int realtime;         /* 1: admit a job no earlier than its arrival time; 0: all arrive at 0 */
    int densify_qubits;   /* circuit targets of at most this many qubits are materialised once on the
                             GPU at admission (2^n x 2^n FP64, plus an FP32 copy) and read as dense
                             rows afterwards; 0: always applied gate by gate per column */
    double progress_s;    /* > 0: after a launch round (or a finished job, one job at a time), print a
                             progress line to stderr if this many seconds passed since the last one */
    int tr_reuse;         /* 1: each job's trust region reuses what it evaluates: H eta of the ratio
                             test is summed from the truncated CG's own products (one HVP per
                             iteration fewer; equal up to rounding); f at a trial point is evaluated
                             with its gradient while iterations remain, and an iteration that starts
                             at gates evaluated last (an accepted trial point, or the same iterate
                             after a rejected step; same precision) takes f + gradient from there (one
                             evaluation per iteration fewer); 0: the control flow of
                             riemannian_trust_region_optimize_hvp */
    double time_limit_s;  /* > 0: no evaluation starts later than this many seconds after the start;
                             a job cut short ends with status -2 (its outputs are not valid), a job
                             not yet started is not run (status -2); 0: no limit */
} mg_serve_options;

typedef struct mg_serve_stats {
    double wall_s;         /* from start to the last job's end */
    long rounds;           /* launches (batched) or evaluations (one job at a time) */
    long requests;         /* evaluations */
    double kernel_ms;      /* sweep kernel time (with mg_set_kernel_timing(1)) */
    int max_active;        /* peak admitted jobs */
    double peak_footprint; /* peak device bytes held by admitted jobs */
    double eval_s;         /* wall time inside evaluations (host preparation, launches, waits) */
    double prep_s;         /* host time preparing batched launches */
    long graph_replays;    /* batched launches replayed from captured HIP graphs */
} mg_serve_stats;

int mg_serve(mg_job* jobs, int njobs, const mg_serve_options* opt, mg_serve_stats* stats);

/* ---- tuning and measurement ---- */
/* force checkpointed (1) or uncomputed (0) reverse sweeps; -1 restores automatic choice */
void mg_set_checkpointing(int mode);
/* force LDS-resident (1) or global-scratch (0) states; -1 automatic */
void mg_set_lds_states(int mode);
/* register-resident sweep kernel (states in VGPRs, LDS as re-layout buffer): mode -1 automatic
 * (used when the states do not fit LDS), 0 never, 1 whenever a configuration exists;
 * threads 0 automatic, 512 or 1024 */
void mg_set_register_kernel(int mode, int threads);
/* threads per workgroup of the LDS/global-state sweep kernel: 0 automatic, 256, 512 or 1024 */
void mg_set_sweep_threads(int threads);
/* workgroups per CU of the sweep launch (0 automatic) */
void mg_set_grid_per_cu(int n);
/* threads per workgroup of the last register-resident sweep, 0 if the last sweep used the LDS/global kernel */
int mg_last_sweep_threads(void);
/* name of the last sweep's kernel: "sweep_win_kernel" (register-resident, gate windows),
 * "sweep_reg_kernel" (register-resident, one re-layout per gate) or "sweep_kernel" (LDS/global states) */
const char* mg_last_sweep_kernel(void);
/* register-resident gate-window kernel (12 qubits, dense targets, unitary gates): -1 automatic
 * (used for FP64 sweeps where it applies), 0 never, 1 also for FP32 sweeps */
void mg_set_window_kernel(int mode);
/* tiled kernel for dense targets whose states exceed the LDS (L >= 13; FP64 L=12 for gates not
 * unitary to 1e-12): mode -1 automatic (default), 0 never (global-scratch or register kernels) */
void mg_set_tile_kernel(int mode);
/* program (circuit or brickwall) targets beyond 12 qubits are evaluated by materialising their
 * dense rows one column chunk at a time next to the tiled sweeps: bytes of rows per chunk (at least
 * one grid of columns; 0 restores the default 4 GiB) */
void mg_set_row_chunk_bytes(long bytes);
/* rows of program targets beyond 12 qubits (mg_target_densify and the column chunks above): mode -1
 * automatic (default: the matrix cores, FP64 MFMA on 4-qubit windows), 0 the 12-qubit register
 * tiles of the tiled window kernel, 1 the matrix cores */
void mg_set_materialize_kernel(int mode);
/* Forward passes of FP32 sweeps on the matrix cores: the tiled f and f + gradient sweeps (L >= 13)
 * and the f + gradient + HVP sweep with its states in LDS (L = 12).  Mode -1 automatic (default: the
 * forward pass over 4-qubit windows, each window's gates as two 8 x 8 parity blocks, FP32 MFMA;
 * for the HVP also dpsi <- M dpsi + N psi), 0 gate by gate on the VALU */
void mg_set_forward_mfma(int mode);
/* mg_optimize with HVPs (and the drivers over it): mode -1 automatic (default: against a dense
 * target, the trust region with its truncated CG on the device, each iteration one HIP graph
 * launch, where the single-request sweeps have plans: LDS-resident states or the L=12 register
 * window kernel), 0 the host loop.  Both follow the reference control flow.  Mode 2: on the device
 * with reuse (as mg_serve_options.tr_reuse): H eta of each ratio test from the truncated CG's
 * products, and the f + gradient sweep at an accepted trial point kept for the next iteration
 * (one sweep per iteration fewer, up to two without a rejection; iterates equal up to rounding). */
void mg_set_device_trust_region(int mode);
/* graph launches of the last mg_optimize on the device (0: it ran on the host loop) */
long mg_last_trust_region_graphs(void);
/* record HIP events around every sweep kernel on the library's stream */
void mg_set_kernel_timing(int on);
/* duration of the last sweep kernel in ms (needs mg_set_kernel_timing(1)) */
float mg_last_kernel_ms(void);
/* details of the last sweep: states in LDS, checkpointing, grid size, dynamic LDS bytes */
void mg_last_sweep_info(int* lds_states, int* ckpt, int* grid, long* lds_bytes);

#ifdef __cplusplus
}
#endif
