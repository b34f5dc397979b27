"""Benchmark of the matchgate-brickwall f + gradient + HVP path on one MI355X.

Workload (the L=12 single-instance configuration): 6-site 1D Fermi-Hubbard chain,
spin-up on even / spin-down on odd qubits, Jordan-Wigner, J=1, U=4, open ends,
target W = exp(-iHt) at t=0.25 (dense complex128, built on the host, resident in
HBM); circuit = open-boundary brickwall of depth 9 (50 independent matchgates,
seeded Haar U(2) x U(2)).  One step = one fused f + Riemannian gradient +
Hessian-vector product over all 4096 basis columns (one trust-region iteration
unit).  The default precision is FP64 sweeps; --precision fp32 runs FP32 sweeps
with FP64 accumulation.

Also measured: 50 trust-region iterations on that instance; the served stream of
trust-region jobs at its stated qubit counts (L in {8,...,16}; continuous batching, p50/p99
latency, against the one-job-at-a-time FP64 path, which is stopped at a time limit and
reported as unfinished when it does not finish) and a served stream of L <= 12 jobs against
three one-job baselines run to the end; the 64-instance batched workload (20 iterations);
the L=16 instance; and the CPU baselines (the reference's own code on a bounded column sample,
the oracle's HVP, the L=8 configuration).

Prints one JSON line.  `python bench.py --gpus 1 --steps K --warmup W`.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rqcopt_hpc-matchgate_amd"))

import numpy as np  # noqa: E402

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(seed=0):
    from rqcopt_mg.workloads import fermi_hubbard_1d_unitary, haar_matchgates, open_brickwall

    L = 12
    W = fermi_hubbard_1d_unitary(6, J=1.0, U=4.0, t=0.25)
    param, wires, G = open_brickwall(L, 9)
    rng = np.random.default_rng(seed)
    V = haar_matchgates(rng, G)
    v = rng.standard_normal(8 * G)
    return L, W, param, wires, G, V, v


def host_info():
    """The host's CPU as this process sees it: cores online, cores it may run on, the OpenMP
    thread setting and the model name."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0))
    return {"nproc": os.cpu_count(), "affinity": aff, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "model": model}


def cpu_threads():
    """Threads for the CPU baselines: OMP_NUM_THREADS where it is set (the host cores that go with
    one GPU on a shared machine), else every core this process may run on."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(aff, int(env))) if env and env.isdigit() else aff


def _grow(run, ncols0, budget_s, limit, step):
    """run(ncols) -> seconds; grow the column sample to about budget_s seconds."""
    dt = run(ncols0)
    n = int(min(limit, max(ncols0, ncols0 * budget_s / max(dt, 1e-6))))
    n = max(step, (n // step) * step)
    if n == ncols0:
        return n, dt
    return n, run(n)


def cpu_reference_algorithm(W, L, nthreads, budget_s=10.0, seed=0):
    """The reference's own code (oracle/_ref: its src/*.c compiled unchanged) on the L=12
    Fermi-Hubbard target: the per-column work of its OpenMP target loop with the O(G^2)
    placeholder Hessian (src/matchgate_target.c:746-864) on a sample of basis columns, scaled
    to the 4096-column trace; the f-only loop likewise; and the truncated CG of the trust region
    (the oracle's restatement of src/trust_region.c:88-158, the reference's needs CBLAS).  The
    reference's brickwall ties one gate to each layer and applies it to all L/2 pairs, so the
    closest form of the open 50-gate brickwall is 9 tied layers with periodic pairs (54 gate
    instances, 72 parameters)."""
    sys.path.insert(0, ROOT)
    import oracle as orc
    from oracle import ref
    from rqcopt_mg.workloads import haar_matchgates

    if not ref.available():
        return None
    S = 1 << L
    nl = 9
    pm = np.stack([np.arange(L) if k % 2 == 0 else np.roll(np.arange(L), -1) for k in range(nl)]).astype(np.int32)
    V = haar_matchgates(np.random.default_rng(seed), nl)
    nh, th = _grow(lambda n: ref.time_target_columns(W, V, pm, n, nthreads, True)[0], nthreads, budget_s, S, nthreads)
    nf, tf = _grow(lambda n: ref.time_target_columns(W, V, pm, n, nthreads, False)[0], 4 * nthreads, 0.2 * budget_s, S,
                   nthreads)
    m = 8 * nl
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, m))
    Hm = A @ A.T / m + np.eye(m)
    g = rng.standard_normal(m)
    t0 = time.perf_counter()
    orc.truncated_cg_hmat(g, Hm, 0.01, 2 * 16 * nl)
    ttcg = time.perf_counter() - t0
    t_eval = th * S / nh
    t_iter = t_eval + tf * S / nf + ttcg
    return {
        "value": 1.0 / t_iter,
        "unit": "trust-region iterations/s",
        "cores": nthreads,
        "kind": "reference",
        "sample": f"{nh} of {S} basis columns of the f+gradient+Hessian loop and {nf} of the f loop, timed and scaled "
                  "to the full trace, plus one truncated CG",
        "eval_s": t_eval,
        "workload": "L=12 Fermi-Hubbard exp(-iHt) dense target (the same matrix), 9 tied brickwall layers with periodic "
                    "pairs (54 gate instances, 72 parameters); one iteration = f + gradient + full Riemannian Hessian "
                    "(O(G^2) placeholder sweeps) + f at the trial point + truncated CG",
        "gpu_counterpart": "rtr.iters_per_s",
    }


def cpu_hvp_port(W, V, v, param, wires, L, nthreads, budget_s=8.0):
    """The oracle's OpenMP f + gradient + HVP (the GPU's unit) on a bounded sample of basis
    columns, scaled to whole evaluations per second."""
    sys.path.insert(0, ROOT)
    import oracle as orc

    S = 1 << L
    Z = orc.real_to_directions(V, v)

    def run(n):
        t0 = time.perf_counter()
        orc.target_gradient_hvp(W, V, Z, param, wires, L, 0, n, nthreads)
        return time.perf_counter() - t0

    n, dt = _grow(run, nthreads, budget_s, S, nthreads)
    return {"value": (n / S) / dt, "unit": "f+grad+HVP evaluations/s", "cores": nthreads, "kind": "port",
            "sample": f"{n} of {S} basis columns (FP64, OpenMP over columns), scaled to the full trace"}


def cpu_small_config(nthreads):
    """The L=8 CPU configuration: open brickwall of 5 layers (4/3/4/3/4 = 18 independent gates),
    seeded Haar gates (seed 0), target a seeded 21-layer brickwall (74 gates) materialised as a
    dense 256 x 256 matrix, 10 trust-region iterations with the explicit Riemannian Hessian (the
    reference's algorithm), on the oracle (the reference's brickwall API ties gates to layers)."""
    sys.path.insert(0, ROOT)
    import oracle as orc
    from rqcopt_mg.workloads import haar_matchgates, open_brickwall

    L = 8
    rng = np.random.default_rng(0)
    param, wires, G = open_brickwall(L, 5)
    V = haar_matchgates(rng, G)
    tp, tw, Gt = open_brickwall(L, 21)
    Ug = haar_matchgates(rng, Gt)
    S = 1 << L
    U = np.stack([orc.apply_program(Ug, tp, tw, L, np.eye(1, S, b).ravel()) for b in range(S)], axis=1)
    t0 = time.perf_counter()
    f_iter, _, _ = orc.rtr_optimize(U, V, param, wires, L, 10, mode="hmat", nthreads=nthreads)
    dt = time.perf_counter() - t0
    return {"s_for_10_iterations": dt, "iters_per_s": 10 / dt, "cores": nthreads, "kind": "port",
            "f_first": float(f_iter[0]), "f_last": float(f_iter[-1]),
            "workload": "L=8, open brickwall 5 layers (18 gates), target a 21-layer brickwall (74 gates, dense), "
                        "10 trust-region iterations, explicit Hessian"}


def measured_traffic(kernel, dtype):
    """HBM bytes per launch of this kernel on this workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE measurement (calibrated against a device copy), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{'fp64' if dtype == 'f64' else 'fp32'}_L12_hubbard_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)["traffic_bytes"]


def copy_calibration(mib):
    """hipMemcpyDtoD of a known size: its FETCH_SIZE / WRITE_SIZE calibrate the sweep kernel's counters."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0 and hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    for _ in range(5):
        assert hip.hipMemcpyDtoD(b, a, ctypes.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"copy_bytes": n, "copies": 5}))


def device_free_bytes():
    """Free device memory (hipMemGetInfo)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value, total.value


ONE_JOB_BASELINES = {
    "plain": ("baseline_fp64_one_job_plain", "speedup_vs_plain", dict(densify_qubits=0),
              "one job at a time, FP64, one launch per evaluation, circuit targets applied per column, the "
              "reference trust-region control flow"),
    "dense": ("baseline_fp64_one_job_dense_targets", "speedup_vs_dense_targets", dict(densify_qubits=16),
              "one job at a time, FP64, one launch per evaluation, targets materialised at admission, best "
              "single-job kernels, the reference control flow"),
    "reuse": ("baseline_fp64_one_job_dense_targets_tr_reuse", "speedup_vs_dense_targets_tr_reuse",
              dict(densify_qubits=16, tr_reuse=True),
              "one job at a time, FP64, one launch per evaluation, targets materialised, the served trust region "
              "(H eta from the truncated CG, f+grad kept)"),
}


def served_bench(mg, njobs, seed=1, Ls=(8, 10, 12, 14, 16), baselines=("plain",), limit_s=60.0):
    """The served stream: the first njobs of the Poisson stream (200 jobs/s, seed 1) of
    trust-region jobs with L uniform in Ls, odd depth 3..15, budget 5..50 iterations, circuit
    targets of 3x the depth; FP32 sweeps with the last 3 iterations in FP64, continuous batching
    (evaluations that fill the GPU alone in the server's second lane), targets materialised at
    admission, each job's trust region reusing its own products (tr_reuse), admission by device
    memory.  Against the same stream run one job at a time in FP64 (see ONE_JOB_BASELINES), each
    baseline stopped after limit_s seconds: a baseline that does not finish is reported as such
    (jobs done, seconds), with the lower bound that its unfinished run puts on the speedup."""
    from rqcopt_mg.workloads import served_stream

    stream = served_stream(njobs, seed=seed, rate=200.0, Ls=Ls)
    jobs = [{"program": mg.Program.create(j["L"], j["G"], j["param"], j["wires"]), "gates": j["gates"],
             "niter": j["niter"], "target_gates": j["target_gates"], "target_wires": j["target_wires"],
             "arrival": j["arrival"]} for j in stream]
    mix = {}
    for j in stream:
        mix[j["L"]] = mix.get(j["L"], 0) + 1
    # autotune the batched kernel on one HVP per job of the first 16 that it runs (L < 12: the
    # dense larger ones go to the single-request kernels), at their initial gates
    rng = np.random.default_rng(seed)
    calib = []
    for j in [j for j in jobs if j["program"].L < 12][:16]:
        t = mg.Target.circuit(j["target_gates"], j["target_wires"], j["program"].L).densify()
        calib.append({"target": t, "program": j["program"], "gates": j["gates"],
                      "v": rng.standard_normal(8 * j["program"].P), "what": mg.EVAL_F_GRAD_HVP})
    tune = None
    if calib:
        tune = {"fp32": mg.batch_autotune(calib, mg.MG_FP32), "fp64": mg.batch_autotune(calib, mg.MG_FP64)}
        log(f"autotuned (threads, window bits, workgroups/CU, ms): {tune}")
    del calib
    free, total = device_free_bytes()
    budget = max(0.0, free - (24 << 30))
    mg.set_kernel_timing(True)
    dq = max(Ls)
    log(f"serving {njobs} jobs (L: {dict(sorted(mix.items()))}) with continuous batching ...")
    res, st = mg.serve(jobs, batched=True, precision=mg.MG_FP32, precision_switch=3, densify_qubits=dq,
                       realtime=True, tr_reuse=True, progress_s=20.0, hbm_budget=budget)
    assert all(r["status"] == 0 for r in res), [r["status"] for r in res]
    jps = njobs / st["wall_s"]
    log(f"  {jps:.3f} jobs/s ({st['wall_s']:.1f} s), {st['rounds']} launches for {st['requests']} evaluations")
    lat = np.array([r["latency"] for r in res]) * 1e3
    out = {"jobs": njobs, "qubits": {str(k): v for k, v in sorted(mix.items())}, "jobs_per_s": jps,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "latency_ms_per_job": [round(float(x), 1) for x in lat],
           "launches": st["rounds"], "evaluations": st["requests"], "peak_admitted": st["max_active"],
           "peak_device_bytes": st["peak_footprint"], "hbm_budget": budget,
           "kernel_s": st["kernel_ms"] * 1e-3, "wall_s": st["wall_s"], "graph_replays": st["graph_replays"],
           "hvp_sweeps": int(sum(r["nhvp"] for r in res))}
    for name in baselines:
        key, skey, kw, note = ONE_JOB_BASELINES[name]
        log(f"the same stream, {note} (stopped after {limit_s:.0f} s) ...")
        r1, s1 = mg.serve(jobs, batched=False, precision=mg.MG_FP64, realtime=True, progress_s=20.0,
                          time_limit_s=limit_s, **kw)
        ok = [r for r in r1 if r["status"] == 0]
        if len(ok) == njobs:
            l1 = np.array([r["latency"] for r in r1]) * 1e3
            out[key] = {"finished": True, "jobs_per_s": njobs / s1["wall_s"], "wall_s": s1["wall_s"],
                        "p50_ms": float(np.percentile(l1, 50)), "p99_ms": float(np.percentile(l1, 99)),
                        "evaluations": s1["requests"], "hvp_sweeps": int(sum(r["nhvp"] for r in r1)), "note": note}
            out[skey] = jps / out[key]["jobs_per_s"]
        else:
            # it stopped after s1["wall_s"] seconds with jobs left: its wall time exceeds that, so
            # its jobs/s is below njobs / wall_s and the speedup above jps * wall_s / njobs
            out[key] = {"finished": False, "jobs_done": len(ok), "stopped_after_s": s1["wall_s"],
                        "time_limit_s": limit_s, "note": note}
            out[skey + "_lower_bound"] = jps * s1["wall_s"] / njobs
    # error of each job's final evaluation (FP64 sweeps in the last iterations) and of an FP32
    # sweep at the same gates, against the FP64 single-request path (pinned to the oracle in
    # tests) on a column subset of the larger jobs
    err = {"grad_final": 0.0, "hvp_final": 0.0, "grad_fp32_sweep": 0.0, "hvp_fp32_sweep": 0.0}
    for j, r in zip(jobs, res):
        L = j["program"].L
        S = 1 << L
        b0, b1 = (0, S) if L <= 12 else (S // 2, S // 2 + 64)
        t = mg.Target.circuit(j["target_gates"], j["target_wires"], L)
        v = rng.standard_normal(8 * j["program"].P)
        req = {"target": t, "program": j["program"], "gates": r["gates"], "v": v, "what": mg.EVAL_F_GRAD_HVP,
               "b0": b0, "b1": b1}
        o64, o32 = mg.evaluate_batch([dict(req, precision=mg.MG_FP64)])[0], mg.evaluate_batch([dict(req, precision=mg.MG_FP32)])[0]
        dq_t = mg.Target.circuit(j["target_gates"], j["target_wires"], L).densify()
        d64 = mg.evaluate_batch([dict(req, target=dq_t, precision=mg.MG_FP64)])[0]
        gn, hn = max(np.max(np.abs(o64[1])), 1e-300), max(np.max(np.abs(o64[2])), 1e-300)
        err["grad_final"] = max(err["grad_final"], float(np.max(np.abs(d64[1] - o64[1])) / gn))
        err["hvp_final"] = max(err["hvp_final"], float(np.max(np.abs(d64[2] - o64[2])) / hn))
        err["grad_fp32_sweep"] = max(err["grad_fp32_sweep"], float(np.max(np.abs(o32[1] - o64[1])) / gn))
        err["hvp_fp32_sweep"] = max(err["hvp_fp32_sweep"], float(np.max(np.abs(o32[2] - o64[2])) / hn))
        del dq_t
    out.update({"rel_err_vs_fp64": err, "autotune": tune,
                "workload": f"the first {njobs} jobs of the Poisson stream (200 jobs/s, seed 1), L uniform in "
                            f"{{{','.join(map(str, Ls))}}}, odd depth 3..15, 5..50 iterations, circuit targets of 3x "
                            "depth materialised at admission; FP32 sweeps, last 3 iterations FP64; trust region with "
                            "tr_reuse; admission by free device memory",
                "rel_err_note": "final gates of each job: the served (dense-row) FP64 path and an FP32 sweep against "
                                "the FP64 circuit-target path, all basis columns up to L=12, 64 columns beyond"})
    return out


def batched_bench(mg, niter, seed=0, tr_reuse=True):
    """The batched workload: 64 independent instances, L uniform in {8, 10, 12, 14}, odd depth
    3..11, targets seeded random brickwalls of 3x the depth (materialised in HBM at admission),
    FP64; every instance runs `niter` trust-region iterations (with tr_reuse: H eta from the
    truncated CG, f with its gradient at trial points) and all of them step together, each round
    of evaluations one persistent launch per precision over all instances and basis columns,
    repeated launch shapes replayed from captured HIP graphs."""
    from rqcopt_mg.workloads import batched_instances

    inst = batched_instances(64, seed=seed, niter=niter)
    jobs = [{"program": mg.Program.create(j["L"], j["G"], j["param"], j["wires"]), "gates": j["gates"],
             "niter": j["niter"], "target_gates": j["target_gates"], "target_wires": j["target_wires"],
             "arrival": 0.0} for j in inst]
    mg.set_kernel_timing(True)
    log(f"batched workload: 64 instances, {niter} trust-region iterations each (FP64) ...")
    res, st = mg.serve(jobs, batched=True, precision=mg.MG_FP64, precision_switch=0, densify_qubits=14,
                       realtime=False, progress_s=30.0, tr_reuse=tr_reuse)
    iters = sum(len(r["f_iter"]) - 1 for r in res)
    log(f"  {iters / st['wall_s']:.1f} trust-region iterations/s, {st['rounds']} launches, "
        f"{st['graph_replays']} from graphs")
    # final gradients of the smaller instances against the FP64 single-request path
    err = 0.0
    rng = np.random.default_rng(seed)
    for j, r in list(zip(jobs, res))[:8]:
        t = mg.Target.circuit(j["target_gates"], j["target_wires"], j["program"].L)
        v = rng.standard_normal(8 * j["program"].P)
        _, g0, _ = mg.riemannian_gradient_hvp(t, j["program"], r["gates"], v, mg.MG_FP64)
        o = mg.evaluate_batch([{"target": t, "program": j["program"], "gates": r["gates"], "v": v,
                                "what": mg.EVAL_F_GRAD_HVP, "precision": mg.MG_FP64}])[0]
        err = max(err, float(np.max(np.abs(o[1] - g0)) / max(np.max(np.abs(g0)), 1e-300)))
    return {"instances": 64, "iterations_per_instance": niter, "tr_iters_per_s": iters / st["wall_s"],
            "evaluations_per_s": st["requests"] / st["wall_s"], "launches": st["rounds"],
            "evaluations": st["requests"], "graph_replays": st["graph_replays"], "kernel_s": st["kernel_ms"] * 1e-3,
            "wall_s": st["wall_s"], "grad_rel_err_vs_fp64_single_path": err,
            "hvp_sweeps": int(sum(r["nhvp"] for r in res)),
            "workload": "64 instances, L uniform in {8,10,12,14}, odd depth 3..11, targets random brickwalls of 3x "
                        "depth (materialised at admission), FP64, seed 0" +
                        ("" if niter == 20 else f"; {niter} of the 20 iterations per instance") +
                        ("; trust region with tr_reuse" if tr_reuse else "; reference trust-region control flow")}


def large_bench(mg, L=16):
    """The large single instance: L=16, open brickwall depth 11 (83 gates), target = 50
    second-order Trotter steps of the XXZ chain with random fields (758 gates); one f + gradient
    + HVP over all 65536 columns with FP32 sweeps and with FP64 sweeps, first with the target
    given as its gates (its rows materialised one 4 GiB column chunk at a time next to the tiled
    sweeps), then with the target materialised once in HBM (69 GB of FP64 rows) and read as dense
    rows by the same tiled kernels.  L=14 runs the same construction at 14 qubits."""
    from rqcopt_mg.workloads import large_instance, sweep_flops

    w = large_instance(L=L)
    prog = mg.Program.create(w["L"], w["G"], w["param"], w["wires"])
    tgt = mg.Target.circuit(w["target_gates"], w["target_wires"], w["L"])
    v = np.random.default_rng(0).standard_normal(8 * w["G"])
    fl = sweep_flops(w["L"], w["G"], w["Gt"])
    out = {"workload": f"L={L} XXZ Trotter-2 target (50 steps, t=1, {w['Gt']} gates), open brickwall depth 11 "
                       f"({w['G']} gates), {1 << L} columns", "flops": fl}
    res = {}
    for name, prec in (("fp32", mg.MG_FP32), ("fp64", mg.MG_FP64)):
        log(f"L={L} f+grad+HVP, {name} sweeps ...")
        req = {"target": tgt, "program": prog, "gates": w["gates"], "v": v, "what": mg.EVAL_F_GRAD_HVP, "precision": prec}
        t0 = time.perf_counter()
        res[name] = mg.evaluate_batch([req])[0]
        dt = time.perf_counter() - t0
        km = mg.batch_last_kernel_ms()
        out[name] = {"s": dt, "kernel_ms": km, "tflops": fl / (km * 1e-3) / 1e12,
                     "kernels": ("materialize_mfma_kernel<float> (FP32 target rows of a column chunk, FP32 MFMA) + "
                                 "sweep_tile_kernel<float,2>") if prec == mg.MG_FP32 else
                                ("materialize_mfma_kernel<double> (FP64 target rows of a column chunk, FP64 MFMA) + "
                                 "sweep_win_tile_kernel<2,2>")}
        log(f"  {dt:.2f} s, kernel {km:.1f} ms, {out[name]['tflops']:.2f} TFLOP/s")
    # the target's rows in HBM: first with the register-tile kernel (timed, then freed), then on
    # the matrix cores (the default), which the evaluations below read
    log("materialising the target in HBM (FP64 rows): register tiles, then matrix cores ...")
    mg.set_materialize_kernel(0)
    t_reg = mg.Target.circuit(w["target_gates"], w["target_wires"], w["L"])
    t0 = time.perf_counter()
    t_reg.densify()
    out["densify_register_tiles_s"] = time.perf_counter() - t0
    del t_reg
    mg.set_materialize_kernel(-1)
    t0 = time.perf_counter()
    tgt.densify()
    out["densify_s"] = time.perf_counter() - t0
    out["densify_flops"] = 8.0 * 4 ** w["L"] * 2 * w["Gt"]
    out["densify_kernel"] = "materialize_mfma_kernel (4-qubit windows, two 8x8 parity blocks, FP64 MFMA)"
    # algorithmic FLOPs (2 complex MACs per amplitude and target gate) over the wall time of the
    # materialisation (one launch), against the FP64 matrix peak
    ach = out["densify_flops"] / out["densify_s"] / 1e12
    out["densify_roofline"] = {"bound": "mfma_fp64", "achieved": ach, "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                               "frac": ach / PEAK_TFLOPS["f64"]}
    log(f"  register tiles {out['densify_register_tiles_s']:.2f} s, matrix cores {out['densify_s']:.2f} s")
    fl_dense = sweep_flops(w["L"], w["G"], 0)
    for name, prec in (("fp32", mg.MG_FP32), ("fp64", mg.MG_FP64)):
        req = {"target": tgt, "program": prog, "gates": w["gates"], "v": v, "what": mg.EVAL_F_GRAD_HVP, "precision": prec}
        t0 = time.perf_counter()
        r = mg.evaluate_batch([req])[0]
        dt = time.perf_counter() - t0
        km = mg.batch_last_kernel_ms()
        out[name + "_dense_target"] = {"s": dt, "kernel_ms": km, "tflops": fl_dense / (km * 1e-3) / 1e12,
                                       "rel_err_vs_in_kernel_target": float(np.max(np.abs(r[1] - res[name][1])) /
                                                                            np.max(np.abs(res[name][1])))}
        log(f"  dense target, {name}: {dt:.2f} s, kernel {km:.1f} ms")
    f32, g32, h32 = res["fp32"]
    f64, g64, h64 = res["fp64"]
    out["rel_err_fp32_vs_fp64"] = {"f": float(abs(f32 - f64) / abs(f64)),
                                   "grad": float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))),
                                   "hvp": float(np.max(np.abs(h32 - h64)) / np.max(np.abs(h64)))}
    out["f"] = [f64.real, f64.imag]
    return out


def time_steps(mg, t, p, V, v, prec, steps, warmup):
    for _ in range(warmup):
        mg.riemannian_gradient_hvp(t, p, V, v, prec)
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        out = mg.riemannian_gradient_hvp(t, p, V, v, prec)
        kms.append(mg.last_kernel_ms())
    dt = time.perf_counter() - t0
    return out, dt, float(np.mean(kms)), mg.last_sweep_info()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--rtr-iters", type=int, default=50, help="full trust-region iterations timed for the extra field")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-sweep variant")
    ap.add_argument("--serve-jobs", type=int, default=6,
                    help="jobs of the served stream at its stated qubit counts (0 skips it); the first 6 (L = 8, "
                         "14, 14, 14, 16, 16) let the one-job FP64 baseline finish within the bench's time")
    ap.add_argument("--serve-Ls", default="8,10,12,14,16", help="qubit counts of the served stream's jobs")
    ap.add_argument("--serve-baselines", default="plain",
                    help="one-job-at-a-time FP64 runs of the same stream to compare with (subset of plain,dense,reuse)")
    ap.add_argument("--serve-baseline-limit-s", type=float, default=300.0,
                    help="each one-job baseline of the served stream stops after this many seconds")
    ap.add_argument("--serve-small-jobs", type=int, default=32,
                    help="jobs of a second served stream with L in {8,10,12}, compared with all three one-job "
                         "baselines run to the end (0 skips it)")
    ap.add_argument("--no-large", action="store_true", help="skip the L=16 instance (about 30 s of GPU time)")
    ap.add_argument("--batched-reference-flow", action="store_true",
                    help="batched workload with the reference trust-region control flow instead of tr_reuse")
    ap.add_argument("--batched-iters", type=int, default=20,
                    help="trust-region iterations per instance of the 64-instance batched workload (20 is the "
                         "full workload; 0 skips it)")
    ap.add_argument("--only-served", action="store_true",
                    help="only the served stream (after the headline steps), e.g. with a long baseline limit")
    ap.add_argument("--only-large", type=int, default=0, metavar="L",
                    help="only the large instance at L qubits (14 or 16) after the headline steps (profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.only_large:
        args.only_served, args.serve_jobs = True, 0
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the column split across GPUs is a design note, "
              "not a measured configuration", file=sys.stderr)
        sys.exit(2)

    import rqcopt_mg as mg
    from rqcopt_mg.workloads import sweep_flops

    log("building the L=12 Fermi-Hubbard target on the host ...")
    L, W, param, wires, G, V, v = build_workload()
    prec = mg.MG_FP64 if args.precision == "fp64" else mg.MG_FP32
    dtype = "f64" if prec == mg.MG_FP64 else "f32"
    t = mg.Target.dense(W)
    p = mg.Program.create(L, G, param, wires)
    mg.set_kernel_timing(True)

    log(f"{args.warmup} warmup + {args.steps} timed steps ({args.precision}) ...")
    (f, g, hv), dt, kernel_ms, info = time_steps(mg, t, p, V, v, prec, args.steps, args.warmup)
    ms_per_step = 1e3 * dt / args.steps
    value = args.steps / dt
    log(f"{ms_per_step:.3f} ms/step, sweep kernel {kernel_ms:.3f} ms, launch {info}")

    flops = sweep_flops(L, G)
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    roofline = {"bound": "valu_" + ("fp64" if dtype == "f64" else "fp32"), "achieved": achieved,
                "peak": PEAK_TFLOPS[dtype], "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS[dtype], "traffic": measured_traffic(info["kernel"], dtype),
                "kernel": info["kernel"],
                "kernel_ms": kernel_ms, "flops_per_launch": flops,
                "note": "the reverse sweep (most of the kernel) is vector-ALU FMA work, the forward pass runs on the "
                        "matrix cores over 4-qubit windows; peak = MI355X FP64 vector rate (equal to its FP64 "
                        "matrix rate) or FP32 vector rate; flops counted per gate as the VALU form; traffic = HBM "
                        "bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/"}

    extra = {}
    if prec == mg.MG_FP64 and not args.no_fp32 and not args.only_served:
        log("FP32-sweep variant ...")
        (f32, g32, hv32), dt32, k32, info32 = time_steps(mg, t, p, V, v, mg.MG_FP32, args.steps, args.warmup)
        a32 = flops / (k32 * 1e-3) / 1e12
        extra["fp32_sweeps"] = {
            "value": args.steps / dt32, "ms_per_step": 1e3 * dt32 / args.steps, "kernel_ms": k32,
            "achieved_tflops": a32, "frac_of_157.3": a32 / PEAK_TFLOPS["f32"],
            "rel_err_vs_fp64": {"f": float(abs(f32 - f) / abs(f)),
                                "grad": float(np.max(np.abs(g32 - g)) / np.max(np.abs(g))),
                                "hvp": float(np.max(np.abs(hv32 - hv)) / np.max(np.abs(hv)))},
            "launch": info32}
        log(f"FP32 sweeps: {1e3 * dt32 / args.steps:.3f} ms/step, kernel {k32:.3f} ms")

    # full trust-region iterations (f+grad, tCG HVPs, f at the candidate): FP64 with the
    # reference control flow, and FP32 sweeps with the last 3 iterations in FP64 and the served
    # trust region (tr_reuse), one job, one launch per evaluation
    if args.rtr_iters > 0 and not args.only_served:
        t0 = time.perf_counter()
        f_iter, _, _, nh = mg.optimize(t, p, V, args.rtr_iters, hvp=True, precision=prec)
        rt = time.perf_counter() - t0
        extra["rtr"] = {"iters": args.rtr_iters, "iters_per_s": args.rtr_iters / rt, "hvp_sweeps": nh,
                        "graph_launches": mg.last_trust_region_graphs(),
                        "f_first": float(f_iter[0]), "f_last": float(f_iter[-1]),
                        "precision": "fp64 sweeps, reference control flow; trust region and truncated CG on the "
                                     "device, one HIP graph launch per iteration"}
        # the job names the resident target (as the FP64 run above uses it): no host-side target
        # setup inside the timed call
        job = {"program": p, "gates": V, "niter": args.rtr_iters, "target": t, "arrival": 0.0}
        r, st = mg.serve([job], batched=False, precision=mg.MG_FP32, precision_switch=3, realtime=False,
                         tr_reuse=True)
        extra["rtr_fp32_sweeps"] = {
            "iters": args.rtr_iters, "iters_per_s": args.rtr_iters / st["wall_s"], "hvp_sweeps": r[0]["nhvp"],
            "evaluations": st["requests"], "f_first": float(r[0]["f_iter"][0]), "f_last": float(r[0]["f_iter"][-1]),
            "f_last_rel_diff_vs_fp64": float(abs(r[0]["f_iter"][-1] - f_iter[-1]) / abs(f_iter[-1])),
            "precision": "fp32 sweeps with fp64 accumulation, last 3 iterations fp64; tr_reuse"}
        # the device trust region with reuse (H eta from the truncated CG, f + gradient at the trial
        # point kept): FP64, and FP32 sweeps with the last 3 iterations in FP64
        mg.set_device_trust_region(2)
        for key, pr, sw in (("rtr_reuse", mg.MG_FP64, 0), ("rtr_fp32_reuse", mg.MG_FP32, 3)):
            t0 = time.perf_counter()
            fr, _, _, nhr = mg.optimize(t, p, V, args.rtr_iters, hvp=True, precision=pr, precision_switch=sw)
            rt = time.perf_counter() - t0
            extra[key] = {"iters": args.rtr_iters, "iters_per_s": args.rtr_iters / rt, "hvp_sweeps": nhr,
                          "graph_launches": mg.last_trust_region_graphs(), "f_last": float(fr[-1]),
                          "f_last_rel_diff_vs_fp64": float(abs(fr[-1] - f_iter[-1]) / abs(f_iter[-1])),
                          "precision": ("fp64 sweeps" if pr == mg.MG_FP64 else
                                        "fp32 sweeps with fp64 accumulation, last 3 iterations fp64") +
                                       "; device trust region with reuse, one HIP graph launch per iteration"}
        mg.set_device_trust_region(-1)
        log(f"trust region: {extra['rtr']['iters_per_s']:.1f} it/s FP64, "
            f"{extra['rtr_reuse']['iters_per_s']:.1f} it/s FP64 device reuse, "
            f"{extra['rtr_fp32_reuse']['iters_per_s']:.1f} it/s FP32 device reuse, "
            f"{extra['rtr_fp32_sweeps']['iters_per_s']:.1f} it/s FP32 sweeps + tr_reuse (served)")

    if args.only_large:
        mg.set_kernel_timing(True)
        extra["large_L%d" % args.only_large] = large_bench(mg, args.only_large)
    if args.serve_jobs > 0:
        extra["served"] = served_bench(mg, args.serve_jobs, Ls=tuple(int(x) for x in args.serve_Ls.split(",")),
                                       baselines=tuple(x for x in args.serve_baselines.split(",") if x),
                                       limit_s=args.serve_baseline_limit_s)
    if args.serve_small_jobs > 0 and not args.only_served:
        extra["served_L8_to_12"] = served_bench(mg, args.serve_small_jobs, Ls=(8, 10, 12),
                                                baselines=("plain", "dense", "reuse"), limit_s=120.0)
    if args.batched_iters > 0 and not args.only_served:
        extra["batched"] = batched_bench(mg, args.batched_iters, tr_reuse=not args.batched_reference_flow)
    if not args.no_large and not args.only_served:
        mg.set_kernel_timing(True)
        extra["large_L16"] = large_bench(mg)

    cpu = None
    if not args.no_cpu_baseline and not args.only_served:
        nt = cpu_threads()
        log(f"CPU baselines on {nt} threads: the reference's own code (bounded column samples), the oracle's HVP, "
            "the L=8 configuration ...")
        hvp_port = cpu_hvp_port(W, V, v, param, wires, L, nt)
        cpu = cpu_reference_algorithm(W, L, nt) or hvp_port
        if cpu is not hvp_port:
            cpu["hvp_port"] = hvp_port
        cpu["host"] = host_info()
        extra["cpu_L8"] = cpu_small_config(nt)

    out = {
        "metric": "trust-region iters/s (f+grad+HVP) at L=12; served jobs/s + p50/p99 latency",
        "value": value,
        "unit": "f+grad+HVP evaluations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded Haar gates; Fermi-Hubbard target built on the host)",
        "config": {"workload": "L=12 Fermi-Hubbard exp(-iHt) t=0.25 J=1 U=4, open brickwall depth 9 (50 gates), "
                               "4096 basis columns", "L": L, "ngates": G, "columns": 1 << L,
                   "precision": args.precision, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "f": [f.real, f.imag],
        **extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
