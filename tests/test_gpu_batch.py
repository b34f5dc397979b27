"""Batched evaluation and serving on the GPU.

- mg_evaluate_batch: every request of a mixed batch (qubit counts 6-12, circuit and
  dense targets, f / f+grad / f+grad+HVP, FP64 and FP32) against the oracle (FP64:
  1e-11 relative; FP32 sweeps: 2e-5) and against the single-request path;
- a request gives the same bits alone and in any mix (the batched kernel's summation
  order depends only on the request's own size);
- mg_serve: continuous batching of trust-region jobs reproduces each job run alone
  bitwise, and the one-job-at-a-time path within 1e-9 (different kernels); with tr_reuse
  the trajectories stay within 1e-9 of the reference control flow with one HVP fewer per
  iteration.
"""
import sys

import numpy as np
import pytest

import oracle as orc
from dense_ref import program_matrix

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mg():
    import rqcopt_mg

    return rqcopt_mg


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def make_instance(mg, rng, L, depth, dense=False):
    from rqcopt_mg.workloads import brickwall_job

    job = brickwall_job(rng, L, depth, 5)
    prog = mg.Program.create(L, job["G"], job["param"], job["wires"])
    if dense:
        W = program_matrix(job["target_gates"], np.arange(job["Gt"]), job["target_wires"], L)
        tgt = mg.Target.dense(W)
    else:
        tgt = mg.Target.circuit(job["target_gates"], job["target_wires"], L)
    v = rng.standard_normal(8 * job["G"])
    return job, prog, tgt, v


def oracle_riemannian(job, v, L):
    """f, Riemannian gradient and HVP from the oracle (dense target, FP64)."""
    W = program_matrix(job["target_gates"], np.arange(job["Gt"]), job["target_wires"], L)
    V = job["gates"]
    Z = orc.real_to_directions(V, v)
    f, dg, dh = orc.target_gradient_hvp(W, V, Z, job["param"], job["wires"], L)
    return f, orc.riemannian_gradient(V, dg), orc.riemannian_hvp(V, dg, Z, dh)


MIX = [(6, 3, False), (8, 5, True), (8, 3, False), (10, 4, False), (6, 5, True), (12, 3, False)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_matches_oracle_and_single_path(mg, prec):
    rng = np.random.default_rng(11)
    P = mg.MG_FP64 if prec == "fp64" else mg.MG_FP32
    tol = 1e-11 if prec == "fp64" else 2e-5
    inst = [make_instance(mg, rng, L, d, dense) for L, d, dense in MIX]
    reqs = []
    for k, (job, prog, tgt, v) in enumerate(inst):
        what = [mg.EVAL_F_GRAD_HVP, mg.EVAL_F_GRAD, mg.EVAL_F][k % 3] if k < 3 else mg.EVAL_F_GRAD_HVP
        reqs.append({"target": tgt, "program": prog, "gates": job["gates"], "v": v, "what": what, "precision": P})
    say(f"case: batch of {len(reqs)} requests, {prec}, L {[m[0] for m in MIX]}")
    out = mg.evaluate_batch(reqs)
    for (job, prog, tgt, v), r, (f, g, hv) in zip(inst, reqs, out):
        L = job["L"]
        fs, gs, hs = mg.riemannian_gradient_hvp(tgt, prog, job["gates"], v, P)
        if L <= 10:
            fo, go, ho = oracle_riemannian(job, v, L)
            assert abs(f - fo) <= tol * abs(fo), (L, f, fo)
            if r["what"] >= mg.EVAL_F_GRAD:
                assert rel(g, go) < tol, (L, rel(g, go))
            if r["what"] == mg.EVAL_F_GRAD_HVP:
                assert rel(hv, ho) < tol, (L, rel(hv, ho))
        # the single-request path (different kernels): same values to rounding
        tol_s = 1e-12 if prec == "fp64" else 2e-5
        assert abs(f - fs) <= tol_s * abs(fs)
        if r["what"] >= mg.EVAL_F_GRAD:
            assert rel(g, gs) < tol_s
        if r["what"] == mg.EVAL_F_GRAD_HVP:
            assert rel(hv, hs) < tol_s


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_alone_equals_mixed_bitwise(mg, prec):
    rng = np.random.default_rng(12)
    P = mg.MG_FP64 if prec == "fp64" else mg.MG_FP32
    inst = [make_instance(mg, rng, L, d, dense) for L, d, dense in MIX]
    reqs = [{"target": t, "program": p, "gates": j["gates"], "v": v, "what": mg.EVAL_F_GRAD_HVP, "precision": P}
            for j, p, t, v in inst]
    say(f"case: alone vs mixed, {prec}")
    mixed = mg.evaluate_batch(reqs)
    mixed_rev = mg.evaluate_batch(reqs[::-1])[::-1]
    for k, r in enumerate(reqs):
        alone = mg.evaluate_batch([r])[0]
        for a, b, c in zip(alone, mixed[k], mixed_rev[k]):
            assert np.array_equal(np.asarray(a), np.asarray(b)) and np.array_equal(np.asarray(a), np.asarray(c)), k


def test_batch_global_states_and_shapes(mg):
    """FP64 at L=12 keeps its states in global scratch; other column-group widths and
    item sizes change only the summation order."""
    rng = np.random.default_rng(13)
    job, prog, tgt, v = make_instance(mg, rng, 12, 3)
    jb, pb, tb, vb = make_instance(mg, rng, 8, 4)
    say("case: batch, L=12 FP64 global states + L=8 with non-default shapes")
    base = mg.evaluate_batch([
        {"target": tgt, "program": prog, "gates": job["gates"], "v": v, "what": 2, "precision": mg.MG_FP64},
        {"target": tb, "program": pb, "gates": jb["gates"], "v": vb, "what": 2, "precision": mg.MG_FP64}])
    fs, gs, hs = mg.riemannian_gradient_hvp(tgt, prog, job["gates"], v, mg.MG_FP64)
    assert abs(base[0][0] - fs) <= 1e-12 * abs(fs) and rel(base[0][1], gs) < 1e-12 and rel(base[0][2], hs) < 1e-12
    try:
        # (threads per column, columns per item): 16 and 8 threads per column put 16 and
        # 32 columns of L=8 FP64 states into one workgroup, beyond the LDS (global scratch)
        for cgw, cpi in [(16, 16), (8, 40), (32, 64), (4, 3)]:
            mg.batch_set_shape(mg.MG_FP64, 8, cgw, cpi)
            o = mg.evaluate_batch([{"target": tb, "program": pb, "gates": jb["gates"], "v": vb, "what": 2,
                                    "precision": mg.MG_FP64}])[0]
            assert abs(o[0] - base[1][0]) <= 1e-12 * abs(base[1][0])
            assert rel(o[1], base[1][1]) < 1e-12 and rel(o[2], base[1][2]) < 1e-12
    finally:
        mg.batch_set_shape(mg.MG_FP64, 8, 0, 0)


def test_densified_target_matches_program(mg):
    rng = np.random.default_rng(14)
    job, prog, tgt, v = make_instance(mg, rng, 10, 5)
    say("case: densified circuit target, L=10, single path and batched")
    ref = mg.riemannian_gradient_hvp(tgt, prog, job["gates"], v, mg.MG_FP64)
    dense = mg.Target.circuit(job["target_gates"], job["target_wires"], 10).densify()
    for out in (mg.riemannian_gradient_hvp(dense, prog, job["gates"], v, mg.MG_FP64),
                mg.evaluate_batch([{"target": dense, "program": prog, "gates": job["gates"], "v": v, "what": 2,
                                    "precision": mg.MG_FP64}])[0]):
        assert abs(out[0] - ref[0]) <= 1e-12 * abs(ref[0])
        assert rel(out[1], ref[1]) < 1e-12 and rel(out[2], ref[2]) < 1e-12
    W = program_matrix(job["target_gates"], np.arange(job["Gt"]), job["target_wires"], 10)
    fo, go, ho = oracle_riemannian(job, v, 10)
    assert abs(ref[0] - fo) <= 1e-11 * abs(fo) and rel(ref[1], go) < 1e-11 and rel(ref[2], ho) < 1e-11


@pytest.mark.parametrize("L", [13, 14])
def test_densify_beyond_lds_matches_circuit_target(mg, L):
    """Targets of L > 12 qubits are materialised on the matrix cores (4-qubit windows as two
    8 x 8 parity blocks, FP64 MFMA, in 12-qubit LDS tiles) by default; the alternatives (12-qubit
    register tiles, each row its own column's state between tile blocks; 3-qubit windows applied
    in place in the rows) must give the same rows.  Checked through f, gradient and HVP over all
    2^L columns (every row enters), FP64, against the circuit target applied per column."""
    rng = np.random.default_rng(40 + L)
    job, prog, tgt, v = make_instance(mg, rng, L, 3)
    say(f"case: densify L={L} ({job['Gt']} target gates): matrix cores vs register tiles vs in-place windows "
        "vs circuit target")
    ref = mg.riemannian_gradient_hvp(tgt, prog, job["gates"], v, mg.MG_FP64)
    outs = []
    try:
        for win, mat in ((-1, -1), (-1, 0), (0, -1)):
            mg.set_window_kernel(win)
            mg.set_materialize_kernel(mat)
            dense = mg.Target.circuit(job["target_gates"], job["target_wires"], L).densify()
            mg.set_window_kernel(-1)
            mg.set_materialize_kernel(-1)
            outs.append(mg.riemannian_gradient_hvp(dense, prog, job["gates"], v, mg.MG_FP64))
            del dense
    finally:
        mg.set_window_kernel(-1)
        mg.set_materialize_kernel(-1)
    for out in outs:
        assert abs(out[0] - ref[0]) <= 1e-12 * abs(ref[0])
        assert rel(out[1], ref[1]) < 1e-12 and rel(out[2], ref[2]) < 1e-12
    assert abs(outs[0][0] - outs[1][0]) <= 1e-13 * abs(ref[0]) and abs(outs[0][0] - outs[2][0]) <= 1e-13 * abs(ref[0])


def _jobs(mg, seed, spec):
    from rqcopt_mg.workloads import brickwall_job

    rng = np.random.default_rng(seed)
    out = []
    for k, (L, d, it) in enumerate(spec):
        j = brickwall_job(rng, L, d, it)
        out.append({"program": mg.Program.create(L, j["G"], j["param"], j["wires"]), "gates": j["gates"],
                    "niter": it, "target_gates": j["target_gates"], "target_wires": j["target_wires"],
                    "arrival": 0.002 * k, "program_wires": j["wires"]})
    return out


SERVE_SPEC = [(6, 3, 6), (8, 5, 4), (8, 3, 7), (10, 3, 3), (6, 5, 5)]


def test_serve_batched_equals_alone_and_single_path(mg):
    jobs = _jobs(mg, 21, SERVE_SPEC)
    say("case: serve 5 jobs batched (FP32 sweeps, last 3 iterations FP64)")
    res, st = mg.serve(jobs, batched=True, precision=mg.MG_FP32, precision_switch=3, realtime=False)
    assert all(r["status"] == 0 for r in res)
    assert st["max_active"] == len(jobs) and st["rounds"] < st["requests"]
    say("case: the same jobs with real-time arrivals")
    rt, _ = mg.serve(jobs, batched=True, precision=mg.MG_FP32, precision_switch=3, realtime=True)
    for a, b in zip(rt, res):
        assert np.array_equal(a["f_iter"], b["f_iter"]) and np.array_equal(a["gates"], b["gates"])
    for k, j in enumerate(jobs):
        say(f"case: serve job {k} alone")
        alone, _ = mg.serve([dict(j, arrival=0.0)], batched=True, precision=mg.MG_FP32, precision_switch=3)
        assert np.array_equal(alone[0]["f_iter"], res[k]["f_iter"]), k
        assert np.array_equal(alone[0]["gates"], res[k]["gates"]), k
    say("case: serve FP64 batched vs one job at a time")
    b64, _ = mg.serve(jobs, batched=True, precision=mg.MG_FP64, realtime=False)
    s64, st1 = mg.serve(jobs, batched=False, precision=mg.MG_FP64, realtime=False)
    assert st1["rounds"] == st1["requests"]
    for a, b in zip(b64, s64):
        assert np.max(np.abs(a["f_iter"] - b["f_iter"])) < 1e-9
        assert np.max(np.abs(a["gates"] - b["gates"])) < 1e-9
    # the FP32-sweep trajectories end on FP64 iterations close to the FP64 ones
    for a, b in zip(res, b64):
        assert abs(a["f_iter"][-1] - b["f_iter"][-1]) < 1e-3 * abs(b["f_iter"][-1])
    say("case: serve FP64 batched with targets materialised at admission")
    d64, _ = mg.serve(jobs, batched=True, precision=mg.MG_FP64, realtime=False, densify_qubits=10)
    for a, b in zip(d64, b64):
        assert np.max(np.abs(a["f_iter"] - b["f_iter"])) < 1e-9 and np.max(np.abs(a["gates"] - b["gates"])) < 1e-9


def test_serve_jobs_sharing_a_resident_target(mg):
    """Jobs that name a resident target (mg_job.target) run exactly as jobs that hand over the
    same dense matrix: the same bits batched and one job at a time, on the batched kernel (L=8)
    and on the single-request kernels (L=12); the target is the caller's, so it adds nothing to
    the jobs' footprint and stays usable after the call."""
    from rqcopt_mg.workloads import brickwall_job

    rng = np.random.default_rng(29)
    for L, specs in ((8, [(3, 5), (5, 4)]), (12, [(3, 3), (5, 2)])):
        jobs = []
        W = None
        for d, it in specs:
            j = brickwall_job(rng, L, d, it)
            if W is None:
                W = program_matrix(j["target_gates"], np.arange(j["Gt"]), j["target_wires"], L)
            jobs.append({"program": mg.Program.create(L, j["G"], j["param"], j["wires"]), "gates": j["gates"],
                         "niter": it, "arrival": 0.0})
        t = mg.Target.dense(W)
        with_u = [dict(j, U=W) for j in jobs]
        with_t = [dict(j, target=t) for j in jobs]
        for batched in (True, False):
            say(f"case: L={L} serve 2 jobs batched={batched}, dense U vs a shared resident target")
            ru, su = mg.serve(with_u, batched=batched, precision=mg.MG_FP32, precision_switch=2, realtime=False,
                              tr_reuse=True)
            rt, st = mg.serve(with_t, batched=batched, precision=mg.MG_FP32, precision_switch=2, realtime=False,
                              tr_reuse=True)
            for a, b in zip(rt, ru):
                assert a["status"] == 0 and b["status"] == 0
                assert np.array_equal(a["f_iter"], b["f_iter"]) and np.array_equal(a["gates"], b["gates"])
            assert st["peak_footprint"] < su["peak_footprint"] - 0.5 * W.size * 16
        f, _, _ = mg.riemannian_gradient_hvp(t, jobs[0]["program"], jobs[0]["gates"], None, mg.MG_FP64)
        assert np.isfinite(f)


def test_serve_tr_reuse_matches_reference_control_flow(mg):
    """tr_reuse: H eta of the ratio test summed from the truncated CG's products, f + gradient
    kept across rejected steps.  Same trajectories as the reference control flow (1e-9, FP64),
    one HVP fewer per iteration, and still the same bits alone and in the mix."""
    jobs = _jobs(mg, 23, SERVE_SPEC)
    say("case: serve FP64 batched, reference control flow vs tr_reuse")
    ref, st0 = mg.serve(jobs, batched=True, precision=mg.MG_FP64, realtime=False)
    reu, st1 = mg.serve(jobs, batched=True, precision=mg.MG_FP64, realtime=False, tr_reuse=True)
    niter = sum(int(j["niter"]) for j in jobs)
    for a, b in zip(reu, ref):
        assert a["status"] == 0
        assert np.max(np.abs(a["f_iter"] - b["f_iter"])) < 1e-9
        assert np.max(np.abs(a["gates"] - b["gates"])) < 1e-9
        assert a["nhvp"] <= b["nhvp"]
    saved_hvp = sum(b["nhvp"] for b in ref) - sum(a["nhvp"] for a in reu)
    assert saved_hvp >= niter - 2, (saved_hvp, niter)
    assert st0["requests"] - st1["requests"] >= saved_hvp
    say("case: tr_reuse one job at a time vs batched, and a job alone vs in the mix")
    one, _ = mg.serve(jobs, batched=False, precision=mg.MG_FP64, realtime=False, tr_reuse=True)
    for a, b in zip(one, reu):
        assert np.max(np.abs(a["f_iter"] - b["f_iter"])) < 1e-9
    mix, _ = mg.serve(jobs, batched=True, precision=mg.MG_FP32, precision_switch=3, realtime=False, tr_reuse=True)
    alone, _ = mg.serve([jobs[1]], batched=True, precision=mg.MG_FP32, precision_switch=3, realtime=False,
                        tr_reuse=True)
    assert np.array_equal(alone[0]["f_iter"], mix[1]["f_iter"]) and np.array_equal(alone[0]["gates"], mix[1]["gates"])


def test_serve_mixed_sizes_with_large_jobs(mg):
    """A served stream with L=12 and L=14 jobs next to small ones, targets materialised at
    admission, FP32 sweeps with the last 3 iterations FP64, tr_reuse (the served stream's
    settings).  The L=12 and L=14 jobs fill the GPU alone and run on the single-request kernels in
    the server's second lane.  Every job's trajectory is the same bits served alone; at each
    job's final gates the served path's FP64 f, gradient and HVP (dense rows, the kernels the
    server ran) match the oracle on a column subset (1e-11), and the last f of the trajectory
    matches the oracle's full trace where that is cheap (L <= 12)."""
    spec = [(12, 3, 4), (8, 5, 4), (14, 3, 3), (10, 3, 3), (14, 5, 2), (6, 5, 5)]
    jobs = _jobs(mg, 24, spec)
    kw = dict(batched=True, precision=mg.MG_FP32, precision_switch=3, densify_qubits=14, tr_reuse=True)
    say("case: serve 6 jobs, L in {6,8,10,12,14,14}")
    res, st = mg.serve(jobs, realtime=True, **kw)
    assert all(r["status"] == 0 for r in res)
    for k, j in enumerate(jobs):
        say(f"case: serve job {k} (L={j['program'].L}) alone")
        alone, _ = mg.serve([dict(j, arrival=0.0)], realtime=False, **kw)
        assert np.array_equal(alone[0]["f_iter"], res[k]["f_iter"]), k
        assert np.array_equal(alone[0]["gates"], res[k]["gates"]), k
    rng = np.random.default_rng(25)
    for k, (j, r) in enumerate(zip(jobs, res)):
        L, prog = j["program"].L, j["program"]
        say(f"case: job {k} (L={L}) final gates against the oracle")
        V = r["gates"]
        v = rng.standard_normal(8 * prog.P)
        t = mg.Target.circuit(j["target_gates"], j["target_wires"], L).densify()
        tp = np.arange(len(j["target_gates"]), dtype=np.int32)
        param, wires = np.arange(prog.G, dtype=np.int32), j["program_wires"]
        S = 1 << L
        b0, b1 = (0, S) if L <= 8 else (S // 3, S // 3 + 16)
        f, g, h = mg.evaluate_batch([{"target": t, "program": prog, "gates": V, "v": v, "what": mg.EVAL_F_GRAD_HVP,
                                      "precision": mg.MG_FP64, "b0": b0, "b1": b1}])[0]
        Z = orc.real_to_directions(V, v)
        fo, dg, dh = orc.target_gradient_hvp_program(j["target_gates"], j["target_wires"], V, Z, param, wires, L,
                                                     b0, b1, nthreads=8)
        go, ho = orc.riemannian_gradient(V, dg), orc.riemannian_hvp(V, dg, Z, dh)
        assert abs(f - fo) <= 1e-11 * abs(fo) and rel(g, go) < 1e-11 and rel(h, ho) < 1e-11, k
        if L <= 12:
            ff, _, _ = orc.target_gradient_hvp_program(j["target_gates"], j["target_wires"], V, 0 * Z, param, wires, L,
                                                       0, S, nthreads=8)
            assert abs(r["f_iter"][-1] - ff.real) <= 1e-11 * abs(ff), k
        del t


def test_serve_time_limit(mg):
    """time_limit_s: a job that would evaluate past the limit ends with status -2 (not an error of
    the call); jobs before the limit are unaffected."""
    jobs = _jobs(mg, 26, SERVE_SPEC[:3])
    say("case: serve with a time limit")
    for batched in (True, False):
        res, _ = mg.serve(jobs, batched=batched, precision=mg.MG_FP64, realtime=False, time_limit_s=1e-9)
        assert all(r["status"] == -2 for r in res)
    full, _ = mg.serve(jobs, batched=False, precision=mg.MG_FP64, realtime=False)
    lim, _ = mg.serve(jobs, batched=False, precision=mg.MG_FP64, realtime=False, time_limit_s=3600.0)
    for a, b in zip(full, lim):
        assert a["status"] == 0 and b["status"] == 0 and np.array_equal(a["f_iter"], b["f_iter"])
    # a limit inside a job's run: that job stops at its next evaluation (status -2, the rest of
    # f_iter NaN), its finished iterations equal the unlimited run's, the later jobs do not run
    say("case: serve with a limit inside the first job")
    long_job = _jobs(mg, 27, [(8, 5, 40)])
    for batched in (False, True):
        ref, _ = mg.serve(long_job, batched=batched, precision=mg.MG_FP64, realtime=False)
        assert ref[0]["status"] == 0
        cut, st = mg.serve(long_job + jobs, batched=batched, precision=mg.MG_FP64, realtime=False,
                           max_active=1, time_limit_s=0.25 * ref[0]["latency"])
        assert cut[0]["status"] == -2 and all(r["status"] == -2 for r in cut[1:])
        fi = cut[0]["f_iter"]
        k = int(np.argmax(np.isnan(fi)))
        assert 0 < k < len(fi) and np.all(np.isnan(fi[k:]))
        assert np.array_equal(fi[:k], ref[0]["f_iter"][:k])


def test_serve_admission_limits(mg):
    jobs = _jobs(mg, 22, SERVE_SPEC[:4])
    say("case: serve with one slot and with a tight memory budget")
    free, _ = mg.serve(jobs, batched=True, precision=mg.MG_FP64, realtime=False)
    one, st = mg.serve(jobs, batched=True, precision=mg.MG_FP64, realtime=False, max_active=1)
    assert st["max_active"] == 1
    tight, st2 = mg.serve(jobs, batched=True, precision=mg.MG_FP64, realtime=False, hbm_budget=1.0)
    assert st2["max_active"] == 1
    for a, b, c in zip(free, one, tight):
        assert np.array_equal(a["f_iter"], b["f_iter"]) and np.array_equal(a["f_iter"], c["f_iter"])


@pytest.mark.parametrize("L", [14, 16])
def test_large_instances_column_subset_match_oracle(mg, L):
    """L=14 (a batched-workload size) and the L=16 XXZ-Trotter instance (758 target gates
    applied per column, states in global scratch): a column subset of the batched kernel
    against the oracle's program-target evaluation, FP64 (1e-11) and FP32 sweeps (2e-5)."""
    from rqcopt_mg.workloads import brickwall_job, large_instance

    rng = np.random.default_rng(30 + L)
    job = large_instance() if L == 16 else brickwall_job(rng, L, 5, 1)
    prog = mg.Program.create(L, job["G"], job["param"], job["wires"])
    tgt = mg.Target.circuit(job["target_gates"], job["target_wires"], L)
    v = rng.standard_normal(8 * job["G"])
    b0, b1 = 1000, 1008
    say(f"case: L={L} column subset [{b0}, {b1}), {job['G']} gates, {job['Gt']} target gates")
    V = job["gates"]
    Z = orc.real_to_directions(V, v)
    fo, dg, dh = orc.target_gradient_hvp_program(job["target_gates"], job["target_wires"], V, Z, job["param"],
                                                 job["wires"], L, b0, b1, nthreads=8)
    go, ho = orc.riemannian_gradient(V, dg), orc.riemannian_hvp(V, dg, Z, dh)
    for prec, tol in ((mg.MG_FP64, 1e-11), (mg.MG_FP32, 2e-5)):
        f, g, h = mg.evaluate_batch([{"target": tgt, "program": prog, "gates": V, "v": v, "what": mg.EVAL_F_GRAD_HVP,
                                      "precision": prec, "b0": b0, "b1": b1}])[0]
        assert abs(f - fo) <= tol * abs(fo) and rel(g, go) < tol and rel(h, ho) < tol, (prec, rel(g, go), rel(h, ho))
    if L == 16:
        say("case: the L=16 target materialised in HBM (2^32 complex rows, in place)")
        tgt.densify()
        for prec, tol in ((mg.MG_FP64, 1e-11), (mg.MG_FP32, 2e-5)):
            f, g, h = mg.evaluate_batch([{"target": tgt, "program": prog, "gates": V, "v": v,
                                          "what": mg.EVAL_F_GRAD_HVP, "precision": prec, "b0": b0, "b1": b1}])[0]
            assert abs(f - fo) <= tol * abs(fo) and rel(g, go) < tol and rel(h, ho) < tol, (prec, rel(g, go))
        del tgt


def test_program_target_beyond_12_qubits_in_materialised_chunks(mg):
    """Program targets beyond 12 qubits are read as dense rows materialised one column chunk at a
    time next to the tiled sweeps.  At L=14 (one chunk by default, 32 chunks of 512 columns with a
    small chunk size) FP64 f, gradient and HVP equal the batched kernel that applies the target's
    gates per column (1e-11 relative), FP32 within 2e-5; a column subset matches the oracle."""
    from rqcopt_mg.workloads import brickwall_job

    rng = np.random.default_rng(41)
    L = 14
    job = brickwall_job(rng, L, 5, 1)
    prog = mg.Program.create(L, job["G"], job["param"], job["wires"])
    tgt = mg.Target.circuit(job["target_gates"], job["target_wires"], L)
    V, v = job["gates"], rng.standard_normal(8 * job["G"])
    req = {"target": tgt, "program": prog, "gates": V, "v": v, "what": mg.EVAL_F_GRAD_HVP}
    say("case: L=14 program target, per-column batched kernel")
    mg.batch_set_single_range(0, 0)
    try:
        ref = mg.evaluate_batch([dict(req, precision=mg.MG_FP64)])[0]
    finally:
        mg.batch_set_single_range(12, 30)
    outs = []
    for cb in (0, 1 << 20):
        say(f"case: L=14 program target, materialised chunks of {cb or 'the default'} bytes")
        mg.lib.mg_set_row_chunk_bytes(cb)
        try:
            outs.append([mg.evaluate_batch([dict(req, precision=pr)])[0] for pr in (mg.MG_FP64, mg.MG_FP32)])
        finally:
            mg.lib.mg_set_row_chunk_bytes(0)
    for o64, o32 in outs:
        assert abs(o64[0] - ref[0]) <= 1e-11 * abs(ref[0])
        assert rel(o64[1], ref[1]) < 1e-11 and rel(o64[2], ref[2]) < 1e-11
        assert abs(o32[0] - ref[0]) <= 2e-5 * abs(ref[0]) and rel(o32[1], ref[1]) < 2e-5 and rel(o32[2], ref[2]) < 2e-5
    b0, b1 = 5000, 5012
    Z = orc.real_to_directions(V, v)
    fo, dg, dh = orc.target_gradient_hvp_program(job["target_gates"], job["target_wires"], V, Z, job["param"],
                                                 job["wires"], L, b0, b1, nthreads=8)
    f, g, h = mg.evaluate_batch([dict(req, precision=mg.MG_FP64, b0=b0, b1=b1)])[0]
    assert abs(f - fo) <= 1e-11 * abs(fo)
    assert rel(g, orc.riemannian_gradient(V, dg)) < 1e-11 and rel(h, orc.riemannian_hvp(V, dg, Z, dh)) < 1e-11


def test_batch_routes_dense_L12_to_single_request_kernels(mg):
    """Dense-target L=12 requests run on the single-request kernels (FP64: register gate
    windows) inside mg_evaluate_batch: bitwise the single path's result; with the routing off
    they run in the persistent batch kernel, equal to rounding."""
    rng = np.random.default_rng(15)
    from rqcopt_mg.workloads import brickwall_job

    j12 = brickwall_job(rng, 12, 5, 1)
    W = program_matrix(j12["target_gates"][:20], np.arange(20), j12["target_wires"][:40], 12)
    t12 = mg.Target.dense(W)
    p12 = mg.Program.create(12, j12["G"], j12["param"], j12["wires"])
    job, prog, tgt, v = make_instance(mg, rng, 8, 3)
    v12 = rng.standard_normal(8 * j12["G"])
    say("case: batch with a dense L=12 request routed to the single-request kernels")
    for prec, tol in ((mg.MG_FP64, 1e-12), (mg.MG_FP32, 1e-5)):
        reqs = [{"target": t12, "program": p12, "gates": j12["gates"], "v": v12, "what": 2, "precision": prec},
                {"target": tgt, "program": prog, "gates": job["gates"], "v": v, "what": 2, "precision": prec}]
        out = mg.evaluate_batch(reqs)
        single = mg.riemannian_gradient_hvp(t12, p12, j12["gates"], v12, prec)
        for a, b in zip(out[0], single):
            assert np.array_equal(np.asarray(a), np.asarray(b))
        mg.batch_set_single_range(0, 0)
        try:
            o2 = mg.evaluate_batch(reqs)
        finally:
            mg.batch_set_single_range(12, 30)
        assert abs(o2[0][0] - single[0]) <= tol * abs(single[0])
        assert rel(o2[0][1], single[1]) < tol and rel(o2[0][2], single[2]) < tol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_graph_replays_equal_eager(mg, prec):
    """A repeated batched launch is captured into a HIP graph and replayed: every replay,
    with new gates and directions each time and single-request launches in between, gives
    the bits of the eager launch of the same requests."""
    from rqcopt_mg.workloads import haar_matchgates

    rng = np.random.default_rng(21)
    P = mg.MG_FP64 if prec == "fp64" else mg.MG_FP32
    inst = [make_instance(mg, rng, L, d, dense) for L, d, dense in MIX]

    def step_requests(step):
        r2 = np.random.default_rng(1000 + step)
        return [{"target": t, "program": p, "gates": haar_matchgates(r2, j["G"]), "v": r2.standard_normal(8 * j["G"]),
                 "what": mg.EVAL_F_GRAD_HVP, "precision": P} for j, p, t, v in inst]

    say(f"case: graph replays, {prec}, 5 steps")
    mg.batch_set_graphs(False)
    try:
        eager = [mg.evaluate_batch(step_requests(s)) for s in range(5)]
    finally:
        mg.batch_set_graphs(True)
    replays = 0
    for s in range(5):
        got = mg.evaluate_batch(step_requests(s))
        r, c = mg.batch_last_graph_stats()
        replays += r
        say(f"  step {s}: replayed {r}, captured {c}")
        for k, (a, b) in enumerate(zip(got, eager[s])):
            for x, y in zip(a, b):
                assert np.array_equal(np.asarray(x), np.asarray(y)), (s, k)
        # eager work on the same stream between replays
        j, p, t, v = inst[1]
        mg.riemannian_gradient_hvp(t, p, j["gates"], v, P)
    assert replays >= 3
