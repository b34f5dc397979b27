"""Column-sharded evaluation (rqcopt_mg.sharded): the per-GPU share of a multi-GPU
evaluation is a contiguous basis-column range, the partial sums are added with one
all-reduce, and the Riemannian projection is applied to the totals.

CPU: the column split (the reference's rule, benchmark_hpc/benchmark_mpi.c:26-45), the host
projection mg_riemannian_project against the oracle, and two gloo ranks whose partial sums come
from the oracle reproducing the one-process oracle evaluation.  GPU: two gloo ranks whose
partial sums come from mg_evaluate_columns on the (one) GPU reproduce the one-process GPU
evaluation.
"""
import os
import socket
import subprocess
import sys

import numpy as np
import pytest

import oracle as orc

WORKER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "workers", "shard_worker.py")


@pytest.fixture(scope="module")
def mg():
    import rqcopt_mg

    return rqcopt_mg


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def test_column_ranges_partition_the_trace(mg):
    from rqcopt_mg.sharded import column_range

    for L in (2, 4, 6, 13):
        for world in (1, 2, 3, 5, 8):
            S = 1 << L
            ranges = [column_range(L, r, world) for r in range(world)]
            assert ranges[0][0] == 0 and ranges[-1][1] == S
            assert all(ranges[r][1] == ranges[r + 1][0] for r in range(world - 1))
            sizes = [b - a for a, b in ranges]
            q, extra = divmod(S, world)
            # the first S mod world ranks take one column more
            assert sizes == [q + 1] * extra + [q] * (world - extra)


def test_host_projection_matches_oracle(mg):
    rng = np.random.default_rng(7)
    from rqcopt_mg.workloads import haar_matchgates

    P = 5
    V = haar_matchgates(rng, P)
    dg = rng.standard_normal((P, 8)) + 1j * rng.standard_normal((P, 8))
    dh = rng.standard_normal((P, 8)) + 1j * rng.standard_normal((P, 8))
    v = rng.standard_normal(8 * P)
    g, hv = mg.riemannian_project(V, dg, v, dh)
    Z = orc.real_to_directions(V, v)
    assert np.array_equal(mg.real_to_directions(V, v), Z) or rel(mg.real_to_directions(V, v), Z) < 1e-15
    assert rel(g, orc.riemannian_gradient(V, dg)) < 1e-14
    assert rel(hv, orc.riemannian_hvp(V, dg, Z, dh)) < 1e-14
    # linear in the partial sums: projecting a sum = summing the projections
    g1, h1 = mg.riemannian_project(V, 0.25 * dg, v, 0.25 * dh)
    g2, h2 = mg.riemannian_project(V, 0.75 * dg, v, 0.75 * dh)
    assert rel(g1 + g2, g) < 1e-14 and rel(h1 + h2, hv) < 1e-14


def free_port():
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


def run_ranks(mode, world, tmp_path):
    port = free_port()
    env = dict(os.environ, MASTER_ADDR="127.0.0.1")
    outs = [tmp_path / f"rank{r}.npz" for r in range(world)]
    procs = [subprocess.Popen([sys.executable, WORKER, str(r), str(world), str(port), mode, str(outs[r])], env=env,
                              stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True) for r in range(world)]
    logs = []
    try:
        for p in procs:
            out, _ = p.communicate(timeout=180)
            logs.append(out)
    finally:
        for p in procs:
            if p.poll() is None:
                p.kill()
    for p, log in zip(procs, logs):
        assert p.returncode == 0, log
    return [np.load(o) for o in outs]


def reference_full(mg, mode):
    sys.path.insert(0, os.path.dirname(WORKER))
    from shard_worker import problem

    job, W, v = problem()
    L, V = job["L"], job["gates"]
    if mode == "oracle":
        Z = orc.real_to_directions(V, v)
        f, dg, dh = orc.target_gradient_hvp(W, V, Z, job["param"], job["wires"], L)
        return f, orc.riemannian_gradient(V, dg), orc.riemannian_hvp(V, dg, Z, dh)
    prog = mg.Program.create(L, job["G"], job["param"], job["wires"])
    return mg.riemannian_gradient_hvp(mg.Target.dense(W), prog, V, v, mg.MG_FP64)


def test_two_gloo_ranks_oracle_partials_match_full(mg, tmp_path):
    res = run_ranks("oracle", 2, tmp_path)
    f, g, hv = reference_full(mg, "oracle")
    for r in res:
        assert abs(r["f"][0] - f) <= 1e-12 * abs(f)
        assert rel(r["g"], g) < 1e-12 and rel(r["hv"], hv) < 1e-12
    # every rank holds the same totals
    assert np.array_equal(res[0]["g"], res[1]["g"]) and np.array_equal(res[0]["hv"], res[1]["hv"])


@pytest.mark.gpu
def test_two_gloo_ranks_gpu_partials_match_full(mg, tmp_path):
    print("case: 2 gloo ranks on one GPU, L=6 column halves, FP64", flush=True)
    res = run_ranks("gpu", 2, tmp_path)
    f, g, hv = reference_full(mg, "gpu")
    for r in res:
        assert abs(r["f"][0] - f) <= 1e-12 * abs(f)
        assert rel(r["g"], g) < 1e-12 and rel(r["hv"], hv) < 1e-12
