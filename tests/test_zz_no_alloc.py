"""mg_evaluate_noalloc (the allocation-free variant of mg_evaluate) against mg_evaluate,
through the C-ABI: two calls in a row, the second on a smaller input, FP64 within 1e-11
relative (same kernels and launch shape, so the results are expected to be identical);
FP32 after the target's FP32 copy exists; requests beyond the static scratch fail cleanly."""
import ctypes
import sys

import numpy as np
import pytest

from golden_io import haar_matchgates

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mg():
    import rqcopt_mg

    return rqcopt_mg


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def open_brickwall(L, depth):
    wires = []
    for k in range(depth):
        for a in range(k % 2, L - 1, 2):
            wires += [a, a + 1]
    G = len(wires) // 2
    return np.arange(G, dtype=np.int32), np.asarray(wires, np.int32), G


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(1e-300, np.max(np.abs(b)))


def noalloc(mg, t, p, V, Z, what, precision):
    P = p.P
    V = mg._c(V)
    Zc = mg._c(Z) if Z is not None else None
    f = np.zeros(1, np.complex128)
    dg = np.zeros((P, 8), np.complex128)
    dh = np.zeros((P, 8), np.complex128)
    ret = mg.lib.mg_evaluate_noalloc(t.h, p.h, mg._p(V), mg._p(Zc) if Zc is not None else None, what, precision,
                                     mg._p(f), mg._p(dg), mg._p(dh))
    return ret, f[0], dg, dh


def case(mg, rng, L, depth, target_kind):
    param, wires, G = open_brickwall(L, depth)
    V = haar_matchgates(rng, G)
    Z = rng.standard_normal((G, 8)) + 1j * rng.standard_normal((G, 8))
    if target_kind == "dense":
        S = 1 << L
        t = mg.Target.dense(rng.standard_normal((S, S)) + 1j * rng.standard_normal((S, S)))
    else:
        pt, wt, Gt = open_brickwall(L, 3 * depth)
        t = mg.Target.circuit(haar_matchgates(rng, Gt), wt, L)
    p = mg.Program.create(L, G, param, wires)
    return t, p, V, Z


@pytest.mark.parametrize("kind", ["dense", "circuit"])
def test_noalloc_equals_evaluate_two_calls(mg, kind):
    rng = np.random.default_rng(7)
    for L, depth in ((12, 9), (8, 5)):  # the second call is smaller
        say(f"case: mg_evaluate_noalloc vs mg_evaluate, {kind} target, L={L}, depth {depth}, FP64")
        t, p, V, Z = case(mg, rng, L, depth, kind)
        for what in (mg.EVAL_F, mg.EVAL_F_GRAD, mg.EVAL_F_GRAD_HVP):
            ret, f, dg, dh = noalloc(mg, t, p, V, Z, what, mg.MG_FP64)
            assert ret == 0, mg.lib.mg_last_error().decode()
            fr, dgr, dhr = mg.evaluate(t, p, V, Z, what, mg.MG_FP64)
            assert abs(f - fr) <= 1e-11 * abs(fr)
            if what >= mg.EVAL_F_GRAD:
                assert rel(dg, dgr) <= 1e-11
            if what == mg.EVAL_F_GRAD_HVP:
                assert rel(dh, dhr) <= 1e-11


def test_noalloc_fp32_after_copy(mg):
    rng = np.random.default_rng(8)
    say("case: mg_evaluate_noalloc FP32, L=10 dense")
    t, p, V, Z = case(mg, rng, 10, 7, "dense")
    ret, *_ = noalloc(mg, t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP32)
    assert ret == -1 and b"FP32 copy" in mg.lib.mg_last_error()
    fr, dgr, dhr = mg.evaluate(t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP32)
    ret, f, dg, dh = noalloc(mg, t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP32)
    assert ret == 0
    assert abs(f - fr) <= 1e-11 * abs(fr) and rel(dg, dgr) <= 1e-11 and rel(dh, dhr) <= 1e-11


def test_noalloc_rejects_sweeps_beyond_the_scratch(mg):
    """L=14: the states live in per-workgroup global slices (hundreds of MiB), beyond the
    32 MiB static scratch; the call fails with an error and the next call still works."""
    rng = np.random.default_rng(9)
    say("case: mg_evaluate_noalloc rejects L=14")
    param, wires, G = open_brickwall(14, 3)
    pt, wt, Gt = open_brickwall(14, 3)
    t = mg.Target.circuit(haar_matchgates(rng, Gt), wt, 14).densify()
    p = mg.Program.create(14, G, param, wires)
    V = haar_matchgates(rng, G)
    ret, *_ = noalloc(mg, t, p, V, None, mg.EVAL_F_GRAD, mg.MG_FP64)
    assert ret == -1 and b"static scratch" in mg.lib.mg_last_error()
    t2, p2, V2, Z2 = case(mg, rng, 8, 5, "dense")
    ret, f, dg, dh = noalloc(mg, t2, p2, V2, Z2, mg.EVAL_F_GRAD_HVP, mg.MG_FP64)
    assert ret == 0
    fr, dgr, dhr = mg.evaluate(t2, p2, V2, Z2, mg.EVAL_F_GRAD_HVP, mg.MG_FP64)
    assert abs(f - fr) <= 1e-11 * abs(fr) and rel(dh, dhr) <= 1e-11
