"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle, the
reference's golden fixtures and size-independent properties.

FP64 sweeps: f, gradient and D^2 f[Z] within 1e-11 relative of the oracle.
FP32 sweeps (FP64 accumulation): within 2e-5 relative.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle as orc
from golden_io import cvec, gate, gates, haar_matchgates, load, perms, random_perms, scramble_target

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mg():
    import rqcopt_mg

    return rqcopt_mg


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def open_brickwall(L, depth):
    """open-boundary brickwall, one independent gate per instance: layer k uses
    pairs (0,1),(2,3),... when k is even and (1,2),(3,4),... when k is odd."""
    wires = []
    for k in range(depth):
        start = k % 2
        for a in range(start, L - 1, 2):
            wires += [a, a + 1]
    G = len(wires) // 2
    return np.arange(G, dtype=np.int32), np.asarray(wires, np.int32), G


def random_target(rng, L, depth=4):
    param, wires, G = open_brickwall(L, depth)
    Ug = haar_matchgates(rng, G)
    from dense_ref import program_matrix

    return program_matrix(Ug, param, wires, L) if L <= 8 else None, (Ug, param, wires)


# ---------------- reference entry points with the reference's fixtures ----------------

def test_reference_api_gradient_vector_hessian_matrix_fixture(mg):
    say("case: matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix, L=6, 5 layers")
    d = load("test_matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix_cplx")
    V, pm = gates(d, "V.", 5), perms(d, 5)
    uf = mg.dense_linear_func(scramble_target(6))
    pm_i, arr = mg._perm_ptrs(pm)
    f = np.zeros(1, np.complex128)
    g = np.zeros(40)
    H = np.zeros((40, 40))
    ret = mg.lib.matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix(
        uf, None, mg._p(V), 5, 6, ctypes.cast(arr, ctypes.c_void_p), mg._p(f), mg._p(g), mg._p(H))
    assert ret == 0
    # reference tolerances (test/test_matchgate_target.c:464-500): f 1e-12, grad 1e-14, H 1e-13
    assert abs(f[0] - cvec(d["f"])[0]) < 1e-12
    assert np.max(np.abs(g - d["grad"])) < 1e-13
    assert np.max(np.abs(H - H.T)) < 1e-14
    assert np.max(np.abs(H - np.asarray(d["H"]).reshape(40, 40))) < 1e-12


def test_reference_api_target_gradient_hessian_nonunitary_fixture(mg):
    say("case: matchgate_brickwall_unitary_target_gradient_hessian, non-unitary gates, L=6")
    d = load("test_matchgate_brickwall_unitary_target_gradient_hessian_cplx")
    V, pm = gates(d, "V.", 5), perms(d, 5)
    uf = mg.dense_linear_func(scramble_target(6))
    emap = [5, 6, 9, 10, 0, 3, 12, 15]
    for k, nl in enumerate((4, 5)):
        pm_i, arr = mg._perm_ptrs(pm[:nl])
        f = np.zeros(1, np.complex128)
        dv = np.zeros((nl, 8), np.complex128)
        hess = np.zeros((8 * nl, 8 * nl), np.complex128)
        ret = mg.lib.matchgate_brickwall_unitary_target_gradient_hessian(
            uf, None, mg._p(V), nl, 6, ctypes.cast(arr, ctypes.c_void_p), mg._p(f), mg._p(dv), mg._p(hess))
        assert ret == 0
        assert abs(f[0] - cvec(d[f"f{k}"])[0]) < 1e-12
        assert np.max(np.abs(dv - gates(d, f"dVlist{k}.", nl))) < 1e-12
        hr = d[f"hess{k}"]
        hr = (hr[..., 0] + 1j * hr[..., 1]).reshape(nl, 16, nl, 16)[:, emap][:, :, :, emap].reshape(8 * nl, 8 * nl)
        assert rel(hess, hr) < 1e-12


def test_reference_api_target_and_circuit_fixtures(mg):
    say("case: matchgate_brickwall_unitary_target / circuit target fixtures")
    d = load("test_matchgate_brickwall_unitary_target_cplx")
    V, pm = gates(d, "V.", 5), perms(d, 5)
    uf = mg.dense_linear_func(scramble_target(8))
    for k, nl in enumerate((4, 5)):
        pm_i, arr = mg._perm_ptrs(pm[:nl])
        f = np.zeros(1, np.complex128)
        assert mg.lib.matchgate_brickwall_unitary_target(uf, None, mg._p(V), nl, 8, ctypes.cast(arr, ctypes.c_void_p), mg._p(f)) == 0
        assert abs(f[0] - cvec(d[f"f{k}"])[0]) < 1e-12
    d = load("test_matchgate_circuit_unitary_target_cplx")
    G = gates(d, "G.", 5)
    wires = np.asarray(d["wires"], np.int32).reshape(-1)
    uf7 = mg.dense_linear_func(scramble_target(7))
    f = np.zeros(1, np.complex128)
    assert mg.lib.matchgate_circuit_unitary_target(uf7, None, mg._p(G), 5, mg._p(wires), 7, mg._p(f)) == 0
    assert abs(f[0] - cvec(d["f"])[0]) < 1e-12
    # circuit gradient against the oracle
    dg = np.zeros((5, 8), np.complex128)
    assert mg.lib.matchgate_circuit_unitary_target_and_gradient(uf7, None, mg._p(G), 5, mg._p(wires), 7, mg._p(f), mg._p(dg)) == 0
    fo, dgo = orc.target_gradient(scramble_target(7), G, np.arange(5, dtype=np.int32), wires, 7)
    assert rel(dg, dgo) < 1e-12


def test_ufunc_failure_returns_minus_two(mg):
    say("case: failing ufunc")

    def bad(psi, udata, out):
        return -5

    cb = mg.LINEAR_FUNC(bad)
    V = haar_matchgates(np.random.default_rng(0), 2)
    pm_i, arr = mg._perm_ptrs(np.stack([np.arange(4), np.arange(4)]))
    f = np.zeros(1, np.complex128)
    assert mg.lib.matchgate_brickwall_unitary_target(cb, None, mg._p(V), 2, 4, ctypes.cast(arr, ctypes.c_void_p), mg._p(f)) == -2


# ---------------- device targets and programs against the oracle ----------------

CASES = [
    # (L, layout, depth)
    (6, "tied", 5),
    (8, "tied", 4),
    (8, "open", 5),
    (10, "open", 7),
]


def build_case(L, layout, depth, seed):
    rng = np.random.default_rng(seed)
    if layout == "tied":
        pm = random_perms(rng, L, depth)
        param, wires = orc.brickwall_program(L, pm)
        P = depth
    else:
        param, wires, P = open_brickwall(L, depth)
    V = haar_matchgates(rng, P)
    Z = rng.standard_normal((P, 8)) + 1j * rng.standard_normal((P, 8))
    S = 1 << L
    U = rng.standard_normal((S, S)) + 1j * rng.standard_normal((S, S))
    return V, Z, param, wires, P, U


@pytest.mark.parametrize("L,layout,depth", CASES)
@pytest.mark.parametrize("ckpt,lds", [(-1, -1), (1, -1), (0, 0), (1, 0)])
def test_evaluate_fp64_matches_oracle(mg, L, layout, depth, ckpt, lds):
    say(f"case: evaluate FP64 L={L} {layout} depth={depth} ckpt={ckpt} lds={lds}")
    V, Z, param, wires, P, U = build_case(L, layout, depth, seed=L * 10 + depth)
    t = mg.Target.dense(U)
    p = mg.Program.create(L, P, param, wires)
    mg.set_checkpointing(ckpt)
    mg.set_lds_states(lds)
    try:
        f, dg, dh = mg.evaluate(t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP64)
    finally:
        mg.set_checkpointing(-1)
        mg.set_lds_states(-1)
    fo, dgo, dho = orc.target_gradient_hvp(U, V, Z, param, wires, L)
    assert abs(f - fo) / abs(fo) < 1e-11
    assert rel(dg, dgo) < 1e-11
    assert rel(dh, dho) < 1e-11
    # modes 0 and 1 agree with mode 2's f and gradient
    f0, _, _ = mg.evaluate(t, p, V, None, mg.EVAL_F, mg.MG_FP64)
    f1, dg1, _ = mg.evaluate(t, p, V, None, mg.EVAL_F_GRAD, mg.MG_FP64)
    assert abs(f0 - fo) / abs(fo) < 1e-11 and abs(f1 - fo) / abs(fo) < 1e-11
    assert rel(dg1, dgo) < 1e-11


@pytest.mark.parametrize("L,layout,depth", [(11, "open", 6), (12, "tied", 3)])
@pytest.mark.parametrize("threads", [512, 1024])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_register_kernel_matches_oracle(mg, L, layout, depth, threads, prec):
    """The register-resident kernel (states in VGPRs, LDS re-layouts) on a column subset."""
    say(f"case: register kernel L={L} {layout} depth={depth} threads={threads} {prec}")
    if L == 11 and threads == 1024:
        pytest.skip("2^11/4 quadruples do not split over 1024 threads with an instantiated count")
    V, Z, param, wires, P, U = build_case(L, layout, depth, seed=L * 7 + depth)
    t = mg.Target.dense(U)
    p = mg.Program.create(L, P, param, wires)
    precision = mg.MG_FP64 if prec == "fp64" else mg.MG_FP32
    mg.set_register_kernel(1, threads)
    mg.set_window_kernel(0)
    try:
        r = mg.evaluate_columns(t, p, V, Z, mg.EVAL_F_GRAD_HVP, precision, 5, 69)
        info = mg.last_sweep_info()
        r1 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F_GRAD, precision, 5, 69)
        r0 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F, precision, 5, 69)
    finally:
        mg.set_register_kernel(-1, 0)
        mg.set_window_kernel(-1)
    assert info["reg_threads"] == threads and info["kernel"] == "sweep_reg_kernel"
    fo, dgo, dho = orc.target_gradient_hvp(U, V, Z, param, wires, L, 5, 69)
    tol = 1e-11 if prec == "fp64" else 2e-5
    assert abs(r[0] - fo) / abs(fo) < tol and rel(r[1], dgo) < tol and rel(r[2], dho) < tol
    assert abs(r1[0] - fo) / abs(fo) < tol and rel(r1[1], dgo) < tol and abs(r0[0] - fo) / abs(fo) < tol


@pytest.mark.parametrize("layout,depth", [("tied", 3), ("open", 5), ("open", 9), ("open", 15)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_window_register_kernel_matches_oracle(mg, layout, depth, prec):
    """The register-resident gate-window kernel (L=12): tied brickwalls with random
    permutations (windows of scattered qubits) and open brickwalls, all three modes; the FP64
    kernel re-lays two states out per round, or one when the accumulator of many parameter gates
    and the staging slots of the largest window block leave no LDS room for two."""
    L = 12
    say(f"case: register window kernel L={L} {layout} depth={depth} {prec}")
    V, Z, param, wires, P, U = build_case(L, layout, depth, seed=depth * 5 + 1)
    t = mg.Target.dense(U)
    p = mg.Program.create(L, P, param, wires)
    precision = mg.MG_FP64 if prec == "fp64" else mg.MG_FP32
    mg.set_window_kernel(1)
    try:
        r = mg.evaluate_columns(t, p, V, Z, mg.EVAL_F_GRAD_HVP, precision, 7, 71)
        info = mg.last_sweep_info()
        assert info["kernel"] == "sweep_win_kernel"
        if prec == "fp64":
            # one state (64 KB) or two (128 KB) per re-layout round: two whenever the accumulator
            # (32 P doubles) and at most 6 staging slots (12 KB) leave room
            print(f"P={P}, LDS bytes {info['lds_bytes']}", flush=True)
            assert info["lds_bytes"] <= 160 * 1024
            if P <= 79:
                assert info["lds_bytes"] > 128 * 1024, (P, info["lds_bytes"])
        r1 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F_GRAD, precision, 7, 71)
        if prec == "fp64":
            # without the D^2 f[Z] half of the accumulator two states fit at every depth
            assert mg.last_sweep_info()["lds_bytes"] > 128 * 1024, (P, mg.last_sweep_info())
        r0 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F, precision, 7, 71)
    finally:
        mg.set_window_kernel(-1)
    fo, dgo, dho = orc.target_gradient_hvp(U, V, Z, param, wires, L, 7, 71)
    tol = 1e-11 if prec == "fp64" else 2e-5
    assert abs(r[0] - fo) / abs(fo) < tol and rel(r[1], dgo) < tol and rel(r[2], dho) < tol
    assert abs(r1[0] - fo) / abs(fo) < tol and rel(r1[1], dgo) < tol and abs(r0[0] - fo) / abs(fo) < tol


@pytest.mark.parametrize("layout,depth", [("tied", 3), ("open", 9), ("open", 15)])
def test_window_kernel_forward_on_matrix_cores(mg, layout, depth):
    """FP64 L=12 f, f + gradient and f + gradient + HVP on the register-window kernel: the forward
    pass over 4-qubit windows on the FP64 matrix cores, the column in the re-layout buffer (the
    default; for the HVP only when two states fit a re-layout round), and gate by gate in registers
    (mg_set_forward_mfma(0)), both against the oracle (1e-11).  Equal bits would mean the matrix
    cores were not used."""
    L = 12
    say(f"case: window kernel FP64 L={L} {layout} depth={depth}, forward pass on the matrix cores")
    V, Z, param, wires, P, U = build_case(L, layout, depth, seed=depth * 13 + 6)
    t = mg.Target.dense(U)
    p = mg.Program.create(L, P, param, wires)

    def run():
        r = mg.evaluate_columns(t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP64, 7, 71)
        info = mg.last_sweep_info()
        assert info["kernel"] == "sweep_win_kernel", info
        r1 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F_GRAD, mg.MG_FP64, 7, 71)
        r0 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F, mg.MG_FP64, 7, 71)
        return r, r1, r0, info

    r, r1, r0, info = run()
    mg.set_forward_mfma(0)
    try:
        v, v1, v0, _ = run()
    finally:
        mg.set_forward_mfma(-1)
    fo, dgo, dho = orc.target_gradient_hvp(U, V, Z, param, wires, L, 7, 71)
    for x, x1, x0 in ((r, r1, r0), (v, v1, v0)):
        assert abs(x[0] - fo) / abs(fo) < 1e-11 and rel(x[1], dgo) < 1e-11 and rel(x[2], dho) < 1e-11
        assert abs(x1[0] - fo) / abs(fo) < 1e-11 and rel(x1[1], dgo) < 1e-11 and abs(x0[0] - fo) / abs(fo) < 1e-11
    assert np.any(r1[1] != v1[1]) and r0[0] != v0[0]
    if info["lds_bytes"] > 128 * 1024:
        # two states per re-layout round: the HVP's forward pass ran on the matrix cores too
        assert np.any(r[2] != v[2])


@pytest.mark.parametrize("layout,depth", [("tied", 3), ("open", 9), ("open", 15)])
def test_lds_kernel_forward_on_matrix_cores(mg, layout, depth):
    """FP32 L=12 f, f + gradient and f + gradient + HVP on the LDS-resident kernel: the forward pass
    (psi, and dpsi for the HVP) over 4-qubit windows on the matrix cores (the default) and gate by
    gate on the VALU (mg_set_forward_mfma(0)), both against the oracle (2e-5); tied brickwalls with
    random permutations give windows of scattered qubits.  The two paths round differently, so
    equal bits would mean the matrix cores were not used."""
    L = 12
    say(f"case: LDS kernel FP32 L={L} {layout} depth={depth}, forward pass on the matrix cores")
    V, Z, param, wires, P, U = build_case(L, layout, depth, seed=depth * 11 + 4)
    t = mg.Target.dense(U)
    p = mg.Program.create(L, P, param, wires)

    def run():
        r = mg.evaluate_columns(t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP32, 7, 71)
        info = mg.last_sweep_info()
        assert info["kernel"] == "sweep_kernel" and info["lds_states"], info
        r1 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F_GRAD, mg.MG_FP32, 7, 71)
        assert mg.last_sweep_info()["kernel"] == "sweep_kernel"
        r0 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F, mg.MG_FP32, 7, 71)
        return r, r1, r0

    r, r1, r0 = run()
    mg.set_forward_mfma(0)
    try:
        v, v1, v0 = run()
    finally:
        mg.set_forward_mfma(-1)
    fo, dgo, dho = orc.target_gradient_hvp(U, V, Z, param, wires, L, 7, 71)
    for x, x1, x0 in ((r, r1, r0), (v, v1, v0)):
        assert abs(x[0] - fo) / abs(fo) < 2e-5 and rel(x[1], dgo) < 2e-5 and rel(x[2], dho) < 2e-5
        assert abs(x1[0] - fo) / abs(fo) < 2e-5 and rel(x1[1], dgo) < 2e-5 and abs(x0[0] - fo) / abs(fo) < 2e-5
    assert np.any(r[2] != v[2]) and np.any(r1[1] != v1[1]) and r0[0] != v0[0]


@pytest.mark.parametrize("depth", [5, 15])
def test_lds_kernel_f_and_gradient_two_workgroups_per_cu(mg, depth):
    """FP32 L=12 f and f + gradient on the LDS-resident kernel: 256-thread workgroups whose
    states (2 x 32 KB) and gradient-only accumulator fit twice per CU, up to depth 15 (83
    parameter gates), against the oracle (2e-5)."""
    L = 12
    say(f"case: LDS kernel FP32 L={L} open depth={depth}, f and f+grad")
    V, Z, param, wires, P, U = build_case(L, "open", depth, seed=depth * 3 + 2)
    t = mg.Target.dense(U)
    p = mg.Program.create(L, P, param, wires)
    fo, dgo, _ = orc.target_gradient_hvp(U, V, Z, param, wires, L, 5, 69)
    r1 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F_GRAD, mg.MG_FP32, 5, 69)
    info = mg.last_sweep_info()
    assert info["kernel"] == "sweep_kernel" and info["lds_states"] and info["lds_bytes"] <= 80 * 1024, info
    r0 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F, mg.MG_FP32, 5, 69)
    assert abs(r1[0] - fo) / abs(fo) < 2e-5 and rel(r1[1], dgo) < 2e-5 and abs(r0[0] - fo) / abs(fo) < 2e-5


@pytest.mark.parametrize("L,layout,depth,prec", [(13, "open", 11, "fp64"), (13, "tied", 4, "fp32"), (14, "open", 7, "fp64"),
                                                  (13, "open", 15, "fp32"), (13, "open", 15, "fp64"),
                                                  (14, "open", 5, "fp32"), (12, "open", 5, "fp64"),
                                                  (13, "open", 17, "fp32"), (13, "open", 21, "fp64")])
def test_tile_kernels_match_oracle(mg, L, layout, depth, prec):
    """The tiled kernels for dense targets whose states exceed the LDS, all three modes on a
    column subset against the oracle: FP64 unitary gates run the tiled register-window kernel
    (12-qubit tiles swept in 3-qubit register windows; one or two states per re-layout round)
    and, with it switched off, the LDS-tile kernel; FP32 and the FP64 gates that are only
    invertible (L=12, unitarity defect 2e-11) run the LDS-tile kernel (2^12 / 2^11 amplitudes).
    Depth 17 FP32 (102 parameter gates) and depth 21 FP64 (126) exceed what the LDS-tile
    kernel's per-parameter accumulator holds next to the tile: it accumulates per tile block
    and adds each block's sums to the partial sums in memory."""
    say(f"case: tiled kernels L={L} {layout} depth={depth} {prec}")
    V, Z, param, wires, P, U = build_case(L, layout, depth, seed=L * 7 + depth)
    if L == 12:
        V = V * (1.0 + 1e-11)  # unitarity defect 2e-11: no register window kernel, no checkpoints
    t = mg.Target.dense(U)
    p = mg.Program.create(L, P, param, wires)
    precision = mg.MG_FP64 if prec == "fp64" else mg.MG_FP32
    c0, c1 = 3, 3 + 24
    fo, dgo, dho = orc.target_gradient_hvp(U, V, Z, param, wires, L, c0, c1)
    tol = 1e-11 if prec == "fp64" else 2e-5
    kernels = ["sweep_win_tile_kernel", "sweep_tile_kernel"] if prec == "fp64" and L > 12 else ["sweep_tile_kernel"]
    results = []
    for kname in kernels:
        mg.set_window_kernel(0 if kname == "sweep_tile_kernel" else -1)
        try:
            r = mg.evaluate_columns(t, p, V, Z, mg.EVAL_F_GRAD_HVP, precision, c0, c1)
            assert mg.last_sweep_info()["kernel"] == kname, (kname, mg.last_sweep_info())
            r1 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F_GRAD, precision, c0, c1)
            # f + gradient needs two states, which fit the LDS at FP64 L=12 and FP32 L=13
            k1 = mg.last_sweep_info()["kernel"]
            assert k1 == kname or L == 12 or (L == 13 and prec == "fp32"), k1
            r0 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F, precision, c0, c1)
        finally:
            mg.set_window_kernel(-1)
        assert abs(r[0] - fo) / abs(fo) < tol and rel(r[1], dgo) < tol and rel(r[2], dho) < tol, kname
        assert abs(r1[0] - fo) / abs(fo) < tol and rel(r1[1], dgo) < tol and abs(r0[0] - fo) / abs(fo) < tol, kname
        results.append(r)
    # the same columns through the kernels without tiles (global-scratch or register states)
    mg.set_lds_states(0)
    try:
        g = mg.evaluate_columns(t, p, V, Z, mg.EVAL_F_GRAD_HVP, precision, c0, c1)
        assert mg.last_sweep_info()["kernel"] in ("sweep_kernel", "sweep_reg_kernel")
    finally:
        mg.set_lds_states(-1)
    tol_g = 1e-12 if prec == "fp64" else 2e-5
    for r in results:
        assert abs(r[0] - g[0]) / abs(g[0]) < tol_g and rel(r[1], g[1]) < tol_g and rel(r[2], g[2]) < tol_g
    if prec == "fp32" and L > 13:
        # FP32 f, f + gradient and f + gradient + HVP: forward pass gate by gate on the VALU against
        # the default (4-qubit windows on the matrix cores), both against the oracle
        mg.set_forward_mfma(0)
        try:
            v2 = mg.evaluate_columns(t, p, V, Z, mg.EVAL_F_GRAD_HVP, precision, c0, c1)
            assert mg.last_sweep_info()["kernel"] == "sweep_tile_kernel"
            v1 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F_GRAD, precision, c0, c1)
            v0 = mg.evaluate_columns(t, p, V, None, mg.EVAL_F, precision, c0, c1)
            assert mg.last_sweep_info()["kernel"] == "sweep_tile_kernel"
        finally:
            mg.set_forward_mfma(-1)
        assert abs(v2[0] - fo) / abs(fo) < tol and rel(v2[1], dgo) < tol and rel(v2[2], dho) < tol
        assert abs(v1[0] - fo) / abs(fo) < tol and rel(v1[1], dgo) < tol and abs(v0[0] - fo) / abs(fo) < tol
        # the default took the matrix cores: the two forward passes round differently
        assert any(np.any(r[2] != v2[2]) for r in results)


def test_register_kernel_brickwall_target(mg):
    """Register kernel with an in-kernel (matrix-free) brickwall target at L=12."""
    say("case: register kernel with brickwall target, L=12")
    rng = np.random.default_rng(21)
    L = 12
    up = random_perms(rng, L, 3)
    Ug = haar_matchgates(rng, 3)
    pm = random_perms(rng, L, 3)
    V = haar_matchgates(rng, 3)
    Z = rng.standard_normal((3, 8)) + 1j * rng.standard_normal((3, 8))
    p = mg.Program.brickwall(L, pm)
    t = mg.Target.brickwall(Ug, up, L)
    mg.set_register_kernel(1, 512)
    try:
        r = mg.evaluate_columns(t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP64, 0, 40)
    finally:
        mg.set_register_kernel(-1, 0)
    tp, tw = orc.brickwall_program(L, up)
    S = 1 << L
    Ud = np.zeros((S, S), np.complex128)
    for b in range(40):
        Ud[:, b] = orc.apply_program(Ug, tp, tw, L, np.eye(1, S, b).ravel())
    param, wires = orc.brickwall_program(L, pm)
    fo, dgo, dho = orc.target_gradient_hvp(Ud, V, Z, param, wires, L, 0, 40)
    assert abs(r[0] - fo) / abs(fo) < 1e-11 and rel(r[1], dgo) < 1e-11 and rel(r[2], dho) < 1e-11


@pytest.mark.parametrize("L,layout,depth", CASES)
def test_evaluate_fp32_matches_oracle(mg, L, layout, depth):
    say(f"case: evaluate FP32 L={L} {layout} depth={depth}")
    V, Z, param, wires, P, U = build_case(L, layout, depth, seed=L * 10 + depth)
    t = mg.Target.dense(U)
    p = mg.Program.create(L, P, param, wires)
    f, dg, dh = mg.evaluate(t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP32)
    fo, dgo, dho = orc.target_gradient_hvp(U, V, Z, param, wires, L)
    assert abs(f - fo) / abs(fo) < 2e-5
    assert rel(dg, dgo) < 2e-5
    assert rel(dh, dho) < 2e-5


def test_nonunitary_gates_use_checkpoints(mg):
    say("case: non-unitary gates (checkpointed reverse sweep)")
    d = load("test_matchgate_brickwall_unitary_target_gradient_hessian_cplx")
    V, pm, Z = gates(d, "V.", 5), perms(d, 5), gates(d, "Z.", 5)
    U = scramble_target(6)
    p = mg.Program.brickwall(6, pm)
    t = mg.Target.dense(U)
    f, dg, dh = mg.evaluate(t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP64)
    assert mg.last_sweep_info()["ckpt"]
    param, wires = orc.brickwall_program(6, pm)
    fo, dgo, dho = orc.target_gradient_hvp(U, V, Z, param, wires, 6)
    assert abs(f - fo) < 1e-12 and rel(dg, dgo) < 1e-12 and rel(dh, dho) < 1e-11


def test_riemannian_hvp_matches_fixture_hessian(mg):
    say("case: Riemannian HVP vs the reference's H matrix")
    d = load("test_matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix_cplx")
    V, pm = gates(d, "V.", 5), perms(d, 5)
    H = np.asarray(d["H"]).reshape(40, 40)
    t = mg.Target.dense(scramble_target(6))
    p = mg.Program.brickwall(6, pm)
    rng = np.random.default_rng(4)
    for _ in range(4):
        v = rng.standard_normal(40)
        f, g, hv = mg.riemannian_gradient_hvp(t, p, V, v, mg.MG_FP64)
        assert rel(hv, H @ v) < 1e-12
        assert np.max(np.abs(g - d["grad"])) < 1e-13


def test_brickwall_and_circuit_targets_match_dense(mg):
    say("case: matrix-free targets vs dense")
    rng = np.random.default_rng(9)
    L = 8
    up = random_perms(rng, L, 6)
    Ug = haar_matchgates(rng, 6)
    tp, tw = orc.brickwall_program(L, up)
    S = 1 << L
    Ud = np.stack([orc.apply_program(Ug, tp, tw, L, np.eye(S)[:, b]) for b in range(S)], axis=1)
    V = haar_matchgates(rng, 4)
    pm = random_perms(rng, L, 4)
    p = mg.Program.brickwall(L, pm)
    Z = rng.standard_normal((4, 8)) + 1j * rng.standard_normal((4, 8))
    r_dense = mg.evaluate(mg.Target.dense(Ud), p, V, Z)
    r_bw = mg.evaluate(mg.Target.brickwall(Ug, up, L), p, V, Z)
    r_circ = mg.evaluate(mg.Target.circuit(Ug[tp], tw, L), p, V, Z)
    for r in (r_bw, r_circ):
        assert abs(r[0] - r_dense[0]) < 1e-11
        assert rel(r[1], r_dense[1]) < 1e-11 and rel(r[2], r_dense[2]) < 1e-11
    r32 = mg.evaluate(mg.Target.brickwall(Ug, up, L), p, V, Z, precision=mg.MG_FP32)
    assert rel(r32[1], r_dense[1]) < 2e-5


def test_column_ranges_add_up_and_results_are_deterministic(mg):
    say("case: column ranges and determinism, L=10")
    V, Z, param, wires, P, U = build_case(10, "open", 7, seed=3)
    t = mg.Target.dense(U)
    p = mg.Program.create(10, P, param, wires)
    full = mg.evaluate(t, p, V, Z)
    again = mg.evaluate(t, p, V, Z)
    assert full[0] == again[0] and np.array_equal(full[1], again[1]) and np.array_equal(full[2], again[2])
    acc = [0, 0, 0]
    for b0, b1 in ((0, 100), (100, 777), (777, 1024)):
        part = mg.evaluate_columns(t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP64, b0, b1)
        acc = [acc[k] + part[k] for k in range(3)]
    assert abs(acc[0] - full[0]) < 1e-10 and rel(acc[1], full[1]) < 1e-12 and rel(acc[2], full[2]) < 1e-12


def test_hvp_linearity_and_symmetry_L12(mg):
    """Full L=12 size (4096 columns, depth-9 open brickwall, 50 gates): the HVP is
    linear in v and the Riemannian Hessian is symmetric, <w, H v> = <v, H w>; FP32
    sweeps agree with FP64; a column subset agrees with the oracle."""
    say("case: L=12 depth 9 open brickwall properties")
    rng = np.random.default_rng(12)
    L = 12
    param, wires, G = open_brickwall(L, 9)
    assert G == 50
    V = haar_matchgates(rng, G)
    Ut = haar_matchgates(rng, G)
    t = mg.Target.circuit(Ut, wires, L)
    p = mg.Program.create(L, G, param, wires)
    v = rng.standard_normal(8 * G)
    w = rng.standard_normal(8 * G)
    f, g, hv = mg.riemannian_gradient_hvp(t, p, V, v)
    _, _, hw = mg.riemannian_gradient_hvp(t, p, V, w)
    _, _, hvw = mg.riemannian_gradient_hvp(t, p, V, 2.0 * v - 3.0 * w)
    assert rel(hvw, 2.0 * hv - 3.0 * hw) < 1e-10
    assert abs(w @ hv - v @ hw) < 1e-9 * max(1.0, abs(w @ hv))
    f32, g32, hv32 = mg.riemannian_gradient_hvp(t, p, V, v, mg.MG_FP32)
    assert rel(g32, g) < 1e-4 and rel(hv32, hv) < 1e-4
    # oracle on 48 columns
    S = 1 << L
    Ud = np.zeros((S, S), np.complex128)
    Z = rng.standard_normal((G, 8)) + 1j * rng.standard_normal((G, 8))
    tp = np.arange(G, dtype=np.int32)
    for b in range(48):
        Ud[:, b] = orc.apply_program(Ut, tp, wires, L, np.eye(1, S, b).ravel())
    part = mg.evaluate_columns(t, p, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP64, 0, 48)
    fo, dgo, dho = orc.target_gradient_hvp(Ud, V, Z, param, wires, L, 0, 48)
    assert abs(part[0] - fo) / abs(fo) < 1e-11
    assert rel(part[1], dgo) < 1e-11 and rel(part[2], dho) < 1e-11


# ---------------- single-column entry points (gate.h, matchgate_brickwall.h) ----------------

def test_single_column_kernels_fixtures(mg):
    say("case: single-column gate kernels")
    d = load("test_apply_matchgate_cplx")
    V, psi = gate(d, "v."), cvec(d["psi"])
    for (i, j), key in (((2, 5), "chi1"), ((4, 1), "chi2"), ((3, 4), "chi3")):
        out = np.zeros(512, np.complex128)
        mg.lib.apply_matchgate(mg._p(V), i, j, ctypes.byref(mg.statevector(psi)), ctypes.byref(mg.statevector(out)))
        assert np.max(np.abs(out - cvec(d[key]))) < 1e-12
    d = load("test_apply_matchgate_backward_cplx")
    V, psi, dout = gate(d, "v."), cvec(d["psi"]), cvec(d["dpsi_out"])
    for i, j in ((2, 5), (3, 4), (4, 1)):
        dpsi = np.zeros(512, np.complex128)
        dV = np.zeros(8, np.complex128)
        sv_d = mg.statevector(dpsi)
        mg.lib.apply_matchgate_backward(mg._p(V), i, j, ctypes.byref(mg.statevector(psi)), ctypes.byref(mg.statevector(dout)),
                                        ctypes.byref(sv_d), mg._p(dV))
        rd, rg = orc.apply_gate_backward(V, i, j, 9, psi, dout)
        assert np.max(np.abs(sv_d._keep - rd)) < 1e-12 and np.max(np.abs(dV - rg)) < 1e-12
    d = load("test_apply_matchgate_placeholder_cplx")
    psi = cvec(d["psi"])
    emap = [5, 6, 9, 10, 0, 3, 12, 15]
    for (i, j), key in (((2, 5), "psi_out1"), ((5, 1), "psi_out2")):
        out = np.zeros((128, 8), np.complex128)
        sa = mg.statevector_array(out)
        mg.lib.apply_matchgate_placeholder(i, j, ctypes.byref(mg.statevector(psi)), ctypes.byref(sa))
        assert np.max(np.abs(sa._keep - cvec(d[key]).reshape(128, 16)[:, emap])) < 1e-12
        # the general two-qubit gate placeholder (test/test_gate.c:376-380, 399-403): all 16 states
        out16 = np.full((128, 16), np.nan + 0j)
        sa16 = mg.statevector_array(out16)
        mg.lib.apply_gate_placeholder(i, j, ctypes.byref(mg.statevector(psi)), ctypes.byref(sa16))
        assert np.max(np.abs(sa16._keep - cvec(d[key]).reshape(128, 16))) < 1e-12
    d = load("test_apply_matchgate_to_array_cplx")
    V = gate(d, "v.")
    a = d["psi"]
    arr = a[..., 0] + 1j * a[..., 1]
    for (i, j), key in (((2, 5), "chi1"), ((4, 1), "chi2"), ((3, 4), "chi3")):
        out = mg.statevector_array(np.zeros_like(arr))
        mg.lib.apply_matchgate_to_array(mg._p(V), i, j, ctypes.byref(mg.statevector_array(arr)), ctypes.byref(out))
        r = d[key]
        assert np.max(np.abs(out._keep - (r[..., 0] + 1j * r[..., 1]))) < 1e-12


def test_single_column_brickwall_backward_hessian_fixture(mg):
    say("case: single-column brickwall forward/backward/backward_hessian")
    d = load("test_matchgate_brickwall_unitary_backward_hessian_cplx")
    M, pm, psi, dout = gates(d, "v.", 4), perms(d, 4), cvec(d["psi"]), cvec(d["dpsi_out"])
    L = 6
    for nl in (2, 4):
        pm_i, arr = mg._perm_ptrs(pm[:nl])
        parr = ctypes.cast(arr, ctypes.c_void_p)
        cache = mg.QuantumCircuitCache()
        assert mg.lib.allocate_quantum_circuit_cache(L, nl * 3, ctypes.byref(cache)) == 0
        out = mg.statevector(np.zeros(64, np.complex128))
        assert mg.lib.matchgate_brickwall_unitary_forward(mg._p(M), nl, parr, ctypes.byref(mg.statevector(psi)),
                                                          ctypes.byref(cache), ctypes.byref(out)) == 0
        assert np.max(np.abs(out._keep - cvec(d[f"psi_out{nl - 1}"]))) < 1e-12
        dpsi = mg.statevector(np.zeros(64, np.complex128))
        dv = np.zeros((nl, 8), np.complex128)
        hess = np.zeros((8 * nl, 8 * nl), np.complex128)
        assert mg.lib.matchgate_brickwall_unitary_backward_hessian(mg._p(M), nl, parr, ctypes.byref(cache),
                                                                   ctypes.byref(mg.statevector(dout)), ctypes.byref(dpsi),
                                                                   mg._p(dv), mg._p(hess)) == 0
        param, wires = orc.brickwall_program(L, pm[:nl])
        c, _ = orc.program_forward(M[:nl], param, wires, L, psi)
        rdpsi, rdg, rh = orc.program_backward_hessian(M[:nl], param, wires, L, c, dout)
        assert np.max(np.abs(dpsi._keep - rdpsi)) < 1e-12
        assert rel(dv, rdg) < 1e-12 and rel(hess, rh) < 1e-12
        mg.lib.free_quantum_circuit_cache(ctypes.byref(cache))


# ---------------- trust region on the GPU path ----------------

def _hubbard_q8(nl):
    U = cvec(load("hubbard1d_q8_unitary_t0.25s_g1.50_init")["expiH"]).reshape(256, 256)
    d = load(f"hubbard1d_suzuki2_n{nl}_q8_init")
    v = np.asarray(d["vlist"]).reshape(nl, 8, 2)
    return U, v[..., 0] + 1j * v[..., 1], perms(d, nl)


def test_rtr_hmat_reproduces_reference_trajectory(mg):
    """optimize_matchgate_brickwall_circuit_hmat on the GPU reproduces the reference-C
    output file of examples/hubbard1d/hubbard1d_opt.c (L=8, 5 layers, 10 iterations)."""
    say("case: RTR hmat, Hubbard L=8 n5, 10 iterations")
    U, V, pm = _hubbard_q8(5)
    out = load("hubbard1d_suzuki2_n5_q8_iter10_opt")
    uf = mg.dense_linear_func(U)
    pm_i, arr = mg._perm_ptrs(pm)
    params = mg.rtr_default_params(5 * 16)
    f_iter = np.zeros(11)
    f_citer = np.zeros(11, np.complex128)
    vopt = np.zeros_like(V)
    mg.lib.optimize_matchgate_brickwall_circuit_hmat(uf, None, mg._p(V), 5, 8, ctypes.cast(arr, ctypes.c_void_p),
                                                     ctypes.byref(params), 10, mg._p(f_iter), mg._p(f_citer), mg._p(vopt))
    assert np.max(np.abs(f_iter - out["f_iter"])) < 1e-9
    vr = np.asarray(out["vlist_opt"])
    assert np.max(np.abs(vopt - (vr[..., 0] + 1j * vr[..., 1]))) < 1e-9
    assert np.max(np.abs(f_citer.real - f_iter)) < 1e-12


def test_rtr_hvp_matches_oracle(mg):
    say("case: RTR hvp, Hubbard L=8 n5, 10 iterations")
    U, V, pm = _hubbard_q8(5)
    t = mg.Target.dense(U)
    p = mg.Program.brickwall(8, pm)
    f_iter, f_citer, vopt, nh = mg.optimize(t, p, V, 10, hvp=True)
    param, wires = orc.brickwall_program(8, pm)
    fo, vo, nho = orc.rtr_optimize(U, V, param, wires, 8, 10, mode="hvp", tcg_maxiter=160)
    assert np.max(np.abs(f_iter - fo)) < 1e-9 and np.max(np.abs(vopt - vo)) < 1e-9
    # a truncated-CG stop decided by a residual at the tolerance may differ by one step
    assert abs(nh - nho) <= 2
    # FP32 sweeps with the last 3 iterations in FP64
    f32, _, v32, _ = mg.optimize(t, p, V, 10, hvp=True, precision=mg.MG_FP32, precision_switch=3)
    assert abs(f32[-1] - fo[-1]) < 1e-4 * abs(fo[-1])


def test_device_trust_region_matches_host_loop(mg):
    """mg_optimize with HVPs runs its trust region on the device, one HIP graph per iteration
    (truncated CG, retraction and ratio test in device kernels between the sweeps): the same
    iterates as the host loop over the same sweeps up to rounding, the same number of HVP sweeps,
    one graph launch per iteration (and one per two further steps of a longer truncated CG).  L=8 tied brickwall (the reference driver's inputs, FP64, LDS
    sweeps), L=12 open brickwall against a materialised circuit target (FP64 register-window
    sweeps; FP32 LDS sweeps with the last two iterations in FP64)."""
    def both(t, p, V, niter, prec=mg.MG_FP64, sw=0):
        res = []
        for mode in (0, -1):
            mg.set_device_trust_region(mode)
            res.append(mg.optimize(t, p, V, niter, hvp=True, precision=prec, precision_switch=sw) +
                       (mg.last_trust_region_graphs(),))
        mg.set_device_trust_region(-1)
        return res

    say("case: device trust region, Hubbard L=8 n5 (tied), FP64, 10 iterations")
    U, V, pm = _hubbard_q8(5)
    host, dev = both(mg.Target.dense(U), mg.Program.brickwall(8, pm), V, 10)
    # one graph per iteration, and one more for each truncated CG longer than a graph's two steps
    assert host[4] == 0 and 10 <= dev[4] <= 10 + dev[3] // 2
    assert np.max(np.abs(dev[0] - host[0])) < 1e-12 and np.max(np.abs(dev[2] - host[2])) < 1e-12
    assert np.max(np.abs(dev[1] - host[1])) < 1e-12 and dev[3] == host[3]
    rng = np.random.default_rng(41)
    L = 12
    param, wires, G = open_brickwall(L, 7)
    V = haar_matchgates(rng, G)
    tparam, twires, Gt = open_brickwall(L, 13)
    t = mg.Target.circuit(haar_matchgates(rng, Gt), twires, L)
    t.densify()
    p = mg.Program.create(L, G, param, wires)
    for prec, sw, tol in ((mg.MG_FP64, 0, 1e-11), (mg.MG_FP32, 2, 1e-9)):
        say(f"case: device trust region, L=12 open depth 7, precision {prec}, switch {sw}, 6 iterations")
        host, dev = both(t, p, V, 6, prec, sw)
        assert host[4] == 0 and 6 <= dev[4] <= 6 + dev[3] // 2
        assert np.max(np.abs(dev[0] - host[0])) < tol * np.max(np.abs(host[0]))
        assert np.max(np.abs(dev[2] - host[2])) < tol and dev[3] == host[3]


def test_device_trust_region_reuse(mg):
    """Mode 2 of the device trust region reuses what it evaluates: H eta of each ratio test is the
    truncated CG's sum of its steps' products (no HVP sweep along eta) and f at the trial point
    comes with its gradient, which starts the next iteration when the step is accepted (and is
    kept after a rejection).  The iterates equal the reference control flow's up to rounding,
    with exactly one HVP sweep per iteration fewer.  L=8 tied Hubbard (FP64) and L=12 open
    brickwall against a materialised circuit target (FP64; FP32 with a two-iteration switch)."""
    def both(t, p, V, niter, prec=mg.MG_FP64, sw=0):
        res = []
        for mode in (0, 2):
            mg.set_device_trust_region(mode)
            res.append(mg.optimize(t, p, V, niter, hvp=True, precision=prec, precision_switch=sw) +
                       (mg.last_trust_region_graphs(),))
        mg.set_device_trust_region(-1)
        return res

    say("case: device trust region with reuse, Hubbard L=8 n5 (tied), FP64, 10 iterations")
    U, V, pm = _hubbard_q8(5)
    host, dev = both(mg.Target.dense(U), mg.Program.brickwall(8, pm), V, 10)
    assert host[4] == 0 and dev[4] >= 10
    assert np.max(np.abs(dev[0] - host[0])) < 1e-10 and np.max(np.abs(dev[2] - host[2])) < 1e-10
    assert dev[3] == host[3] - 10
    rng = np.random.default_rng(41)
    L = 12
    param, wires, G = open_brickwall(L, 7)
    V = haar_matchgates(rng, G)
    tparam, twires, Gt = open_brickwall(L, 13)
    t = mg.Target.circuit(haar_matchgates(rng, Gt), twires, L)
    t.densify()
    p = mg.Program.create(L, G, param, wires)
    for prec, sw, tol in ((mg.MG_FP64, 0, 1e-9), (mg.MG_FP32, 2, 1e-4)):
        say(f"case: device trust region with reuse, L=12 open depth 7, precision {prec}, switch {sw}, 6 iterations")
        host, dev = both(t, p, V, 6, prec, sw)
        assert host[4] == 0 and dev[4] >= 6
        assert np.max(np.abs(dev[0] - host[0])) < tol * np.max(np.abs(host[0]))
        assert np.max(np.abs(dev[2] - host[2])) < tol and dev[3] == host[3] - 6


def test_device_reuse_matches_served_reuse(mg):
    """Two implementations of the reuse trust region side by side: mg_optimize on the device
    (mode 2: loop state, truncated CG and ratio test in device kernels, one graph per iteration)
    and a served job with tr_reuse (trust region on a host thread of the server, its evaluations
    through the server's lanes), both with FP32 sweeps and the last 3 iterations in FP64, on the
    same resident target: the same iterates to 1e-9 and the same HVP count.  L=12 open depth 7
    against a materialised circuit target, 8 iterations."""
    say("case: device reuse vs served tr_reuse, L=12 open depth 7, FP32 sweeps, switch 3, 8 iterations")
    rng = np.random.default_rng(43)
    L = 12
    param, wires, G = open_brickwall(L, 7)
    V = haar_matchgates(rng, G)
    tparam, twires, Gt = open_brickwall(L, 13)
    t = mg.Target.circuit(haar_matchgates(rng, Gt), twires, L)
    t.densify()
    p = mg.Program.create(L, G, param, wires)
    mg.set_device_trust_region(2)
    try:
        f_dev, _, v_dev, nh_dev = mg.optimize(t, p, V, 8, hvp=True, precision=mg.MG_FP32, precision_switch=3)
        assert mg.last_trust_region_graphs() >= 8
    finally:
        mg.set_device_trust_region(-1)
    job = {"program": p, "gates": V, "niter": 8, "target": t, "arrival": 0.0}
    r, _ = mg.serve([job], batched=False, precision=mg.MG_FP32, precision_switch=3, realtime=False, tr_reuse=True)
    assert r[0]["status"] == 0
    assert np.max(np.abs(f_dev - r[0]["f_iter"])) < 1e-9 * np.max(np.abs(f_dev))
    assert np.max(np.abs(v_dev - r[0]["gates"])) < 1e-9 and nh_dev == r[0]["nhvp"]


def test_ti_hessian_matches_oracle(mg):
    say("case: translation-invariant Hessian shortcut")
    rng = np.random.default_rng(2)
    L, nl = 6, 3
    shifts = [np.arange(L), np.roll(np.arange(L), -1)]
    pm = np.stack([shifts[k % 2] for k in range(nl)]).astype(np.int32)
    V = haar_matchgates(rng, nl)
    U = rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64))
    uf = mg.dense_linear_func(U)
    pm_i, arr = mg._perm_ptrs(pm)
    f = np.zeros(1, np.complex128)
    dv = np.zeros((nl, 8), np.complex128)
    hess = np.zeros((8 * nl, 8 * nl), np.complex128)
    assert mg.lib.matchgate_ti_brickwall_unitary_target_gradient_hessian(uf, None, mg._p(V), nl, L, ctypes.cast(arr, ctypes.c_void_p),
                                                                         mg._p(f), mg._p(dv), mg._p(hess)) == 0
    param, wires = orc.brickwall_program(L, pm)
    fo, dgo, ho = orc.target_gradient_hessian(U, V, param, wires, L, ti=True)
    assert abs(f[0] - fo) < 1e-11 and rel(dv, dgo) < 1e-11 and rel(hess, ho) < 1e-11
