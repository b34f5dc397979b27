"""Pin the CPU oracle against the reference's own golden fixtures, finite
differences, brute-force dense matrices and closed forms.  Each test names the
reference test it mirrors (test/test_*.c)."""
import numpy as np
import pytest

import oracle as orc
from dense_ref import apply_dense_gate, fd_gradient, matchgate_matrix, program_matrix, target_f
from golden_io import cvec, gate, gates, haar_matchgates, load, perms, random_perms, scramble_target


def udist(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b)))


# ---------------- gate kernels (test/test_gate.c) ----------------

def test_apply_gate_fixture():
    d = load("test_apply_matchgate_cplx")  # test/test_gate.c:14-70, L=9
    V, psi = gate(d, "v."), cvec(d["psi"])
    for (i, j), key in (((2, 5), "chi1"), ((4, 1), "chi2"), ((3, 4), "chi3")):
        assert udist(orc.apply_gate(V, i, j, 9, psi), cvec(d[key])) < 1e-12


def test_apply_gate_matches_dense_embedding():
    rng = np.random.default_rng(7)
    L = 5
    g = rng.standard_normal(8) + 1j * rng.standard_normal(8)
    psi = rng.standard_normal(1 << L) + 1j * rng.standard_normal(1 << L)
    for i, j in ((0, 1), (1, 0), (0, 4), (4, 0), (2, 3), (3, 1)):
        assert udist(orc.apply_gate(g, i, j, L, psi), apply_dense_gate(matchgate_matrix(g), i, j, L, psi)) < 1e-13


def test_apply_gate_backward_fd():
    d = load("test_apply_matchgate_backward_cplx")  # test/test_gate.c:123-205, L=9
    V, psi, dout = gate(d, "v."), cvec(d["psi"]), cvec(d["dpsi_out"])
    L = 9
    for i, j in ((2, 5), (3, 4), (4, 1)):
        dpsi, dV = orc.apply_gate_backward(V, i, j, L, psi, dout)
        # the map is linear in psi and in V: dpsi = M^T dout, dV[e] = dout . (dM/dV_e) psi
        M = matchgate_matrix(V)
        ref = np.stack([apply_dense_gate(M, i, j, L, np.eye(1 << L)[:, b]) for b in range(1 << L)], axis=1)
        assert udist(dpsi, ref.T @ dout) < 1e-12
        for e in range(8):
            E = np.zeros(8, np.complex128); E[e] = 1
            assert abs(dV[e] - dout @ orc.apply_gate(E, i, j, L, psi)) < 1e-12


def test_apply_gate_backward_array_fixture():
    d = load("test_apply_matchgate_backward_array_cplx")  # test/test_gate.c:209-282
    V, dout = gate(d, "v."), cvec(d["dpsi_out"])
    arr = d["psi_array"]
    psi = arr[..., 0] + 1j * arr[..., 1]
    for i, j in ((2, 5), (3, 4), (4, 1), (7, 6), (5, 8)):
        dgs = orc.apply_gate_backward_array(V, i, j, 9, psi, dout)
        for k in range(8):
            _, dref = orc.apply_gate_backward(V, i, j, 9, cvec(d[f"psi{k}"]), dout)
            assert udist(dgs[k], dref) < 1e-12


def test_apply_gate_to_array_fixture():
    d = load("test_apply_matchgate_to_array_cplx")  # test/test_gate.c:286-350, L=6, 5 states
    V = gate(d, "v.")
    a = d["psi"]
    psi = a[..., 0] + 1j * a[..., 1]
    for (i, j), key in (((2, 5), "chi1"), ((4, 1), "chi2"), ((3, 4), "chi3")):
        r = d[key]
        assert udist(orc.apply_gate_to_array(V, i, j, 6, psi), r[..., 0] + 1j * r[..., 1]) < 1e-12


def test_apply_gate_placeholder_fixture():
    d = load("test_apply_matchgate_placeholder_cplx")  # test/test_gate.c:352-410, L=7
    psi = cvec(d["psi"])
    emap = [5, 6, 9, 10, 0, 3, 12, 15]  # matchgate entry -> 4x4 entry (test_gate.c:382)
    for (i, j), key in (((2, 5), "psi_out1"), ((5, 1), "psi_out2")):
        out = orc.apply_gate_placeholder(i, j, 7, psi)
        ref = cvec(d[key]).reshape(1 << 7, 16)
        assert udist(out, ref[:, emap]) < 1e-12


# ---------------- brickwall sweeps (test/test_matchgate_brickwall.c) ----------------

def test_apply_brickwall_fixture():
    d = load("test_apply_matchgate_brickwall_unitary_cplx")  # :23-95, L=8
    M, pm, psi = gates(d, "v.", 4), perms(d, 4), cvec(d["psi"])
    for nl in (3, 4):
        param, wires = orc.brickwall_program(8, pm[:nl])
        assert udist(orc.apply_program(M, param, wires, 8, psi), cvec(d[f"chi{nl - 3}"])) < 1e-12


def test_brickwall_forward_backward_fixture():
    d = load("test_matchgate_brickwall_unitary_backward_cplx")  # :153-280, L=6
    M, pm, psi, dout = gates(d, "v.", 4), perms(d, 4), cvec(d["psi"]), cvec(d["dpsi_out"])
    L = 6
    for nl in (3, 4):
        param, wires = orc.brickwall_program(L, pm[:nl])
        cache, out = orc.program_forward(M[:nl], param, wires, L, psi)
        assert udist(out, cvec(d[f"psi_out{nl - 3}"])) < 1e-12
        dpsi, dgi = orc.program_backward(M[:nl], param, wires, L, cache, dout)
        W = program_matrix(M[:nl], param, wires, L)
        assert udist(dpsi, W.T @ dout) < 1e-11
        dV = np.zeros((nl, 8), np.complex128)
        np.add.at(dV, param, dgi)
        # FD of  dout . W(V) psi  w.r.t. the tied layer gates
        h = 1e-5
        for p in range(nl):
            for e in range(8):
                Mp = M[:nl].copy(); Mp[p, e] += h
                Mm = M[:nl].copy(); Mm[p, e] -= h
                fd = (dout @ orc.apply_program(Mp, param, wires, L, psi) - dout @ orc.apply_program(Mm, param, wires, L, psi)) / (2 * h)
                assert abs(dV[p, e] - fd) < 1e-8


def test_brickwall_backward_hessian_fixture():
    d = load("test_matchgate_brickwall_unitary_backward_hessian_cplx")  # :315-480, L=6
    M, pm, psi, dout = gates(d, "v.", 4), perms(d, 4), cvec(d["psi"]), cvec(d["dpsi_out"])
    Z = gates(d, "z.", 4)
    L = 6
    for nl in (1, 2, 3, 4):
        param, wires = orc.brickwall_program(L, pm[:nl])
        cache, out = orc.program_forward(M[:nl], param, wires, L, psi)
        assert udist(out, cvec(d[f"psi_out{nl - 1}"])) < 1e-12
        dpsi, dgrad, hess = orc.program_backward_hessian(M[:nl], param, wires, L, cache, dout)
        assert udist(hess, hess.T) < 1e-12
        # hess . Z against a central difference of the gradient along Z
        h = 1e-5

        def grad_at(Mx):
            c, _ = orc.program_forward(Mx, param, wires, L, psi)
            _, dgi = orc.program_backward(Mx, param, wires, L, c, dout)
            g = np.zeros((nl, 8), np.complex128)
            np.add.at(g, param, dgi)
            return g.reshape(-1)

        fd = (grad_at(M[:nl] + h * Z[:nl]) - grad_at(M[:nl] - h * Z[:nl])) / (2 * h)
        # central-difference truncation error scales with the entries (|fd| ~ 1e2 here)
        assert udist(hess @ Z[:nl].reshape(-1), fd) < 1e-8 * max(1.0, np.abs(fd).max())


# ---------------- targets (test/test_matchgate_target.c) ----------------

def test_circuit_target_fixture():
    d = load("test_matchgate_circuit_unitary_target_cplx")  # :70-121, L=7, 5 gates
    G = gates(d, "G.", 5)
    wires = np.asarray(d["wires"], np.int32).reshape(-1)
    U = scramble_target(7)
    f = orc.target(U, G, np.arange(5, dtype=np.int32), wires, 7)
    assert abs(f - cvec(d["f"])[0]) < 1e-12


def test_brickwall_target_fixture():
    d = load("test_matchgate_brickwall_unitary_target_cplx")  # :123-200, L=8
    V, pm = gates(d, "V.", 5), perms(d, 5)
    U = scramble_target(8)
    for k, nl in enumerate((4, 5)):
        param, wires = orc.brickwall_program(8, pm[:nl])
        f = orc.target(U, V[:nl], param, wires, 8)
        assert abs(f - cvec(d[f"f{k}"])[0]) < 1e-12


def test_target_gradient_hessian_fixture():
    d = load("test_matchgate_brickwall_unitary_target_gradient_hessian_cplx")  # :219-380, L=6
    V, pm, Z = gates(d, "V.", 5), perms(d, 5), gates(d, "Z.", 5)
    U = scramble_target(6)
    emap = [5, 6, 9, 10, 0, 3, 12, 15]
    for k, nl in enumerate((4, 5)):
        param, wires = orc.brickwall_program(6, pm[:nl])
        f, dg, hess = orc.target_gradient_hessian(U, V[:nl], param, wires, 6)
        assert abs(f - cvec(d[f"f{k}"])[0]) < 1e-12
        assert udist(dg, gates(d, f"dVlist{k}.", nl)) < 1e-12
        assert udist(hess, hess.T) < 1e-12
        # the fixture Hessian is in the general 4x4 layout (16 entries per gate)
        hr = d[f"hess{k}"]
        hr = (hr[..., 0] + 1j * hr[..., 1]).reshape(nl, 16, nl, 16)[:, emap][:, :, :, emap].reshape(8 * nl, 8 * nl)
        assert udist(hess, hr) < 1e-11
        # and the brute-force dense gradient
        assert udist(dg, fd_gradient(U, V[:nl], param, wires, 6)) < 1e-7


def test_gradient_vector_hessian_matrix_fixture():
    d = load("test_matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix_cplx")  # :390-500
    V, pm = gates(d, "V.", 5), perms(d, 5)
    U = scramble_target(6)
    param, wires = orc.brickwall_program(6, pm)
    f, dg, hess = orc.target_gradient_hessian(U, V, param, wires, 6)
    g, H = orc.gradient_vector_hessian_matrix(V, dg, hess)
    assert abs(f - cvec(d["f"])[0]) < 1e-12
    assert udist(g, d["grad"]) < 1e-14
    assert udist(H, H.T) < 1e-14
    assert udist(H, np.asarray(d["H"]).reshape(40, 40)) < 1e-13


def test_hvp_matches_fixture_hessian():
    """H v by forward-over-reverse equals the reference's Riemannian Hessian matrix
    (Haar-unitary gates) for random directions."""
    d = load("test_matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix_cplx")
    V, pm = gates(d, "V.", 5), perms(d, 5)
    H = np.asarray(d["H"]).reshape(40, 40)
    U = scramble_target(6)
    param, wires = orc.brickwall_program(6, pm)
    rng = np.random.default_rng(3)
    for _ in range(3):
        v = rng.standard_normal(40)
        hv = orc.hvp(U, V, param, wires, 6, v)
        assert udist(hv, H @ v) < 1e-12 * max(1, np.abs(H @ v).max())
    f, dg, _ = orc.target_gradient_hvp(U, V, np.zeros_like(V), param, wires, 6)
    assert udist(orc.riemannian_gradient(V, dg), d["grad"]) < 1e-14


def test_hvp_complex_matches_oracle_hessian_nonunitary():
    """D^2 f[Z] equals hess . Z for arbitrary (non-unitary) gates and complex Z."""
    d = load("test_matchgate_brickwall_unitary_target_gradient_hessian_cplx")
    V, pm, Z = gates(d, "V.", 5), perms(d, 5), gates(d, "Z.", 5)
    U = scramble_target(6)
    param, wires = orc.brickwall_program(6, pm)
    f, dg, hess = orc.target_gradient_hessian(U, V, param, wires, 6)
    f2, dg2, dhz = orc.target_gradient_hvp(U, V, Z, param, wires, 6)
    assert abs(f - f2) < 1e-12 and udist(dg, dg2) < 1e-12
    assert udist(dhz.reshape(-1), hess @ Z.reshape(-1)) < 1e-11


def test_untied_program_reduces_to_tied():
    """Per-gate parameters, summed per layer, reproduce the tied brickwall."""
    rng = np.random.default_rng(11)
    L, nl = 6, 3
    pm = random_perms(rng, L, nl)
    V = haar_matchgates(rng, nl)
    U = scramble_target(L)
    param, wires = orc.brickwall_program(L, pm)
    G = len(param)
    Vu = V[param]
    Zt = rng.standard_normal((nl, 8)) + 1j * rng.standard_normal((nl, 8))
    ft, dgt, dhzt = orc.target_gradient_hvp(U, V, Zt, param, wires, L)
    fu, dgu, dhzu = orc.target_gradient_hvp(U, Vu, Zt[param], np.arange(G, dtype=np.int32), wires, L)
    assert abs(ft - fu) < 1e-12
    sg = np.zeros_like(dgt); np.add.at(sg, param, dgu)
    sh = np.zeros_like(dhzt); np.add.at(sh, param, dhzu)
    assert udist(sg, dgt) < 1e-12 and udist(sh, dhzt) < 1e-12
    # Hessian blocks sum the same way
    _, _, Ht = orc.target_gradient_hessian(U, V, param, wires, L)
    _, _, Hu = orc.target_gradient_hessian(U, Vu, np.arange(G, dtype=np.int32), wires, L)
    Ared = np.zeros((nl, 8, nl, 8), np.complex128)
    Hu4 = Hu.reshape(G, 8, G, 8)
    for a in range(G):
        for b in range(G):
            Ared[param[a], :, param[b], :] += Hu4[a, :, b, :]
    assert udist(Ared.reshape(8 * nl, 8 * nl), Ht) < 1e-11


def test_closed_forms():
    """f(W = U) = -2^L for unitary U; f(U = I) = -Tr W; ||W - U||^2 = 2 f + 2 * 2^L."""
    rng = np.random.default_rng(5)
    L, nl = 6, 3
    pm = random_perms(rng, L, nl)
    V = haar_matchgates(rng, nl)
    param, wires = orc.brickwall_program(L, pm)
    W = program_matrix(V, param, wires, L)
    assert abs(orc.target(W, V, param, wires, L) + (1 << L)) < 1e-11
    assert abs(orc.target(np.eye(1 << L), V, param, wires, L) + np.trace(W)) < 1e-11
    U = program_matrix(haar_matchgates(rng, nl), param, wires, L)
    f = orc.target(U, V, param, wires, L)
    assert abs(f - target_f(U, V, param, wires, L)) < 1e-11
    assert abs(np.linalg.norm(W - U) ** 2 - (2 * f.real + 2 * (1 << L))) < 1e-10


def test_ti_hessian_equals_full_for_translation_invariant_setup():
    """TI shortcut (reference matchgate_brickwall.c:308-435) equals the full Hessian
    for a translation-invariant target and nearest-neighbour periodic brickwall."""
    rng = np.random.default_rng(2)
    L, nl = 6, 3
    shifts = [np.arange(L), np.roll(np.arange(L), -1)]
    pm = np.stack([shifts[k % 2] for k in range(nl)]).astype(np.int32)
    V = haar_matchgates(rng, nl)
    Ut = haar_matchgates(rng, 2)
    up = np.stack([shifts[k % 2] for k in range(2)]).astype(np.int32)
    tp, tw = orc.brickwall_program(L, up)
    U = program_matrix(Ut, tp, tw, L)
    param, wires = orc.brickwall_program(L, pm)
    f1, g1, h1 = orc.target_gradient_hessian(U, V, param, wires, L)
    f2, g2, h2 = orc.target_gradient_hessian(U, V, param, wires, L, ti=True)
    assert abs(f1 - f2) < 1e-12 and udist(g1, g2) < 1e-12
    assert udist(h1, h2) < 1e-10


# ---------------- algebra (test/test_matchgate.c) ----------------

def test_algebra_fixtures():
    d = load("test_multiply_matchgate_cplx")
    assert udist(orc.multiply(gate(d, "a."), gate(d, "b.")), gate(d, "c.")) < 1e-12
    d = load("test_project_tangent_matchgate_cplx")
    assert udist(orc.project_tangent(gate(d, "u."), gate(d, "z.")), gate(d, "p.")) < 1e-12
    d = load("test_polar_matchgate_factor_cplx")
    assert udist(orc.polar(gate(d, "a.")), gate(d, "u.")) < 1e-12
    d = load("test_real_to_tangent_matchgate_cplx")
    v, t = gate(d, "v."), gate(d, "t.")  # test/test_matchgate.c:150-208
    r = orc.tangent_to_real(v, t)
    assert udist(orc.project_tangent(v, t), orc.real_to_tangent(r, v)) < 1e-14


def test_algebra_invariants():
    rng = np.random.default_rng(1)
    r = rng.standard_normal(8)
    w = orc.real_to_antisymm(r)
    assert udist(orc.antisymm_to_real(w), r) < 1e-15
    for blk in (w[:4].reshape(2, 2), w[4:].reshape(2, 2)):
        assert udist(blk, -blk.conj().T) < 1e-15
    a = rng.standard_normal(8) + 1j * rng.standard_normal(8)
    u = orc.polar(a)
    for blk, ab in ((u[:4].reshape(2, 2), a[:4].reshape(2, 2)), (u[4:].reshape(2, 2), a[4:].reshape(2, 2))):
        X, s, Yh = np.linalg.svd(ab)
        assert udist(blk, X @ Yh) < 1e-12
    ret, inv = orc.inverse_mat2x2(a[:4])
    assert ret == 0 and udist(inv.reshape(2, 2) @ a[:4].reshape(2, 2), np.eye(2)) < 1e-12
    assert orc.inverse_mat2x2(np.zeros(4))[0] == -1
    V = haar_matchgates(rng, 1)[0]
    vr = orc.retract(V, 0.1 * rng.standard_normal(8))
    for blk in (vr[:4].reshape(2, 2), vr[4:].reshape(2, 2)):
        assert udist(blk.conj().T @ blk, np.eye(2)) < 1e-13


# ---------------- trust region (examples/hubbard1d, reference-C outputs) ----------------

def _hubbard_q8(nl):
    U = cvec(load("hubbard1d_q8_unitary_t0.25s_g1.50_init")["expiH"]).reshape(256, 256)
    d = load(f"hubbard1d_suzuki2_n{nl}_q8_init")
    v = np.asarray(d["vlist"]).reshape(nl, 8, 2)
    V = v[..., 0] + 1j * v[..., 1]
    return U, V, perms(d, nl)


@pytest.mark.parametrize("nl", [5])
def test_rtr_hmat_matches_reference_trajectory(nl):
    """10 RTR iterations (explicit Hessian, reference trust_region.c:235) reproduce
    the reference-C output file (MKL, 112 OpenMP threads) of hubbard1d_opt.c."""
    U, V, pm = _hubbard_q8(nl)
    out = load(f"hubbard1d_suzuki2_n{nl}_q8_iter10_opt")
    param, wires = orc.brickwall_program(8, pm)
    f_iter, vopt, _ = orc.rtr_optimize(U, V, param, wires, 8, 10, mode="hmat", tcg_maxiter=2 * 16 * nl)
    assert udist(f_iter, out["f_iter"]) < 1e-9
    vr = np.asarray(out["vlist_opt"])
    assert udist(vopt, vr[..., 0] + 1j * vr[..., 1]) < 1e-9


def test_rtr_hvp_matches_hmat():
    U, V, pm = _hubbard_q8(5)
    param, wires = orc.brickwall_program(8, pm)
    f1, v1, _ = orc.rtr_optimize(U, V, param, wires, 8, 10, mode="hmat", tcg_maxiter=160)
    f2, v2, nh = orc.rtr_optimize(U, V, param, wires, 8, 10, mode="hvp", tcg_maxiter=160)
    assert udist(f1, f2) < 1e-9 and udist(v1, v2) < 1e-9
    assert nh > 10
