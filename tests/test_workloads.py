"""Synthetic workloads (CPU): the XXZ Trotter target of the L=16 instance against the
dense matrix exponential, the Fermi-Hubbard target's symmetries, the batched and served
generators' shapes, and the oracle's program-target entry against its dense-target one."""
import numpy as np
import scipy.linalg as sl

import oracle as orc
from dense_ref import program_matrix
from rqcopt_mg.workloads import (batched_instances, brickwall_job, large_instance, matchgate_from_matrix,
                                 served_stream, xxz_bond_matrix, xxz_hamiltonian, xxz_trotter_target)


def test_xxz_layer_is_exact_and_trotter_is_second_order():
    L = 6
    g, w, h = xxz_trotter_target(L, 50, 1.0, seed=3)
    assert len(g) == 51 * 3 + 50 * 2  # 2 steps + 1 merged layers of 3 / 2 bonds
    E = sl.expm(-1j * xxz_hamiltonian(L, h))
    errs = []
    for steps in (25, 50, 100):
        gs, ws, _ = xxz_trotter_target(L, steps, 1.0, seed=3)
        errs.append(np.max(np.abs(program_matrix(gs, np.arange(len(gs)), ws, L) - E)))
    assert errs[1] < 2e-3 and 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5
    # one layer of commuting bonds is exact: exp(-i tau H_odd) with the odd bonds' field shares
    tau = 0.37
    nb = np.array([2.0] * L); nb[0] = nb[-1] = 1.0
    Ho = np.zeros((1 << L, 1 << L), np.complex128)
    gates, wires = [], []
    for a in range(1, L - 1, 2):
        hb = np.zeros(L); hb[a] = h[a] / nb[a]; hb[a + 1] = h[a + 1] / nb[a + 1]
        Ho += bond_dense(L, a, hb)
        gates.append(matchgate_from_matrix(xxz_bond_matrix(tau, hb[a], hb[a + 1])))
        wires += [a, a + 1]
    U = program_matrix(np.stack(gates), np.arange(len(gates)), np.asarray(wires), L)
    assert np.max(np.abs(U - sl.expm(-1j * tau * Ho))) < 1e-13


def bond_dense(L, a, hb):
    """(XX + YY)/2 + 2 ZZ on bond (a, a+1) plus the fields hb (dense, qubit 0 = MSB)."""
    S = 1 << L
    idx = np.arange(S)
    ba, bb = (idx >> (L - 1 - a)) & 1, (idx >> (L - 2 - a)) & 1
    H = np.zeros((S, S), np.complex128)
    H[idx, idx] = 2.0 * (1 - 2.0 * ba) * (1 - 2.0 * bb) + hb[a] * (1 - 2.0 * ba) + hb[a + 1] * (1 - 2.0 * bb)
    flip = ba != bb
    s2 = idx ^ (1 << (L - 1 - a)) ^ (1 << (L - 2 - a))
    H[s2[flip], idx[flip]] += 1.0
    return H


def test_large_instance_shape():
    w = large_instance()
    assert (w["L"], w["G"], w["Gt"]) == (16, 83, 758)
    assert np.all(np.abs(w["fields"]) <= 1.0)
    # every target gate is unitary (center and corner blocks)
    for g in w["target_gates"][:20]:
        for blk in (g[:4].reshape(2, 2), g[4:].reshape(2, 2)):
            assert np.allclose(blk.conj().T @ blk, np.eye(2), atol=1e-13)


def test_generators_are_seeded_and_shaped():
    a, b = batched_instances(64, seed=0), batched_instances(64, seed=0)
    assert all(np.array_equal(x["gates"], y["gates"]) for x, y in zip(a, b))
    assert {j["L"] for j in a} <= {8, 10, 12, 14} and all(j["depth"] % 2 == 1 for j in a)
    assert all(j["Gt"] > j["G"] for j in a)
    s = served_stream(50, seed=1)
    t = np.array([j["arrival"] for j in s])
    assert np.all(np.diff(t) > 0) and 0.1 < t[-1] < 0.5  # Poisson, mean 200 jobs/s
    assert all(5 <= j["niter"] <= 50 for j in s)


def test_oracle_program_target_equals_dense_target():
    rng = np.random.default_rng(5)
    job = brickwall_job(rng, 6, 3, 1)
    U = program_matrix(job["target_gates"], np.arange(job["Gt"]), job["target_wires"], 6)
    Z = rng.standard_normal((job["G"], 8)) + 1j * rng.standard_normal((job["G"], 8))
    d = orc.target_gradient_hvp(U, job["gates"], Z, job["param"], job["wires"], 6, 3, 40)
    p = orc.target_gradient_hvp_program(job["target_gates"], job["target_wires"], job["gates"], Z, job["param"],
                                        job["wires"], 6, 3, 40)
    for x, y in zip(d, p):
        assert np.max(np.abs(np.asarray(x) - np.asarray(y))) < 1e-12 * max(1.0, np.max(np.abs(np.asarray(x))))


def test_fermi_hubbard_target_matches_jordan_wigner_expm():
    """The L=12 bench target's generator (sector-block eigendecomposition, bit arithmetic)
    against exp(-iHt) of H built independently from Jordan-Wigner fermion operators
    (c_q = Z_0 ... Z_{q-1} a_q, qubit 0 most significant; spin-up of site j on qubit 2j,
    spin-down on 2j+1), for 3 sites."""
    from rqcopt_mg.workloads import fermi_hubbard_1d_unitary

    nsites, J, Uc, t = 3, 1.0, 4.0, 0.25
    L = 2 * nsites
    Zm, Im = np.diag([1.0, -1.0]), np.eye(2)
    am = np.array([[0.0, 1.0], [0.0, 0.0]])  # |1> -> |0>

    def op(q):
        mats = [Zm] * q + [am] + [Im] * (L - q - 1)
        out = mats[0]
        for m in mats[1:]:
            out = np.kron(out, m)
        return out

    c = [op(q) for q in range(L)]
    n = [ci.T @ ci for ci in c]
    H = np.zeros((1 << L, 1 << L))
    for j in range(nsites - 1):
        for s in range(2):
            a, b = 2 * j + s, 2 * j + 2 + s
            H += -J * (c[a].T @ c[b] + c[b].T @ c[a])
    for j in range(nsites):
        H += Uc * n[2 * j] @ n[2 * j + 1]
    W = fermi_hubbard_1d_unitary(nsites, J=J, U=Uc, t=t)
    assert np.max(np.abs(W - sl.expm(-1j * t * H))) < 1e-12
