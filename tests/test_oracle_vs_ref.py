"""The oracle (oracle/mg_oracle.c, a restatement) against the reference's own code
(oracle/_ref/libmg_ref.so: the reference's src/*.c compiled unchanged) on random inputs:
gate kernels, brickwall sweeps including the O(G^2) placeholder Hessian and both
translation-invariant variants, and the OpenMP target loops.  CPU only; skipped where the
reference library was not built (no reference checkout or HDF5 at build time)."""
import numpy as np
import pytest

import oracle as orc
from oracle import ref
from golden_io import random_perms

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (needs the reference checkout)")


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(1e-300, np.max(np.abs(b)))


def test_gate_kernels():
    rng = np.random.default_rng(7)
    L = 7
    for i, j in ((0, 1), (1, 0), (2, 5), (6, 3), (5, 6), (3, 0)):
        v = 0.5 * crandn(rng, 8)
        psi, dout = crandn(rng, 1 << L), crandn(rng, 1 << L)
        assert rel(orc.apply_gate(v, i, j, L, psi), ref.apply_matchgate(v, i, j, psi)) < 1e-15
        dpsi, dv = orc.apply_gate_backward(v, i, j, L, psi, dout)
        rdpsi, rdv = ref.apply_matchgate_backward(v, i, j, psi, dout)
        assert rel(dpsi, rdpsi) < 1e-15 and rel(dv, rdv) < 1e-14
        arr = crandn(rng, 1 << L, 8)
        assert rel(orc.apply_gate_backward_array(v, i, j, L, arr, dout),
                   ref.apply_matchgate_backward_array(v, i, j, arr, dout)) < 1e-14


@pytest.mark.parametrize("L,nl", [(6, 4), (8, 3)])
def test_brickwall_sweeps_and_placeholder_hessian(L, nl):
    rng = np.random.default_rng(L * 10 + nl)
    V = 0.5 * crandn(rng, nl, 8)  # not unitary, as the reference's fixtures
    pm = random_perms(rng, L, nl)
    psi, dout = crandn(rng, 1 << L), crandn(rng, 1 << L)
    param, wires = orc.brickwall_program(L, pm)
    cache, out = orc.program_forward(V, param, wires, L, psi)
    rout, rcache = ref.brickwall_forward(V, pm, psi)
    assert rel(out, rout) < 1e-14 and rel(cache, rcache.states()) < 1e-14
    rdpsi, rdv = ref.brickwall_backward(V, pm, rcache, dout)
    dpsi, dgi = orc.program_backward(V, param, wires, L, cache, dout)
    dv = np.zeros((nl, 8), np.complex128)
    np.add.at(dv, param, dgi)
    assert rel(dpsi, rdpsi) < 1e-14 and rel(dv, rdv) < 1e-13
    hd, hg, hh = orc.program_backward_hessian(V, param, wires, L, cache, dout, P=nl)
    rd, rg, rh = ref.brickwall_backward_hessian(V, pm, rcache, dout)
    assert rel(hd, rd) < 1e-14 and rel(hg, rg) < 1e-13 and rel(hh, rh) < 1e-13


@pytest.mark.parametrize("L", [6, 10])
def test_ti_sweeps(L):
    """The TI gradient sweep for odd L/2 (where the reference's two-buffer alternation
    chains the layers; for even L/2 it reads a buffer the previous layer did not write,
    so it is compared at odd L/2 only) and the TI placeholder Hessian."""
    rng = np.random.default_rng(L)
    nl = 3
    V = 0.5 * crandn(rng, nl, 8)
    shift = [np.arange(L), np.roll(np.arange(L), -1)]
    pm = np.stack([shift[k % 2] for k in range(nl)]).astype(np.int32)
    psi, dout = crandn(rng, 1 << L), crandn(rng, 1 << L)
    param, wires = orc.brickwall_program(L, pm)
    cache, _ = orc.program_forward(V, param, wires, L, psi)
    _, rcache = ref.brickwall_forward(V, pm, psi)
    dpsi, dv = orc.ti_brickwall_backward(V, pm, L, cache, dout)
    rdpsi, rdv = ref.ti_brickwall_backward(V, pm, rcache, dout)
    assert rel(dpsi, rdpsi) < 1e-14 and rel(dv, rdv) < 1e-13
    if L <= 8:
        hd, hg, hh = orc.ti_brickwall_backward_hessian(V, pm, L, cache, dout)
        rd, rg, rh = ref.ti_brickwall_backward_hessian(V, pm, rcache, dout)
        assert rel(hd, rd) < 1e-14 and rel(hg, rg) < 1e-13 and rel(hh, rh) < 1e-13


def test_target_loops():
    """Serial, OpenMP and TI target loops (src/matchgate_target.c:397-869) and the
    Riemannian projection (:877-973) with a dense map, Haar-unitary gates."""
    from golden_io import haar_matchgates

    rng = np.random.default_rng(3)
    L, nl = 6, 4
    V = haar_matchgates(rng, nl)
    pm = random_perms(rng, L, nl)
    U = crandn(rng, 1 << L, 1 << L)
    param, wires = orc.brickwall_program(L, pm)
    fo, go, ho = orc.target_gradient_hessian(U, V, param, wires, L)
    for parallel in (False, True):
        f, g, h = ref.brickwall_target_gradient_hessian(U, V, pm, parallel=parallel)
        assert abs(f - fo) / abs(fo) < 1e-13 and rel(g, go) < 1e-13 and rel(h, ho) < 1e-12
    assert abs(ref.brickwall_target(U, V, pm) - fo) / abs(fo) < 1e-13
    f, g = ref.brickwall_target_and_gradient(U, V, pm)
    assert abs(f - fo) / abs(fo) < 1e-13 and rel(g, go) < 1e-13
    fti, gti, hti = orc.target_gradient_hessian(U, V, param, wires, L, ti=True)
    rf, rg, rh = ref.brickwall_target_gradient_hessian(U, V, pm, ti=True)
    assert abs(rf - fti) / abs(fti) < 1e-13 and rel(rg, gti) < 1e-13 and rel(rh, hti) < 1e-12
    gv, H = orc.gradient_vector_hessian_matrix(V, go, ho)
    rf, rgv, rH = ref.brickwall_target_gradient_vector_hessian_matrix(U, V, pm)
    assert rel(rgv, gv) < 1e-12 and rel(rH, H) < 1e-12


def test_circuit_target_gradient():
    rng = np.random.default_rng(5)
    L, n = 7, 6
    gates = 0.5 * crandn(rng, n, 8)
    wires = np.concatenate([rng.choice(L, 2, replace=False) for _ in range(n)]).astype(np.int32)
    U = crandn(rng, 1 << L, 1 << L)
    fo, go = orc.target_gradient(U, gates, np.arange(n, dtype=np.int32), wires, L)
    f, g = ref.circuit_target_and_gradient(U, gates, wires)
    assert abs(f - fo) / abs(fo) < 1e-13 and rel(g, go) < 1e-13
