"""ctypes binding of oracle/_ref/libmg_ref.so: the reference's own kernel, sweep and target
code (test and baseline infrastructure only).

The library is built by ``make -C oracle ref`` from the reference's sources where they lie
(its own files compiled unchanged, plus ``oracle/ref_glue.c``); it is absent where the
reference checkout or HDF5 is missing, and then ``available()`` is False and the tests that
use it skip.  Gates are complex128 arrays of shape (n, 8), states complex128 vectors of
length 2^L, exactly as in the oracle's binding.
"""
import ctypes
import os

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "_ref", "libmg_ref.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


class Statevector(ctypes.Structure):
    _fields_ = [("data", _P), ("nqubits", _I)]


class StatevectorArray(ctypes.Structure):
    _fields_ = [("data", _P), ("nqubits", _I), ("nstates", _I)]


class CircuitCache(ctypes.Structure):
    _fields_ = [("psi_list", ctypes.POINTER(Statevector)), ("nqubits", _I), ("ngates", _I)]


class DenseMap(ctypes.Structure):
    _fields_ = [("U", _P), ("n", ctypes.c_long)]


class BrickwallMap(ctypes.Structure):
    _fields_ = [("ulist", _P), ("ulayers", _I), ("uperms", _P)]


_lib = None
if os.path.exists(LIB_PATH):
    try:
        _lib = ctypes.CDLL(LIB_PATH)
    except OSError:
        _lib = None


def available():
    return _lib is not None


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _p(a):
    return a.ctypes.data_as(_P)


def _sv(a):
    return Statevector(a.ctypes.data, int(np.log2(a.shape[0])))


def _perm_ptrs(perms):
    perms = np.ascontiguousarray(perms, dtype=np.int32)
    arr = (_P * perms.shape[0])(*[perms[k].ctypes.data for k in range(perms.shape[0])])
    return perms, arr


def _fn(name):
    f = getattr(_lib, name)
    f.restype = _I
    return f


def _vfn(name):
    f = getattr(_lib, name)
    f.restype = None
    return f


# ---- gate kernels (src/gate.c) ----

def apply_matchgate(v, i, j, psi):
    psi = _c(psi)
    out = np.zeros_like(psi)
    _vfn("apply_matchgate")(_p(_c(v)), i, j, ctypes.byref(_sv(psi)), ctypes.byref(_sv(out)))
    return out


def apply_matchgate_backward(v, i, j, psi, dout):
    psi, dout = _c(psi), _c(dout)
    dpsi = np.zeros_like(psi)
    dv = np.zeros(8, np.complex128)
    _vfn("apply_matchgate_backward")(_p(_c(v)), i, j, ctypes.byref(_sv(psi)), ctypes.byref(_sv(dout)),
                                     ctypes.byref(_sv(dpsi)), _p(dv))
    return dpsi, dv


def apply_matchgate_backward_array(v, i, j, arr, dout):
    """arr: (2^L, nstates) interleaved states; returns (nstates, 8)"""
    arr, dout = _c(arr), _c(dout)
    S, ns = arr.shape
    sa = StatevectorArray(arr.ctypes.data, int(np.log2(S)), ns)
    dv = np.zeros((ns, 8), np.complex128)
    _vfn("apply_matchgate_backward_array")(_p(_c(v)), i, j, ctypes.byref(sa), ctypes.byref(_sv(dout)), _p(dv))
    return dv


# ---- sweeps (src/matchgate_brickwall.c, src/quantum_circuit.c) ----

class _Cache:
    def __init__(self, L, ngates):
        self.c = CircuitCache()
        if _fn("allocate_quantum_circuit_cache")(L, ngates, ctypes.byref(self.c)) < 0:
            raise MemoryError("allocate_quantum_circuit_cache failed")
        self.L, self.G = L, ngates

    def states(self):
        S = 1 << self.L
        return np.stack([np.ctypeslib.as_array(ctypes.cast(self.c.psi_list[k].data, ctypes.POINTER(ctypes.c_double)),
                                               shape=(2 * S,)).view(np.complex128).copy() for k in range(self.G)])

    def __del__(self):
        if _lib is not None and self.c.psi_list:
            _vfn("free_quantum_circuit_cache")(ctypes.byref(self.c))


def brickwall_forward(vlist, perms, psi):
    """matchgate_brickwall_unitary_forward: (psi_out, cache object)"""
    vlist, psi = _c(vlist), _c(psi)
    L = int(np.log2(psi.shape[0]))
    pm, arr = _perm_ptrs(perms)
    cache = _Cache(L, pm.shape[0] * (L // 2))
    out = np.zeros_like(psi)
    if _fn("matchgate_brickwall_unitary_forward")(_p(vlist), pm.shape[0], ctypes.cast(arr, _P), ctypes.byref(_sv(psi)),
                                                  ctypes.byref(cache.c), ctypes.byref(_sv(out))) < 0:
        raise RuntimeError("matchgate_brickwall_unitary_forward failed")
    return out, cache


def _backward(name, vlist, perms, cache, dout, hessian):
    vlist, dout = _c(vlist), _c(dout)
    pm, arr = _perm_ptrs(perms)
    nl = pm.shape[0]
    dpsi = np.zeros_like(dout)
    dv = np.zeros((nl, 8), np.complex128)
    args = [_p(vlist), nl, ctypes.cast(arr, _P), ctypes.byref(cache.c), ctypes.byref(_sv(dout)), ctypes.byref(_sv(dpsi)),
            _p(dv)]
    hess = None
    if hessian:
        hess = np.zeros((8 * nl, 8 * nl), np.complex128)
        args.append(_p(hess))
    if _fn(name)(*args) < 0:
        raise RuntimeError(name + " failed")
    return dpsi, dv, hess


def brickwall_backward(vlist, perms, cache, dout):
    return _backward("matchgate_brickwall_unitary_backward", vlist, perms, cache, dout, False)[:2]


def brickwall_backward_hessian(vlist, perms, cache, dout):
    return _backward("matchgate_brickwall_unitary_backward_hessian", vlist, perms, cache, dout, True)


def ti_brickwall_backward(vlist, perms, cache, dout):
    return _backward("matchgate_ti_brickwall_unitary_backward", vlist, perms, cache, dout, False)[:2]


def ti_brickwall_backward_hessian(vlist, perms, cache, dout):
    return _backward("matchgate_ti_brickwall_unitary_backward_hessian", vlist, perms, cache, dout, True)


# ---- targets (src/matchgate_target.c) with a dense map U ----

def _dense_map(U):
    U = _c(U)
    m = DenseMap(U.ctypes.data, U.shape[0])
    m._keep = U
    return m, ctypes.cast(_lib.ref_dense_ufunc, _P)


def brickwall_target_gradient_hessian(U, vlist, perms, ti=False, parallel=True):
    """f, per-layer complex gradient (nl, 8), complex Hessian (8 nl)^2"""
    vlist = _c(vlist)
    m, uf = _dense_map(U)
    L = int(np.log2(m.n))
    pm, arr = _perm_ptrs(perms)
    nl = pm.shape[0]
    f = np.zeros(1, np.complex128)
    dv = np.zeros((nl, 8), np.complex128)
    hess = np.zeros((8 * nl, 8 * nl), np.complex128)
    name = ("matchgate_ti_brickwall_unitary_target_gradient_hessian" if ti else
            "parallel_matchgate_brickwall_unitary_target_gradient_hessian" if parallel else
            "matchgate_brickwall_unitary_target_gradient_hessian")
    if _fn(name)(uf, ctypes.byref(m), _p(vlist), nl, L, ctypes.cast(arr, _P), _p(f), _p(dv), _p(hess)) < 0:
        raise RuntimeError(name + " failed")
    return f[0], dv, hess


def brickwall_target_gradient_vector_hessian_matrix(U, vlist, perms):
    """f, real Riemannian gradient (8 nl), real Riemannian Hessian (8 nl)^2"""
    vlist = _c(vlist)
    m, uf = _dense_map(U)
    L = int(np.log2(m.n))
    pm, arr = _perm_ptrs(perms)
    nl = pm.shape[0]
    f = np.zeros(1, np.complex128)
    g = np.zeros(8 * nl)
    H = np.zeros((8 * nl, 8 * nl))
    if _fn("matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix")(
            uf, ctypes.byref(m), _p(vlist), nl, L, ctypes.cast(arr, _P), _p(f), _p(g), _p(H)) < 0:
        raise RuntimeError("matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix failed")
    return f[0], g, H


def brickwall_target(U, vlist, perms):
    vlist = _c(vlist)
    m, uf = _dense_map(U)
    L = int(np.log2(m.n))
    pm, arr = _perm_ptrs(perms)
    f = np.zeros(1, np.complex128)
    if _fn("matchgate_brickwall_unitary_target")(uf, ctypes.byref(m), _p(vlist), pm.shape[0], L, ctypes.cast(arr, _P),
                                                 _p(f)) < 0:
        raise RuntimeError("matchgate_brickwall_unitary_target failed")
    return f[0]


def brickwall_target_and_gradient(U, vlist, perms):
    vlist = _c(vlist)
    m, uf = _dense_map(U)
    L = int(np.log2(m.n))
    pm, arr = _perm_ptrs(perms)
    nl = pm.shape[0]
    f = np.zeros(1, np.complex128)
    dv = np.zeros((nl, 8), np.complex128)
    if _fn("matchgate_brickwall_unitary_target_and_gradient")(uf, ctypes.byref(m), _p(vlist), nl, L,
                                                              ctypes.cast(arr, _P), _p(f), _p(dv)) < 0:
        raise RuntimeError("matchgate_brickwall_unitary_target_and_gradient failed")
    return f[0], dv


def circuit_target_and_gradient(U, gates, wires):
    gates = _c(gates)
    wires = np.ascontiguousarray(wires, dtype=np.int32)
    m, uf = _dense_map(U)
    L = int(np.log2(m.n))
    n = gates.shape[0]
    f = np.zeros(1, np.complex128)
    dg = np.zeros((n, 8), np.complex128)
    if _fn("matchgate_circuit_unitary_target_and_gradient")(uf, ctypes.byref(m), _p(gates), n, _p(wires), L, _p(f),
                                                            _p(dg)) < 0:
        raise RuntimeError("matchgate_circuit_unitary_target_and_gradient failed")
    return f[0], dg


def time_target_columns(U, vlist, perms, ncols, nthreads, hessian=True):
    """Wall seconds of the reference's per-column target work (ufunc seed, forward sweep, trace
    term and, with hessian, the backward sweep with the placeholder Hessian and accumulation;
    src/matchgate_target.c:746-864) over the columns [0, ncols) on nthreads OpenMP threads,
    and the trace over those columns."""
    vlist = _c(vlist)
    m, uf = _dense_map(U)
    L = int(np.log2(m.n))
    pm, arr = _perm_ptrs(perms)
    f = np.zeros(1, np.complex128)
    fn = _lib.ref_time_target_columns
    fn.restype = ctypes.c_double
    fn.argtypes = [_P, _P, _P, _I, _I, _P, ctypes.c_long, _I, _I, _P]
    dt = fn(uf, ctypes.cast(ctypes.byref(m), _P), _p(vlist), pm.shape[0], L, ctypes.cast(arr, _P), int(ncols),
            int(nthreads), int(bool(hessian)), _p(f))
    return dt, f[0]
