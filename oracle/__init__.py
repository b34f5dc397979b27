"""CPU oracle for the matchgate-brickwall hot path (test infrastructure only).

Thin ctypes binding of ``oracle/mg_oracle.c`` plus numpy helpers.  Only the
tests, the smoke check and bench.py's CPU-baseline leg import this package;
the product library (``rqcopt_hpc-matchgate_amd``) never does.

Gates are numpy ``complex128`` arrays of shape ``(P, 8)``: entries 0-3 are the
center block (odd parity), 4-7 the corner block (even parity), each 2x2
row-major (reference src/matchgate.h:5-14).  Targets are dense row-major
``(2^L, 2^L)`` complex matrices U; f = -Tr(U^dagger W).
"""
import ctypes
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_DIR, "lib", "libmg_oracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _DIR], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    return ctypes.CDLL(_LIB)


_lib = _load()
_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_long
_D = ctypes.c_double


def _sig(name, args, res=None):
    fn = getattr(_lib, name)
    fn.argtypes = args
    fn.restype = res
    return fn


_sig("orc_apply_gate", [_P, _I, _I, _I, _P, _P])
_sig("orc_apply_gate_backward", [_P, _I, _I, _I, _P, _P, _P, _P])
_sig("orc_apply_gate_to_array", [_P, _I, _I, _I, _I, _P, _P])
_sig("orc_apply_gate_backward_array", [_P, _I, _I, _I, _I, _P, _P, _P])
_sig("orc_apply_gate_placeholder", [_I, _I, _I, _P, _P])
_sig("orc_brickwall_program", [_I, _I, _P, _P, _P])
_sig("orc_apply_program", [_P, _I, _P, _P, _I, _P, _P])
_sig("orc_program_forward", [_P, _I, _P, _P, _I, _P, _P, _P])
_sig("orc_program_backward", [_P, _I, _P, _P, _I, _P, _P, _P, _P])
_sig("orc_program_backward_hessian", [_P, _I, _I, _P, _P, _I, _P, _P, _P, _P, _P])
_sig("orc_ti_brickwall_backward", [_P, _I, _P, _I, _P, _P, _P, _P])
_sig("orc_ti_brickwall_backward_hessian", [_P, _I, _P, _I, _P, _P, _P, _P, _P])
_sig("orc_target", [_P, _P, _I, _P, _P, _I, _I, _P])
_sig("orc_target_gradient", [_P, _P, _I, _I, _P, _P, _I, _I, _P, _P])
_sig("orc_target_gradient_hessian", [_P, _P, _I, _I, _P, _P, _I, _I, _I, _P, _P, _P])
_sig("orc_target_gradient_hvp", [_P, _P, _P, _I, _I, _P, _P, _I, _L, _L, _I, _P, _P, _P])
_sig("orc_target_gradient_hvp_program", [_P, _I, _P, _P, _P, _I, _I, _P, _P, _I, _L, _L, _I, _P, _P, _P])
for _n in ("orc_multiply",):
    _sig(_n, [_P, _P, _P])
_sig("orc_adjoint", [_P, _P])
_sig("orc_symm", [_P, _P])
_sig("orc_antisymm", [_P, _P])
_sig("orc_real_to_antisymm", [_P, _P])
_sig("orc_antisymm_to_real", [_P, _P])
_sig("orc_real_to_tangent", [_P, _P, _P])
_sig("orc_tangent_to_real", [_P, _P, _P])
_sig("orc_project_tangent", [_P, _P, _P])
_sig("orc_symmetric_triple", [_P, _P, _P, _P])
_sig("orc_inverse_mat2x2", [_P, _P], _I)
_sig("orc_polar", [_P, _P])
_sig("orc_retract", [_P, _P, _P])
_sig("orc_gradient_vector_hessian_matrix", [_P, _I, _P, _P, _P, _P])
_sig("orc_riemannian_gradient", [_P, _I, _P, _P])
_sig("orc_real_to_directions", [_P, _I, _P, _P])
_sig("orc_riemannian_hvp", [_P, _I, _P, _P, _P, _P])
_sig("orc_truncated_cg_hmat", [_P, _P, _I, _D, _I, _D, _D, _P], _I)
_sig("orc_rtr_optimize", [_P, _P, _I, _I, _P, _P, _I, _I, _I, _I, _D, _D, _D, _D, _D, _I, _P, _P, _P])


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _r(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------- gate kernels ----------------

def apply_gate(g, i, j, L, psi):
    g, psi = _c(g), _c(psi)
    out = np.empty(1 << L, np.complex128)
    _lib.orc_apply_gate(_p(g), i, j, L, _p(psi), _p(out))
    return out


def apply_gate_backward(g, i, j, L, psi, dout):
    g, psi, dout = _c(g), _c(psi), _c(dout)
    dpsi = np.empty(1 << L, np.complex128)
    dg = np.empty(8, np.complex128)
    _lib.orc_apply_gate_backward(_p(g), i, j, L, _p(psi), _p(dout), _p(dpsi), _p(dg))
    return dpsi, dg


def apply_gate_to_array(g, i, j, L, psi):
    """psi: (2^L, nstates) interleaved."""
    g, psi = _c(g), _c(psi)
    out = np.empty_like(psi)
    _lib.orc_apply_gate_to_array(_p(g), i, j, L, psi.shape[1], _p(psi), _p(out))
    return out


def apply_gate_backward_array(g, i, j, L, psi, dout):
    g, psi, dout = _c(g), _c(psi), _c(dout)
    dgs = np.empty((psi.shape[1], 8), np.complex128)
    _lib.orc_apply_gate_backward_array(_p(g), i, j, L, psi.shape[1], _p(psi), _p(dout), _p(dgs))
    return dgs


def apply_gate_placeholder(i, j, L, psi):
    psi = _c(psi)
    out = np.empty((1 << L, 8), np.complex128)
    _lib.orc_apply_gate_placeholder(i, j, L, _p(psi), _p(out))
    return out


# ---------------- gate programs ----------------

def brickwall_program(L, perms):
    """Tied brickwall (reference src/matchgate_target.c:308-324): returns (param, wires)."""
    perms = _i(perms)
    nl = perms.shape[0]
    G = nl * (L // 2)
    param = np.empty(G, np.int32)
    wires = np.empty(2 * G, np.int32)
    _lib.orc_brickwall_program(L, nl, _p(perms), _p(param), _p(wires))
    return param, wires


def apply_program(gates, param, wires, L, psi):
    gates, param, wires, psi = _c(gates), _i(param), _i(wires), _c(psi)
    out = np.empty(1 << L, np.complex128)
    _lib.orc_apply_program(_p(gates), len(param), _p(param), _p(wires), L, _p(psi), _p(out))
    return out


def program_forward(gates, param, wires, L, psi):
    gates, param, wires, psi = _c(gates), _i(param), _i(wires), _c(psi)
    G = len(param)
    cache = np.empty((G, 1 << L), np.complex128)
    out = np.empty(1 << L, np.complex128)
    _lib.orc_program_forward(_p(gates), G, _p(param), _p(wires), L, _p(psi), _p(cache), _p(out))
    return cache, out


def program_backward(gates, param, wires, L, cache, dout):
    gates, param, wires, cache, dout = _c(gates), _i(param), _i(wires), _c(cache), _c(dout)
    G = len(param)
    dpsi = np.empty(1 << L, np.complex128)
    dgi = np.empty((G, 8), np.complex128)
    _lib.orc_program_backward(_p(gates), G, _p(param), _p(wires), L, _p(cache), _p(dout), _p(dpsi), _p(dgi))
    return dpsi, dgi


def program_backward_hessian(gates, param, wires, L, cache, dout, P=None):
    gates, param, wires, cache, dout = _c(gates), _i(param), _i(wires), _c(cache), _c(dout)
    P = gates.shape[0] if P is None else P
    m = 8 * P
    dpsi = np.empty(1 << L, np.complex128)
    dgrad = np.empty(m, np.complex128)
    hess = np.empty((m, m), np.complex128)
    _lib.orc_program_backward_hessian(_p(gates), P, len(param), _p(param), _p(wires), L, _p(cache), _p(dout),
                                      _p(dpsi), _p(dgrad), _p(hess))
    return dpsi, dgrad.reshape(P, 8), hess


def ti_brickwall_backward(vlist, perms, L, cache, dout):
    """TI gradient sweep of one column (reference src/matchgate_brickwall.c:266-306)"""
    vlist, perms, cache, dout = _c(vlist), _i(perms), _c(cache), _c(dout)
    nl = perms.shape[0]
    dpsi = np.empty(1 << L, np.complex128)
    dgrad = np.empty(8 * nl, np.complex128)
    _lib.orc_ti_brickwall_backward(_p(vlist), nl, _p(perms), L, _p(cache), _p(dout), _p(dpsi), _p(dgrad))
    return dpsi, dgrad.reshape(nl, 8)


def ti_brickwall_backward_hessian(vlist, perms, L, cache, dout):
    vlist, perms, cache, dout = _c(vlist), _i(perms), _c(cache), _c(dout)
    nl = perms.shape[0]
    m = 8 * nl
    dpsi = np.empty(1 << L, np.complex128)
    dgrad = np.empty(m, np.complex128)
    hess = np.empty((m, m), np.complex128)
    _lib.orc_ti_brickwall_backward_hessian(_p(vlist), nl, _p(perms), L, _p(cache), _p(dout), _p(dpsi), _p(dgrad), _p(hess))
    return dpsi, dgrad.reshape(nl, 8), hess


# ---------------- targets ----------------

def target(U, gates, param, wires, L, nthreads=0):
    U, gates, param, wires = _c(U), _c(gates), _i(param), _i(wires)
    f = np.zeros(1, np.complex128)
    _lib.orc_target(_p(U), _p(gates), len(param), _p(param), _p(wires), L, nthreads, _p(f))
    return f[0]


def target_gradient(U, gates, param, wires, L, nthreads=0):
    U, gates, param, wires = _c(U), _c(gates), _i(param), _i(wires)
    P = gates.shape[0]
    f = np.zeros(1, np.complex128)
    dg = np.empty(8 * P, np.complex128)
    _lib.orc_target_gradient(_p(U), _p(gates), P, len(param), _p(param), _p(wires), L, nthreads, _p(f), _p(dg))
    return f[0], dg.reshape(P, 8)


def target_gradient_hessian(U, gates, param, wires, L, ti=False, nthreads=0):
    U, gates, param, wires = _c(U), _c(gates), _i(param), _i(wires)
    P = gates.shape[0]
    m = 8 * P
    f = np.zeros(1, np.complex128)
    dg = np.empty(m, np.complex128)
    hess = np.empty((m, m), np.complex128)
    _lib.orc_target_gradient_hessian(_p(U), _p(gates), P, len(param), _p(param), _p(wires), L, int(ti), nthreads,
                                     _p(f), _p(dg), _p(hess))
    return f[0], dg.reshape(P, 8), hess


def target_gradient_hvp_program(tgates, twires, gates, Z, param, wires, L, b0=0, b1=None, nthreads=0):
    """As target_gradient_hvp with the target a gate program applied to e_b per column."""
    tgates, twires = _c(tgates), _i(twires)
    gates, Z, param, wires = _c(gates), _c(Z), _i(param), _i(wires)
    P = gates.shape[0]
    b1 = (1 << L) if b1 is None else b1
    f = np.zeros(1, np.complex128)
    dg = np.empty(8 * P, np.complex128)
    dhz = np.empty(8 * P, np.complex128)
    _lib.orc_target_gradient_hvp_program(_p(tgates), tgates.shape[0], _p(twires), _p(gates), _p(Z), P, len(param),
                                         _p(param), _p(wires), L, b0, b1, nthreads, _p(f), _p(dg), _p(dhz))
    return f[0], dg.reshape(P, 8), dhz.reshape(P, 8)


def target_gradient_hvp(U, gates, Z, param, wires, L, b0=0, b1=None, nthreads=0):
    """f, complex gradient (P,8) and D^2 f[Z] (P,8) summed over columns [b0, b1)."""
    U, gates, Z, param, wires = _c(U), _c(gates), _c(Z), _i(param), _i(wires)
    P = gates.shape[0]
    b1 = (1 << L) if b1 is None else b1
    f = np.zeros(1, np.complex128)
    dg = np.empty(8 * P, np.complex128)
    dhz = np.empty(8 * P, np.complex128)
    _lib.orc_target_gradient_hvp(_p(U), _p(gates), _p(Z), P, len(param), _p(param), _p(wires), L, b0, b1, nthreads,
                                 _p(f), _p(dg), _p(dhz))
    return f[0], dg.reshape(P, 8), dhz.reshape(P, 8)


# ---------------- algebra ----------------

def _un(fn, a):
    a = _c(a)
    out = np.empty(8, np.complex128)
    getattr(_lib, fn)(_p(a), _p(out))
    return out


def multiply(a, b):
    a, b = _c(a), _c(b)
    c = np.empty(8, np.complex128)
    _lib.orc_multiply(_p(a), _p(b), _p(c))
    return c


def adjoint(a):
    return _un("orc_adjoint", a)


def symm(a):
    return _un("orc_symm", a)


def antisymm(a):
    return _un("orc_antisymm", a)


def polar(a):
    return _un("orc_polar", a)


def real_to_antisymm(r):
    r = _r(r)
    w = np.empty(8, np.complex128)
    _lib.orc_real_to_antisymm(_p(r), _p(w))
    return w


def antisymm_to_real(w):
    w = _c(w)
    r = np.empty(8)
    _lib.orc_antisymm_to_real(_p(w), _p(r))
    return r


def real_to_tangent(r, v):
    r, v = _r(r), _c(v)
    z = np.empty(8, np.complex128)
    _lib.orc_real_to_tangent(_p(r), _p(v), _p(z))
    return z


def tangent_to_real(v, z):
    v, z = _c(v), _c(z)
    r = np.empty(8)
    _lib.orc_tangent_to_real(_p(v), _p(z), _p(r))
    return r


def project_tangent(u, z):
    u, z = _c(u), _c(z)
    p = np.empty(8, np.complex128)
    _lib.orc_project_tangent(_p(u), _p(z), _p(p))
    return p


def inverse_mat2x2(a):
    a = _c(a)
    out = np.empty(4, np.complex128)
    ret = _lib.orc_inverse_mat2x2(_p(a), _p(out))
    return ret, out


def retract(u, eta):
    u, eta = _c(u), _r(eta)
    v = np.empty(8, np.complex128)
    _lib.orc_retract(_p(u), _p(eta), _p(v))
    return v


# ---------------- projection, HVP, trust region ----------------

def gradient_vector_hessian_matrix(vlist, dgrad, hess):
    vlist, dgrad, hess = _c(vlist), _c(dgrad), _c(hess)
    P = vlist.shape[0]
    g = np.empty(8 * P)
    H = np.empty((8 * P, 8 * P))
    _lib.orc_gradient_vector_hessian_matrix(_p(vlist), P, _p(dgrad), _p(hess), _p(g), _p(H))
    return g, H


def riemannian_gradient(vlist, dgrad):
    vlist, dgrad = _c(vlist), _c(dgrad)
    g = np.empty(8 * vlist.shape[0])
    _lib.orc_riemannian_gradient(_p(vlist), vlist.shape[0], _p(dgrad), _p(g))
    return g


def real_to_directions(vlist, v):
    vlist, v = _c(vlist), _r(v)
    Z = np.empty_like(vlist)
    _lib.orc_real_to_directions(_p(vlist), vlist.shape[0], _p(v), _p(Z))
    return Z


def riemannian_hvp(vlist, dgrad, Z, dhz):
    vlist, dgrad, Z, dhz = _c(vlist), _c(dgrad), _c(Z), _c(dhz)
    hv = np.empty(8 * vlist.shape[0])
    _lib.orc_riemannian_hvp(_p(vlist), vlist.shape[0], _p(dgrad), _p(Z), _p(dhz), _p(hv))
    return hv


def hvp(U, vlist, param, wires, L, v, nthreads=0):
    """Riemannian H v (real, length 8P) by forward-over-reverse."""
    Z = real_to_directions(vlist, v)
    f, dg, dhz = target_gradient_hvp(U, vlist, Z, param, wires, L, nthreads=nthreads)
    return riemannian_hvp(vlist, dg, Z, dhz)


def truncated_cg_hmat(grad, H, radius, maxiter, abstol=1e-8, reltol=1e-6):
    grad, H = _r(grad), _r(H)
    z = np.empty_like(grad)
    b = _lib.orc_truncated_cg_hmat(_p(grad), _p(H), len(grad), radius, maxiter, abstol, reltol, _p(z))
    return z, bool(b)


def rtr_optimize(U, vstart, param, wires, L, niter, mode="hmat", tcg_maxiter=None, rho_trust=0.125,
                 radius_init=0.01, maxradius=0.1, abstol=1e-8, reltol=1e-6, nthreads=0):
    U, vstart, param, wires = _c(U), _c(vstart), _i(param), _i(wires)
    P = vstart.shape[0]
    tcg_maxiter = 2 * 16 * P if tcg_maxiter is None else tcg_maxiter
    f_iter = np.empty(niter + 1)
    vopt = np.empty_like(vstart)
    nh = np.zeros(1, np.int32)
    _lib.orc_rtr_optimize(_p(U), _p(vstart), P, len(param), _p(param), _p(wires), L, 0 if mode == "hmat" else 1,
                          niter, tcg_maxiter, rho_trust, radius_init, maxradius, abstol, reltol, nthreads,
                          _p(f_iter), _p(vopt), _p(nh))
    return f_iter, vopt, int(nh[0])
