"""The library's host trust region on the CPU (no device work).

- truncated_cg_hmat (reference src/trust_region.c:88-159) against the oracle's Steihaug-Toint
  truncated CG on random symmetric matrices, positive definite and indefinite, with interior and
  boundary solutions;
- riemannian_trust_region_optimize_hmat / _hvp (reference src/trust_region.c:235-306 and the
  declared HVP optimiser) and mg_trust_region_optimize_hvp with H eta taken from the truncated CG,
  on a smooth non-quadratic function of R^n with the retraction x + eta: the three give the same
  iterates, the last one with exactly one Hessian-vector product per iteration fewer, and the
  accepted steps decrease f.
"""
import ctypes

import numpy as np
import pytest

import oracle as orc

_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
TARGET = ctypes.CFUNCTYPE(_D, _PD, ctypes.c_void_p, _PD)
TARGET_GRAD = ctypes.CFUNCTYPE(_D, _PD, ctypes.c_void_p, _PD)
TARGET_GRAD_HESS = ctypes.CFUNCTYPE(_D, _PD, ctypes.c_void_p, _PD, _PD, _PD)
HVP = ctypes.CFUNCTYPE(None, _PD, ctypes.c_void_p, _PD, _PD)
RETRACT = ctypes.CFUNCTYPE(None, _PD, _PD, ctypes.c_void_p, _PD)


@pytest.fixture(scope="module")
def mg():
    import rqcopt_mg

    lib = rqcopt_mg.lib
    lib.truncated_cg_hmat.restype = ctypes.c_bool
    lib.truncated_cg_hmat.argtypes = [_PD, _PD, ctypes.c_int, _D, ctypes.c_void_p, _PD]
    lib.riemannian_trust_region_optimize_hmat.restype = None
    lib.riemannian_trust_region_optimize_hmat.argtypes = [TARGET, TARGET_GRAD_HESS, ctypes.c_void_p, RETRACT,
                                                          ctypes.c_void_p, ctypes.c_int, _PD, ctypes.c_int,
                                                          ctypes.c_void_p, ctypes.c_int, _PD, _PD, _PD]
    lib.riemannian_trust_region_optimize_hvp.restype = None
    lib.riemannian_trust_region_optimize_hvp.argtypes = [TARGET, TARGET_GRAD, HVP, ctypes.c_void_p, RETRACT,
                                                         ctypes.c_void_p, ctypes.c_int, _PD, ctypes.c_int,
                                                         ctypes.c_void_p, ctypes.c_int, _PD, _PD]
    lib.mg_trust_region_optimize_hvp.restype = None
    lib.mg_trust_region_optimize_hvp.argtypes = lib.riemannian_trust_region_optimize_hvp.argtypes + [ctypes.c_int]
    return rqcopt_mg


def ptr(a):
    return a.ctypes.data_as(_PD)


def params(mg, n, radius_init, maxradius):
    p = mg.RTRParams()
    p.tcg_params.maxiter = 2 * n
    p.tcg_params.abstol = 1e-8
    p.tcg_params.reltol = 1e-6
    p.rho_trust = 0.125
    p.radius_init = radius_init
    p.maxradius = maxradius
    return p


@pytest.mark.parametrize("kind", ["spd", "indefinite"])
def test_truncated_cg_matches_oracle(mg, kind):
    rng = np.random.default_rng(11 if kind == "spd" else 12)
    n = 24
    for trial in range(12):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        ev = rng.uniform(0.1, 3.0, n) if kind == "spd" else rng.uniform(-1.0, 3.0, n)
        H = np.ascontiguousarray((Q * ev) @ Q.T)
        H = 0.5 * (H + H.T)
        g = rng.standard_normal(n)
        radius = [0.05, 0.5, 5.0, 50.0][trial % 4]
        p = params(mg, n, radius, radius)
        z = np.zeros(n)
        bnd = mg.lib.truncated_cg_hmat(ptr(g), ptr(H), n, radius, ctypes.byref(p.tcg_params), ptr(z))
        zo, bo = orc.truncated_cg_hmat(g, H, radius, 2 * n)
        assert bool(bnd) == bo, (kind, trial)
        assert np.max(np.abs(z - zo)) <= 1e-12 * max(1.0, np.max(np.abs(zo))), (kind, trial)
        assert np.linalg.norm(z) <= radius * (1 + 1e-12)
        if bnd:
            assert abs(np.linalg.norm(z) - radius) <= 1e-10 * radius


class Problem:
    """f(x) = x^T A x / 2 + b^T x + c sum x^4 on R^n (A indefinite, bounded below by the quartic)."""

    def __init__(self, n, seed):
        rng = np.random.default_rng(seed)
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        self.A = (Q * rng.uniform(-0.5, 2.0, n)) @ Q.T
        self.A = 0.5 * (self.A + self.A.T)
        self.b = rng.standard_normal(n)
        self.c = 0.1
        self.n = n
        self.nhvp = 0

    def f(self, x):
        return 0.5 * x @ self.A @ x + self.b @ x + self.c * np.sum(x ** 4)

    def grad(self, x):
        return self.A @ x + self.b + 4 * self.c * x ** 3

    def hess(self, x):
        return self.A + np.diag(12 * self.c * x ** 2)

    def callbacks(self):
        n = self.n

        def arr(p):
            return np.ctypeslib.as_array(p, shape=(n,))

        def f(x, _, fc):
            v = self.f(arr(x).copy())
            fc[0], fc[1] = v, 0.0
            return v

        def fgh(x, _, fc, g, h):
            xv = arr(x).copy()
            v = self.f(xv)
            fc[0], fc[1] = v, 0.0
            arr(g)[:] = self.grad(xv)
            np.ctypeslib.as_array(h, shape=(n * n,))[:] = self.hess(xv).ravel()
            return v

        def fg(x, _, g):
            xv = arr(x).copy()
            arr(g)[:] = self.grad(xv)
            return self.f(xv)

        def hvp(x, _, vec, out):
            self.nhvp += 1
            arr(out)[:] = self.hess(arr(x).copy()) @ arr(vec)

        def retract(x, eta, _, xs):
            arr(xs)[:] = arr(x) + arr(eta)

        return TARGET(f), TARGET_GRAD(fg), TARGET_GRAD_HESS(fgh), HVP(hvp), RETRACT(retract)


@pytest.mark.parametrize("radius", [(0.01, 0.1), (0.5, 2.0)])
def test_trust_region_control_flows_agree(mg, radius):
    n, niter = 12, 25
    prob = Problem(n, 21)
    f, fg, fgh, hvp, retract = prob.callbacks()
    x0 = np.random.default_rng(22).standard_normal(n)
    p = params(mg, n, *radius)
    runs = {}
    for name in ("hmat", "hvp", "reuse"):
        prob.nhvp = 0
        f_iter, x_final = np.zeros(niter + 1), np.zeros(n)
        if name == "hmat":
            fc = np.zeros(2 * (niter + 1))
            mg.lib.riemannian_trust_region_optimize_hmat(f, fgh, None, retract, None, n, ptr(x0), n, ctypes.byref(p),
                                                         niter, ptr(f_iter), ptr(fc), ptr(x_final))
        elif name == "hvp":
            mg.lib.riemannian_trust_region_optimize_hvp(f, fg, hvp, None, retract, None, n, ptr(x0), n,
                                                        ctypes.byref(p), niter, ptr(f_iter), ptr(x_final))
        else:
            mg.lib.mg_trust_region_optimize_hvp(f, fg, hvp, None, retract, None, n, ptr(x0), n, ctypes.byref(p),
                                                niter, ptr(f_iter), ptr(x_final), 1)
        runs[name] = (f_iter, x_final, prob.nhvp)
    fh, xh, _ = runs["hmat"]
    for name in ("hvp", "reuse"):
        fi, xf, _ = runs[name]
        assert np.max(np.abs(fi - fh)) <= 1e-11 * np.max(np.abs(fh)), name
        # near the minimiser f is flat: products rounded in another order move the iterates by
        # ~1e-9 while f agrees to 1e-11
        assert np.max(np.abs(xf - xh)) <= 1e-8, name
    # one product fewer per iteration: H eta comes from the truncated CG
    assert runs["hvp"][2] - runs["reuse"][2] == niter
    # accepted steps decrease f; rejected ones keep it
    assert np.all(np.diff(fh) <= 1e-12 * np.max(np.abs(fh)))
    assert fh[-1] < fh[0]
