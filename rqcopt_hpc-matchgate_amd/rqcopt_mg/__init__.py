"""ctypes binding of librqcopt_mg.so, the MI355X implementation of the
matchgate-brickwall target / gradient / Hessian-vector-product path.

The C-ABI (include/*.h) is the product; this module mirrors it for Python callers,
tests and the benchmark.  Gates are numpy complex128 arrays of shape (P, 8)
(center block entries 0-3, corner block 4-7, each 2x2 row-major), byte-identical
to ``struct matchgate[P]``.

Importing fails loudly if the HIP library has not been built (no CPU fallback).
"""
import ctypes
import os

import numpy as np

_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("RQCOPT_MG_LIB") or os.path.join(_DIR, "lib", "librqcopt_mg.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {_DIR}` (or __graft_entry__.build())")

lib = ctypes.CDLL(LIB_PATH)

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_L = ctypes.c_long

MG_FP64, MG_FP32 = 0, 1
EVAL_F, EVAL_F_GRAD, EVAL_F_GRAD_HVP = 0, 1, 2

# typedef int (*linear_func)(const struct statevector*, void*, struct statevector*)
LINEAR_FUNC = ctypes.CFUNCTYPE(ctypes.c_int, _P, _P, _P)


class Statevector(ctypes.Structure):
    _fields_ = [("data", _P), ("nqubits", ctypes.c_int)]


class StatevectorArray(ctypes.Structure):
    _fields_ = [("data", _P), ("nqubits", ctypes.c_int), ("nstates", ctypes.c_int)]


class QuantumCircuitCache(ctypes.Structure):
    _fields_ = [("psi_list", ctypes.POINTER(Statevector)), ("nqubits", ctypes.c_int), ("ngates", ctypes.c_int)]


class TruncatedCGParams(ctypes.Structure):
    _fields_ = [("maxiter", ctypes.c_int), ("abstol", _D), ("reltol", _D)]


class RTRParams(ctypes.Structure):
    _fields_ = [("tcg_params", TruncatedCGParams), ("rho_trust", _D), ("radius_init", _D), ("maxradius", _D),
                ("g_func", _P), ("g_data", _P), ("g_iter", _P)]


class Request(ctypes.Structure):
    """struct mg_request (include/mg_gpu.h)"""
    _fields_ = [("target", _P), ("program", _P), ("gates", _P), ("v", _P), ("what", ctypes.c_int),
                ("precision", ctypes.c_int), ("f", _P), ("grad_vec", _P), ("hvp", _P), ("b0", ctypes.c_long),
                ("b1", ctypes.c_long)]


class Job(ctypes.Structure):
    """struct mg_job (include/mg_gpu.h)"""
    _fields_ = [("nqubits", ctypes.c_int), ("program", _P), ("gates_start", _P), ("U", _P),
                ("target_ngates", ctypes.c_int), ("target_gates", _P), ("target_wires", _P),
                ("niter", ctypes.c_int), ("arrival", _D), ("f_iter", _P), ("gates_opt", _P),
                ("t_admit", _D), ("t_done", _D), ("nhvp", ctypes.c_int), ("status", ctypes.c_int),
                ("target", _P)]


class ServeOptions(ctypes.Structure):
    """struct mg_serve_options (include/mg_gpu.h)"""
    _fields_ = [("batched", ctypes.c_int), ("precision", ctypes.c_int), ("precision_switch", ctypes.c_int),
                ("max_active", ctypes.c_int), ("hbm_budget", _D), ("realtime", ctypes.c_int),
                ("densify_qubits", ctypes.c_int), ("progress_s", _D), ("tr_reuse", ctypes.c_int),
                ("time_limit_s", _D)]


class ServeStats(ctypes.Structure):
    """struct mg_serve_stats (include/mg_gpu.h)"""
    _fields_ = [("wall_s", _D), ("rounds", ctypes.c_long), ("requests", ctypes.c_long), ("kernel_ms", _D),
                ("max_active", ctypes.c_int), ("peak_footprint", _D), ("eval_s", _D), ("prep_s", _D),
                ("graph_replays", ctypes.c_long)]


def rtr_default_params(n):
    """set_rtr_default_params (reference src/trust_region.h:69-78)"""
    return RTRParams(TruncatedCGParams(2 * n, 1e-8, 1e-6), 0.125, 0.01, 0.1, None, None, None)


def _sig(name, args, res=ctypes.c_int):
    fn = getattr(lib, name)
    fn.argtypes = args
    fn.restype = res
    return fn


_sig("mg_last_error", [], ctypes.c_char_p)
_sig("mg_target_dense", [_P, _I], _P)
_sig("mg_target_brickwall", [_P, _I, _P, _I], _P)
_sig("mg_target_circuit", [_P, _I, _P, _I], _P)
_sig("mg_target_from_linear_func", [LINEAR_FUNC, _P, _I], _P)
_sig("mg_target_free", [_P], None)
_sig("mg_target_densify", [_P])
_sig("mg_program_create", [_I, _I, _I, _P, _P], _P)
_sig("mg_program_brickwall", [_I, _I, _P], _P)
_sig("mg_program_free", [_P], None)
_sig("mg_program_nparams", [_P])
_sig("mg_program_ninst", [_P])
_sig("mg_evaluate", [_P, _P, _P, _P, _I, _I, _P, _P, _P])
_sig("mg_evaluate_columns", [_P, _P, _P, _P, _I, _I, _L, _L, _P, _P, _P])
_sig("mg_evaluate_noalloc", [_P, _P, _P, _P, _I, _I, _P, _P, _P])
_sig("mg_riemannian_gradient_hvp", [_P, _P, _P, _P, _I, _P, _P, _P])
_sig("mg_riemannian_project", [_P, _I, _P, _P, _P, _P, _P])
_sig("mg_hessian", [_P, _P, _P, _I, _P, _P, _P])
_sig("mg_gradient_vector_hessian_matrix", [_P, _P, _P, _I, _P, _P, _P])
_sig("mg_optimize", [_P, _P, _P, ctypes.POINTER(RTRParams), _I, _I, _I, _I, _P, _P, _P, _P])
_sig("mg_set_checkpointing", [_I], None)
_sig("mg_set_lds_states", [_I], None)
_sig("mg_set_kernel_timing", [_I], None)
_sig("mg_set_register_kernel", [_I, _I], None)
_sig("mg_set_grid_per_cu", [_I], None)
_sig("mg_set_sweep_threads", [_I], None)
_sig("mg_last_sweep_threads", [])
_sig("mg_last_sweep_kernel", [], ctypes.c_char_p)
_sig("mg_set_window_kernel", [_I], None)
_sig("mg_set_tile_kernel", [_I], None)
_sig("mg_set_row_chunk_bytes", [_L], None)
_sig("mg_set_materialize_kernel", [_I], None)
_sig("mg_set_forward_mfma", [_I], None)
_sig("mg_set_device_trust_region", [_I], None)
_sig("mg_last_trust_region_graphs", [], ctypes.c_long)
_sig("mg_last_kernel_ms", [], ctypes.c_float)
_sig("mg_last_sweep_info", [_P, _P, _P, _P], None)
_sig("mg_target_cache_clear", [], None)
_sig("mg_target_cache_reuse", [_I], None)
_sig("mg_evaluate_batch", [ctypes.POINTER(Request), _I])
_sig("mg_batch_set_kernel", [_I, _I, _I])
_sig("mg_batch_get_kernel", [_I, _P, _P], None)
_sig("mg_batch_set_single_range", [_I, _I], None)
_sig("mg_batch_set_grid_per_cu", [_I], None)
_sig("mg_batch_set_shape", [_I, _I, _I, _I], None)
_sig("mg_batch_get_shape", [_I, _I, _P, _P], None)
_sig("mg_batch_last_kernel_ms", [], ctypes.c_float)
_sig("mg_batch_set_graphs", [ctypes.c_int], None)
_sig("mg_batch_last_graph_stats", [_P, _P], None)
_sig("mg_batch_autotune", [ctypes.POINTER(Request), _I, _I, _I, _P, _P, _P, _P])
_sig("mg_serve", [ctypes.POINTER(Job), _I, ctypes.POINTER(ServeOptions), ctypes.POINTER(ServeStats)])
# reference entry points (include/matchgate_target.h, mg_brickwall_opt.h)
_sig("matchgate_circuit_unitary_target", [LINEAR_FUNC, _P, _P, _I, _P, _I, _P])
_sig("parallel_matchgate_circuit_unitary_target", [LINEAR_FUNC, _P, _P, _I, _P, _I, _P])
_sig("matchgate_circuit_unitary_target_and_gradient", [LINEAR_FUNC, _P, _P, _I, _P, _I, _P, _P])
_sig("matchgate_brickwall_unitary_target", [LINEAR_FUNC, _P, _P, _I, _I, _P, _P])
_sig("matchgate_brickwall_unitary_target_and_gradient", [LINEAR_FUNC, _P, _P, _I, _I, _P, _P, _P])
_sig("matchgate_brickwall_unitary_target_gradient_hessian", [LINEAR_FUNC, _P, _P, _I, _I, _P, _P, _P, _P])
_sig("parallel_matchgate_brickwall_unitary_target_gradient_hessian", [LINEAR_FUNC, _P, _P, _I, _I, _P, _P, _P, _P])
_sig("matchgate_ti_brickwall_unitary_target_gradient_hessian", [LINEAR_FUNC, _P, _P, _I, _I, _P, _P, _P, _P])
_sig("matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix", [LINEAR_FUNC, _P, _P, _I, _I, _P, _P, _P, _P])
_sig("optimize_matchgate_brickwall_circuit_hmat", [LINEAR_FUNC, _P, _P, _I, _I, _P, ctypes.POINTER(RTRParams), _I, _P, _P, _P], None)
_sig("optimize_matchgate_brickwall_circuit_hvp", [LINEAR_FUNC, _P, _P, _I, _I, _P, ctypes.POINTER(RTRParams), _I, _P, _P, _P], None)
# single-column kernels (include/gate.h, quantum_circuit.h, matchgate_brickwall.h)
_SV = ctypes.POINTER(Statevector)
_SA = ctypes.POINTER(StatevectorArray)
_QC = ctypes.POINTER(QuantumCircuitCache)
_sig("apply_matchgate", [_P, _I, _I, _SV, _SV], None)
_sig("apply_matchgate_backward", [_P, _I, _I, _SV, _SV, _SV, _P], None)
_sig("apply_matchgate_to_array", [_P, _I, _I, _SA, _SA], None)
_sig("apply_matchgate_backward_array", [_P, _I, _I, _SA, _SV, _P], None)
_sig("apply_matchgate_placeholder", [_I, _I, _SV, _SA], None)
_sig("apply_gate_placeholder", [_I, _I, _SV, _SA], None)
_sig("apply_quantum_matchgate_circuit", [_P, _I, _P, _SV, _SV])
_sig("quantum_matchgate_circuit_forward", [_P, _I, _P, _SV, _QC, _SV])
_sig("quantum_matchgate_circuit_backward", [_P, _I, _P, _QC, _SV, _SV, _P])
_sig("allocate_quantum_circuit_cache", [_I, _I, _QC])
_sig("free_quantum_circuit_cache", [_QC], None)
_sig("apply_matchgate_brickwall_unitary", [_P, _I, _P, _SV, _SV])
_sig("matchgate_brickwall_unitary_forward", [_P, _I, _P, _SV, _QC, _SV])
_sig("matchgate_brickwall_unitary_backward", [_P, _I, _P, _QC, _SV, _SV, _P])
_sig("matchgate_brickwall_unitary_backward_hessian", [_P, _I, _P, _QC, _SV, _SV, _P, _P])
_sig("matchgate_ti_brickwall_unitary_backward_hessian", [_P, _I, _P, _QC, _SV, _SV, _P, _P])
_sig("matchgate_ti_brickwall_unitary_backward", [_P, _I, _P, _QC, _SV, _SV, _P])
# host algebra and trust region
for _n in ("multiply_matchgates", "real_to_tangent_matchgate", "project_tangent_matchgate"):
    _sig(_n, [_P, _P, _P], None)
_sig("tangent_matchgate_to_real", [_P, _P, _P], None)
_sig("polar_matchgate_factor", [_P, _P], None)
_sig("retract_matchgate", [_P, _P, _P], None)
_sig("inverse_mat2x2", [_P, _P])
_sig("truncated_cg_hmat", [_P, _P, _I, _D, ctypes.POINTER(TruncatedCGParams), _P], ctypes.c_bool)


class GPUError(RuntimeError):
    pass


def _check(ret, what):
    if ret != 0:
        raise GPUError(f"{what} failed ({ret}): {lib.mg_last_error().decode(errors='replace')}")


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return a.ctypes.data_as(_P)


def _perm_ptrs(perms):
    """const int* perms[] from an (nlayers, L) int array (keeps the buffers alive)."""
    perms = _i(perms)
    arr = (_P * perms.shape[0])(*[perms[k].ctypes.data for k in range(perms.shape[0])])
    return perms, arr


# ---------------- device targets and programs ----------------

class Target:
    def __init__(self, handle, keep=()):
        if not handle:
            raise GPUError("target creation failed: " + lib.mg_last_error().decode(errors="replace"))
        self.h = handle
        self._keep = keep

    def __del__(self):
        if getattr(self, "h", None):
            lib.mg_target_free(self.h)
            self.h = None

    @classmethod
    def dense(cls, U):
        U = _c(U)
        L = int(np.log2(U.shape[0]))
        return cls(lib.mg_target_dense(_p(U), L))

    @classmethod
    def brickwall(cls, ulist, uperms, L):
        ulist = _c(ulist)
        perms, arr = _perm_ptrs(uperms)
        return cls(lib.mg_target_brickwall(_p(ulist), perms.shape[0], ctypes.cast(arr, _P), L))

    def densify(self):
        """Materialise a circuit/brickwall target as dense rows on the GPU (mg_target_densify)."""
        _check(lib.mg_target_densify(self.h), "mg_target_densify")
        return self

    @classmethod
    def circuit(cls, gates, wires, L):
        gates, wires = _c(gates), _i(wires)
        return cls(lib.mg_target_circuit(_p(gates), gates.shape[0], _p(wires), L))


class Program:
    def __init__(self, handle, L):
        if not handle:
            raise GPUError("program creation failed: " + lib.mg_last_error().decode(errors="replace"))
        self.h = handle
        self.L = L
        self.P = lib.mg_program_nparams(handle)
        self.G = lib.mg_program_ninst(handle)

    def __del__(self):
        if getattr(self, "h", None):
            lib.mg_program_free(self.h)
            self.h = None

    @classmethod
    def create(cls, L, P, param, wires):
        param, wires = _i(param), _i(wires)
        return cls(lib.mg_program_create(L, P, len(param), _p(param), _p(wires)), L)

    @classmethod
    def brickwall(cls, L, perms):
        perms, arr = _perm_ptrs(perms)
        return cls(lib.mg_program_brickwall(L, perms.shape[0], ctypes.cast(arr, _P)), L)


def evaluate(target, program, gates, Z=None, what=EVAL_F_GRAD_HVP, precision=MG_FP64):
    """f, complex gradient (P,8), D^2 f[Z] (P,8) over all basis columns."""
    gates = _c(gates)
    P = program.P
    f = np.zeros(1, np.complex128)
    dg = np.zeros((P, 8), np.complex128)
    dh = np.zeros((P, 8), np.complex128)
    Zc = _c(Z) if Z is not None else None
    _check(lib.mg_evaluate(target.h, program.h, _p(gates), _p(Zc) if Zc is not None else None, what, precision,
                           _p(f), _p(dg), _p(dh)), "mg_evaluate")
    return f[0], dg, dh


def evaluate_columns(target, program, gates, Z, what, precision, b0, b1):
    """Partial sums over the basis columns [b0, b1)."""
    gates = _c(gates)
    P = program.P
    f = np.zeros(1, np.complex128)
    dg = np.zeros((P, 8), np.complex128)
    dh = np.zeros((P, 8), np.complex128)
    Zc = _c(Z) if Z is not None else None
    _check(lib.mg_evaluate_columns(target.h, program.h, _p(gates), _p(Zc) if Zc is not None else None, what,
                                   precision, b0, b1, _p(f), _p(dg), _p(dh)), "mg_evaluate_columns")
    return f[0], dg, dh


def real_to_directions(gates, v):
    """complex tangent directions Z_i = V_i A(v_i) (P, 8) of real parameters v (8P)
    (real_to_tangent_matchgate per gate, reference src/matchgate.c:296)"""
    gates = _c(gates)
    v = np.ascontiguousarray(v, dtype=np.float64)
    Z = np.empty_like(gates)
    for k in range(gates.shape[0]):
        lib.real_to_tangent_matchgate(_p(v[8 * k:8 * k + 8]), _p(gates[k]), _p(Z[k]))
    return Z


def riemannian_project(gates, dgrad, v=None, dhz=None):
    """Riemannian gradient (8P) and, with v, H v (8P) from summed partial sums (P, 8) of
    evaluate_columns (mg_riemannian_project; host FP64, no device work)."""
    gates = _c(gates)
    P = gates.shape[0]
    dg = _c(dgrad).reshape(P, 8)
    g = np.zeros(8 * P)
    hv = np.zeros(8 * P)
    vv = np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
    dh = _c(dhz).reshape(P, 8) if v is not None else None
    _check(lib.mg_riemannian_project(_p(gates), P, _p(dg), _p(vv) if vv is not None else None,
                                     _p(dh) if dh is not None else None, _p(g), _p(hv)), "mg_riemannian_project")
    return g, (hv if v is not None else None)


def riemannian_gradient_hvp(target, program, gates, v=None, precision=MG_FP64):
    gates = _c(gates)
    P = program.P
    f = np.zeros(1, np.complex128)
    g = np.zeros(8 * P)
    hv = np.zeros(8 * P)
    vv = np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
    _check(lib.mg_riemannian_gradient_hvp(target.h, program.h, _p(gates), _p(vv) if vv is not None else None,
                                          precision, _p(f), _p(g), _p(hv)), "mg_riemannian_gradient_hvp")
    return f[0], g, (hv if v is not None else None)


def hessian(target, program, gates, precision=MG_FP64):
    gates = _c(gates)
    m = 8 * program.P
    f = np.zeros(1, np.complex128)
    dg = np.zeros((program.P, 8), np.complex128)
    H = np.zeros((m, m), np.complex128)
    _check(lib.mg_hessian(target.h, program.h, _p(gates), precision, _p(f), _p(dg), _p(H)), "mg_hessian")
    return f[0], dg, H


def gradient_vector_hessian_matrix(target, program, gates, precision=MG_FP64):
    gates = _c(gates)
    m = 8 * program.P
    f = np.zeros(1, np.complex128)
    g = np.zeros(m)
    H = np.zeros((m, m))
    _check(lib.mg_gradient_vector_hessian_matrix(target.h, program.h, _p(gates), precision, _p(f), _p(g), _p(H)),
           "mg_gradient_vector_hessian_matrix")
    return f[0], g, H


def optimize(target, program, gates, niter, hvp=True, precision=MG_FP64, precision_switch=0, params=None):
    gates = _c(gates)
    params = params if params is not None else rtr_default_params(16 * program.P)
    f_iter = np.zeros(niter + 1)
    f_citer = np.zeros(niter + 1, np.complex128)
    out = np.zeros_like(gates)
    nh = np.zeros(1, np.int32)
    _check(lib.mg_optimize(target.h, program.h, _p(gates), ctypes.byref(params), niter, int(hvp), precision,
                           precision_switch, _p(f_iter), _p(f_citer), _p(out), _p(nh)), "mg_optimize")
    return f_iter, f_citer, out, int(nh[0])


# ---------------- batched evaluation and serving ----------------

def _requests(requests):
    n = len(requests)
    arr = (Request * n)()
    keep, outs = [], []
    for k, r in enumerate(requests):
        prog = r["program"]
        gates = _c(r["gates"])
        what = int(r.get("what", EVAL_F_GRAD_HVP))
        v = r.get("v")
        vv = np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
        f = np.zeros(1, np.complex128)
        g = np.zeros(8 * prog.P)
        hv = np.zeros(8 * prog.P)
        keep += [gates, vv, f, g, hv]
        arr[k] = Request(r["target"].h, prog.h, gates.ctypes.data, vv.ctypes.data if vv is not None else None, what,
                         int(r.get("precision", MG_FP64)), f.ctypes.data, g.ctypes.data, hv.ctypes.data,
                         int(r.get("b0", 0)), int(r.get("b1", 0)))
        outs.append((f, g, hv, what))
    return arr, keep, outs


def batch_autotune(requests, precision, reps=2):
    """Pick the fastest batched-kernel variant for `precision` on calibration requests
    (mg_batch_autotune).  Returns (threads, window_bits, workgroups_per_cu, kernel_ms)."""
    arr, keep, _ = _requests(requests)
    t, w, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    ms = ctypes.c_float()
    _check(lib.mg_batch_autotune(arr, len(requests), int(precision), int(reps), ctypes.byref(t), ctypes.byref(w),
                                 ctypes.byref(p), ctypes.byref(ms)), "mg_batch_autotune")
    return t.value, w.value, p.value, ms.value


def evaluate_batch(requests):
    """Evaluate many requests in shared launches.

    requests: list of dicts with keys target, program, gates, v (or None), what,
    precision and optionally b0, b1 (a column range: partial sums).  Returns a list of
    (f, grad_vec, hvp) like riemannian_gradient_hvp
    (grad_vec None for what == EVAL_F, hvp None unless what == EVAL_F_GRAD_HVP).
    """
    arr, keep, outs = _requests(requests)
    _check(lib.mg_evaluate_batch(arr, len(requests)), "mg_evaluate_batch")
    return [(f[0], g if what >= EVAL_F_GRAD else None, hv if what == EVAL_F_GRAD_HVP else None)
            for f, g, hv, what in outs]


BATCH_VARIANTS = {MG_FP32: [(256, 4), (256, 3), (256, 2), (512, 3), (512, 2)],
                  MG_FP64: [(256, 3), (256, 2), (512, 2)]}


def batch_set_kernel(precision, threads, window_bits):
    _check(lib.mg_batch_set_kernel(int(precision), int(threads), int(window_bits)), "mg_batch_set_kernel")


def batch_set_single_range(min_qubits, max_qubits):
    lib.mg_batch_set_single_range(int(min_qubits), int(max_qubits))


def batch_get_kernel(precision):
    a, b = ctypes.c_int(), ctypes.c_int()
    lib.mg_batch_get_kernel(int(precision), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def batch_set_grid_per_cu(n):
    lib.mg_batch_set_grid_per_cu(int(n))


def batch_set_shape(precision, L, threads_per_column=0, columns_per_item=0):
    lib.mg_batch_set_shape(int(precision), int(L), int(threads_per_column), int(columns_per_item))


def batch_get_shape(precision, L):
    a, b = ctypes.c_int(), ctypes.c_int()
    lib.mg_batch_get_shape(int(precision), int(L), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def batch_last_kernel_ms():
    return float(lib.mg_batch_last_kernel_ms())


def batch_set_graphs(on):
    """replay repeated batched launches from captured HIP graphs (default on)"""
    lib.mg_batch_set_graphs(int(bool(on)))


def batch_last_graph_stats():
    """(replayed, newly captured) launches of the last evaluate_batch call"""
    r, c = ctypes.c_int(0), ctypes.c_int(0)
    lib.mg_batch_last_graph_stats(ctypes.byref(r), ctypes.byref(c))
    return r.value, c.value


def serve(jobs, batched=True, precision=MG_FP32, precision_switch=3, max_active=0, hbm_budget=0.0,
          realtime=True, densify_qubits=0, progress_s=0.0, tr_reuse=False, time_limit_s=0.0):
    """Run a stream of trust-region optimisation jobs (mg_serve).

    jobs: list of dicts with keys program (Program), gates (P, 8), niter, arrival (s),
    and either target (a resident Target, shared and kept by the caller), U (dense target)
    or target_gates + target_wires (circuit target).
    tr_reuse: H eta of each ratio test from the truncated CG's own products and f + gradient
    kept across rejected steps (mg_serve_options.tr_reuse); False keeps the control flow of
    riemannian_trust_region_optimize_hvp.
    time_limit_s > 0: no evaluation starts later; jobs cut short or not started end with status -2.
    Returns (results, stats): per job a dict with f_iter, gates, t_admit, t_done,
    latency (t_done - arrival), nhvp, status; stats a dict of mg_serve_stats.
    """
    n = len(jobs)
    arr = (Job * n)()
    keep, outs = [], []
    for k, j in enumerate(jobs):
        prog = j["program"]
        g0 = _c(j["gates"])
        niter = int(j["niter"])
        f_iter = np.zeros(niter + 1)
        gout = np.zeros_like(g0)
        res_t = j.get("target")
        U = _c(j["U"]) if res_t is None and j.get("U") is not None else None
        tg = _c(j["target_gates"]) if res_t is None and U is None else None
        tw = _i(j["target_wires"]) if res_t is None and U is None else None
        keep += [prog, g0, U, tg, tw, res_t]
        arr[k] = Job(prog.L, prog.h, g0.ctypes.data, U.ctypes.data if U is not None else None,
                     tg.shape[0] if tg is not None else 0, tg.ctypes.data if tg is not None else None,
                     tw.ctypes.data if tw is not None else None, niter, float(j.get("arrival", 0.0)),
                     f_iter.ctypes.data, gout.ctypes.data, 0.0, 0.0, 0, 0,
                     res_t.h if res_t is not None else None)
        outs.append((f_iter, gout))
    opt = ServeOptions(int(bool(batched)), int(precision), int(precision_switch), int(max_active), float(hbm_budget),
                       int(bool(realtime)), int(densify_qubits), float(progress_s), int(bool(tr_reuse)),
                       float(time_limit_s))
    st = ServeStats()
    _check(lib.mg_serve(arr, n, ctypes.byref(opt), ctypes.byref(st)), "mg_serve")
    res = []
    for k, (f_iter, gout) in enumerate(outs):
        a = arr[k]
        arrival = float(jobs[k].get("arrival", 0.0)) if realtime else 0.0
        res.append({"f_iter": f_iter, "gates": gout, "t_admit": a.t_admit, "t_done": a.t_done,
                    "latency": a.t_done - arrival, "nhvp": a.nhvp, "status": a.status})
    stats = {name: getattr(st, name) for name, _ in ServeStats._fields_}
    return res, stats


def set_checkpointing(mode):
    lib.mg_set_checkpointing(int(mode))


def set_lds_states(mode):
    lib.mg_set_lds_states(int(mode))


def set_register_kernel(mode, threads=0):
    lib.mg_set_register_kernel(int(mode), int(threads))


def set_window_kernel(mode):
    lib.mg_set_window_kernel(int(mode))


def set_tile_kernel(mode):
    """-1: tiled kernel where it applies (default), 0: never"""
    lib.mg_set_tile_kernel(int(mode))


def set_forward_mfma(mode):
    """FP32 tiled sweeps: -1 forward pass on the matrix cores (default), 0 gate by gate"""
    lib.mg_set_forward_mfma(int(mode))


def set_device_trust_region(mode):
    """optimize with hvp=True: -1 trust region on the device, one graph per iteration (default,
    where it applies), 0 the host loop, 2 on the device with reuse (H eta of the ratio test from
    the truncated CG's products, f + gradient at an accepted trial point kept)"""
    lib.mg_set_device_trust_region(int(mode))


def last_trust_region_graphs():
    """graph launches of the last optimize on the device (0: it ran on the host loop)"""
    return int(lib.mg_last_trust_region_graphs())


def set_materialize_kernel(mode):
    """rows of program targets beyond 12 qubits: -1 automatic (default, matrix cores), 0 register
    tiles, 1 matrix cores"""
    lib.mg_set_materialize_kernel(int(mode))


def last_sweep_kernel():
    return lib.mg_last_sweep_kernel().decode()


def set_sweep_threads(n):
    lib.mg_set_sweep_threads(int(n))


def set_grid_per_cu(n):
    lib.mg_set_grid_per_cu(int(n))


def set_kernel_timing(on):
    lib.mg_set_kernel_timing(int(on))


def last_kernel_ms():
    return float(lib.mg_last_kernel_ms())


def last_sweep_info():
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    d = ctypes.c_long()
    lib.mg_last_sweep_info(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d))
    return {"lds_states": bool(a.value), "ckpt": bool(b.value), "grid": c.value, "lds_bytes": d.value,
            "reg_threads": int(lib.mg_last_sweep_threads()), "kernel": last_sweep_kernel()}


# ---------------- reference-API helpers ----------------

def dense_linear_func(U):
    """A linear_func callback applying the dense matrix U (the zgemv 'ufunc' of
    examples/hubbard1d/hubbard1d_opt.c:41-52)."""
    U = _c(U)
    S = U.shape[0]

    def fn(psi_p, udata, out_p):
        psi = ctypes.cast(psi_p, ctypes.POINTER(Statevector)).contents
        out = ctypes.cast(out_p, ctypes.POINTER(Statevector)).contents
        x = np.ctypeslib.as_array(ctypes.cast(psi.data, ctypes.POINTER(ctypes.c_double)), shape=(2 * S,)).view(np.complex128)
        y = np.ctypeslib.as_array(ctypes.cast(out.data, ctypes.POINTER(ctypes.c_double)), shape=(2 * S,)).view(np.complex128)
        y[:] = U @ x
        return 0

    cb = LINEAR_FUNC(fn)
    cb._U = U
    return cb


def statevector(data):
    data = _c(data)
    sv = Statevector(data.ctypes.data, int(np.log2(data.shape[0])))
    sv._keep = data
    return sv


def statevector_array(data):
    """data: (2^L, nstates) interleaved"""
    data = _c(data)
    sa = StatevectorArray(data.ctypes.data, int(np.log2(data.shape[0])), data.shape[1])
    sa._keep = data
    return sa


# ---- HDF5 files of the reference drivers (include/mg_hdf5.h, util.h) -------------------------
_H5T = {np.dtype(np.float64): 1, np.dtype(np.float32): 2, np.dtype(np.int32): 3, np.dtype(np.int64): 4,
        np.dtype(np.uint8): 5, np.dtype(np.int8): 6, np.dtype(np.int16): 7, np.dtype(np.uint32): 8,
        np.dtype(np.uint64): 9}
_H5T_BASE = 0x4d47485400000000
_HID = ctypes.c_int64
_sig("mg_h5_fopen", [ctypes.c_char_p, ctypes.c_uint], _HID)
_sig("mg_h5_fcreate", [ctypes.c_char_p, ctypes.c_uint], _HID)
_sig("mg_h5_fclose", [_HID])
_sig("get_hdf5_dataset_ndims", [_HID, ctypes.c_char_p, _P])
_sig("get_hdf5_dataset_dims", [_HID, ctypes.c_char_p, _P])
_sig("read_hdf5_dataset", [_HID, ctypes.c_char_p, _HID, _P])
_sig("read_hdf5_attribute", [_HID, ctypes.c_char_p, _HID, _P])
_sig("write_hdf5_dataset", [_HID, ctypes.c_char_p, ctypes.c_int, _P, _HID, _HID, _P])
_sig("write_hdf5_scalar_attribute", [_HID, ctypes.c_char_p, _HID, _HID, _P])


class H5Error(RuntimeError):
    pass


def _h5t(dtype):
    return _HID(_H5T_BASE + _H5T[np.dtype(dtype)])


def h5_read(path, datasets=(), attributes=(), dtype=np.float64, attr_dtype=None):
    """Read named datasets and root attributes of an HDF5 file with the library's reader
    (the reference's read_hdf5_dataset / read_hdf5_attribute, src/util.c:115-161).
    Returns (dict name -> ndarray of `dtype`, dict name -> scalar of `attr_dtype` or `dtype`)."""
    f = lib.mg_h5_fopen(os.fsencode(path), 0)
    if f < 0:
        raise H5Error(f"cannot open {path}")
    try:
        out, att = {}, {}
        for name in datasets:
            nd = ctypes.c_int(0)
            if lib.get_hdf5_dataset_ndims(f, name.encode(), ctypes.byref(nd)) != 0:
                raise H5Error(f"{path}: no dataset {name}")
            dims = (ctypes.c_ulonglong * max(1, nd.value))()
            lib.get_hdf5_dataset_dims(f, name.encode(), dims)
            a = np.zeros(tuple(dims[i] for i in range(nd.value)), dtype)
            if lib.read_hdf5_dataset(f, name.encode(), _h5t(dtype), _p(a)) != 0:
                raise H5Error(f"{path}: reading {name} failed")
            out[name] = a
        for name in attributes:
            a = np.zeros(1, attr_dtype or dtype)
            if lib.read_hdf5_attribute(f, name.encode(), _h5t(a.dtype), _p(a)) != 0:
                raise H5Error(f"{path}: reading attribute {name} failed")
            att[name] = a[0]
        return out, att
    finally:
        lib.mg_h5_fclose(f)


def h5_write(path, datasets=None, attributes=None, store=None):
    """Write an HDF5 file with the library's writer (the reference's write_hdf5_dataset /
    write_hdf5_scalar_attribute, src/util.c:168-236).  datasets / attributes: dict name ->
    array / scalar (numpy dtypes); `store` maps a name to its file type (default: its own)."""
    store = store or {}
    f = lib.mg_h5_fcreate(os.fsencode(path), 2)
    if f < 0:
        raise H5Error(f"cannot create {path}")
    ok = True
    try:
        for name, a in (datasets or {}).items():
            a = np.ascontiguousarray(a)
            dims = (ctypes.c_ulonglong * max(1, a.ndim))(*a.shape)
            ok &= lib.write_hdf5_dataset(f, name.encode(), a.ndim, dims, _h5t(store.get(name, a.dtype)), _h5t(a.dtype),
                                         _p(a)) == 0
        for name, v in (attributes or {}).items():
            a = np.asarray([v])
            ok &= lib.write_hdf5_scalar_attribute(f, name.encode(), _h5t(store.get(name, a.dtype)), _h5t(a.dtype),
                                                  _p(a)) == 0
    finally:
        ok &= lib.mg_h5_fclose(f) == 0
    if not ok:
        raise H5Error(f"writing {path} failed")
