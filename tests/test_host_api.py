"""CPU checks of the C-ABI's host-side behaviour: argument validation and error
returns (before any device work), and the host FP64 matchgate algebra against the
oracle and the reference's conventions (src/matchgate.c)."""
import ctypes
import os

import numpy as np
import pytest

import oracle as orc
from golden_io import haar_matchgates


def lib():
    import rqcopt_mg

    return rqcopt_mg


def test_program_validation_errors():
    mg = lib()
    P = ctypes.c_void_p
    for param, wires, L, nparams in (([0], [0, 0], 4, 1),      # same wire twice
                                     ([0], [0, 4], 4, 1),      # wire out of range
                                     ([1], [0, 1], 4, 1),      # parameter index out of range
                                     ([0], [0, -1], 4, 1)):
        pa, wi = np.asarray(param, np.int32), np.asarray(wires, np.int32)
        h = mg.lib.mg_program_create(L, nparams, len(pa), pa.ctypes.data, wi.ctypes.data)
        assert not h
        assert mg.lib.mg_last_error()
    assert not mg.lib.mg_program_create(4, 1, 0, None, None)
    pa, wi = np.zeros(1, np.int32), np.asarray([1, 2], np.int32)
    h = mg.lib.mg_program_create(4, 1, 1, pa.ctypes.data, wi.ctypes.data)
    assert h and mg.lib.mg_program_nparams(h) == 1 and mg.lib.mg_program_ninst(h) == 1
    mg.lib.mg_program_free(h)


def test_batch_and_serve_reject_bad_arguments_before_device_work():
    mg = lib()
    assert mg.lib.mg_evaluate_batch(None, 3) == -1
    req = mg.Request()  # null target and program
    assert mg.lib.mg_evaluate_batch(ctypes.byref(req), 1) == -1
    job = mg.Job()      # null program
    opt = mg.ServeOptions(1, mg.MG_FP32, 3, 0, 0.0, 0, 0)
    assert mg.lib.mg_serve(ctypes.byref(job), 1, ctypes.byref(opt), None) == -1
    assert mg.lib.mg_serve(None, 1, ctypes.byref(opt), None) == -1
    # kernel variants that are not compiled are refused; compiled ones accepted
    assert mg.lib.mg_batch_set_kernel(mg.MG_FP64, 512, 3) == -1
    assert mg.lib.mg_batch_set_kernel(mg.MG_FP32, 1024, 4) == -1
    t, w = mg.batch_get_kernel(mg.MG_FP32)
    assert mg.lib.mg_batch_set_kernel(mg.MG_FP32, t, w) == 0


def test_host_algebra_matches_oracle():
    mg = lib()
    rng = np.random.default_rng(4)
    V = haar_matchgates(rng, 3)
    for u in V:
        r = rng.standard_normal(8)
        z = np.zeros(8, np.complex128)
        mg.lib.real_to_tangent_matchgate(r.ctypes.data, u.ctypes.data, z.ctypes.data)
        assert np.max(np.abs(z - orc.real_to_tangent(r, u))) < 1e-14
        back = np.zeros(8)
        mg.lib.tangent_matchgate_to_real(u.ctypes.data, z.ctypes.data, back.ctypes.data)
        assert np.max(np.abs(back - r)) < 1e-13
        w = np.zeros(8, np.complex128)
        mg.lib.retract_matchgate(u.ctypes.data, r.ctypes.data, w.ctypes.data)
        for blk in (w[:4].reshape(2, 2), w[4:].reshape(2, 2)):  # retraction lands on U(2) x U(2)
            assert np.max(np.abs(blk.conj().T @ blk - np.eye(2))) < 1e-13
        assert np.max(np.abs(w - orc.retract(u, r))) < 1e-13


def test_inverse_of_singular_block_fails_like_the_reference():
    mg = lib()
    a = np.asarray([1, 2, 2, 4], np.complex128)   # det = 0 (reference src/matchgate.c:433-436)
    inv = np.zeros(4, np.complex128)
    assert mg.lib.inverse_mat2x2(a.ctypes.data, inv.ctypes.data) == -1
    b = np.asarray([2, 1j, 0, 1], np.complex128)
    assert mg.lib.inverse_mat2x2(b.ctypes.data, inv.ctypes.data) == 0
    assert np.max(np.abs(inv.reshape(2, 2) @ b.reshape(2, 2) - np.eye(2))) < 1e-15


def test_ctypes_structs_match_the_c_header(tmp_path):
    """The Python binding's ctypes mirrors of the public structs have the C header's layout
    (every field offset and the size), compiled here with gcc against include/mg_gpu.h."""
    import shutil
    import subprocess

    import rqcopt_mg as mg

    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = {"mg_request": mg.Request, "mg_job": mg.Job, "mg_serve_options": mg.ServeOptions,
               "mg_serve_stats": mg.ServeStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mg_gpu.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "abi"
    inc = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "include")
    r = subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_reldist_in_the_header(tmp_path):
    """reldist (src/util.h:13-16) is usable from C and C++ callers of include/util.h."""
    import subprocess

    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    src = """#include <stdio.h>
#include "util.h"
int main(void) {
    numeric x = 3.0, y = 4.0;
    printf("%.17g %.17g\\n", reldist(x, y, 1e-14), reldist(0.0, 0.0, 2.0));
    return 0;
}
"""
    for cc, ext in (("gcc", "c"), ("g++", "cpp")):
        f = tmp_path / f"r.{ext}"
        f.write_text(src)
        exe = tmp_path / f"r_{ext}"
        subprocess.run([cc, "-I", inc, str(f), "-o", str(exe), "-lm"], check=True)
        a, b = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
        assert abs(float(a) - 1.0 / 7.0) < 1e-16 and float(b) == 0.0
