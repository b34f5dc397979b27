"""The C-ABI from C: tests/c/brickwall_opt_driver.c includes only include/*.h, links only
librqcopt_mg.so, and runs the reference driver's call sequence (examples/hubbard1d/
hubbard1d_opt.c: linear_func target, optimize_matchgate_brickwall_circuit_hmat, then the
HVP optimiser) on the reference's q8 Fermi-Hubbard input; its trajectory must match the
reference-C output file to 1e-9."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from golden_io import cvec, load, perms

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "brickwall_opt_driver")


def test_c_driver_reproduces_reference_trajectory(tmp_path):
    assert os.path.exists(DRIVER), "build the C driver first (__graft_entry__.build())"
    nl, L, niter = 5, 8, 10
    U = cvec(load("hubbard1d_q8_unitary_t0.25s_g1.50_init")["expiH"]).reshape(256, 256)
    d = load(f"hubbard1d_suzuki2_n{nl}_q8_init")
    v = np.asarray(d["vlist"]).reshape(nl, 8, 2)
    V = np.ascontiguousarray(v[..., 0] + 1j * v[..., 1])
    pm = np.ascontiguousarray(perms(d, nl), np.int32)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fin.write_bytes(struct.pack("3i", L, nl, niter) + np.ascontiguousarray(U).tobytes() + V.tobytes() + pm.tobytes())
    print("case: C driver, Hubbard q8, 5 layers, 10 iterations (hmat, then hvp)", flush=True)
    sys.stdout.flush()
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    raw = fout.read_bytes()
    k = 0
    f_hmat = np.frombuffer(raw, np.float64, niter + 1, k); k += 8 * (niter + 1)
    v_hmat = np.frombuffer(raw, np.complex128, nl * 8, k).reshape(nl, 8); k += 16 * nl * 8
    f_hvp = np.frombuffer(raw, np.float64, niter + 1, k); k += 8 * (niter + 1)
    v_hvp = np.frombuffer(raw, np.complex128, nl * 8, k).reshape(nl, 8)
    ref = load(f"hubbard1d_suzuki2_n{nl}_q8_iter10_opt")
    vr = np.asarray(ref["vlist_opt"])
    vref = vr[..., 0] + 1j * vr[..., 1]
    assert np.max(np.abs(f_hmat - ref["f_iter"])) < 1e-9 and np.max(np.abs(v_hmat - vref)) < 1e-9
    # the HVP-driven trust region follows the same iterates (tCG on exact Hessian-vector products)
    assert np.max(np.abs(f_hvp - ref["f_iter"])) < 1e-9 and np.max(np.abs(v_hvp - vref)) < 1e-9
