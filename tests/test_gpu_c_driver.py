"""The C-ABI from C: tests/c/brickwall_opt_driver.c includes only the headers in include/,
links only librqcopt_mg.so, and runs the reference driver's call sequence and file I/O
(examples/hubbard1d/hubbard1d_opt.c: HDF5 input files, linear_func target,
optimize_matchgate_brickwall_circuit_hmat, then the HVP optimiser, HDF5 output files) on the
reference's q8 Fermi-Hubbard input file; both trajectories must match the reference-C output
file to 1e-9."""
import os
import sys

import numpy as np
import pytest

from golden_io import GOLDEN, cvec, load

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "brickwall_opt_driver")
H5DIR = os.path.join(GOLDEN, "hdf5")


def test_c_driver_reproduces_reference_trajectory(tmp_path):
    import subprocess

    import rqcopt_mg as mg

    assert os.path.exists(DRIVER), "build the C driver first (__graft_entry__.build())"
    niter = 10
    # the target unitary file in the driver's input layout (expiH: (256, 256, 2) doubles)
    U = cvec(load("hubbard1d_q8_unitary_t0.25s_g1.50_init")["expiH"]).reshape(256, 256)
    unitary = tmp_path / "hubbard1d_q8_unitary_t0.25s_g1.50_init.hdf5"
    mg.h5_write(unitary, {"expiH": np.stack([U.real, U.imag], axis=-1)})
    init = os.path.join(H5DIR, "hubbard1d_suzuki2_n5_q8_u29_t0.25s_g1.50_init.hdf5")
    out_hmat, out_hvp = tmp_path / "opt_hmat.hdf5", tmp_path / "opt_hvp.hdf5"
    print("case: C driver, Hubbard q8, 5 layers, 10 iterations (hmat, then hvp), HDF5 in and out", flush=True)
    sys.stdout.flush()
    r = subprocess.run([DRIVER, str(unitary), init, str(niter), str(out_hmat), str(out_hvp)], capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    ref, ref_at = mg.h5_read(os.path.join(H5DIR, "hubbard1d_suzuki2_n5_q8_u0_t0.25s_g1.50_iter10_inv0_opt.hdf5"),
                             ["f_iter", "vlist_opt"], ["nqubits", "nlayers"], attr_dtype=np.int32)
    vref = ref["vlist_opt"][..., 0] + 1j * ref["vlist_opt"][..., 1]
    for path in (out_hmat, out_hvp):
        got, at = mg.h5_read(path, ["f_iter", "vlist_opt"], ["nqubits", "nlayers"], attr_dtype=np.int32)
        assert at == ref_at
        assert got["vlist_opt"].shape == ref["vlist_opt"].shape
        v = got["vlist_opt"][..., 0] + 1j * got["vlist_opt"][..., 1]
        # the HVP-driven trust region follows the same iterates (tCG on exact Hessian-vector products)
        assert np.max(np.abs(got["f_iter"] - ref["f_iter"])) < 1e-9, path
        assert np.max(np.abs(v - vref)) < 1e-9, path
