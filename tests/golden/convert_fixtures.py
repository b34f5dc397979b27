"""Convert the reference's HDF5 golden fixtures into HDF5-free ``.npz`` files.

Run once in a container that holds the reference checkout (``/root/reference``)
and the HDF5 command-line tools (``/opt/conda/bin/h5dump``).  Machines that run
the tests need neither: the converted fixtures are committed next to this script
and the tests read only the ``.npz`` files (``numpy.load`` with
``allow_pickle=False``).

Every dataset of each listed file is exported with ``h5dump -b LE`` (raw
little-endian bytes) and stored under its HDF5 path with ``/`` replaced by
``.`` (``V/center0`` -> ``V.center0``).  Scalar attributes of the trust-region
output files (``nlayers``, ``nqubits``, ``Walltime``, ...) are stored as
``attr.<name>``.

Usage:  python tests/golden/convert_fixtures.py
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

H5DUMP = "/opt/conda/bin/h5dump"
REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))

# reference fixture files used by the tests (test/test_*.c read them)
TEST_FILES = [
    "test_apply_matchgate_cplx",
    "test_apply_matchgate_backward_cplx",
    "test_apply_matchgate_backward_array_cplx",
    "test_apply_matchgate_to_array_cplx",
    "test_apply_matchgate_placeholder_cplx",
    "test_apply_matchgate_brickwall_unitary_cplx",
    "test_matchgate_brickwall_unitary_backward_cplx",
    "test_matchgate_brickwall_unitary_backward_hessian_cplx",
    "test_matchgate_circuit_unitary_target_cplx",
    "test_matchgate_brickwall_unitary_target_cplx",
    "test_matchgate_brickwall_unitary_target_gradient_hessian_cplx",
    "test_matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix_cplx",
    "test_multiply_matchgate_cplx",
    "test_project_tangent_matchgate_cplx",
    "test_polar_matchgate_factor_cplx",
    "test_real_to_tangent_matchgate_cplx",
    # translation-invariant Hessian shortcut (test/test_matchgate_target.c:506-772)
    "test_matchgate_ti_brickwall_unitary_target_gradient_hessian_cplx",
    "test_matchgate_ti_brickwall_unitary_target_gradient_hessian2_cplx",
    "test_matchgate_ti_brickwall_unitary_target_gradient_hessian3_cplx",
    "test_matchgate_ti_brickwall_unitary_target_gradient_hessian4_cplx",
]

# trust-region inputs and outputs of the L=8 Fermi-Hubbard example
# (examples/hubbard1d/hubbard1d_opt.c reads the inputs and writes the outputs)
HUB = "examples/hubbard1d"
EXAMPLE_FILES = [
    (f"{HUB}/opt_in/q8/hubbard1d_q8_unitary_t0.25s_g1.50_init.hdf5", "hubbard1d_q8_unitary_t0.25s_g1.50_init"),
    (f"{HUB}/opt_in/q8/hubbard1d_suzuki2_n5_q8_u29_t0.25s_g1.50_init.hdf5", "hubbard1d_suzuki2_n5_q8_init"),
    (f"{HUB}/opt_in/q8/hubbard1d_suzuki2_n9_q8_u29_t0.25s_g1.50_init.hdf5", "hubbard1d_suzuki2_n9_q8_init"),
    (f"{HUB}/opt_out/q8/hubbard1d_suzuki2_n5_q8_u0_t0.25s_g1.50_iter10_inv0_opt.hdf5", "hubbard1d_suzuki2_n5_q8_iter10_opt"),
    (f"{HUB}/opt_out/q8/hubbard1d_suzuki2_n5_q8_u0_t0.25s_g1.50_iter30_inv0_opt.hdf5", "hubbard1d_suzuki2_n5_q8_iter30_opt"),
    (f"{HUB}/opt_out/q8/hubbard1d_suzuki2_n9_q8_u0_t0.25s_g1.50_iter10_inv0_opt.hdf5", "hubbard1d_suzuki2_n9_q8_iter10_opt"),
    (f"{HUB}/opt_in/q8/hubbard1d_suzuki2_n13_q8_u29_t0.25s_g1.50_init.hdf5", "hubbard1d_suzuki2_n13_q8_init"),
    (f"{HUB}/opt_out/q8/hubbard1d_suzuki2_n13_q8_u0_t0.25s_g1.50_iter10_inv0_opt.hdf5", "hubbard1d_suzuki2_n13_q8_iter10_opt"),
    (f"{HUB}/opt_in/q8/hubbard1d_suzuki2_n21_q8_u29_t0.25s_g1.50_init.hdf5", "hubbard1d_suzuki2_n21_q8_init"),
    (f"{HUB}/opt_out/q8/hubbard1d_suzuki2_n21_q8_u0_t0.25s_g1.50_iter10_inv0_opt.hdf5", "hubbard1d_suzuki2_n21_q8_iter10_opt"),
    (f"{HUB}/opt_in/q8/hubbard1d_suzuki4_n21_q8_u29_t0.25s_g1.50_init.hdf5", "hubbard1d_suzuki4_n21_q8_init"),
    (f"{HUB}/opt_out/q8/hubbard1d_suzuki4_n21_q8_u0_t0.25s_g1.50_iter10_inv0_opt.hdf5", "hubbard1d_suzuki4_n21_q8_iter10_opt"),
]

_DTYPES = {
    "H5T_IEEE_F64LE": "<f8",
    "H5T_IEEE_F32LE": "<f4",
    "H5T_STD_I64LE": "<i8",
    "H5T_STD_I32LE": "<i4",
    "H5T_STD_U8LE": "<u1",
}


def _run(args):
    return subprocess.run(args, check=True, capture_output=True, text=True).stdout


def _datasets(path):
    out = _run([H5DUMP, "-n", path])
    return [m.group(1) for m in re.finditer(r"^\s*dataset\s+(\S+)\s*$", out, re.M)]


def _header(path, ds):
    out = _run([H5DUMP, "-H", "-d", ds, path])
    dt = re.search(r"DATATYPE\s+(\S+)", out).group(1)
    sp = re.search(r"DATASPACE\s+(SCALAR|SIMPLE\s*\{\s*\(([^)]*)\))", out)
    shape = () if sp.group(1) == "SCALAR" else tuple(int(v) for v in sp.group(2).split(","))
    return _DTYPES[dt], shape


def _attributes(path):
    """Scalar numeric attributes of the root group."""
    out = _run([H5DUMP, "-A", "-H", path])
    names = re.findall(r'ATTRIBUTE\s+"([^"]+)"', out)
    attrs = {}
    for name in names:
        txt = _run([H5DUMP, "-a", "/" + name, path])
        dt = re.search(r"DATATYPE\s+(\S+)", txt).group(1)
        if dt not in _DTYPES:
            continue
        val = re.search(r"\(0\):\s*([-+0-9.eE]+)", txt)
        if val is not None:
            attrs["attr." + name] = np.array(float(val.group(1)) if "F" in dt else int(val.group(1)))
    return attrs


def convert(path, name):
    arrays = {}
    with tempfile.TemporaryDirectory() as tmp:
        for ds in _datasets(path):
            dtype, shape = _header(path, ds)
            binf = os.path.join(tmp, "x.bin")
            _run([H5DUMP, "-d", ds, "-b", "LE", "-o", binf, path])
            data = np.fromfile(binf, dtype=dtype)
            arrays[ds.lstrip("/").replace("/", ".")] = data.reshape(shape)
    arrays.update(_attributes(path))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrays)
    return arrays


def main():
    if not os.path.isdir(REF) or not os.path.exists(H5DUMP):
        sys.exit("needs the reference checkout and h5dump")
    for name in TEST_FILES:
        a = convert(os.path.join(REF, "test", "data", name + ".hdf5"), name)
        print(f"{name}: {len(a)} datasets")
    for rel, name in EXAMPLE_FILES:
        a = convert(os.path.join(REF, rel), name)
        print(f"{name}: {len(a)} datasets")


if __name__ == "__main__":
    main()
