"""The library's HDF5-free file I/O (include/mg_hdf5.h, util.h): the reference drivers'
helpers read_hdf5_dataset / read_hdf5_attribute / write_hdf5_dataset /
write_hdf5_scalar_attribute (reference src/util.c:56-236) over the library's own reader and
writer of the HDF5 file format.  No GPU is involved.

- Reading: the reference's own HDF5 files (driver input and output of examples/hubbard1d,
  two test fixtures of test/data, one with nested groups) give exactly the values of the
  ``.npz`` conversions made with h5dump (an independent reader, libhdf5).
- Writing: round trips of every supported type, scalar attributes, more datasets than one
  symbol-table node holds; files in the drivers' output layout are read back by h5dump where
  that tool is installed.
- Errors: missing files and names, float-to-integer and out-of-range conversions, and every
  truncation of a valid file return an error instead of crashing.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from golden_io import GOLDEN, load

H5DIR = os.path.join(GOLDEN, "hdf5")
H5DUMP = "/opt/conda/bin/h5dump"


@pytest.fixture(scope="module")
def mg():
    import rqcopt_mg

    return rqcopt_mg


# (hdf5 file, npz conversion, datasets, attributes)
FILES = [
    ("hubbard1d_suzuki2_n5_q8_u29_t0.25s_g1.50_init", "hubbard1d_suzuki2_n5_q8_init",
     ["vlist", "ulist"] + [f"perm{i}" for i in range(5)] + [f"uperm{i}" for i in range(29)],
     ["J", "g", "t", "nlayers", "nqubits", "ulayers"]),
    ("hubbard1d_suzuki2_n5_q8_u0_t0.25s_g1.50_iter10_inv0_opt", "hubbard1d_suzuki2_n5_q8_iter10_opt",
     ["f_iter", "vlist_opt"], ["nqubits", "nlayers", "Walltime", "FULL_TARGET", "TRANSLATIONAL_INVARIANCE", "MPI"]),
    ("test_matchgate_brickwall_unitary_target_cplx", "test_matchgate_brickwall_unitary_target_cplx",
     ["f0", "f1"] + [f"perm{i}" for i in range(5)] + [f"V/center{i}" for i in range(5)]
     + [f"V/corner{i}" for i in range(5)], []),
    ("test_polar_matchgate_factor_cplx", "test_polar_matchgate_factor_cplx",
     ["a/center", "a/corner", "u/center", "u/corner"], []),
]


@pytest.mark.parametrize("h5,npz,dsets,attrs", FILES, ids=[f[0][:40] for f in FILES])
def test_reader_matches_h5dump_conversion(mg, h5, npz, dsets, attrs):
    ref = load(npz)
    path = os.path.join(H5DIR, h5 + ".hdf5")
    for name in dsets:
        want = np.asarray(ref[name.replace("/", ".")])
        got, _ = mg.h5_read(path, [name], dtype=want.dtype)
        assert got[name].shape == want.shape, name
        assert np.array_equal(got[name], want), name
        if want.dtype.kind == "i":
            # the drivers read int64 files into int (H5T_NATIVE_INT)
            g32, _ = mg.h5_read(path, [name], dtype=np.int32)
            assert np.array_equal(g32[name], want), name
    for name in attrs:
        want = np.asarray(ref["attr." + name])
        _, got = mg.h5_read(path, attributes=[name], attr_dtype=want.dtype)
        if want.dtype.kind == "f":
            # the conversion took float attributes from h5dump's text output (6 significant digits)
            assert abs(got[name] - want[()]) <= 5e-6 * abs(want[()]), name
        else:
            assert got[name] == want[()], name


def test_round_trip_types_and_many_datasets(mg, tmp_path):
    rng = np.random.default_rng(3)
    ds = {
        "vlist_opt": rng.standard_normal((5, 8, 2)),
        "f_iter": -np.arange(11.0) / 7,
        "single": rng.standard_normal((3, 4)).astype(np.float32),
        "bytes": np.arange(17, dtype=np.uint8),
        "i16": np.array([-32768, 0, 32767], np.int16),
        "u64": np.array([0, 2**63 + 5], np.uint64),
        "empty": np.zeros((0, 3)),
    }
    for k in range(30):  # several symbol-table nodes (8 entries each)
        ds[f"perm{k}"] = (np.arange(8) + k) % 8
    at = {"nqubits": np.int32(8), "nlayers": np.int32(5), "Walltime": 1.25, "big": np.int64(-2**40)}
    path = tmp_path / "out.hdf5"
    mg.h5_write(path, ds, at, store={f"perm{k}": np.int64 for k in range(30)})
    for name, a in ds.items():
        got, _ = mg.h5_read(path, [name], dtype=a.dtype)
        assert got[name].shape == a.shape and np.array_equal(got[name], a), name
    for name, v in at.items():
        _, got = mg.h5_read(path, attributes=[name], attr_dtype=np.asarray(v).dtype)
        assert got[name] == v, name
    # stored as int64, read as int32 and as double
    got, _ = mg.h5_read(path, ["perm3"], dtype=np.int32)
    assert np.array_equal(got["perm3"], ds["perm3"])
    got, _ = mg.h5_read(path, ["perm3"], dtype=np.float64)
    assert np.array_equal(got["perm3"], ds["perm3"].astype(float))


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump (libhdf5) not installed")
def test_written_file_reads_in_libhdf5(mg, tmp_path):
    """A file in the drivers' output layout (hubbard1d_opt.c:238-310) written by the library is
    read by libhdf5's h5dump with the same structure, types and values."""
    rng = np.random.default_rng(4)
    v = rng.standard_normal((5, 8, 2))
    f = -np.linspace(1, 2, 11)
    path = str(tmp_path / "opt.hdf5")
    mg.h5_write(path, {"vlist_opt": v, "f_iter": f},
                {"nqubits": np.int32(8), "nlayers": np.int32(5), "Walltime": 0.5, "FULL_TARGET": np.int32(1)})
    hdr = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, timeout=60)
    assert hdr.returncode == 0, hdr.stderr
    assert 'DATASET "vlist_opt"' in hdr.stdout and "SIMPLE { ( 5, 8, 2 ) / ( 5, 8, 2 ) }" in hdr.stdout
    assert 'ATTRIBUTE "nqubits"' in hdr.stdout and "H5T_STD_I32LE" in hdr.stdout and "H5T_IEEE_F64LE" in hdr.stdout
    for name, a in (("vlist_opt", v), ("f_iter", f)):
        raw = tmp_path / (name + ".bin")
        r = subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", str(raw), path], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 0, r.stderr
        assert np.array_equal(np.fromfile(raw, np.float64).reshape(a.shape), a)
    att = subprocess.run([H5DUMP, "-a", "nqubits", path], capture_output=True, text=True, timeout=60)
    assert att.returncode == 0 and "(0): 8" in att.stdout


def test_errors_are_reported_not_crashes(mg, tmp_path):
    with pytest.raises(mg.H5Error):
        mg.h5_read(tmp_path / "missing.hdf5", ["x"])
    path = tmp_path / "e.hdf5"
    mg.h5_write(path, {"x": np.arange(4.0), "big": np.array([2**40], np.int64)}, {"a": 1.5})
    with pytest.raises(mg.H5Error):
        mg.h5_read(path, ["nope"])
    with pytest.raises(mg.H5Error):
        mg.h5_read(path, ["x"], dtype=np.int32)  # no float -> integer conversion
    with pytest.raises(mg.H5Error):
        mg.h5_read(path, ["big"], dtype=np.int32)  # out of range
    with pytest.raises(mg.H5Error):
        mg.h5_read(path, attributes=["nope"])
    with pytest.raises(mg.H5Error):
        mg.h5_write(tmp_path / "dup.hdf5", {"a/b": np.zeros(2)})  # groups are not written
    # every truncation of a valid file fails cleanly
    data = path.read_bytes()
    cut = tmp_path / "cut.hdf5"
    for n in range(0, len(data), 7):
        cut.write_bytes(data[:n])
        try:
            got, att = mg.h5_read(cut, ["x"], ["a"])
        except mg.H5Error:
            continue
        # a truncation after everything the reads need may still succeed, with the same values
        assert np.array_equal(got["x"], np.arange(4.0)) and att["a"] == 1.5


def test_driver_input_file_gives_the_optimiser_inputs(mg):
    """The reference driver's input file read as the driver reads it (hubbard1d_opt.c:104-151:
    vlist as doubles into struct matchgate, perm<i> as int) gives the inputs of the fixtures."""
    path = os.path.join(H5DIR, "hubbard1d_suzuki2_n5_q8_u29_t0.25s_g1.50_init.hdf5")
    d, at = mg.h5_read(path, ["vlist"] + [f"perm{i}" for i in range(5)], ["nlayers", "nqubits"], attr_dtype=np.int32)
    assert at["nqubits"] == 8 and at["nlayers"] == 5
    v = d["vlist"].reshape(5, 8, 2)
    V = v[..., 0] + 1j * v[..., 1]
    # each gate's center and corner blocks are unitary
    for g in V:
        for blk in (g[:4].reshape(2, 2), g[4:].reshape(2, 2)):
            assert np.allclose(blk.conj().T @ blk, np.eye(2), atol=1e-12)
    for i in range(5):
        assert sorted(d[f"perm{i}"].astype(int).tolist()) == list(range(8))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_headers_compile_the_reference_driver_io_calls(tmp_path):
    """The HDF5 names the reference drivers use compile against include/ alone."""
    src = tmp_path / "io.c"
    src.write_text(r'''
#include "util.h"
int main(void) {
    hid_t f = H5Fopen("x.hdf5", H5F_ACC_RDONLY, H5P_DEFAULT);
    int p[8]; double d[4]; hsize_t dims[1] = {4}; int nd;
    if (f >= 0) { read_hdf5_dataset(f, "perm0", H5T_NATIVE_INT, p); get_hdf5_dataset_ndims(f, "x", &nd);
                  get_hdf5_dataset_dims(f, "x", dims); read_hdf5_attribute(f, "t", H5T_NATIVE_DOUBLE, d); H5Fclose(f); }
    hid_t o = H5Fcreate("y.hdf5", H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    write_hdf5_dataset(o, "f_iter", 1, dims, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, d);
    write_hdf5_scalar_attribute(o, "nqubits", H5T_STD_I32LE, H5T_NATIVE_INT, p);
    return H5Fclose(o);
}
''')
    inc = os.path.join(os.path.dirname(GOLDEN), "..", "include")
    r = subprocess.run(["gcc", "-fsyntax-only", "-Wall", "-Werror", "-I", inc, str(src)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
