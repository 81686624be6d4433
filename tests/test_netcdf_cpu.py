"""The host API's snapshot files (ScaFES::DataFile, classic NetCDF written
without libnetcdf) checked with an independent implementation of the
format, scipy.io.netcdf_file, in both directions:

 * files the engine writes: the reference's layout (dimensions nNodes_0 ..
   nNodes_{d-1} and unlimited "time", an int "version" = 3, double
   variables (time, nNodes_{d-1}, .., nNodes_0)), one record per write,
   fill values for variables a record does not write;
 * files scipy writes (CDF-1 and CDF-2, with attributes, float / short / int
   variables and extra dimensions) read back through the engine's reader.
"""
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

import oracle_lib as O

TOOL = os.path.join(O.ROOT, "masterthesis_amd", "bin", "netcdf_tool")
FILL = 9.969209968386869e+36


@pytest.fixture(scope="module", autouse=True)
def tool():
    if not os.path.exists(TOOL):
        subprocess.check_call(["make", "-s", "-C", os.path.join(O.ROOT, "masterthesis_amd", "host")])


def _run(*args):
    return subprocess.run([TOOL] + [str(a) for a in args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("n", [(7,), (5, 4), (4, 3, 2)])
def test_written_file_layout_and_records(tmp_path, n):
    base = str(tmp_path / "snap")
    p = _run("write", base, 3, *n)
    assert p.returncode == 0, p.stderr
    f = netcdf_file(base + ".nc", "r", mmap=False)
    dims = list(f.dimensions.items())
    assert dims == [(f"nNodes_{q}", n[q]) for q in range(len(n))] + [("time", None)]
    assert f.variables["version"].getValue() == 3
    shape = (3,) + tuple(reversed(n))
    cnt = int(np.prod(n))
    idx = np.arange(cnt, dtype=np.float64).reshape(tuple(reversed(n)))
    A, B = f.variables["A"], f.variables["B"]
    assert A.dimensions == ("time",) + tuple(f"nNodes_{q}" for q in reversed(range(len(n))))
    assert A.shape == shape and A.typecode() == "d"
    for r in range(3):
        np.testing.assert_array_equal(A[r], idx + 1000.0 * r)
        if r % 2 == 0:
            np.testing.assert_array_equal(B[r], -(idx + r))
        else:
            assert np.all(B[r] == FILL)
    f.close()


@pytest.mark.parametrize("version", [1, 2])
def test_reads_files_scipy_writes(tmp_path, version):
    n = (6, 5, 3)
    path = str(tmp_path / "init.nc")
    f = netcdf_file(path, "w", version=version)
    f.history = b"written by scipy"
    f.createDimension("time", None)  # scipy wants the record dimension first
    f.createDimension("extra", 2)
    for q in range(3):
        f.createDimension(f"nNodes_{q}", n[q])
    rng = np.random.default_rng(1)
    u = rng.standard_normal((2,) + tuple(reversed(n)))
    w = rng.standard_normal(tuple(reversed(n))).astype(np.float32)
    k = rng.integers(-300, 300, tuple(reversed(n))).astype(np.int16)
    v = f.createVariable("U", "d", ("time", "nNodes_2", "nNodes_1", "nNodes_0"))
    v.units = b"K"
    v[:] = u
    g = f.createVariable("W", "f", ("nNodes_2", "nNodes_1", "nNodes_0"))
    g[:] = w
    h = f.createVariable("K", "h", ("nNodes_2", "nNodes_1", "nNodes_0"))
    h.scale = np.array([2.0])
    h[:] = k
    e = f.createVariable("E", "i", ("extra",))
    e[:] = [7, 8]
    f.close()
    p = _run("read", str(tmp_path / "init"), *n, "--", "U", "W", "K")
    assert p.returncode == 0, p.stderr
    vals = np.array([float(x) for x in p.stdout.split()])
    cnt = int(np.prod(n))
    np.testing.assert_array_equal(vals[:cnt], u[0].ravel())  # record 0
    np.testing.assert_array_equal(vals[cnt:2 * cnt], w.astype(np.float64).ravel())
    np.testing.assert_array_equal(vals[2 * cnt:], k.astype(np.float64).ravel())


def test_read_errors(tmp_path):
    base = str(tmp_path / "snap")
    assert _run("write", base, 1, 4, 3).returncode == 0
    p = _run("read", base, 4, 3, "--", "missing")
    assert p.returncode != 0 and "not found" in p.stderr
    p = _run("read", base, 5, 3, "--", "A")
    assert p.returncode != 0 and "do not fit" in p.stderr
    p = _run("read", str(tmp_path / "absent"), 4, 3, "--", "A")
    assert p.returncode != 0 and "cannot open" in p.stderr
