"""The HDF5 writer of the distribution files (fortran/sp_h5.c): datasets
declared with Fortran-order dimensions, as the reference's HDF5 Fortran calls
declare them (diagnostics.f90:1244-1650), read back through the HDF5 C
library with the reversed (C-order) shape over the same values."""
import ctypes as C
import os

import numpy as np
import pytest

from h5read import read

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "stochastic-parker_amd", "lib", "libsp_h5.so")


def h5lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (make -C stochastic-parker_amd)")
    h = C.CDLL(LIB)
    h.sp_h5_open.restype = C.c_int64
    h.sp_h5_open.argtypes = [C.c_char_p, C.c_int32]
    h.sp_h5_write.restype = C.c_int32
    h.sp_h5_write.argtypes = [C.c_int64, C.c_char_p, C.c_int32, C.POINTER(C.c_int64),
                              C.POINTER(C.c_double)]
    h.sp_h5_close.restype = C.c_int32
    h.sp_h5_close.argtypes = [C.c_int64]
    return h


def write(h, fid, name, fdims, data):
    dims = (C.c_int64 * len(fdims))(*fdims)
    a = np.ascontiguousarray(data, dtype=np.float64)
    return h.sp_h5_write(fid, name.encode(), len(fdims), dims, a.ctypes.data_as(C.POINTER(C.c_double)))


def test_fortran_order_datasets_read_back_reversed(tmp_path):
    h = h5lib()
    path = str(tmp_path / "fdists_0001.h5")
    fid = h.sp_h5_open(path.encode(), 1)
    assert fid > 0
    # fglobal (nmu, npp) = (2, 5): Fortran order, nmu fastest
    fg = np.arange(10, dtype=np.float64) * 0.5
    assert write(h, fid, "fglobal", [2, 5], fg) == 0
    # flocal1 (nmu, npbins, nrx, nry, nrz) = (1, 3, 4, 2, 1)
    fl = np.arange(24, dtype=np.float64) ** 2
    assert write(h, fid, "flocal1", [1, 3, 4, 2, 1], fl) == 0
    edges = 10.0 ** np.linspace(-2, 4, 6)
    assert write(h, fid, "pbins_edges_global", [6], edges) == 0
    assert h.sp_h5_close(fid) == 0
    # reopen read-write and add a dataset, as the local sets are appended
    fid = h.sp_h5_open(path.encode(), 0)
    assert fid > 0
    assert write(h, fid, "fescaped1_x", [1, 3, 2, 1, 2], np.arange(12.0)) == 0
    assert h.sp_h5_close(fid) == 0
    got = read(path)
    assert sorted(got) == ["fescaped1_x", "fglobal", "flocal1", "pbins_edges_global"]
    assert got["fglobal"].shape == (5, 2)
    np.testing.assert_array_equal(got["fglobal"], fg.reshape(5, 2))
    # element (imu, ip) of the Fortran array is [ip, imu] in C order
    f = fg.reshape(5, 2).T
    assert got["fglobal"][3, 1] == f[1, 3]
    assert got["flocal1"].shape == (1, 2, 4, 3, 1)
    np.testing.assert_array_equal(got["flocal1"].ravel(), fl)
    np.testing.assert_array_equal(got["pbins_edges_global"], edges)
    assert got["fescaped1_x"].shape == (2, 1, 2, 3, 1)


def test_errors_are_reported(tmp_path):
    h = h5lib()
    assert h.sp_h5_open(str(tmp_path / "missing" / "x.h5").encode(), 1) < 0
    fid = h.sp_h5_open(str(tmp_path / "y.h5").encode(), 1)
    assert write(h, fid, "a", [3], np.zeros(3)) == 0
    # the same name twice, and an invalid rank
    assert write(h, fid, "a", [3], np.zeros(3)) == -1
    assert write(h, fid, "b", [], np.zeros(1)) == -1
    assert h.sp_h5_close(fid) == 0
