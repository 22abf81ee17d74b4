"""Minimal HDF5 reader over the image's HDF5 C library (ctypes; h5py is not
installed): dataset names of the root group and FP64 datasets as numpy
arrays in the file's C-order shape."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

H5_LIB = os.environ.get("SP_HDF5_LIB", "/opt/conda/lib/libhdf5.so")
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_ACC_RDONLY = 0
_lib = None


def lib():
    global _lib
    if _lib is None:
        h = C.CDLL(H5_LIB)
        h.H5open.restype = C.c_int
        for fn, res, args in (
            ("H5Fopen", C.c_int64, [C.c_char_p, C.c_uint, C.c_int64]),
            ("H5Fclose", C.c_int, [C.c_int64]),
            ("H5Dopen2", C.c_int64, [C.c_int64, C.c_char_p, C.c_int64]),
            ("H5Dclose", C.c_int, [C.c_int64]),
            ("H5Dget_space", C.c_int64, [C.c_int64]),
            ("H5Dread", C.c_int, [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                  C.c_void_p]),
            ("H5Sget_simple_extent_ndims", C.c_int, [C.c_int64]),
            ("H5Sget_simple_extent_dims", C.c_int, [C.c_int64, C.POINTER(C.c_uint64), C.c_void_p]),
            ("H5Sclose", C.c_int, [C.c_int64]),
            ("H5Lget_name_by_idx", C.c_ssize_t, [C.c_int64, C.c_char_p, C.c_int, C.c_int,
                                                 C.c_uint64, C.c_char_p, C.c_size_t, C.c_int64]),
        ):
            f = getattr(h, fn)
            f.restype, f.argtypes = res, args
        if h.H5open() < 0:
            raise RuntimeError("H5open failed")
        # errors are reported by return codes; no stack printing (the name
        # listing below probes one index past the last link)
        h.H5Eset_auto2.argtypes = [C.c_int64, C.c_void_p, C.c_void_p]
        h.H5Eset_auto2(0, None, None)
        _lib = h
    return _lib


def read(path: str) -> dict:
    """{dataset name: array} of every dataset in the root group of path."""
    h = lib()
    native_double = C.c_int64.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
    fid = h.H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT)
    if fid < 0:
        raise OSError(f"cannot open {path}")
    out = {}
    try:
        names, i = [], 0
        while True:
            buf = C.create_string_buffer(256)
            # H5_INDEX_NAME = 0, H5_ITER_INC = 0
            n = h.H5Lget_name_by_idx(fid, b".", 0, 0, i, buf, 256, H5P_DEFAULT)
            if n < 0:
                break
            names.append(buf.value.decode())
            i += 1
        for name in names:
            d = h.H5Dopen2(fid, name.encode(), H5P_DEFAULT)
            sp = h.H5Dget_space(d)
            rank = h.H5Sget_simple_extent_ndims(sp)
            dims = (C.c_uint64 * max(rank, 1))()
            h.H5Sget_simple_extent_dims(sp, dims, None)
            arr = np.empty(tuple(dims[:rank]), dtype=np.float64)
            if h.H5Dread(d, native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                         arr.ctypes.data_as(C.c_void_p)) < 0:
                raise OSError(f"cannot read {name} from {path}")
            h.H5Sclose(sp)
            h.H5Dclose(d)
            out[name] = arr
    finally:
        h.H5Fclose(fid)
    return out
