"""Minimal HDF5 access for the tests (h5py is not installed): ctypes over the
same libhdf5 C library the product loads, 1-d datasets only.  Used to write
initial-state files in the reference's layouts and to read back the files
the library writes."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_CANDIDATES = ("libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so",
               "/opt/conda/lib/libhdf5.so.103")

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    names = ([os.environ["SHM_HDF5_LIB"]] if os.environ.get("SHM_HDF5_LIB") else []) \
        + list(_CANDIDATES)
    for n in names:
        try:
            h = C.CDLL(n)
            break
        except OSError:
            continue
    else:
        return None
    hid = C.c_int64
    h.H5open.restype = C.c_int
    h.H5open()
    h.H5Fcreate.restype = hid
    h.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, hid, hid]
    h.H5Fopen.restype = hid
    h.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid]
    h.H5Fclose.argtypes = [hid]
    h.H5Lexists.restype = C.c_int
    h.H5Lexists.argtypes = [hid, C.c_char_p, hid]
    h.H5Screate_simple.restype = hid
    h.H5Screate_simple.argtypes = [C.c_int, C.POINTER(C.c_ulonglong), C.c_void_p]
    h.H5Sclose.argtypes = [hid]
    h.H5Dcreate2.restype = hid
    h.H5Dcreate2.argtypes = [hid, C.c_char_p, hid, hid, hid, hid, hid]
    h.H5Dopen2.restype = hid
    h.H5Dopen2.argtypes = [hid, C.c_char_p, hid]
    h.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    h.H5Dread.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    h.H5Dclose.argtypes = [hid]
    h.H5Dget_space.restype = hid
    h.H5Dget_space.argtypes = [hid]
    h.H5Sget_simple_extent_npoints.restype = C.c_longlong
    h.H5Sget_simple_extent_npoints.argtypes = [hid]
    h.H5Eset_auto2.argtypes = [hid, C.c_void_p, C.c_void_p]
    h.H5Eset_auto2(0, None, None)
    h.types = {
        np.dtype(np.float64): (hid.in_dll(h, "H5T_NATIVE_DOUBLE_g").value,
                               hid.in_dll(h, "H5T_IEEE_F64LE_g").value),
        np.dtype(np.float32): (hid.in_dll(h, "H5T_NATIVE_FLOAT_g").value,
                               hid.in_dll(h, "H5T_IEEE_F32LE_g").value),
        np.dtype(np.int32): (hid.in_dll(h, "H5T_NATIVE_INT_g").value,
                             hid.in_dll(h, "H5T_STD_I32LE_g").value),
        np.dtype(np.int64): (hid.in_dll(h, "H5T_NATIVE_LLONG_g").value,
                             hid.in_dll(h, "H5T_STD_I64LE_g").value),
    }
    _lib = h
    return h


def available() -> bool:
    return lib() is not None


def write(path: str, datasets: dict) -> None:
    """Creates (truncates) path with one 1-d dataset per entry; the file
    type follows the array's dtype (float64, float32, int32, int64)."""
    h = lib()
    f = h.H5Fcreate(path.encode(), 2, 0, 0)
    assert f >= 0, path
    try:
        for name, arr in datasets.items():
            a = np.ascontiguousarray(np.asarray(arr).ravel())
            mt, ft = h.types[a.dtype]
            dims = (C.c_ulonglong * 1)(a.size)
            s = h.H5Screate_simple(1, dims, None)
            d = h.H5Dcreate2(f, name.encode(), ft, s, 0, 0, 0)
            h.H5Sclose(s)
            assert d >= 0, name
            assert h.H5Dwrite(d, mt, 0, 0, 0, a.ctypes.data) >= 0
            h.H5Dclose(d)
    finally:
        h.H5Fclose(f)


def has(path: str, name: str) -> bool:
    h = lib()
    f = h.H5Fopen(path.encode(), 0, 0)
    assert f >= 0, path
    try:
        return h.H5Lexists(f, name.encode(), 0) > 0
    finally:
        h.H5Fclose(f)


def read(path: str, name: str, dtype=np.float64) -> np.ndarray:
    """The whole dataset, converted to dtype by the library."""
    h = lib()
    f = h.H5Fopen(path.encode(), 0, 0)
    assert f >= 0, path
    try:
        d = h.H5Dopen2(f, name.encode(), 0)
        assert d >= 0, name
        s = h.H5Dget_space(d)
        n = h.H5Sget_simple_extent_npoints(s)
        h.H5Sclose(s)
        out = np.zeros(n, dtype=dtype)
        assert h.H5Dread(d, h.types[np.dtype(dtype)][0], 0, 0, 0, out.ctypes.data) >= 0
        h.H5Dclose(d)
        return out
    finally:
        h.H5Fclose(f)
