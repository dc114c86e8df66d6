"""Minimal HDF5 file access for the tests, through the HDF5 C library over ctypes (the image
has the library but no h5py).  It is independent of the product's C++ writer and reader:
the tests write input fixtures with it and read the product's output files back with it.

Arrays follow the HDF5 C convention: the dataset's dimensions are the numpy array's shape,
so a Fortran array of shape (lx1, lx2, lx3, 7) is the numpy array of shape (7, lx3, lx2, lx1)
holding the same bytes (``a.T`` of a Fortran-ordered array).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

hid_t = ctypes.c_int64
hsize_t = ctypes.c_ulonglong

_L = None


def lib():
    global _L
    if _L is None:
        last = None
        for p in (os.environ.get("GEMHIP_HDF5_LIB"), "libhdf5.so", "/opt/conda/lib/libhdf5.so"):
            if not p:
                continue
            try:
                _L = ctypes.CDLL(p)
                break
            except OSError as e:
                last = e
        if _L is None:
            raise OSError(f"libhdf5 not loadable: {last}")
        _L.H5open()
        for name in ("H5Fcreate", "H5Fopen", "H5Dcreate2", "H5Dopen2", "H5Screate_simple",
                     "H5Screate", "H5Dget_space", "H5Dget_type", "H5Gcreate2", "H5Pcreate"):
            getattr(_L, name).restype = hid_t
        _L.H5Eset_auto2(hid_t(0), None, None)
    return _L


def _g(name):
    return hid_t.in_dll(lib(), name).value


def _native(dtype):
    dt = np.dtype(dtype)
    return _g({"f4": "H5T_NATIVE_FLOAT_g", "f8": "H5T_NATIVE_DOUBLE_g", "i4": "H5T_NATIVE_INT_g",
               "i8": "H5T_NATIVE_LLONG_g"}[dt.kind + str(dt.itemsize)])


def write(path: str, datasets: dict, file_types: dict | None = None):
    """Create path with one dataset per item (name -> array or numpy scalar); intermediate
    groups are created.  file_types optionally maps a name to the numpy dtype stored in the
    file (default: the array's)."""
    L = lib()
    f = L.H5Fcreate(path.encode(), ctypes.c_uint(2), hid_t(0), hid_t(0))
    assert f >= 0, path
    groups = set()
    try:
        for name, value in datasets.items():
            parts = name.strip("/").split("/")
            for k in range(1, len(parts)):
                gname = "/" + "/".join(parts[:k])
                if gname not in groups:
                    gid = L.H5Gcreate2(hid_t(f), gname.encode(), hid_t(0), hid_t(0), hid_t(0))
                    assert gid >= 0, gname
                    L.H5Gclose(hid_t(gid))
                    groups.add(gname)
            a = np.array(value)          # a 0-d array stays a scalar dataset
            if a.ndim == 0:
                space = L.H5Screate(ctypes.c_int(0))
            else:
                dims = (hsize_t * a.ndim)(*a.shape)
                space = L.H5Screate_simple(ctypes.c_int(a.ndim), dims, None)
            ftype = _native((file_types or {}).get(name, a.dtype))
            d = L.H5Dcreate2(hid_t(f), ("/" + name.strip("/")).encode(), hid_t(ftype), hid_t(space),
                             hid_t(0), hid_t(0), hid_t(0))
            assert d >= 0, name
            assert L.H5Dwrite(hid_t(d), hid_t(_native(a.dtype)), hid_t(0), hid_t(0), hid_t(0),
                              a.ctypes.data_as(ctypes.c_void_p)) >= 0, name
            L.H5Dclose(hid_t(d))
            L.H5Sclose(hid_t(space))
    finally:
        L.H5Fclose(hid_t(f))


class File:
    """Read-only file: names(), info(name) -> (kind, itemsize, little_endian, shape), read(name)."""

    def __init__(self, path: str):
        self.f = lib().H5Fopen(path.encode(), ctypes.c_uint(0), hid_t(0))
        if self.f < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        if self.f >= 0:
            lib().H5Fclose(hid_t(self.f))
            self.f = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def names(self):
        """Every dataset path in the file (groups are walked, not listed)."""
        L = lib()
        found = []
        CB = ctypes.CFUNCTYPE(ctypes.c_int, hid_t, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p)

        def visit(_g, name, _info, _data):
            found.append("/" + name.decode())
            return 0

        cb = CB(visit)
        assert L.H5Lvisit(hid_t(self.f), ctypes.c_int(0), ctypes.c_int(0), cb, None) >= 0
        out = []
        for n in found:
            d = L.H5Dopen2(hid_t(self.f), n.encode(), hid_t(0))
            if d >= 0:
                out.append(n)
                L.H5Dclose(hid_t(d))
        return sorted(out)

    def _open(self, name):
        d = lib().H5Dopen2(hid_t(self.f), name.encode(), hid_t(0))
        if d < 0:
            raise KeyError(name)
        return d

    def info(self, name):
        L = lib()
        d = self._open(name)
        try:
            t = L.H5Dget_type(hid_t(d))
            cls = L.H5Tget_class(hid_t(t))
            size = L.H5Tget_size(hid_t(t))
            order = L.H5Tget_order(hid_t(t))
            L.H5Tclose(hid_t(t))
            s = L.H5Dget_space(hid_t(d))
            nd = L.H5Sget_simple_extent_ndims(hid_t(s))
            dims = (hsize_t * max(nd, 1))()
            L.H5Sget_simple_extent_dims(hid_t(s), dims, None)
            L.H5Sclose(hid_t(s))
            kind = {0: "i", 1: "f"}.get(cls, "?")
            return kind, int(size), order == 0, tuple(int(dims[k]) for k in range(nd))
        finally:
            L.H5Dclose(hid_t(d))

    def read(self, name):
        kind, size, _, shape = self.info(name)
        a = np.empty(shape, dtype=np.dtype(f"{kind}{size}"))
        d = self._open(name)
        try:
            assert lib().H5Dread(hid_t(d), hid_t(_native(a.dtype)), hid_t(0), hid_t(0), hid_t(0),
                                 a.ctypes.data_as(ctypes.c_void_p)) >= 0, name
        finally:
            lib().H5Dclose(hid_t(d))
        return a
