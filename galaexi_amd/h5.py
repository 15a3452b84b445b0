"""Minimal HDF5 access through the HDF5 C library (ctypes, HDF5 1.10 ABI: hid_t = int64).

The reference reads and writes its mesh and state files with the HDF5 Fortran API (src/io_hdf5/hdf5_input.f90,
src/io_hdf5/hdf5_output.f90).  This module covers what the mesh ingest and the state I/O need: integer, double
and fixed-length string datasets and attributes, read whole, and written whole.

Layout: HDF5 stores dimensions in C order.  A Fortran array A(5,0:N,0:N,0:N,nElems) written by the reference
appears with dims (nElems,N+1,N+1,N+1,5); `read` returns it with those dims, and `.T` of that array is the
Fortran-ordered view with the reference's index order.

The library is looked up as $GALX_HDF5_LIB, then by the dynamic linker ("libhdf5.so"), then in the usual
install prefixes; `available()` says whether one was found.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os

import numpy as np

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
hsize_t = ctypes.c_ulonglong

H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5E_DEFAULT = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5T_SGN_NONE = 0
H5T_STR_NULLTERM, H5T_STR_SPACEPAD = 0, 2

_CANDIDATES = ("/opt/conda/lib/libhdf5.so", "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
               "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so", "/usr/lib64/libhdf5.so")


class H5Error(RuntimeError):
    pass


class _Lib:
    def __init__(self, path):
        L = ctypes.CDLL(path)
        self.L = L
        self.path = path
        for name, res, args in (
                ("H5open", herr_t, []),
                ("H5Eset_auto2", herr_t, [hid_t, ctypes.c_void_p, ctypes.c_void_p]),
                ("H5Fopen", hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
                ("H5Fcreate", hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]),
                ("H5Fclose", herr_t, [hid_t]),
                ("H5Dopen2", hid_t, [hid_t, ctypes.c_char_p, hid_t]),
                ("H5Dcreate2", hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
                ("H5Dget_space", hid_t, [hid_t]),
                ("H5Dget_type", hid_t, [hid_t]),
                ("H5Dread", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
                ("H5Dwrite", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
                ("H5Dclose", herr_t, [hid_t]),
                ("H5Aexists", ctypes.c_int, [hid_t, ctypes.c_char_p]),
                ("H5Aopen", hid_t, [hid_t, ctypes.c_char_p, hid_t]),
                ("H5Acreate2", hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t]),
                ("H5Aget_space", hid_t, [hid_t]),
                ("H5Aget_type", hid_t, [hid_t]),
                ("H5Aread", herr_t, [hid_t, hid_t, ctypes.c_void_p]),
                ("H5Awrite", herr_t, [hid_t, hid_t, ctypes.c_void_p]),
                ("H5Aclose", herr_t, [hid_t]),
                ("H5Lexists", ctypes.c_int, [hid_t, ctypes.c_char_p, hid_t]),
                ("H5Screate_simple", hid_t, [ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
                ("H5Sget_simple_extent_ndims", ctypes.c_int, [hid_t]),
                ("H5Sget_simple_extent_dims", ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t),
                                                             ctypes.POINTER(hsize_t)]),
                ("H5Sclose", herr_t, [hid_t]),
                ("H5Tcopy", hid_t, [hid_t]),
                ("H5Tget_class", ctypes.c_int, [hid_t]),
                ("H5Tget_size", ctypes.c_size_t, [hid_t]),
                ("H5Tget_sign", ctypes.c_int, [hid_t]),
                ("H5Tis_variable_str", ctypes.c_int, [hid_t]),
                ("H5Tset_size", herr_t, [hid_t, ctypes.c_size_t]),
                ("H5Tset_strpad", herr_t, [hid_t, ctypes.c_int]),
                ("H5Tclose", herr_t, [hid_t])):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
            setattr(self, name, f)
        if self.H5open() < 0:
            raise H5Error(f"H5open failed ({path})")
        self.H5Eset_auto2(H5E_DEFAULT, None, None)  # errors come back as return codes, not as printed stacks

        def g(sym):
            return hid_t.in_dll(L, sym).value

        self.NATIVE_INT = g("H5T_NATIVE_INT_g")
        self.NATIVE_LLONG = g("H5T_NATIVE_LLONG_g")
        self.NATIVE_DOUBLE = g("H5T_NATIVE_DOUBLE_g")
        self.STD_I32LE = g("H5T_STD_I32LE_g")
        self.IEEE_F64LE = g("H5T_IEEE_F64LE_g")
        self.C_S1 = g("H5T_C_S1_g")


_lib = None


def _find():
    env = os.environ.get("GALX_HDF5_LIB")
    if env:
        return [env]
    out = []
    found = ctypes.util.find_library("hdf5")
    if found:
        out.append(found)
    out.extend(p for p in _CANDIDATES if os.path.exists(p))
    return out


def lib():
    """The loaded HDF5 C library (raises H5Error if none is found)."""
    global _lib
    if _lib is None:
        errs = []
        for p in _find():
            try:
                _lib = _Lib(p)
                break
            except (OSError, AttributeError, H5Error) as ex:
                errs.append(f"{p}: {ex}")
        if _lib is None:
            raise H5Error("no usable HDF5 C library (set GALX_HDF5_LIB): " + ("; ".join(errs) or "none found"))
    return _lib


def available():
    try:
        lib()
        return True
    except H5Error:
        return False


def _check(rc, what):
    if rc < 0:
        raise H5Error(f"HDF5: {what} failed")
    return rc


def _dims(L, space):
    nd = _check(L.H5Sget_simple_extent_ndims(space), "get_simple_extent_ndims")
    d = (hsize_t * max(nd, 1))()
    _check(L.H5Sget_simple_extent_dims(space, d, None), "get_simple_extent_dims")
    return tuple(int(d[i]) for i in range(nd))


def _read_any(L, obj, is_attr):
    space = _check((L.H5Aget_space if is_attr else L.H5Dget_space)(obj), "get_space")
    ftype = _check((L.H5Aget_type if is_attr else L.H5Dget_type)(obj), "get_type")
    try:
        dims = _dims(L, space)
        count = int(np.prod(dims)) if dims else 1
        cls = L.H5Tget_class(ftype)
        if cls == H5T_STRING:
            if L.H5Tis_variable_str(ftype) > 0:
                raise H5Error("variable-length strings are not supported")
            # memory strings are null-terminated: one byte more than the file's fixed length
            size = int(L.H5Tget_size(ftype)) + 1
            mtype = _check(L.H5Tcopy(L.C_S1), "Tcopy")
            try:
                _check(L.H5Tset_size(mtype, size), "Tset_size")
                buf = ctypes.create_string_buffer(size * count)
                rd = L.H5Aread(obj, mtype, buf) if is_attr else L.H5Dread(obj, mtype, H5S_ALL, H5S_ALL,
                                                                           H5P_DEFAULT, buf)
                _check(rd, "read")
            finally:
                L.H5Tclose(mtype)
            raw = buf.raw
            vals = [raw[i * size:(i + 1) * size].split(b"\0", 1)[0].decode("ascii", "replace").rstrip()
                    for i in range(count)]
            return np.array(vals, dtype=object).reshape(dims) if dims else vals[0]
        if cls == H5T_INTEGER:
            big = L.H5Tget_size(ftype) > 4 or L.H5Tget_sign(ftype) == H5T_SGN_NONE
            dt, mtype = (np.int64, L.NATIVE_LLONG) if big else (np.int32, L.NATIVE_INT)
        elif cls == H5T_FLOAT:
            dt, mtype = np.float64, L.NATIVE_DOUBLE
        else:
            raise H5Error(f"unsupported HDF5 type class {cls}")
        arr = np.empty(dims if dims else (1,), dtype=dt)
        ptr = arr.ctypes.data_as(ctypes.c_void_p)
        rd = L.H5Aread(obj, mtype, ptr) if is_attr else L.H5Dread(obj, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, ptr)
        _check(rd, "read")
        return arr
    finally:
        L.H5Tclose(ftype)
        L.H5Sclose(space)


class File:
    """An HDF5 file opened read-only (mode 'r') or created/truncated (mode 'w'); use as a context manager."""

    def __init__(self, path, mode="r"):
        L = lib()
        self._L = L
        self.path = str(path)
        b = self.path.encode()
        if mode == "r":
            self.id = L.H5Fopen(b, H5F_ACC_RDONLY, H5P_DEFAULT)
        elif mode == "w":
            self.id = L.H5Fcreate(b, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
        else:
            raise ValueError(mode)
        if self.id < 0:
            raise H5Error(f"cannot open {self.path!r} (mode {mode})")

    def close(self):
        if self.id >= 0:
            self._L.H5Fclose(self.id)
            self.id = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- reading ----
    def has(self, name):
        return self._L.H5Lexists(self.id, name.encode(), H5P_DEFAULT) > 0

    def has_attr(self, name):
        return self._L.H5Aexists(self.id, name.encode()) > 0

    def read(self, name):
        """Dataset `name` as a numpy array with the file's (C-order) dims; strings as an object array."""
        L = self._L
        d = L.H5Dopen2(self.id, name.encode(), H5P_DEFAULT)
        if d < 0:
            raise H5Error(f"{self.path}: no dataset {name!r}")
        try:
            return _read_any(L, d, False)
        finally:
            L.H5Dclose(d)

    def attr(self, name):
        """Root attribute `name`; a one-element numeric attribute is returned as a scalar."""
        L = self._L
        a = L.H5Aopen(self.id, name.encode(), H5P_DEFAULT)
        if a < 0:
            raise H5Error(f"{self.path}: no attribute {name!r}")
        try:
            v = _read_any(L, a, True)
        finally:
            L.H5Aclose(a)
        if isinstance(v, np.ndarray) and v.size == 1:
            x = v.reshape(-1)[0]
            return x if v.dtype == object else x.item()
        return v

    # ---- writing ----
    def _space(self, shape):
        shape = tuple(int(s) for s in shape)
        d = (hsize_t * max(len(shape), 1))(*shape)
        return _check(self._L.H5Screate_simple(len(shape), d, None), "Screate_simple")

    @staticmethod
    def _numeric(value, what):
        a = np.asarray(value)
        if a.dtype.kind == "f":
            return np.ascontiguousarray(a, dtype=np.float64), "f"
        if a.dtype.kind in "iub":
            if a.size and (a.max() > np.iinfo(np.int32).max or a.min() < np.iinfo(np.int32).min):
                raise H5Error(f"{what}: integer values exceed 32 bits")
            return np.ascontiguousarray(a, dtype=np.int32), "i"
        raise H5Error(f"{what}: unsupported dtype {a.dtype}")

    def write(self, name, value):
        """Dataset `name` from a numeric array, stored with the array's C-order dims (pass A.T of a
        Fortran-ordered array to get the reference's on-disk dims)."""
        L = self._L
        a, k = self._numeric(value, name)
        ftype, mtype = (L.IEEE_F64LE, L.NATIVE_DOUBLE) if k == "f" else (L.STD_I32LE, L.NATIVE_INT)
        sp = self._space(a.shape)
        try:
            d = _check(L.H5Dcreate2(self.id, name.encode(), ftype, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                       f"create dataset {name}")
            try:
                _check(L.H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data_as(ctypes.c_void_p)),
                       f"write dataset {name}")
            finally:
                L.H5Dclose(d)
        finally:
            L.H5Sclose(sp)

    def write_attr(self, name, value, strlen=None):
        """Root attribute `name`: int/float scalars or 1-D arrays, or strings (a str or a list of str), the latter
        as space-padded fixed-length strings of length `strlen` (default: the longest value) as the reference
        writes them."""
        L = self._L
        if isinstance(value, str) or (isinstance(value, (list, tuple)) and value and isinstance(value[0], str)):
            vals = [value] if isinstance(value, str) else list(value)
            size = strlen or max(1, max(len(v) for v in vals))
            raw = b"".join(v.encode("ascii").ljust(size, b" ")[:size] for v in vals)
            ftype = _check(L.H5Tcopy(L.C_S1), "Tcopy")
            try:
                _check(L.H5Tset_size(ftype, size), "Tset_size")
                _check(L.H5Tset_strpad(ftype, H5T_STR_SPACEPAD), "Tset_strpad")
                sp = self._space((len(vals),))
                try:
                    a = _check(L.H5Acreate2(self.id, name.encode(), ftype, sp, H5P_DEFAULT, H5P_DEFAULT),
                               f"create attribute {name}")
                    try:
                        buf = ctypes.create_string_buffer(raw, len(raw))
                        _check(L.H5Awrite(a, ftype, buf), f"write attribute {name}")
                    finally:
                        L.H5Aclose(a)
                finally:
                    L.H5Sclose(sp)
            finally:
                L.H5Tclose(ftype)
            return
        arr, k = self._numeric(np.atleast_1d(value), name)
        ftype, mtype = (L.IEEE_F64LE, L.NATIVE_DOUBLE) if k == "f" else (L.STD_I32LE, L.NATIVE_INT)
        sp = self._space(arr.shape)
        try:
            a = _check(L.H5Acreate2(self.id, name.encode(), ftype, sp, H5P_DEFAULT, H5P_DEFAULT),
                       f"create attribute {name}")
            try:
                _check(L.H5Awrite(a, mtype, arr.ctypes.data_as(ctypes.c_void_p)), f"write attribute {name}")
            finally:
                L.H5Aclose(a)
        finally:
            L.H5Sclose(sp)
