"""MAT-file version 7.3 writer and reader: HDF5 with MATLAB's conventions, through the HDF5 C library
(ctypes).

inversion/start_inversion.m:215 saves the solution with ``save(..., '-v7.3')``. A MAT 7.3 file is an
HDF5 file whose 512-byte user block begins with MATLAB's 128-byte header (116 bytes of text, the
8-byte subsystem offset, version 0x0200 and the endian mark 'IM'); every variable is an object at
the root:

  * numeric arrays: a dataset of the class's HDF5 type with the dimensions reversed (HDF5 lists the
    slowest dimension first, MATLAB's column-major data has the first dimension fastest), attribute
    MATLAB_class ('double', 'single', 'int32', ...); complex arrays as the compound {real, imag};
  * logical arrays: uint8, MATLAB_class 'logical', MATLAB_int_decode 1;
  * char arrays: uint16 character codes, MATLAB_class 'char', MATLAB_int_decode 2;
  * empty arrays: a uint64 dataset holding the dimensions, attribute MATLAB_empty 1;
  * scalar structs: a group, MATLAB_class 'struct', MATLAB_fields (the field names as variable-length
    character arrays), one member per field;
  * cell arrays: a dataset of object references to members of the group '/#refs#', MATLAB_class
    'cell'; struct arrays: a struct group whose fields are such reference datasets.

Python values map as scipy.io does for MAT v5: float / int -> double scalar, bool -> logical, str ->
char row, None -> [] (0 x 0 double), dict -> struct, a list of dicts with the same keys -> 1 x N struct
array, a list of numbers -> double row, any other list -> 1 x N cell, numpy arrays by dtype (1-D as
1 x n rows). The library is found through the environment variable FD2D_HDF5_LIB, the system's
loader path, or a conda installation; ``available()`` says whether one was found.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import glob
import os
import time
from typing import Dict, Optional

import numpy as np

hid_t, herr_t, hsize_t = ctypes.c_int64, ctypes.c_int, ctypes.c_uint64

_CLASS = {np.dtype(np.float64): "double", np.dtype(np.float32): "single", np.dtype(np.int8): "int8",
          np.dtype(np.int16): "int16", np.dtype(np.int32): "int32", np.dtype(np.int64): "int64",
          np.dtype(np.uint8): "uint8", np.dtype(np.uint16): "uint16", np.dtype(np.uint32): "uint32",
          np.dtype(np.uint64): "uint64"}
_NATIVE = {"double": "H5T_NATIVE_DOUBLE", "single": "H5T_NATIVE_FLOAT", "int8": "H5T_NATIVE_SCHAR",
           "int16": "H5T_NATIVE_SHORT", "int32": "H5T_NATIVE_INT32", "int64": "H5T_NATIVE_INT64",
           "uint8": "H5T_NATIVE_UINT8", "uint16": "H5T_NATIVE_UINT16", "uint32": "H5T_NATIVE_UINT32",
           "uint64": "H5T_NATIVE_UINT64"}

H5P_DEFAULT, H5S_ALL, H5F_ACC_TRUNC, H5F_ACC_RDONLY = 0, 0, 2, 0
H5S_SCALAR, H5T_COMPOUND, H5R_OBJECT = 0, 6, 0


class _HVL(ctypes.Structure):
    _fields_ = [("len", ctypes.c_size_t), ("p", ctypes.c_void_p)]


def _find() -> Optional[str]:
    cands = [os.environ.get("FD2D_HDF5_LIB", "")]
    name = ctypes.util.find_library("hdf5")
    if name:
        cands.append(name)
    cands += sorted(glob.glob("/opt/conda/lib/libhdf5.so*")) + sorted(glob.glob("/usr/lib/x86_64-linux-gnu/libhdf5*.so*"))
    for c in cands:
        if not c or os.path.basename(c).startswith(("libhdf5_hl", "libhdf5hl", "libhdf5_fortran")):
            continue
        try:
            lib = ctypes.CDLL(c)
            lib.H5open()
            return c
        except OSError:
            continue
    return None


_lib = None


def _h5():
    global _lib
    if _lib is not None:
        return _lib
    path = _find()
    if path is None:
        raise OSError("MAT v7.3 needs the HDF5 C library (set FD2D_HDF5_LIB to libhdf5.so)")
    lib = ctypes.CDLL(path)
    if lib.H5open() < 0:
        raise OSError(f"H5open failed in {path}")
    sig = {
        "H5Pcreate": (hid_t, [hid_t]), "H5Pset_userblock": (herr_t, [hid_t, hsize_t]), "H5Pclose": (herr_t, [hid_t]),
        "H5Fcreate": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]), "H5Fclose": (herr_t, [hid_t]),
        "H5Fopen": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]), "H5Fis_hdf5": (ctypes.c_int, [ctypes.c_char_p]),
        "H5Screate_simple": (hid_t, [ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.c_void_p]),
        "H5Screate": (hid_t, [ctypes.c_int]), "H5Sclose": (herr_t, [hid_t]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t), ctypes.c_void_p]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid_t]),
        "H5Dcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]), "H5Dget_space": (hid_t, [hid_t]),
        "H5Dget_type": (hid_t, [hid_t]),
        "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]), "H5Dclose": (herr_t, [hid_t]),
        "H5Gcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t]),
        "H5Gopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]), "H5Gclose": (herr_t, [hid_t]),
        "H5Acreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t]),
        "H5Aopen": (hid_t, [hid_t, ctypes.c_char_p, hid_t]), "H5Aexists": (ctypes.c_int, [hid_t, ctypes.c_char_p]),
        "H5Awrite": (herr_t, [hid_t, hid_t, ctypes.c_void_p]), "H5Aread": (herr_t, [hid_t, hid_t, ctypes.c_void_p]),
        "H5Aget_type": (hid_t, [hid_t]), "H5Aget_space": (hid_t, [hid_t]), "H5Aclose": (herr_t, [hid_t]),
        "H5Tcopy": (hid_t, [hid_t]), "H5Tset_size": (herr_t, [hid_t, ctypes.c_size_t]),
        "H5Tget_size": (ctypes.c_size_t, [hid_t]), "H5Tget_class": (ctypes.c_int, [hid_t]),
        "H5Tcreate": (hid_t, [ctypes.c_int, ctypes.c_size_t]),
        "H5Tinsert": (herr_t, [hid_t, ctypes.c_char_p, ctypes.c_size_t, hid_t]),
        "H5Tvlen_create": (hid_t, [hid_t]), "H5Tclose": (herr_t, [hid_t]),
        "H5Rcreate": (herr_t, [ctypes.c_void_p, hid_t, ctypes.c_char_p, ctypes.c_int, hid_t]),
        "H5Rdereference2": (hid_t, [hid_t, hid_t, ctypes.c_int, ctypes.c_void_p]),
        "H5Iget_name": (ctypes.c_ssize_t, [hid_t, ctypes.c_char_p, ctypes.c_size_t]),
        "H5Oclose": (herr_t, [hid_t]), "H5Oopen": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Iget_type": (ctypes.c_int, [hid_t]),
        "H5Gget_info": (herr_t, [hid_t, ctypes.c_void_p]),
        "H5Lget_name_by_idx": (ctypes.c_ssize_t, [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t,
                                                  ctypes.c_char_p, ctypes.c_size_t, hid_t]),
        "H5Dvlen_reclaim": (herr_t, [hid_t, hid_t, hid_t, ctypes.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    lib.T = {n: hid_t.in_dll(lib, n + "_g").value for n in
             list(_NATIVE.values()) + ["H5T_C_S1", "H5T_STD_REF_OBJ", "H5P_CLS_FILE_CREATE_ID"]}
    _lib = lib
    return lib


def available() -> bool:
    try:
        _h5()
        return True
    except OSError:
        return False


def _ck(v, what):
    if v < 0:
        raise OSError(f"HDF5: {what} failed")
    return v


class _Writer:
    def __init__(self, lib, fid):
        self.lib, self.fid, self.nref = lib, fid, 0
        self.refs = None

    # ---- attributes --------------------------------------------------------------------------
    def _attr_str(self, obj, name, text):
        L = self.lib
        t = _ck(L.H5Tcopy(L.T["H5T_C_S1"]), "H5Tcopy")
        L.H5Tset_size(t, max(len(text), 1))
        sp = _ck(L.H5Screate(H5S_SCALAR), "H5Screate")
        a = _ck(L.H5Acreate2(obj, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT), f"attribute {name}")
        buf = ctypes.create_string_buffer(text.encode(), max(len(text), 1))
        _ck(L.H5Awrite(a, t, buf), "H5Awrite")
        L.H5Aclose(a), L.H5Sclose(sp), L.H5Tclose(t)

    def _attr_int(self, obj, name, value, tname):
        L = self.lib
        sp = _ck(L.H5Screate(H5S_SCALAR), "H5Screate")
        a = _ck(L.H5Acreate2(obj, name.encode(), L.T[tname], sp, H5P_DEFAULT, H5P_DEFAULT), f"attribute {name}")
        v = (ctypes.c_int32 if tname == "H5T_NATIVE_INT32" else ctypes.c_uint8)(value)
        _ck(L.H5Awrite(a, L.T[tname], ctypes.byref(v)), "H5Awrite")
        L.H5Aclose(a), L.H5Sclose(sp)

    def _attr_fields(self, obj, names):
        L = self.lib
        base = _ck(L.H5Tcopy(L.T["H5T_C_S1"]), "H5Tcopy")
        L.H5Tset_size(base, 1)
        vt = _ck(L.H5Tvlen_create(base), "H5Tvlen_create")
        n = len(names)
        dims = (hsize_t * 1)(n)
        sp = _ck(L.H5Screate_simple(1, dims, None), "H5Screate_simple")
        a = _ck(L.H5Acreate2(obj, b"MATLAB_fields", vt, sp, H5P_DEFAULT, H5P_DEFAULT), "attribute MATLAB_fields")
        keep = [ctypes.create_string_buffer(s.encode(), len(s)) for s in names]
        arr = (_HVL * n)(*[_HVL(len(s), ctypes.cast(b, ctypes.c_void_p)) for s, b in zip(names, keep)])
        _ck(L.H5Awrite(a, vt, arr), "H5Awrite")
        L.H5Aclose(a), L.H5Sclose(sp), L.H5Tclose(vt), L.H5Tclose(base)

    # ---- values --------------------------------------------------------------------------------
    def _dataset(self, loc, name, tid, mdims, buf):
        L = self.lib
        hd = list(reversed([int(d) for d in mdims]))
        dims = (hsize_t * len(hd))(*hd)
        sp = _ck(L.H5Screate_simple(len(hd), dims, None), "H5Screate_simple")
        d = _ck(L.H5Dcreate2(loc, name.encode(), tid, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"dataset {name}")
        if buf is not None:
            _ck(L.H5Dwrite(d, tid, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), f"writing {name}")
        L.H5Sclose(sp)
        return d

    def _empty(self, loc, name, cls, mdims):
        L = self.lib
        dims = np.asarray(mdims, dtype=np.uint64)
        hd = (hsize_t * 1)(dims.size)
        sp = _ck(L.H5Screate_simple(1, hd, None), "H5Screate_simple")
        d = _ck(L.H5Dcreate2(loc, name.encode(), L.T["H5T_NATIVE_UINT64"], sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                f"dataset {name}")
        _ck(L.H5Dwrite(d, L.T["H5T_NATIVE_UINT64"], H5S_ALL, H5S_ALL, H5P_DEFAULT, dims.ctypes.data), "writing")
        self._attr_str(d, "MATLAB_class", cls)
        self._attr_int(d, "MATLAB_empty", 1, "H5T_NATIVE_UINT8")
        L.H5Sclose(sp), L.H5Dclose(d)

    def _array(self, loc, name, a: np.ndarray):
        L = self.lib
        if a.ndim < 2:
            a = a.reshape(1, -1)
        if a.size == 0:
            cls = "logical" if a.dtype == np.bool_ else ("char" if a.dtype.kind == "U" else _CLASS.get(
                np.dtype(a.real.dtype) if a.dtype.kind == "c" else a.dtype, "double"))
            self._empty(loc, name, cls, a.shape)
            return
        if a.dtype == np.bool_:
            data = np.asfortranarray(a.astype(np.uint8))
            d = self._dataset(loc, name, L.T["H5T_NATIVE_UINT8"], a.shape, data.ctypes.data)
            self._attr_str(d, "MATLAB_class", "logical")
            self._attr_int(d, "MATLAB_int_decode", 1, "H5T_NATIVE_INT32")
        elif a.dtype.kind == "c":
            real = np.dtype(np.float64) if a.dtype == np.complex128 else np.dtype(np.float32)
            cls = _CLASS[real]
            base = L.T[_NATIVE[cls]]
            t = _ck(L.H5Tcreate(H5T_COMPOUND, 2 * real.itemsize), "H5Tcreate")
            L.H5Tinsert(t, b"real", 0, base)
            L.H5Tinsert(t, b"imag", real.itemsize, base)
            data = np.asfortranarray(a)
            d = self._dataset(loc, name, t, a.shape, data.ctypes.data)
            self._attr_str(d, "MATLAB_class", cls)
            L.H5Tclose(t)
        else:
            if a.dtype not in _CLASS:
                a = a.astype(np.float64)
            cls = _CLASS[a.dtype]
            data = np.asfortranarray(a)
            d = self._dataset(loc, name, L.T[_NATIVE[cls]], a.shape, data.ctypes.data)
            self._attr_str(d, "MATLAB_class", cls)
        L.H5Dclose(d)

    def _char(self, loc, name, s: str):
        L = self.lib
        if not s:
            self._empty(loc, name, "char", (0, 0))
            return
        codes = np.frombuffer(s.encode("utf-16-le"), dtype=np.uint16).copy()
        d = self._dataset(loc, name, L.T["H5T_NATIVE_UINT16"], (1, codes.size), codes.ctypes.data)
        self._attr_str(d, "MATLAB_class", "char")
        self._attr_int(d, "MATLAB_int_decode", 2, "H5T_NATIVE_INT32")
        L.H5Dclose(d)

    def _struct(self, loc, name, fields: dict):
        L = self.lib
        g = _ck(L.H5Gcreate2(loc, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"group {name}")
        self._attr_str(g, "MATLAB_class", "struct")
        self._attr_fields(g, list(fields))
        for k, v in fields.items():
            self.value(g, k, v)
        L.H5Gclose(g)

    def _ref(self, v) -> int:
        """Writes v as a member of /#refs# and returns an object reference to it."""
        L = self.lib
        if self.refs is None:
            self.refs = _ck(L.H5Gcreate2(self.fid, b"#refs#", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "group #refs#")
        name = _ref_name(self.nref)
        self.nref += 1
        self.value(self.refs, name, v)
        r = ctypes.c_uint64(0)
        _ck(L.H5Rcreate(ctypes.byref(r), self.refs, name.encode(), H5R_OBJECT, -1), "H5Rcreate")
        return r.value

    def _refs(self, loc, name, items, cls):
        L = self.lib
        refs = np.array([self._ref(v) for v in items], dtype=np.uint64).reshape(1, -1)
        d = self._dataset(loc, name, L.T["H5T_STD_REF_OBJ"], refs.shape, np.asfortranarray(refs).ctypes.data)
        if cls:
            self._attr_str(d, "MATLAB_class", cls)
        L.H5Dclose(d)

    def _struct_array(self, loc, name, items):
        L = self.lib
        keys = list(items[0])
        g = _ck(L.H5Gcreate2(loc, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"group {name}")
        self._attr_str(g, "MATLAB_class", "struct")
        self._attr_fields(g, keys)
        for k in keys:
            self._refs(g, k, [it[k] for it in items], None)
        L.H5Gclose(g)

    def value(self, loc, name, v):
        if v is None:
            self._empty(loc, name, "double", (0, 0))
        elif isinstance(v, str):
            self._char(loc, name, v)
        elif isinstance(v, dict):
            self._struct(loc, name, v)
        elif isinstance(v, (bool, np.bool_)):
            self._array(loc, name, np.array([[bool(v)]]))
        elif isinstance(v, (int, float, np.integer, np.floating)):
            self._array(loc, name, np.array([[float(v)]]))
        elif isinstance(v, (complex, np.complexfloating)):
            self._array(loc, name, np.array([[complex(v)]]))
        elif isinstance(v, (list, tuple)):
            if v and all(isinstance(x, dict) for x in v) and all(list(x) == list(v[0]) for x in v):
                self._struct_array(loc, name, list(v))
            elif v and all(isinstance(x, (int, float, np.integer, np.floating)) and not isinstance(x, bool)
                           for x in v):
                self._array(loc, name, np.asarray(v, dtype=np.float64).reshape(1, -1))
            elif not v:
                self._empty(loc, name, "cell", (0, 0))
            else:
                self._refs(loc, name, list(v), "cell")
        else:
            a = np.asarray(v)
            if a.dtype == object:
                self._refs(loc, name, list(a.ravel(order="F")), "cell")
            elif a.dtype.kind == "U":
                self._char(loc, name, "".join(a.ravel().tolist()))
            else:
                self._array(loc, name, a)


def _ref_name(n: int) -> str:
    """Names of the members of /#refs#: a, b, ..., z, ba, bb, ... (base 26)."""
    s = ""
    while True:
        s = chr(ord("a") + n % 26) + s
        n //= 26
        if n == 0:
            return s


def header(created: Optional[str] = None) -> bytes:
    """The 128-byte MATLAB header at the start of the user block."""
    created = created or time.strftime("%a %b %d %H:%M:%S %Y")
    text = f"MATLAB 7.3 MAT-file, Platform: GLNXA64, Created on: {created} HDF5 schema 1.00 .".encode()
    return text[:116].ljust(116, b" ") + b"\x00" * 8 + b"\x00\x02" + b"IM"


class _GInfo(ctypes.Structure):
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", hsize_t), ("max_corder", ctypes.c_int64),
                ("mounted", ctypes.c_bool)]


H5I_GROUP, H5I_DATASET, H5T_STRING = 2, 5, 3


class _Reader:
    """Reads back the conventions the writer uses (for checking files and for callers that read
    solution files): numeric, complex, logical, char, empty, struct, struct-array and cell values."""

    def __init__(self, lib):
        self.lib = lib

    def _attr(self, obj, name):
        L = self.lib
        if L.H5Aexists(obj, name.encode()) <= 0:
            return None
        a = _ck(L.H5Aopen(obj, name.encode(), H5P_DEFAULT), f"attribute {name}")
        t = L.H5Aget_type(a)
        try:
            if L.H5Tget_class(t) == H5T_STRING:
                n = L.H5Tget_size(t)
                buf = ctypes.create_string_buffer(n + 1)
                _ck(L.H5Aread(a, t, buf), "H5Aread")
                return buf.raw[:n].rstrip(b"\x00").decode()
            v = ctypes.c_int64(0)
            _ck(L.H5Aread(a, L.T["H5T_NATIVE_INT64"], ctypes.byref(v)), "H5Aread")
            return v.value
        finally:
            L.H5Tclose(t), L.H5Aclose(a)

    def _fields(self, g):
        L = self.lib
        a = _ck(L.H5Aopen(g, b"MATLAB_fields", H5P_DEFAULT), "attribute MATLAB_fields")
        t, sp = L.H5Aget_type(a), L.H5Aget_space(a)
        dims = (hsize_t * 1)()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        arr = (_HVL * int(dims[0]))()
        _ck(L.H5Aread(a, t, arr), "H5Aread")
        names = [ctypes.string_at(h.p, h.len).decode() for h in arr]
        L.H5Dvlen_reclaim(t, sp, H5P_DEFAULT, arr)
        L.H5Sclose(sp), L.H5Tclose(t), L.H5Aclose(a)
        return names

    def _dims(self, d):
        L = self.lib
        sp = L.H5Dget_space(d)
        n = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * max(n, 1))()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        L.H5Sclose(sp)
        return tuple(reversed([int(dims[i]) for i in range(n)]))

    def _deref(self, d, refs):
        L = self.lib
        out = []
        for r in refs:
            rv = ctypes.c_uint64(int(r))
            o = _ck(L.H5Rdereference2(d, H5P_DEFAULT, H5R_OBJECT, ctypes.byref(rv)), "H5Rdereference2")
            out.append(self.obj(o))
            L.H5Oclose(o)
        return out

    def _refs_of(self, d):
        L = self.lib
        mdims = self._dims(d)
        refs = np.empty(int(np.prod(mdims)), dtype=np.uint64)
        _ck(L.H5Dread(d, L.T["H5T_STD_REF_OBJ"], H5S_ALL, H5S_ALL, H5P_DEFAULT, refs.ctypes.data), "H5Dread")
        return refs

    def obj(self, o):
        L = self.lib
        cls = self._attr(o, "MATLAB_class")
        if L.H5Iget_type(o) == H5I_GROUP:
            names = self._fields(o)
            out = {}
            for k in names:
                c = _ck(L.H5Oopen(o, k.encode(), H5P_DEFAULT), f"member {k}")
                if L.H5Iget_type(c) == H5I_DATASET and self._attr(c, "MATLAB_class") is None:
                    out[k] = self._deref(c, self._refs_of(c))  # a field of a struct array
                else:
                    out[k] = self.obj(c)
                L.H5Oclose(c)
            if out and all(isinstance(v, list) for v in out.values()) and \
                    len({len(v) for v in out.values()}) == 1 and self._is_array(o, names):
                n = len(next(iter(out.values())))
                return [{k: out[k][i] for k in names} for i in range(n)]
            return out
        mdims = self._dims(o)
        if self._attr(o, "MATLAB_empty"):
            dims = np.empty(int(np.prod(self._dims(o))), dtype=np.uint64)
            _ck(L.H5Dread(o, L.T["H5T_NATIVE_UINT64"], H5S_ALL, H5S_ALL, H5P_DEFAULT, dims.ctypes.data), "H5Dread")
            if cls == "char":
                return ""
            if cls == "cell":
                return []
            dt = {"logical": np.bool_}.get(cls, next((k for k, v in _CLASS.items() if v == cls), np.float64))
            return np.zeros(tuple(int(x) for x in dims), dtype=dt)
        if cls == "cell":
            return self._deref(o, self._refs_of(o))
        n = int(np.prod(mdims))
        if cls == "char":
            codes = np.empty(n, dtype=np.uint16)
            _ck(L.H5Dread(o, L.T["H5T_NATIVE_UINT16"], H5S_ALL, H5S_ALL, H5P_DEFAULT, codes.ctypes.data), "H5Dread")
            return codes.tobytes().decode("utf-16-le")
        if cls == "logical":
            b = np.empty(n, dtype=np.uint8)
            _ck(L.H5Dread(o, L.T["H5T_NATIVE_UINT8"], H5S_ALL, H5S_ALL, H5P_DEFAULT, b.ctypes.data), "H5Dread")
            return b.astype(np.bool_).reshape(mdims, order="F")
        dt = next(k for k, v in _CLASS.items() if v == cls)
        t = L.H5Dget_type(o)
        cplx = L.H5Tget_class(t) == H5T_COMPOUND
        L.H5Tclose(t)
        if cplx:
            ct = np.complex128 if dt == np.float64 else np.complex64
            base = L.T[_NATIVE[cls]]
            mt = _ck(L.H5Tcreate(H5T_COMPOUND, 2 * np.dtype(dt).itemsize), "H5Tcreate")
            L.H5Tinsert(mt, b"real", 0, base)
            L.H5Tinsert(mt, b"imag", np.dtype(dt).itemsize, base)
            a = np.empty(n, dtype=ct)
            _ck(L.H5Dread(o, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data), "H5Dread")
            L.H5Tclose(mt)
        else:
            a = np.empty(n, dtype=dt)
            _ck(L.H5Dread(o, L.T[_NATIVE[cls]], H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data), "H5Dread")
        return a.reshape(mdims, order="F")

    def _is_array(self, g, names):
        L = self.lib
        for k in names:
            c = L.H5Oopen(g, k.encode(), H5P_DEFAULT)
            plain = L.H5Iget_type(c) == H5I_DATASET and self._attr(c, "MATLAB_class") is None
            L.H5Oclose(c)
            if not plain:
                return False
        return bool(names)


def load(path: str) -> Dict[str, object]:
    """The variables of a MAT 7.3 file written by ``save`` (name -> value; structs as dicts, struct
    arrays as lists of dicts, cells as lists, char as str, arrays in MATLAB shape)."""
    L = _h5()
    fid = _ck(L.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, H5P_DEFAULT), f"opening {path}")
    r = _Reader(L)
    out = {}
    try:
        info = _GInfo()
        _ck(L.H5Gget_info(fid, ctypes.byref(info)), "H5Gget_info")
        for i in range(int(info.nlinks)):
            buf = ctypes.create_string_buffer(1024)
            L.H5Lget_name_by_idx(fid, b".", 0, 0, i, buf, 1024, H5P_DEFAULT)
            name = buf.value.decode()
            if name.startswith("#"):
                continue
            o = _ck(L.H5Oopen(fid, name.encode(), H5P_DEFAULT), f"variable {name}")
            out[name] = r.obj(o)
            L.H5Oclose(o)
    finally:
        L.H5Fclose(fid)
    return out


def save(path: str, variables: Dict[str, object]) -> None:
    """Writes `variables` (name -> value) as a MAT 7.3 file at `path`."""
    L = _h5()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fcpl = _ck(L.H5Pcreate(L.T["H5P_CLS_FILE_CREATE_ID"]), "H5Pcreate")
    _ck(L.H5Pset_userblock(fcpl, 512), "H5Pset_userblock")
    fid = _ck(L.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, fcpl, H5P_DEFAULT), f"creating {path}")
    L.H5Pclose(fcpl)
    w = _Writer(L, fid)
    try:
        for k, v in variables.items():
            w.value(fid, k, v)
    finally:
        if w.refs is not None:
            L.H5Gclose(w.refs)
        L.H5Fclose(fid)
    with open(path, "r+b") as f:
        f.write(header())
