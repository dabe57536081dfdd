"""MPS checkpoints in the reference's HDF5 layout, written and read through the HDF5 C
library (ctypes; the h5py package is not required).

HDF5.write(parent, name, M::GMPS) and HDF5.read(parent, name, ::Type{GMPS}),
src/structures/mps/gmps.jl:359-383, store a group `name` with the attributes
type = "MPS" and version = 1, the scalar datasets "length" and "center", and one dataset
"MPS[i]" per site. The datasets follow HDF5.jl's conventions, so that TensorNetworks.jl
reads these files and writes files this module reads: a Julia array of size (a, s, b) is
stored with the HDF5 dims reversed, (b, s, a), and its data in column-major order; complex
entries are a compound of two doubles named "r" and "i".

The device DMRG of this package resumes from a checkpoint by reading the GMPS back and
starting dmrg from it (dmrg(..., checkpoint=path) writes one after every pass).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .models import GMPS

_LIBHDF5 = os.environ.get("TNHIP_LIBHDF5", "")


def _load():
    cands = [_LIBHDF5] if _LIBHDF5 else []
    cands += ["libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.103"]
    err = None
    for c in cands:
        if not c:
            continue
        try:
            return C.CDLL(c)
        except OSError as e:  # try the next name
            err = e
    raise ImportError(f"the HDF5 C library was not found ({err}); set TNHIP_LIBHDF5 to its path")


_h = None
hid_t = C.c_int64
hsize_t = C.c_ulonglong
H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0, 2
H5P_DEFAULT = H5S_ALL = 0
H5S_SCALAR = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_COMPOUND = 0, 1, 3, 6
H5T_STR_NULLPAD, H5T_CSET_UTF8 = 1, 1


def _lib():
    global _h
    if _h is None:
        h = _load()
        h.H5open()
        for name, res, args in [
            ("H5Fcreate", hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
            ("H5Fopen", hid_t, [C.c_char_p, C.c_uint, hid_t]),
            ("H5Fclose", C.c_int, [hid_t]),
            ("H5Gcreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t]),
            ("H5Gopen2", hid_t, [hid_t, C.c_char_p, hid_t]),
            ("H5Gclose", C.c_int, [hid_t]),
            ("H5Screate", hid_t, [C.c_int]),
            ("H5Screate_simple", hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            ("H5Sget_simple_extent_ndims", C.c_int, [hid_t]),
            ("H5Sget_simple_extent_dims", C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            ("H5Sclose", C.c_int, [hid_t]),
            ("H5Tcopy", hid_t, [hid_t]),
            ("H5Tcreate", hid_t, [C.c_int, C.c_size_t]),
            ("H5Tinsert", C.c_int, [hid_t, C.c_char_p, C.c_size_t, hid_t]),
            ("H5Tset_size", C.c_int, [hid_t, C.c_size_t]),
            ("H5Tget_size", C.c_size_t, [hid_t]),
            ("H5Tset_strpad", C.c_int, [hid_t, C.c_int]),
            ("H5Tset_cset", C.c_int, [hid_t, C.c_int]),
            ("H5Tget_class", C.c_int, [hid_t]),
            ("H5Tis_variable_str", C.c_int, [hid_t]),
            ("H5Tclose", C.c_int, [hid_t]),
            ("H5Dcreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
            ("H5Dopen2", hid_t, [hid_t, C.c_char_p, hid_t]),
            ("H5Dwrite", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            ("H5Dread", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            ("H5Dget_space", hid_t, [hid_t]),
            ("H5Dget_type", hid_t, [hid_t]),
            ("H5Dclose", C.c_int, [hid_t]),
            ("H5Acreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t]),
            ("H5Aopen", hid_t, [hid_t, C.c_char_p, hid_t]),
            ("H5Awrite", C.c_int, [hid_t, hid_t, C.c_void_p]),
            ("H5Aread", C.c_int, [hid_t, hid_t, C.c_void_p]),
            ("H5Aget_type", hid_t, [hid_t]),
            ("H5Aclose", C.c_int, [hid_t]),
        ]:
            f = getattr(h, name)
            f.restype, f.argtypes = res, args
        _h = h
    return _h


def _g(name):
    return hid_t.in_dll(_lib(), name).value


def _ok(v, what):
    if v < 0:
        raise OSError(f"HDF5: {what} failed")
    return v


def _complex_type():
    h = _lib()
    t = _ok(h.H5Tcreate(H5T_COMPOUND, 16), "H5Tcreate")
    _ok(h.H5Tinsert(t, b"r", 0, _g("H5T_NATIVE_DOUBLE_g")), "H5Tinsert")
    _ok(h.H5Tinsert(t, b"i", 8, _g("H5T_NATIVE_DOUBLE_g")), "H5Tinsert")
    return t


def _write_scalar(h, loc, name, value, attribute):
    sid = h.H5Screate(H5S_SCALAR)
    tid = _g("H5T_NATIVE_LLONG_g")
    v = C.c_int64(int(value))
    if attribute:
        aid = _ok(h.H5Acreate2(loc, name.encode(), tid, sid, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate2")
        _ok(h.H5Awrite(aid, tid, C.byref(v)), "H5Awrite")
        h.H5Aclose(aid)
    else:
        did = _ok(h.H5Dcreate2(loc, name.encode(), tid, sid, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Dcreate2")
        _ok(h.H5Dwrite(did, tid, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.byref(v)), "H5Dwrite")
        h.H5Dclose(did)
    h.H5Sclose(sid)


def _write_string_attr(h, loc, name, text):
    data = text.encode("utf-8")
    tid = h.H5Tcopy(_g("H5T_C_S1_g"))
    h.H5Tset_size(tid, max(len(data), 1))
    h.H5Tset_strpad(tid, H5T_STR_NULLPAD)
    h.H5Tset_cset(tid, H5T_CSET_UTF8)
    sid = h.H5Screate(H5S_SCALAR)
    aid = _ok(h.H5Acreate2(loc, name.encode(), tid, sid, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate2")
    buf = C.create_string_buffer(data, max(len(data), 1))
    _ok(h.H5Awrite(aid, tid, buf), "H5Awrite")
    h.H5Aclose(aid)
    h.H5Sclose(sid)
    h.H5Tclose(tid)


def write_mps(path, psi: GMPS, name: str = "psi"):
    """HDF5.write(h5open(path, "w"), name, psi), gmps.jl:359-370."""
    h = _lib()
    fid = _ok(h.H5Fcreate(str(path).encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate")
    ctype = _complex_type()
    try:
        g = _ok(h.H5Gcreate2(fid, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2")
        _write_string_attr(h, g, "type", "MPS")
        _write_scalar(h, g, "version", 1, attribute=True)
        _write_scalar(h, g, "length", len(psi), attribute=False)
        _write_scalar(h, g, "center", psi.center, attribute=False)
        for i, A in enumerate(psi.tensors, start=1):
            A = np.asarray(A, dtype=np.complex128)
            dims = (hsize_t * A.ndim)(*reversed(A.shape))  # HDF5.jl stores Julia dims reversed
            sid = h.H5Screate_simple(A.ndim, dims, None)
            did = _ok(h.H5Dcreate2(g, f"MPS[{i}]".encode(), ctype, sid, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                      "H5Dcreate2")
            data = np.ascontiguousarray(A.reshape(-1, order="F"))  # column-major, (r, i) pairs
            _ok(h.H5Dwrite(did, ctype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data.ctypes.data), "H5Dwrite")
            h.H5Dclose(did)
            h.H5Sclose(sid)
        h.H5Gclose(g)
    finally:
        h.H5Tclose(ctype)
        h.H5Fclose(fid)


def _read_scalar(h, loc, name, attribute):
    v = C.c_int64(0)
    if attribute:
        aid = _ok(h.H5Aopen(loc, name.encode(), H5P_DEFAULT), f"attribute {name}")
        _ok(h.H5Aread(aid, _g("H5T_NATIVE_LLONG_g"), C.byref(v)), "H5Aread")
        h.H5Aclose(aid)
    else:
        did = _ok(h.H5Dopen2(loc, name.encode(), H5P_DEFAULT), f"dataset {name}")
        _ok(h.H5Dread(did, _g("H5T_NATIVE_LLONG_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, C.byref(v)), "H5Dread")
        h.H5Dclose(did)
    return int(v.value)


def _read_string_attr(h, loc, name):
    aid = _ok(h.H5Aopen(loc, name.encode(), H5P_DEFAULT), f"attribute {name}")
    ftype = h.H5Aget_type(aid)
    try:
        if h.H5Tis_variable_str(ftype) > 0:
            mtype = h.H5Tcopy(_g("H5T_C_S1_g"))
            h.H5Tset_size(mtype, C.c_size_t(-1).value)  # H5T_VARIABLE
            p = C.c_char_p()
            _ok(h.H5Aread(aid, mtype, C.byref(p)), "H5Aread")
            h.H5Tclose(mtype)
            return (p.value or b"").decode("utf-8")
        n = h.H5Tget_size(ftype)
        buf = C.create_string_buffer(n + 1)
        _ok(h.H5Aread(aid, ftype, buf), "H5Aread")
        return buf.raw[:n].rstrip(b"\0").decode("utf-8")
    finally:
        h.H5Tclose(ftype)
        h.H5Aclose(aid)


def read_mps(path, name: str = "psi") -> GMPS:
    """HDF5.read(h5open(path, "r"), name, GMPS), gmps.jl:372-383."""
    h = _lib()
    if not Path(path).exists():
        raise FileNotFoundError(path)
    fid = _ok(h.H5Fopen(str(path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen")
    ctype = _complex_type()
    try:
        g = _ok(h.H5Gopen2(fid, name.encode(), H5P_DEFAULT), f"group {name}")
        if _read_string_attr(h, g, "type") != "MPS":
            raise ValueError("HDF5 group of file does not contain MPS data.")
        N = _read_scalar(h, g, "length", attribute=False)
        center = _read_scalar(h, g, "center", attribute=False)
        tensors = []
        for i in range(1, N + 1):
            did = _ok(h.H5Dopen2(g, f"MPS[{i}]".encode(), H5P_DEFAULT), f"dataset MPS[{i}]")
            sid = h.H5Dget_space(did)
            nd = h.H5Sget_simple_extent_ndims(sid)
            dims = (hsize_t * nd)()
            h.H5Sget_simple_extent_dims(sid, dims, None)
            shape = tuple(int(x) for x in reversed(list(dims)))  # back to Julia's order
            ftype = h.H5Dget_type(did)
            if h.H5Tget_class(ftype) == H5T_COMPOUND:
                data = np.empty(int(np.prod(shape)), dtype=np.complex128)
                _ok(h.H5Dread(did, ctype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data.ctypes.data), "H5Dread")
            else:
                real = np.empty(int(np.prod(shape)), dtype=np.float64)
                _ok(h.H5Dread(did, _g("H5T_NATIVE_DOUBLE_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, real.ctypes.data),
                    "H5Dread")
                data = real.astype(np.complex128)
            h.H5Tclose(ftype)
            h.H5Sclose(sid)
            h.H5Dclose(did)
            tensors.append(np.asfortranarray(data.reshape(shape, order="F")))
        h.H5Gclose(g)
    finally:
        h.H5Tclose(ctype)
        h.H5Fclose(fid)
    rank = tensors[0].ndim - 2
    return GMPS(rank, tensors[0].shape[1], tensors, center)
