"""Device context and tensors over the libtnhip C-ABI.

Arrays cross the boundary column-major (Fortran order), like Julia arrays: a numpy
array of shape (chi, d, chi) in Fortran order is the same bytes Julia would pass.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib
from ._lib import lib, TnhError

_NP2DT = {
    np.dtype(np.float64): _lib.F64,
    np.dtype(np.float32): _lib.F32,
    np.dtype(np.complex128): _lib.C128,
    np.dtype(np.complex64): _lib.C64,
}
_DT2NP = {v: k for k, v in _NP2DT.items()}

PRECISIONS = {"f64": _lib.PREC_F64, "f32": _lib.PREC_F32, "bf16x3": _lib.PREC_BF16X3}


def precision_code(p) -> int:
    if isinstance(p, int):
        return p
    try:
        return PRECISIONS[str(p).lower()]
    except KeyError:
        raise ValueError(f"unknown precision {p!r}; expected one of {sorted(PRECISIONS)}")


def dtype_code(dt) -> int:
    if isinstance(dt, int):
        return dt
    return _NP2DT[np.dtype(dt)]


class Context:
    """One HIP device + stream (tnh_ctx). Calls on a context are serialised."""

    _default: "Context | None" = None
    _lock = threading.Lock()

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib.tnh_ctx_create(device, C.byref(h))
        if rc != 0:
            raise TnhError(rc, "could not create a HIP context (is a GPU visible?)", "tnh_ctx_create")
        self.h = h
        self.device = device

    @classmethod
    def default(cls) -> "Context":
        with cls._lock:
            if cls._default is None:
                cls._default = Context(0)
            return cls._default

    def check(self, rc: int, fn: str):
        if rc != 0:
            msg = lib.tnh_last_error(self.h)
            raise TnhError(rc, msg.decode() if msg else "", fn)

    def call(self, fn: str, *args):
        self.check(getattr(lib, fn)(self.h, *args), fn)

    def synchronize(self):
        self.call("tnh_ctx_synchronize")

    # kernel timing with HIP events on this context's stream
    def profile(self, on: bool = True):
        self.call("tnh_profile_enable", int(bool(on)))

    def profile_reset(self):
        self.call("tnh_profile_reset")

    def profile_read(self, family: str) -> dict:
        n = C.c_int64()
        ms, fl, by = C.c_double(), C.c_double(), C.c_double()
        self.call("tnh_profile_read", family.encode(), C.byref(n), C.byref(ms), C.byref(fl), C.byref(by))
        return {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}

    # GEMM tile selection: True = autotune per size class (default), False = size rule
    def gemm_autotune(self, on: bool = True):
        self.call("tnh_gemm_autotune", 1 if on else 0)

    def gemm_autotune_reset(self):
        self.call("tnh_gemm_autotune", 2)

    def gemm_tuning(self) -> list[dict]:
        """The tuned size classes: mode, transposes, the timed M N K, chosen and rule tile,
        milliseconds per candidate tile."""
        need = C.c_int64()
        self.call("tnh_gemm_tuning_report", None, 0, C.byref(need))
        buf = C.create_string_buffer(int(need.value))
        self.call("tnh_gemm_tuning_report", buf, int(need.value), C.byref(need))
        out = []
        for line in buf.value.decode().splitlines():
            f = line.split()
            kv = dict(x.split("=", 1) for x in f[2:])
            out.append({"mode": f[0], "trans": f[1], "M": int(kv["M"]), "N": int(kv["N"]), "K": int(kv["K"]),
                        "tile": kv["tile"], "rule": kv["rule"], "ms": [float(x) for x in kv["ms"].split(",")]})
        return out

    def close(self):
        if self.h:
            lib.tnh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dims(shape) -> C.Array:
    arr = (C.c_int64 * max(1, len(shape)))(*[int(s) for s in shape])
    return arr


class Tensor:
    """A device tensor (tnh_buf) owned by a Context. Column-major dims."""

    __slots__ = ("ctx", "h", "_owner")

    def __init__(self, ctx: Context, handle: C.c_void_p, owner=None):
        self.ctx = ctx
        self.h = handle
        self._owner = owner  # keeps wrapped device memory alive

    # ---------------------------------------------------------- construction
    @classmethod
    def empty(cls, shape, dtype=np.float64, ctx: Context | None = None) -> "Tensor":
        ctx = ctx or Context.default()
        h = C.c_void_p()
        ctx.call("tnh_buf_create", dtype_code(dtype), len(shape), _dims(shape), C.byref(h))
        return cls(ctx, h)

    @classmethod
    def from_numpy(cls, a, dtype=None, ctx: Context | None = None) -> "Tensor":
        """Upload a host array (converted to Fortran order) as device dtype `dtype`."""
        ctx = ctx or Context.default()
        a = np.asarray(a)
        if a.dtype not in _NP2DT:
            a = a.astype(np.complex128 if np.iscomplexobj(a) else np.float64)
        a = np.asfortranarray(a)
        dev = dtype_code(dtype if dtype is not None else a.dtype)
        h = C.c_void_p()
        ctx.call("tnh_buf_upload", a.ctypes.data_as(C.c_void_p), _NP2DT[a.dtype], dev, a.ndim,
                 _dims(a.shape), C.byref(h))
        return cls(ctx, h)

    @classmethod
    def wrap(cls, ptr: int, shape, dtype, owner=None, ctx: Context | None = None) -> "Tensor":
        """Wrap caller-owned device memory (e.g. a torch tensor's data_ptr())."""
        ctx = ctx or Context.default()
        h = C.c_void_p()
        ctx.call("tnh_buf_wrap", C.c_void_p(ptr), dtype_code(dtype), len(shape), _dims(shape), C.byref(h))
        return cls(ctx, h, owner)

    # ------------------------------------------------------------ properties
    def _shape(self):
        dt = C.c_int()
        nd = C.c_int()
        dims = (C.c_int64 * 8)()
        rc = lib.tnh_buf_shape(self.h, C.byref(dt), C.byref(nd), dims)
        if rc != 0:
            raise TnhError(rc, "invalid buffer", "tnh_buf_shape")
        return dt.value, tuple(dims[i] for i in range(nd.value))

    @property
    def shape(self):
        return self._shape()[1]

    @property
    def dtype(self):
        return _DT2NP[self._shape()[0]]

    @property
    def code(self) -> int:
        return self._shape()[0]

    @property
    def ndim(self):
        return len(self.shape)

    def device_ptr(self) -> int:
        return lib.tnh_buf_device_ptr(self.h) or 0

    def numpy(self, dtype=None) -> np.ndarray:
        dt = np.dtype(dtype) if dtype is not None else self.dtype
        out = np.empty(self.shape, dtype=dt, order="F")
        if out.size:
            self.ctx.call("tnh_buf_download", self.h, out.ctypes.data_as(C.c_void_p), _NP2DT[dt])
        return out

    def astype(self, dtype) -> "Tensor":
        h = C.c_void_p()
        self.ctx.call("tnh_buf_convert", self.h, dtype_code(dtype), C.byref(h))
        return Tensor(self.ctx, h)

    def free(self):
        if self.h:
            lib.tnh_buf_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                lib.tnh_buf_free(self.ctx.h, self.h)
        except Exception:
            pass

    def __repr__(self):
        return f"Tensor(shape={self.shape}, dtype={self.dtype})"


def handle(t: "Tensor | None"):
    return t.h if t is not None else None
