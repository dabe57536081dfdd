"""Fine-grained hot-path operators over the C-ABI: the pieces of the reference's
AbstractProjMPS seam (product, buildleft!, buildright!, calculate), the local
eigensolver and the truncated split. Every call runs on the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .core import Context, Tensor, handle, precision_code


def _ctx(*ts):
    for t in ts:
        if t is not None:
            return t.ctx
    return Context.default()


_OPS = {"N": 0, "T": 1, "C": 2}


def gemm(opA, opB, M, N, K, alpha, A, lda, B, ldb, beta, Cmat, ldc, precision="f64"):
    """C = alpha op(A) op(B) + beta C on device buffers viewed as column-major matrices
    (the contraction primitive behind contract(), tensors.jl:13-18)."""
    ctx = _ctx(A)
    ctx.call("tnh_gemm", _OPS[opA], _OPS[opB], int(M), int(N), int(K), float(alpha), A.h, int(lda),
             B.h, int(ldb), float(beta), Cmat.h, int(ldc), precision_code(precision))
    return Cmat


def heff_apply(L, W1, W2, R, x, out=None, coeff=1.0, precision="f64") -> Tensor:
    """H_eff * x, replaces product(::ProjMPS, x, direction, 2), projmps.jl:103-144.
    L or R = None stands for the edge block ones(1,1,1)."""
    ctx = _ctx(x)
    if out is None:
        out = Tensor.empty(x.shape, x.dtype, ctx)
    c = complex(coeff)
    ctx.call("tnh_heff_apply", handle(L), W1.h, W2.h, handle(R), x.h, out.h, c.real, c.imag,
             precision_code(precision))
    return out


def heff_apply_noalloc(L, W1, W2, R, x, out, coeff=1.0, precision="f64") -> Tensor:
    """heff_apply into a caller-provided `out` without any device allocation: the library
    takes its temporaries from a static per-GPU scratch (tnh_heff_apply_noalloc) and
    raises TnhError (TNH_ERR_NOMEM) for shapes whose scratch would not fit."""
    ctx = _ctx(x)
    c = complex(coeff)
    ctx.call("tnh_heff_apply_noalloc", handle(L), W1.h, W2.h, handle(R), x.h, out.h, c.real, c.imag,
             precision_code(precision))
    return out


def env_left(Lprev, A, W, out=None, precision="f64") -> Tensor:
    """buildleft!, projmps.jl:50-66. Lprev = None is the edge block."""
    ctx = _ctx(A)
    h = C.c_void_p(out.h.value if out is not None else None)
    ctx.call("tnh_env_left", handle(Lprev), A.h, W.h, C.byref(h), precision_code(precision))
    if out is not None and out.h.value == h.value:
        return out
    if out is not None:
        out.h = None  # the library replaced (freed) the old buffer
    return Tensor(ctx, h)


def env_right(Rprev, A, W, out=None, precision="f64") -> Tensor:
    """buildright!, projmps.jl:74-95. Rprev = None is the edge block."""
    ctx = _ctx(A)
    h = C.c_void_p(out.h.value if out is not None else None)
    ctx.call("tnh_env_right", handle(Rprev), A.h, W.h, C.byref(h), precision_code(precision))
    if out is not None and out.h.value == h.value:
        return out
    if out is not None:
        out.h = None
    return Tensor(ctx, h)


def calculate(L, A, W, R, coeff=1.0) -> complex:
    """calculate(::ProjMPS), projmps.jl:192-216."""
    ctx = _ctx(A)
    re, im = C.c_double(), C.c_double()
    ctx.call("tnh_env_calculate", handle(L), A.h, W.h, handle(R), float(coeff), C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def two_site_theta(A1, A2) -> Tensor:
    """theta = contract(A1, A2, 3, 1), dmrg.jl:44-47."""
    ctx = _ctx(A1)
    h = C.c_void_p()
    ctx.call("tnh_two_site_theta", A1.h, A2.h, C.byref(h))
    return Tensor(ctx, h)


def eigsolve_ground(L, W1, W2, R, theta, krylovdim=3, maxiter=2, tol=1e-14, coeff=1.0,
                    precision="f64"):
    """Lowest Ritz pair of H_eff (replaces the KrylovKit eigsolve at dmrg.jl:51-53).
    theta is overwritten with the Ritz vector. Returns (energy, nmatvec)."""
    ctx = _ctx(theta)
    e = C.c_double()
    nmv = C.c_int()
    ctx.call("tnh_eigsolve_ground", handle(L), W1.h, W2.h, handle(R), theta.h, int(krylovdim),
             int(maxiter), float(tol), float(coeff), precision_code(precision), C.byref(e), C.byref(nmv))
    return e.value, nmv.value


def svd_trunc(M, cutoff=0.0, maxdim=0, mindim=1, force_fallback=False, want=(True, True, True)):
    """Truncated SVD with the rule of svd(x, idx; cutoff, maxdim, mindim), tensors.jl:168-227.
    Returns (U, S, Vt, discarded, used_fallback) with device tensors (None if not wanted)."""
    ctx = _ctx(M)
    hs = [C.c_void_p() for _ in range(3)]
    ptrs = [C.byref(h) if w else None for h, w in zip(hs, want)]
    k = C.c_int()
    disc = C.c_double()
    fb = C.c_int()
    ctx.call("tnh_svd_trunc", M.h, float(cutoff), int(maxdim), int(mindim), int(bool(force_fallback)),
             ptrs[0], ptrs[1], ptrs[2], C.byref(k), C.byref(disc), C.byref(fb))
    outs = [Tensor(ctx, h) if w else None for h, w in zip(hs, want)]
    return outs[0], outs[1], outs[2], disc.value, bool(fb.value)


def replacesites2(theta, direction, normalize=True, cutoff=0.0, maxdim=0, mindim=1):
    """replacesites!(psi, theta, site, direction, normalize; ...) for two sites,
    gmps.jl:199-267. Returns (A1, A2, schmidt values (host), discarded)."""
    ctx = _ctx(theta)
    h1, h2 = C.c_void_p(), C.c_void_p()
    k = C.c_int()
    disc = C.c_double()
    cap = int(min(theta.shape[0] * theta.shape[1], theta.shape[2] * theta.shape[3]))
    s = np.zeros(max(cap, 1))
    ctx.call("tnh_replacesites2", theta.h, int(bool(direction)), int(bool(normalize)), float(cutoff),
             int(maxdim), int(mindim), C.byref(h1), C.byref(h2), C.byref(k), C.byref(disc),
             s.ctypes.data_as(C.POINTER(C.c_double)), cap)
    return Tensor(ctx, h1), Tensor(ctx, h2), s[: k.value].copy(), disc.value
