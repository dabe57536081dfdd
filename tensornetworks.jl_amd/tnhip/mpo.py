"""MPOs on the GPU: construction from an operator list with the reference's compression,
MPO * MPS and MPO * MPO products, and the time-evolution MPO of trotterMPO.

Mirrors MPO(st, H; kwargs), applyMPO / MPOMPSProduct / MPOMPOProduct and trotterMPO of
src/structures/mps/mpo.jl (:323-459, :105-169, :474-518). Tensors stay on the device
inside the products; the GMPS-level functions take and return host GMPS objects like the
reference's."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .core import Context, Tensor
from .models import GMPS, OpList, Sitetypes, mpo_tensors
from .tebd import DeviceMPS, _oplist_payload, expm_tensor


def _handles(ts):
    return (C.c_void_p * len(ts))(*[t.h.value for t in ts])


def _adopt(ctx, arr, n):
    return [Tensor(ctx, C.c_void_p(arr[i])) for i in range(n)]


def mpo_device(st: Sitetypes, H: OpList, cutoff=1e-15, maxdim=0, mindim=1, ctx: Context | None = None):
    """The MPO of H as N device tensors: the automaton built on the host by the library
    and compressed on the device (tnh_mpo_from_oplist); cutoff < 0 skips the compression."""
    ctx = ctx or Context.default()
    N, n = H.length, len(H.ops)
    firsts, ranges, flat = _oplist_payload(st, H)
    coeffs = np.zeros(2 * max(n, 1))
    for t, c in enumerate(H.coeffs):
        coeffs[2 * t], coeffs[2 * t + 1] = complex(c).real, complex(c).imag
    out = (C.c_void_p * N)()
    ctx.call("tnh_mpo_from_oplist", N, st.dim, n, (C.c_int * max(n, 1))(*firsts), (C.c_int * max(n, 1))(*ranges),
             flat.ctypes.data_as(C.POINTER(C.c_double)), coeffs.ctypes.data_as(C.POINTER(C.c_double)),
             float(cutoff), int(maxdim), int(mindim), out)
    return _adopt(ctx, out, N)


def MPO(st: Sitetypes, H: OpList, cutoff=1e-15, maxdim=0, mindim=1, compress=True) -> GMPS:
    """MPO(st, H; cutoff=1e-15, maxdim=0, mindim=1), src/structures/mps/mpo.jl:323-459: the
    automaton MPO and the SVD compression sweeps (on the GPU). compress=False returns the
    automaton alone, built on the host (mpo_tensors)."""
    if not compress:
        return GMPS(2, st.dim, mpo_tensors(st, H), 0)
    ts = mpo_device(st, H, cutoff, maxdim, mindim)
    return GMPS(2, st.dim, [t.numpy() for t in ts], 0)


def _device_sites(x, ctx):
    ts = x.sites if isinstance(x, DeviceMPS) else x.tensors
    return [t if isinstance(t, Tensor) else Tensor.from_numpy(np.asarray(t), ctx=ctx) for t in ts]


def MPOMPSProduct(O, psi, cutoff=0.0, maxdim=0, mindim=1, ctx: Context | None = None) -> DeviceMPS:
    """MPOMPSProduct(O, psi; kwargs), mpo.jl:125-147, on the GPU (tnh_mpo_apply): O a rank-2
    GMPS (or list of MPO tensors), psi a GMPS or DeviceMPS. Returns a DeviceMPS, centre 1."""
    ctx = ctx or (psi.ctx if isinstance(psi, DeviceMPS) else Context.default())
    Os = _device_sites(O, ctx) if not isinstance(O, list) else [Tensor.from_numpy(np.asarray(w), ctx=ctx) for w in O]
    ps = _device_sites(psi, ctx)
    if len(Os) != len(ps):
        raise ValueError("GMPS must share the same length.")
    N = len(ps)
    out = (C.c_void_p * N)()
    ctx.call("tnh_mpo_apply", N, _handles(Os), _handles(ps), float(cutoff), int(maxdim), int(mindim), out)
    return DeviceMPS(_adopt(ctx, out, N), 1, ps[0].shape[1], ctx)


def MPOMPOProduct(O1, O2, cutoff=0.0, maxdim=0, mindim=1, ctx: Context | None = None) -> list:
    """MPOMPOProduct(O1, O2; kwargs), mpo.jl:149-169, on the GPU (tnh_mpo_mpo_product).
    O1, O2: rank-2 GMPS or lists of MPO tensors (host or device). Returns the N device
    tensors of O1 O2 (centre 1)."""
    ctx = ctx or Context.default()

    def sites(O):
        ts = O.tensors if isinstance(O, GMPS) else O
        return [t if isinstance(t, Tensor) else Tensor.from_numpy(np.asarray(t), ctx=ctx) for t in ts]

    a, b = sites(O1), sites(O2)
    if len(a) != len(b):
        raise ValueError("GMPS must share the same length.")
    N = len(a)
    out = (C.c_void_p * N)()
    ctx.call("tnh_mpo_mpo_product", N, _handles(a), _handles(b), float(cutoff), int(maxdim), int(mindim), out)
    return _adopt(ctx, out, N)


def applyMPO(arg1, arg2, cutoff=0.0, maxdim=0, mindim=1):
    """applyMPO(arg1, arg2; kwargs), mpo.jl:105-123: O |psi> (rank 2 then rank 1), <psi| O
    (rank 1 then rank 2: the product with adjoint(O)), or O1 O2. Returns a host GMPS."""
    r1, r2 = arg1.rank, arg2.rank
    if len(arg1) != len(arg2):
        raise ValueError("GMPS must share the same length.")
    if arg1.dim != arg2.dim:
        raise ValueError("GMPS must share the same physical dims.")
    kw = dict(cutoff=cutoff, maxdim=maxdim, mindim=mindim)
    if r1 == 2 and r2 == 1:
        return MPOMPSProduct(arg1, arg2, **kw).to_gmps()
    if r1 == 1 and r2 == 2:  # adjoint(O): conjugate and swap the physical indices
        adj = [np.conj(np.transpose(np.asarray(w), (0, 2, 1, 3))) for w in arg2.tensors]
        return MPOMPSProduct(adj, arg1, **kw).to_gmps()
    if r1 == 2 and r2 == 2:
        return GMPS(2, arg1.dim, [t.numpy() for t in MPOMPOProduct(arg1, arg2, **kw)], 1)
    raise ValueError("Unallowed combinations of MPS ranks.")


def _qo_to_mpo(T, cutoff, maxdim, mindim, ctx):
    """GMPS(psi::GQS; kwargs), gmps.jl:294-311, for an operator tensor T[o1, i1, o2, i2, ...]:
    splits from the left by truncated SVDs (on the GPU)."""
    from .ops import svd_trunc

    n = T.ndim // 2
    d = T.shape[0]
    rest = np.asfortranarray(np.asarray(T, dtype=np.complex128).reshape((1,) + T.shape, order="F"))
    ts = []
    for _ in range(n - 1):
        left = rest.shape[0]
        M = np.asfortranarray(rest.reshape(left * d * d, -1, order="F"))
        U, S, Vt, _, _ = svd_trunc(Tensor.from_numpy(M, ctx=ctx), cutoff, maxdim, mindim)
        k = S.shape[0]
        ts.append(np.asfortranarray(U.numpy().reshape(left, d, d, k, order="F")))
        SV = S.numpy()[:, None] * Vt.numpy()
        rest = np.asfortranarray(SV.reshape((k,) + rest.shape[3:], order="F"))
    ts.append(np.asfortranarray(rest.reshape(rest.shape + (1,), order="F")))
    return ts


def trotterMPO(st: Sitetypes, H: OpList, t, cutoff=0.0, maxdim=0, mindim=1, ctx: Context | None = None) -> GMPS:
    """trotterMPO(st, H, t; kwargs), src/structures/mps/mpo.jl:474-518: from the identity MPO,
    for every term in order, exp(t c O) of the term's operators (productQO + exp, set up on
    the host), split into an MPO on its sites, identities in its gaps (of the incoming bond
    dimension) and elsewhere, and multiplied on by the GPU MPO product with truncation."""
    ctx = ctx or Context.default()
    d, N = st.dim, H.length
    Id = np.asarray(st.op("id"), dtype=np.complex128)
    O = [Tensor.from_numpy(Id.reshape(1, d, d, 1, order="F"), ctx=ctx) for _ in range(N)]
    for ops, sites, c in zip(H.ops, H.sites, H.coeffs):
        P = np.asarray(st.op(ops[0]), dtype=np.complex128)
        for o in ops[1:]:
            P = np.multiply.outer(P, np.asarray(st.op(o), dtype=np.complex128))
        E = expm_tensor(t * c * P, [2 * k for k in range(1, len(ops) + 1)])
        O2 = _qo_to_mpo(E, cutoff, maxdim, mindim, ctx)
        O3 = [None] * N
        for j, s in enumerate(sites):
            start = 1 if j == 0 else sites[j - 1] + 1
            for k in range(start, s):
                D = O2[j].shape[0]
                ten = np.zeros((D, d, d, D), dtype=np.complex128, order="F")
                for q in range(D):
                    ten[q, :, :, q] = Id
                O3[k - 1] = ten
            O3[s - 1] = O2[j]
        for k in range(sites[-1] + 1, N + 1):
            O3[k - 1] = Id.reshape(1, d, d, 1, order="F")
        O = MPOMPOProduct(O3, O, cutoff=cutoff, maxdim=maxdim, mindim=mindim, ctx=ctx)
    return GMPS(2, d, [w.numpy() for w in O], 1)


def adjoint(O: GMPS) -> GMPS:
    """adjoint(O::GMPS), mpo.jl:32-39: conj with the physical indices swapped."""
    if O.rank != 2:
        raise ValueError("The generalised MPS must be of rank 2 (an MPO).")
    return GMPS(2, O.dim, [np.conj(np.transpose(np.asarray(w), (0, 2, 1, 3))) for w in O.tensors],
                O.center)


def inner(*args: GMPS, ctx: Context | None = None) -> complex:
    """inner(args::GMPS...), mpo.jl:181-217, on the GPU: <bra| O_1 ... O_k |ket> for rank-1
    (states) and rank-2 (MPOs) arguments, with exactly two states, the last one the ket. The
    bra moves to the front past any operators before it, each becoming its adjoint
    (:195-204); the operators between bra and ket are folded into one MPO by exact
    MPOMPOProduct (cutoff 0) and the chain is contracted by tnh_inner_mpo."""
    if len(args) < 2:
        raise ValueError("There must be atleast 2 MPS arguments (rank 1).")
    if any(a.dim != args[0].dim for a in args):
        raise ValueError("GMPS must share the same physical dim.")
    if any(len(a) != len(args[0]) for a in args):
        raise ValueError("GMPS must share the same length.")
    ranks = [a.rank for a in args]
    if sum(r == 1 for r in ranks) != 2:
        raise ValueError("The inner product must have a braket structure.")
    if ranks[-1] != 1:
        raise ValueError("The final GMPS must be rank 1 (MPS).")
    ctx = ctx or Context.default()
    args = list(args)
    idx = ranks.index(1)
    for j in range(idx - 1, -1, -1):
        psi, O = args[j + 1], adjoint(args[j])
        args[j], args[j + 1] = psi, O
    bra, ops, ket = args[0], args[1:-1], args[-1]

    def up(ts):
        return [Tensor.from_numpy(np.asarray(t, dtype=np.result_type(t, np.float64)), ctx=ctx) for t in ts]

    dbra, dket = up(bra.tensors), up(ket.tensors)
    W = None
    if ops:
        W = up(ops[0].tensors)
        for O in ops[1:]:
            W = MPOMPOProduct(W, O, cutoff=0.0, ctx=ctx)
    re, im = C.c_double(), C.c_double()
    N = len(bra)
    ctx.call("tnh_inner_mpo", N, _handles(dbra), _handles(W) if W is not None else None,
             _handles(dket), C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def dot(phi: GMPS, psi: GMPS, ctx: Context | None = None) -> complex:
    """dot(phi, psi) = inner(phi, psi), mpo.jl:218."""
    return inner(phi, psi, ctx=ctx)
