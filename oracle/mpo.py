"""MPO construction and MPO products of src/structures/mps/mpo.jl restated in numpy (test
infrastructure): the finite-state automaton of MPO(st, H) with its channel bookkeeping and
`expand`, the SVD compression sweeps that follow it, MPOMPSProduct, MPOMPOProduct and
trotterMPO. Site tensors of an MPO are W[w, s_out, s_in, w'] (rank-2 GMPS); column-major
reshapes throughout, as in Julia."""
from __future__ import annotations

import numpy as np

from .exact import OPS
from .tensors import contract, moveidx, svd


def _op(o, st_ops):
    return np.asarray(st_ops[o] if isinstance(o, str) else o, dtype=np.complex128)


def combineidxs(x, idxs):
    """combineidxs(x, idxs), src/tensors.jl:86-101: the listed indices are moved to the end,
    one after the other, and joined into one (the first listed index fastest). Returns the
    tensor and (idxs, their dims)."""
    y = x
    dims = tuple(x.shape[i - 1] for i in idxs)
    for i, idx in enumerate(idxs, start=1):
        y = moveidx(y, idx + 1 - i, -1)
    shape = y.shape
    keep = shape[: len(shape) - len(idxs)]
    return y.reshape(keep + (int(np.prod(shape[len(keep):])),), order="F"), (list(idxs), dims)


def expand(O, D1, D2):
    """expand(O, D1, D2), src/structures/mps/mpo.jl:462-470: D1 new rows and D2 new
    columns before the last row and column."""
    dims = O.shape
    new = np.zeros((dims[0] + D1, dims[1], dims[2], dims[3] + D2), dtype=np.complex128)
    d1 = 1 if dims[0] == 1 else dims[0] - 1
    new[:d1, :, :, : dims[3] - 1] = O[:d1, :, :, : dims[3] - 1]
    new[:d1, :, :, -1] = O[:d1, :, :, -1]
    new[dims[0] + D1 - 1, :, :, dims[3] + D2 - 1] = O[dims[0] - 1, :, :, dims[3] - 1]
    return new


def automaton(terms, N, d=2, st_ops=OPS):
    """The automaton part of MPO(st, H), src/structures/mps/mpo.jl:323-440, before the SVD
    compression. terms: (ops, sites (1-based, ascending), coeff) as in OpList."""
    Id = _op("id", st_ops)
    ten = np.zeros((2, d, d, 2), dtype=np.complex128)
    ten[0, :, :, 0] = Id
    ten[1, :, :, 1] = Id
    O = [ten[0:1, :, :, 0:2].copy()] + [ten.copy() for _ in range(2, N)] + [ten[0:2, :, :, 1:2].copy()]
    maxrng = max([1] + [s[-1] - s[0] + 1 for _, s, _ in terms])
    rngs = [[] for _ in range(maxrng)]
    for t, (_, sites, _) in enumerate(terms):
        rngs[sites[-1] - sites[0]].append(t)
    for i in range(1, maxrng + 1):
        nextterms = [[] for _ in range(i)]
        coeffs = [[] for _ in range(i)]
        ingoings = [[] for _ in range(i)]
        outgoings = [[] for _ in range(i)]
        for site in range(1, N + 1):
            idxs = [t for t in rngs[i - 1] if terms[t][1][0] == site]
            if i == 1:
                for t in idxs:  # one-site terms: top-right corner (:363-367)
                    O[site - 1][0, :, :, -1] += terms[t][2] * _op(terms[t][0][0], st_ops)
                continue
            for t in idxs:  # schedule the term's entries on its sites (:369-395)
                ops, sites, coeff = terms[t]
                outgoing = 0
                for k in range(1, i + 1):
                    ingoing = outgoing
                    for l in range(1, len(outgoings[k - 1]) + 2):
                        outgoing = l
                        if l not in outgoings[k - 1]:
                            break
                    outgoing = 0 if k == i else outgoing
                    s = site + k - 1
                    op = ops[sites.index(s)] if s in sites else "id"
                    nextterms[k - 1].append(op)
                    coeffs[k - 1].append(coeff if k == 1 else 1)
                    ingoings[k - 1].append(ingoing)
                    outgoings[k - 1].append(outgoing)
            # pull the entries of this site (:397-412)
            terms_here, ins, outs, cos = nextterms[0], ingoings[0], outgoings[0], coeffs[0]
            for j in range(i - 1):
                nextterms[j], ingoings[j] = nextterms[j + 1], ingoings[j + 1]
                outgoings[j], coeffs[j] = outgoings[j + 1], coeffs[j + 1]
            nextterms[i - 1], ingoings[i - 1], outgoings[i - 1], coeffs[i - 1] = [], [], [], []
            if terms_here:  # (:414-431)
                ingoinglen = sum(1 for x in ins if x != 0)
                outgoinglen = sum(1 for x in outs if x != 0)
                ingoingsrt = O[site - 1].shape[0] - 1
                outgoingsrt = O[site - 1].shape[3] - 1
                O[site - 1] = expand(O[site - 1], ingoinglen, outgoinglen)
                for j, op in enumerate(terms_here):
                    x = 1 if ins[j] == 0 else ingoingsrt + ins[j]
                    y = outgoingsrt + 1 + outgoinglen if outs[j] == 0 else outgoingsrt + outs[j]
                    O[site - 1][x - 1, :, :, y - 1] = cos[j] * _op(op, st_ops)
    return O


def compress(O, cutoff=1e-15, maxdim=0, mindim=1):
    """The SVD sweeps closing MPO(st, H), src/structures/mps/mpo.jl:442-457: left to right,
    O[s] <- U S and O[s+1] <- V O[s+1] (SVD over the right bond); right to left, O[s] <- S U
    and O[s-1] <- O[s-1] V^T (SVD over the left bond); each SVD truncated by
    (cutoff, maxdim, mindim)."""
    O = [w.copy() for w in O]
    N = len(O)
    kw = dict(cutoff=cutoff, maxdim=maxdim, mindim=mindim)
    for s in range(N - 1):
        U, S, V, _ = svd(O[s], 4, **kw)
        O[s] = U * S[None, None, None, :]
        O[s + 1] = contract(V, O[s + 1], 2, 1)
    for s in range(N - 1, 0, -1):
        U, S, V, _ = svd(O[s], 1, **kw)
        O[s] = S[:, None, None, None] * U
        O[s - 1] = contract(O[s - 1], V, 4, 2)
    return O


def mpo(terms, N, d=2, cutoff=1e-15, maxdim=0, mindim=1, st_ops=OPS):
    """MPO(st, H; cutoff, maxdim, mindim), src/structures/mps/mpo.jl:323-459."""
    return compress(automaton(terms, N, d, st_ops), cutoff, maxdim, mindim)


def _moveright(ts, i, rank):
    """moveright!(psi, i) without truncation, gmps.jl:75-82, for rank-1 or rank-2 sites."""
    A = ts[i - 1]
    U, S, V, _ = svd(A, A.ndim)
    ts[i - 1] = U
    ts[i] = contract(S[:, None] * V, ts[i], 2, 1)


def mpo_mps_product(O, psi, cutoff=0.0, maxdim=0, mindim=1):
    """MPOMPSProduct(O, psi; kwargs), src/structures/mps/mpo.jl:125-147: per site
    B = contract(M, A, 3, 2), the bonds (w, a) and (w', b) combined (w fastest), the centre
    moved right through every new site without truncation, then movecenter!(phi, 1; kwargs)
    from an unset centre (gmps.jl:90-112). Returns the oracle MPS (centre 1)."""
    from .mps import MPS, movecenter

    N = len(psi)
    ts = []
    for i in range(1, N + 1):
        B = contract(O[i - 1], psi[i], 3, 2)        # (w, so, w', a, b)
        B, _ = combineidxs(B, [1, 4])               # (so, w', b, (w, a))
        B, _ = combineidxs(B, [2, 3])               # (so, (w, a), (w', b))
        B = moveidx(B, 1, 2)                        # ((w, a), so, (w', b))
        ts.append(B)
        if i > 1:
            _moveright(ts, i - 1, 1)
    phi = MPS(ts, 0)
    movecenter(phi, 1, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
    return phi


def _as_rank1(O):
    """View MPO sites (Dl, d, d, Dr) as rank-1 sites (Dl, d*d, Dr): the SVD moves group
    every index except the bond, so both give the same decompositions."""
    return [w.reshape(w.shape[0], -1, w.shape[3], order="F") for w in O]


def mpo_mpo_product(O1, O2, cutoff=0.0, maxdim=0, mindim=1):
    """MPOMPOProduct(O1, O2; kwargs), src/structures/mps/mpo.jl:149-169: per site
    M = contract(M1, M2, 3, 2) with the bonds (w1, w2) and (w1', w2') combined (the first
    fastest), the centre moved right without truncation, then movecenter!(O, 1; kwargs)."""
    from .mps import MPS, movecenter

    N = len(O1)
    ts = []
    for i in range(N):
        M = contract(O1[i], O2[i], 3, 2)            # (w1, o, w1', w2, i2, w2')
        M, _ = combineidxs(M, [1, 4])               # (o, w1', i2, w2', (w1, w2))
        M, _ = combineidxs(M, [2, 4])               # (o, i2, (w1, w2), (w1', w2'))
        M = moveidx(M, 3, 1)                        # ((w1, w2), o, i2, (w1', w2'))
        ts.append(M)
        if i > 0:
            _moveright(ts, i, 2)
    d = O1[0].shape[1]
    phi = MPS(_as_rank1(ts), 0)
    movecenter(phi, 1, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
    return [t.reshape(t.shape[0], d, d, t.shape[2], order="F") for t in phi.tensors]


def qo_to_mpo(T, cutoff=0.0, maxdim=0, mindim=1):
    """GMPS(psi::GQS; kwargs) for an operator (rank 2), src/structures/mps/gmps.jl:294-311: the
    tensor T[o1, i1, o2, i2, ...] (column-major, as productQO and exp(::GQS) build it,
    qo.jl:72-78, :247-251) padded with unit bonds and split from the left by truncated SVDs
    over the last index after the remaining sites are combined; U of each split is a site."""
    T = np.asarray(T, dtype=np.complex128)
    n = T.ndim // 2
    tensor = T.reshape((1,) + T.shape + (1,), order="F")
    ts = []
    for _ in range(n - 1):
        tensor, cmb = combineidxs(tensor, list(range(4, tensor.ndim + 1)))
        U, S, V, _ = svd(tensor, -1, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
        tensor = S[:, None] * V
        ts.append(U)
        tensor = tensor.reshape((tensor.shape[0],) + tuple(cmb[1]), order="F")
    ts.append(tensor)
    return ts


def trotter_mpo(terms, N, t, d=2, cutoff=0.0, maxdim=0, mindim=1, st_ops=OPS):
    """trotterMPO(st, H, t; kwargs), src/structures/mps/mpo.jl:474-518: starting from the
    identity MPO, for each term in order, the exact exp(t c O_term) on the term's own sites
    (productQO of its operators) split into an MPO, identities of that bond dimension in the
    gaps between its sites and of dimension 1 elsewhere, multiplied onto the running MPO by
    applyMPO(O3, O; kwargs) (MPOMPOProduct with truncation)."""
    from .tensors import expm_tensor

    Id = _op("id", st_ops)
    O = [Id.reshape(1, d, d, 1).copy() for _ in range(N)]
    for ops, sites, c in terms:
        P = _op(ops[0], st_ops)
        for o in ops[1:]:  # tensorproduct: P[o1, i1, o2, i2, ...]
            P = np.multiply.outer(P, _op(o, st_ops))
        E = expm_tensor(t * c * P, [2 * k for k in range(1, len(ops) + 1)])  # exp(::GQS)
        O2 = qo_to_mpo(E, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
        O3 = [None] * N
        for j, s in enumerate(sites):
            start = 1 if j == 0 else sites[j - 1] + 1
            for k in range(start, s):  # gaps: identity of the incoming bond dimension
                D = O2[j].shape[0]
                ten = np.zeros((D, d, d, D), dtype=np.complex128)
                for q in range(D):
                    ten[q, :, :, q] = Id
                O3[k - 1] = ten
            O3[s - 1] = O2[j].copy()
        for k in range(sites[-1] + 1, N + 1):
            O3[k - 1] = Id.reshape(1, d, d, 1).copy()
        O = mpo_mpo_product(O3, O, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
    return O


def adjoint(O):
    """adjoint(O::GMPS), src/structures/mps/mpo.jl:32-39: conj with the physical indices
    swapped, per site."""
    return [np.conj(moveidx(w, 2, 3)) for w in O]


def inner(*args):
    """inner(args::GMPS...), src/structures/mps/mpo.jl:181-217: <bra| O_1 ... O_k |ket> for
    site-tensor lists of states (a, s, b) and MPOs (w, s_out, s_in, w'); the bra is moved to
    the front past any operators before it, each of which becomes its adjoint (:195-204)."""
    args = [list(a) for a in args]
    if len(args) < 2:
        raise ValueError("There must be atleast 2 MPS arguments (rank 1).")
    if any(len(a) != len(args[0]) for a in args):
        raise ValueError("GMPS must share the same length.")
    ranks = [np.ndim(a[0]) - 2 for a in args]
    if sum(r == 1 for r in ranks) != 2:
        raise ValueError("The inner product must have a braket structure.")
    if ranks[-1] != 1:
        raise ValueError("The final GMPS must be rank 1 (MPS).")
    idx = next(j for j, r in enumerate(ranks) if r == 1)    # :195-199
    for j in range(idx - 1, -1, -1):                         # :200-205
        psi = args[j + 1]
        O = adjoint(args[j])
        args[j] = psi
        args[j + 1] = O
    prod = np.ones([1] * len(args), dtype=np.complex128)     # :208
    for site in range(len(args[0])):                         # :209-214
        prod = contract(prod, np.conj(args[0][site]), 1, 1)
        for j in range(1, len(args)):
            prod = contract(prod, args[j][site], [1, prod.ndim - 1], [1, 2])
    return prod.reshape(-1)[0]                               # :216
