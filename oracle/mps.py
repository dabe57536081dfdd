"""MPS containers and canonical-form moves of src/structures/mps/{gmps,mps}.jl, restated
for rank-1 GMPS (site tensors A[a, s, b]) held as a Python list plus a centre."""
from __future__ import annotations

import numpy as np

from .tensors import contract, moveidx, svd


class MPS:
    """GMPS(rank=1, dim, tensors, center), src/structures/mps/gmps.jl:8-13 (1-based centre)."""

    def __init__(self, tensors, center=0):
        self.tensors = [np.asarray(t) for t in tensors]
        self.center = center

    def __len__(self):
        return len(self.tensors)

    def __getitem__(self, i):  # 1-based
        return self.tensors[i - 1]

    def __setitem__(self, i, v):
        self.tensors[i - 1] = v

    def copy(self):
        return MPS([t.copy() for t in self.tensors], self.center)

    def bonddims(self):
        return [t.shape[2] for t in self.tensors[:-1]]

    def maxbonddim(self):
        """maxbonddim(psi), src/structures/mps/abstractmps.jl:75-81."""
        return max(self.bonddims()) if len(self) > 1 else 1


def norm(psi):
    """norm(psi), gmps.jl:29-37 (at the centre)."""
    if psi.center == 0:
        movecenter(psi, 1)
    A = psi[psi.center]
    return np.sqrt(np.vdot(A, A).real)


def normalize(psi):
    """normalize!(psi), gmps.jl:46-51."""
    if psi.center == 0:
        movecenter(psi, 1)
    psi[psi.center] = psi[psi.center] / norm(psi)


def moveleft(psi, idx, **kw):
    """moveleft!(psi, idx), gmps.jl:60-67."""
    if 1 < idx <= len(psi):
        U, S, V, _ = svd(psi[idx], 1, **kw)
        V = S[:, None] * V
        psi[idx] = U
        psi[idx - 1] = contract(psi[idx - 1], V, 3, 2)


def moveright(psi, idx, **kw):
    """moveright!(psi, idx), gmps.jl:75-82."""
    if 0 < idx < len(psi):
        U, S, V, _ = svd(psi[idx], 3, **kw)
        V = S[:, None] * V
        psi[idx] = U
        psi[idx + 1] = contract(V, psi[idx + 1], 2, 1)


def movecenter(psi, idx, **kw):
    """movecenter!(psi, idx; kwargs...), gmps.jl:90-112."""
    N = len(psi)
    if idx < 1 or idx > N:
        raise ValueError("The index is out of range.")
    if psi.center == 0:
        for i in range(1, idx):
            moveright(psi, i, **kw)
        for i in range(1, N - idx + 1):
            moveleft(psi, N + 1 - i, **kw)
    elif idx > psi.center:
        for i in range(psi.center, idx):
            moveright(psi, i, **kw)
    elif idx < psi.center:
        c = psi.center
        for i in range(1, c - idx + 1):
            moveleft(psi, c + 1 - i, **kw)
    psi.center = idx


def replacesites_n(A, direction, cutoff=0.0, maxdim=0, mindim=1):
    """replacesites!(psi, A, site, direction, false; kw) for nsites >= 2, rank 1,
    src/structures/mps/gmps.jl:215-261: A[cl, d, ..., d, cr] is split one site at a time by
    the truncated SVD. Sweeping right (direction false) every site but the last gets V of
    svd(U, -1) over the combined (left bond, d) index and the rest carries U S; sweeping left
    the mirror image. Returns (site tensors, largest discarded weight)."""
    from .mpo import combineidxs

    n = A.ndim - 2
    d = A.shape[1]
    out = [None] * n
    U = A
    dmax = 0.0
    for i in range(1, n):
        if direction:  # :220-238
            U, _ = combineidxs(U, [U.ndim - 1, U.ndim])
            site1 = n - i + 1
            U, S, V, disc = svd(U, -1, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
            U = U * S                                             # :229
            D = A.shape[-1] if site1 == n else out[site1].shape[0]  # :230
            out[site1 - 1] = V.reshape((S.shape[0], d, D), order="F")  # :231-232
        else:  # :239-256
            U, _ = combineidxs(U, [1, 2])
            site1 = i
            U, S, V, disc = svd(U, -1, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
            U = moveidx(U * S, U.ndim, 1)                         # :247-248
            D = A.shape[0] if site1 == 1 else out[site1 - 2].shape[2]  # :249
            V = V.reshape((S.shape[0], D, d), order="F")          # :250-251
            out[site1 - 1] = moveidx(V, 1, -1)                    # :252
        dmax = max(dmax, disc)
    out[0 if direction else n - 1] = U                            # :259-261
    return out, dmax


def replacesites2(theta, direction, normalize_=False, cutoff=0.0, maxdim=0, mindim=1):
    """replacesites!(psi, A, site, direction, normalize; kw) for nsites=2, rank 1,
    src/structures/mps/gmps.jl:199-267. Returns (A1, A2, S, discarded)."""
    cl, d1, d2, cr = theta.shape
    if direction:  # sweeping left, :218-235
        U = theta.reshape((cl, d1, d2 * cr), order="F")           # combineidxs(U, [3, 4])
        U, S, V, disc = svd(U, -1, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
        U = U * S[None, None, :]                                 # :229
        A2 = V.reshape((S.shape[0], d2, cr), order="F")          # :231-232
        A1 = U
        centre = 1
    else:  # sweeping right, :236-256
        U = np.transpose(theta, (2, 3, 0, 1)).reshape((d2, cr, cl * d1), order="F")
        U, S, V, disc = svd(U, -1, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
        U = U * S[None, None, :]                                 # :247
        A2 = moveidx(U, 3, 1)                                    # :248
        V = V.reshape((S.shape[0], cl, d1), order="F")           # :250-251
        A1 = moveidx(V, 1, -1)                                   # :252
        centre = 2
    if normalize_:
        if centre == 1:
            A1 = A1 / np.linalg.norm(A1)
        else:
            A2 = A2 / np.linalg.norm(A2)
    return A1, A2, S, disc


def random_mps(N, d, chi, rng):
    """randomGMPS(1, d, N, chi), src/structures/mps/gmps.jl:276-291, drawing N(0,1)
    entries from `rng` (a numpy Generator) instead of Julia's global RNG."""
    ts = []
    for i in range(1, N + 1):
        D1 = 1 if i == 1 else chi
        D2 = 1 if i == N else chi
        ts.append(rng.standard_normal((D1, d, D2)))
    psi = MPS(ts, 0)
    movecenter(psi, N)
    movecenter(psi, 1)
    psi[1] = rng.standard_normal((1, d, min(d, chi)))
    normalize(psi)
    return psi


def product_mps(states):
    """productMPS(st, names), src/structures/mps/mps.jl:68-76 (local state vectors)."""
    return MPS([np.asarray(s, dtype=np.complex128).reshape((1, -1, 1)) for s in states], 0)


def entropy(psi, site):
    """entropy(psi, site), gmps.jl:184-189: -sum s^2 ln s^2 across bond (site, site+1)."""
    movecenter(psi, site)
    _, S, _, _ = svd(psi[site], -1)
    S2 = S ** 2
    S2 = S2[S2 > 0]
    return float(-(S2 * np.log(S2)).sum())


def to_dense(psi):
    """Dense state vector, site 1 most significant (kron order)."""
    v = psi[1]
    for i in range(2, len(psi) + 1):
        v = np.tensordot(v, psi[i], axes=([v.ndim - 1], [0]))
    v = v.reshape(v.shape[1:-1])  # drop the unit edge bonds
    return v.reshape(-1)  # C order: site 1 slowest


def inner_oplist(psi, oplist_terms, phi=None):
    """inner(st, psi, oplist, phi), src/structures/mps/mps.jl:87-134: the list of
    coeff * <psi| O_term |phi> for terms (ops matrices, sites (1-based), coeff)."""
    phi = psi if phi is None else phi
    out = []
    N = len(psi)
    for mats, sites, coeff in oplist_terms:
        left = np.ones((1, 1))
        for s in range(1, N + 1):
            A = np.conj(psi[s])
            B = phi[s]
            if s in sites:
                O = mats[sites.index(s)]
                B = np.einsum("ts,asb->atb", O, B)
            left = np.einsum("ab,asc,bsd->cd", left, A, B)
        out.append(coeff * left[0, 0])
    return np.array(out)


def inner_oplist_env(psi, oplist_terms, phi=None):
    """inner(st, psi, oplist, phi), src/structures/mps/mps.jl:87-134, with the reference's
    environment blocks (ProjMPS(psi, phi; rank=1), :90): left blocks of sites 1..i and right
    blocks of sites i..N are built once, and each term contracts only its own sites between
    them. Returns coeff * <psi| O_term |phi> per term (terms as in inner_oplist)."""
    phi = psi if phi is None else phi
    N = len(psi)
    left = [np.ones((1, 1))]
    for s in range(1, N + 1):
        left.append(np.einsum("ab,asc,bsd->cd", left[-1], np.conj(psi[s]), phi[s], optimize=True))
    right = [None] * (N + 2)
    right[N + 1] = np.ones((1, 1))
    for s in range(N, 0, -1):
        right[s] = np.einsum("asc,bsd,cd->ab", np.conj(psi[s]), phi[s], right[s + 1], optimize=True)
    out = []
    for mats, sites, coeff in oplist_terms:
        env = left[sites[0] - 1]
        for s in range(sites[0], sites[-1] + 1):
            B = phi[s]
            if s in sites:
                B = np.einsum("ts,asb->atb", mats[sites.index(s)], B)
            env = np.einsum("ab,asc,bsd->cd", env, np.conj(psi[s]), B, optimize=True)
        out.append(coeff * np.einsum("ab,ab->", env, right[sites[-1] + 1]))
    return np.array(out)
