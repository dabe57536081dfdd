"""Environments and the H_eff matvec of src/structures/mps/projmps.jl, restated.

Conventions (column-major, as in the reference): site tensor A[a, s, b];
MPO tensor W[w, s_out(bra), s_in(ket), w']; left block L[a_bra, w, a_ket];
right block R[b_bra, w, b_ket].
"""
from __future__ import annotations

import numpy as np

from .tensors import contract, moveidx


def edgeblock():
    """edgeblock(3), src/structures/mps/abstractprojmps.jl:33-36."""
    return np.ones((1, 1, 1), dtype=np.float64)


def buildleft(L, A, W):
    """buildleft!(projV, idx), src/structures/mps/projmps.jl:50-66, for [psi, H, psi]."""
    prod = contract(L, np.conj(A), 1, 1)              # :57
    prod = contract(prod, W, [1, prod.ndim - 1], [1, 2])  # :58-61
    prod = contract(prod, A, [1, prod.ndim - 1], [1, 2])  # :62
    return prod


def buildright(R, A, W):
    """buildright!(projV, idx), src/structures/mps/projmps.jl:74-95, for [psi, H, psi]."""
    prod = contract(A, R, 3, R.ndim)                    # :86
    prod = contract(W, prod, [3, 4], [2, prod.ndim])    # :87-90
    prod = contract(np.conj(A), prod, [2, 3], [2, prod.ndim])  # :91
    return prod


def product_reference(L, W1, W2, R, A, coeff=1.0):
    """product(projV, A, direction, nsites=2), rank-2 branch, in the reference's own
    contraction order, src/structures/mps/projmps.jl:107-134 and :144."""
    return product_reference_n(L, [W1, W2], R, A, coeff)


def product_reference_n(L, Ws, R, A, coeff=1.0):
    """The same for nsites = len(Ws) sites (the one-site update: one MPO tensor)."""
    nsites = len(Ws)
    prod = moveidx(L, L.ndim, 2)                       # :116
    for i in range(1, nsites + 1):                     # :119-129
        prod = contract(prod, Ws[i - 1], 1 + 2 * i, 1)  # :122 (one MPO object)
        prod = moveidx(prod, prod.ndim - 1, 2 * i + 2)  # :128
    prod = contract(prod, A, [2 * k for k in range(1, nsites + 2)],
                    [k for k in range(1, nsites + 2)])  # :132
    prod = contract(prod, R, list(range(2 + nsites, prod.ndim + 1)),
                    list(range(2, R.ndim + 1)))        # :133-134
    return prod * coeff                                # :144


def product_optimal(L, W1, W2, R, A, coeff=1.0):
    """The same matvec in the optimal order used by the device chain:
    T1 = L x, T2 = T1 (W1 W2), out = T2 R^T (2 chi^3 D d^2 + 2 chi^2 d^3 D^2 MACs)."""
    T1 = np.einsum("awx,xstb->awstb", L, A, optimize=True)
    Wc = np.einsum("wosv,vpty->wstopy", W1, W2)
    T2 = np.einsum("awstb,wstopy->aopyb", T1, Wc, optimize=True)
    return coeff * np.einsum("aopyb,cyb->aopc", T2, R, optimize=True)


def calculate(L, A, W, R, coeff=1.0):
    """calculate(projV), src/structures/mps/projmps.jl:192-216."""
    prod = buildleft(L, A, W)
    prod = contract(prod, R, list(range(1, prod.ndim + 1)), list(range(1, prod.ndim + 1)))
    return coeff * prod.reshape(-1)[0]


def heff_dense(L, W1, W2, R):
    """H_eff as a dense (chi_l d d chi_r)^2 matrix acting on column-major vec(theta)."""
    cl, d, cr = L.shape[0], W1.shape[1], R.shape[0]
    n = cl * d * d * cr
    H = np.zeros((n, n), dtype=np.result_type(L, W1, W2, R))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        x = e.reshape((cl, d, d, cr), order="F")
        H[:, k] = product_reference(L, W1, W2, R, x).reshape(-1, order="F")
    return H


class ProjMPS:
    """ProjMPS(psi, H, psi; rank=2) with a block per site and a movable centre,
    src/structures/mps/projmps.jl:1-42 and abstractprojmps.jl:33-81.
    `psi` is a list of site tensors shared by reference (it is updated in place
    by the DMRG driver, as the reference's GMPS is)."""

    def __init__(self, psi, Ws, center=1, coeff=1.0):
        self.psi = psi
        self.Ws = Ws
        self.coeff = coeff
        self.blocks = [edgeblock() for _ in psi]
        self.center = 0
        self.movecenter(center)

    def __len__(self):
        return len(self.psi)

    def block(self, idx):
        """block(projV, idx), abstractprojmps.jl:33-46 (1-based; out of range -> edge)."""
        if idx < 1 or idx > len(self):
            return edgeblock()
        return self.blocks[idx - 1]

    def buildleft(self, idx):
        self.blocks[idx - 1] = buildleft(self.block(idx - 1), self.psi[idx - 1], self.Ws[idx - 1])

    def buildright(self, idx):
        self.blocks[idx - 1] = buildright(self.block(idx + 1), self.psi[idx - 1], self.Ws[idx - 1])

    def movecenter(self, idx):
        """movecenter!(projV, idx), abstractprojmps.jl:60-81."""
        N = len(self)
        if self.center == 0:
            for i in range(1, idx):
                self.buildleft(i)
            for i in range(1, N - idx + 1):
                self.buildright(N + 1 - i)
        elif idx > self.center:
            for i in range(1, idx - self.center + 1):
                self.buildleft(self.center - 1 + i)
        elif idx < self.center:
            for i in range(1, self.center - idx + 1):
                self.buildright(self.center + 1 - i)
        self.center = idx

    def product(self, A, direction=False, nsites=2):
        """product(projV, A, direction, nsites), projmps.jl:103-144 (optimal order for two
        sites, the reference's order for one)."""
        site = self.center - nsites + 1 if direction else self.center
        L = self.block(site - 1)
        R = self.block(site + nsites)
        if nsites == 1:
            return product_reference_n(L, [self.Ws[site - 1]], R, A, self.coeff)
        return product_optimal(L, self.Ws[site - 1], self.Ws[site], R, A, self.coeff)

    def calculate(self):
        site = self.center
        return calculate(self.block(site - 1), self.psi[site - 1], self.Ws[site - 1],
                         self.block(site + 1), self.coeff)


# ----------------------------------------------- rank-1 ("squared") terms and sums
# dmrg(psi, Hs...) turns a rank-1 argument phi into ProjMPS(phi, psi; rank=2, squared=true,
# coeff), src/algorithms/mps/dmrg.jl:144-145: a bra-ket pair with no MPO between, whose
# blocks are 2-index overlaps L[a_phi, a_psi] and whose product is coeff |phi_p><phi_p| A.


def edgeblock2():
    """edgeblock(2), src/structures/mps/abstractprojmps.jl:33-36."""
    return np.ones((1, 1), dtype=np.float64)


def buildleft_overlap(L, phi, psi):
    """buildleft!(projV, idx) for objects [phi, psi], src/structures/mps/projmps.jl:50-66
    (the MPO loop of :58-61 is empty)."""
    prod = contract(L, np.conj(phi), 1, 1)                 # :57
    return contract(prod, psi, [1, prod.ndim - 1], [1, 2])  # :62


def buildright_overlap(R, phi, psi):
    """buildright!(projV, idx) for objects [phi, psi], src/structures/mps/projmps.jl:74-95."""
    prod = contract(psi, R, 3, R.ndim)                     # :86
    return contract(np.conj(phi), prod, [2, 3], [2, prod.ndim])  # :91


def project(L, phi1, phi2, R):
    """project(projV, A, direction, nsites=2) for objects [phi, psi],
    src/structures/mps/projmps.jl:152-179: phi's two sites, conjugated, between the
    overlap blocks; the result has theta's shape (chi_l, d, d, chi_r) of psi."""
    return project_n(L, [phi1, phi2], R)


def project_n(L, phis, R):
    """The same for nsites = len(phis) sites."""
    nsites = len(phis)
    prod = moveidx(L, L.ndim, 1)                            # :160
    for i, ph in enumerate(phis, start=1):                  # :163-173
        prod = contract(prod, np.conj(ph), 1 + i, 1)        # :165
        prod = moveidx(prod, prod.ndim - 1, 1 + i)          # :172
    return contract(prod, R, list(range(2 + nsites, prod.ndim + 1)),
                    list(range(1, R.ndim)))                 # :176-177


def product_squared(L, phi1, phi2, R, A, coeff=1.0):
    """product(projV, A, direction, nsites=2), squared branch, projmps.jl:135-144:
    coeff * conj(P) * sum(P .* A) with P = project(...)."""
    return product_squared_n(L, [phi1, phi2], R, A, coeff)


def product_squared_n(L, phis, R, A, coeff=1.0):
    """The same for nsites = len(phis) sites."""
    P = project_n(L, phis, R)                               # :137
    s = contract(P, A, list(range(1, P.ndim + 1)), list(range(1, A.ndim + 1)))[()]  # :140
    return coeff * (s * np.conj(P))                         # :139, :141, :144


def calculate_overlap(L, phi, psi, R, coeff=1.0):
    """calculate(projV) for objects [phi, psi], projmps.jl:192-216: coeff <phi|psi>."""
    prod = buildleft_overlap(L, phi, psi)
    prod = contract(prod, R, list(range(1, prod.ndim + 1)), list(range(1, prod.ndim + 1)))
    return coeff * prod.reshape(-1)[0]


class ProjMPSSquared:
    """ProjMPS(phi, psi; rank=2, squared=true, coeff), src/structures/mps/projmps.jl:16-42,
    with overlap blocks per site and a movable centre (abstractprojmps.jl:33-81). `psi` is
    shared by reference with the DMRG driver; `phi` is fixed."""

    def __init__(self, phi, psi, center=1, coeff=1.0):
        if len(phi) != len(psi):
            raise ValueError("GMPS must share the same length.")
        self.phi = phi
        self.psi = psi
        self.coeff = coeff
        self.blocks = [edgeblock2() for _ in psi]
        self.center = 0
        self.movecenter(center)

    def __len__(self):
        return len(self.psi)

    def block(self, idx):
        if idx < 1 or idx > len(self):
            return edgeblock2()
        return self.blocks[idx - 1]

    def buildleft(self, idx):
        self.blocks[idx - 1] = buildleft_overlap(self.block(idx - 1), self.phi[idx - 1],
                                                 self.psi[idx - 1])

    def buildright(self, idx):
        self.blocks[idx - 1] = buildright_overlap(self.block(idx + 1), self.phi[idx - 1],
                                                  self.psi[idx - 1])

    movecenter = ProjMPS.movecenter

    def product(self, A, direction=False, nsites=2):
        site = self.center - nsites + 1 if direction else self.center
        return product_squared_n(self.block(site - 1), self.phi[site - 1:site - 1 + nsites],
                                 self.block(site + nsites), A, self.coeff)

    def calculate(self):
        site = self.center
        return calculate_overlap(self.block(site - 1), self.phi[site - 1], self.psi[site - 1],
                                 self.block(site + 1), self.coeff)


class ProjMPSSum:
    """ProjMPSSum(projs), src/structures/mps/projmpssum.jl: movecenter!, product and
    calculate summed over the terms in order (:63-73)."""

    def __init__(self, projs):
        self.projs = list(projs)

    def movecenter(self, idx):
        for p in self.projs:
            p.movecenter(idx)

    def product(self, A, direction=False, nsites=2):
        out = self.projs[0].product(A, direction, nsites)
        for p in self.projs[1:]:
            out = out + p.product(A, direction, nsites)
        return out

    def calculate(self):
        return sum(p.calculate() for p in self.projs)
