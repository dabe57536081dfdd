"""Two-site DMRG driver of src/algorithms/mps/dmrg.jl restated on the oracle pieces."""
from __future__ import annotations

import numpy as np

from .krylov import eigsolve_arnoldi, eigsolve_ground
from .mps import movecenter, normalize, replacesites2
from .projmps import ProjMPS, ProjMPSSquared, ProjMPSSum


def dmrg(psi, Ws, nsites=2, krylovdim=3, kryloviter=2, minsweeps=1, maxsweeps=1000,
         tol=1e-10, tolgrad=1e-5, numconverges=4, cutoff=1e-12, maxdim=1000, mindim=1,
         verbose=False, coeff=1.0, history=None, terms=(), ishermitian=True):
    """dmrg(psi, H::GMPS...; kwargs...), src/algorithms/mps/dmrg.jl:128-154 and :1-99.

    psi is an oracle MPS updated in place; Ws the MPO site tensors (with `coeff`).
    `terms` holds further Hamiltonian terms as (site tensors, coeff): MPO tensors (4
    indices) give ProjMPS(psi, H, psi; rank=2), MPS tensors (3 indices) the rank-1 term
    ProjMPS(phi, psi; rank=2, squared=true) (:140-147); all are summed as ProjMPSSum.
    Returns (psi, cost). `history` (a list) receives (sweep, cost, maxbonddim, max
    discarded). A "sweep" is one directional pass, as in the reference (:66).
    ishermitian=False solves each local problem with eigsolve_arnoldi (the eigenvalue of
    smallest real part; psi is made complex) and returns a complex cost.
    """
    assert nsites in (1, 2), "the one- and two-site updates"
    N = len(psi)
    if not ishermitian:
        for i in range(1, N + 1):
            psi[i] = psi[i].astype(np.complex128)
    solve = eigsolve_ground if ishermitian else eigsolve_arnoldi
    movecenter(psi, 1)                                       # :133
    projs = [ProjMPS(psi.tensors, Ws, center=1, coeff=coeff)]  # :143, projmps.jl:16-42
    for T, c in terms:                                       # :139-149
        if len(T) != N:
            raise ValueError("GMPS must share the same properties.")
        if T[0].ndim == 4:
            projs.append(ProjMPS(psi.tensors, T, center=1, coeff=c))
        elif T[0].ndim == 3:
            projs.append(ProjMPSSquared(T, psi.tensors, center=1, coeff=c))
        else:
            raise ValueError("Hamiltonian must be composed of MPOs (rank 2) or MPSs (rank 1).")
    Hs = projs[0] if len(projs) == 1 else ProjMPSSum(projs)  # :151-152
    cost = Hs.calculate()                                    # :22
    lastcost = cost
    D = psi.maxbonddim()
    lastD = D
    grad = 0.0
    direction = False
    converged = False
    convergedsweeps = convergedgrad = 0
    sweeps = 0
    while not converged:
        maxdisc = 0.0
        for j in range(1, N + 2 - nsites):                   # :35
            site = N + 1 - j if direction else j
            site1 = site + 1 - nsites if direction else site
            Hs.movecenter(site)                              # :41
            A0 = psi[site1]                                  # :44-47
            if nsites == 2:
                A0 = np.tensordot(A0, psi[site1 + 1], axes=([2], [0]))
            e, vec, _ = solve(lambda x: Hs.product(x, direction, nsites), A0,
                              krylovdim, kryloviter, 1e-14)                 # :50-53
            cost = e                                         # :55
            if nsites == 1:  # replacesites!'s one-site branch, gmps.jl:204-213
                psi[site1] = vec
                nxt = site1 + 1 - 2 * int(direction)
                if 0 < nxt <= N:
                    movecenter(psi, nxt)
                normalize(psi)
                continue
            A1, A2, _, disc = replacesites2(vec, direction, True, cutoff, maxdim, mindim)  # :57
            psi[site1] = A1
            psi[site1 + 1] = A2
            psi.center = site1 if direction else site1 + 1
            maxdisc = max(maxdisc, disc)
        Hs.movecenter(1 if direction else N)                 # :62
        direction = not direction                            # :63
        sweeps += 1                                          # :66
        D = psi.maxbonddim()

        def diff(x, y):
            return abs(x - y) if abs(x) < 1e-10 else abs((x - y) / x)

        if sweeps >= minsweeps:                               # :69-86
            if diff(cost, lastcost) < tol and lastD == D:
                convergedsweeps += 1
            else:
                convergedsweeps = 0
            dc = diff(cost, lastcost)
            if (dc + grad) != 0 and abs((dc - grad) / (dc + grad)) < tolgrad and lastD == D:
                convergedgrad += 1
            else:
                convergedgrad = 0
            if max(convergedsweeps, convergedgrad) >= numconverges:
                converged = True
            if sweeps >= maxsweeps and maxsweeps != 0:
                converged = True
        grad = abs(diff(cost, lastcost))
        lastcost = cost
        lastD = D
        if history is not None:
            history.append((sweeps, cost, D, maxdisc))
        if verbose:
            print(f"Sweep={sweeps}, energy={cost:.12f}, maxbonddim={D} ")
    return psi, cost
