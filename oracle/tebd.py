"""TEBD / QJMC references: the Trotter gate sequence of src/structures/mps/gatelist.jl
applied to a dense state vector (exact when the MPS is not truncated), and a dense
quantum-jump trajectory following src/algorithms/mps/qjmc.jl:28-167 with caller-supplied
uniforms."""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from .exact import OPS


def local_tensor(terms, st_ops, site, N, d=2):
    """sitetensor(oplist, st, site), src/structures/mps/oplist.jl:164-183: the terms starting
    at `site`, padded with identities to the list's site range r (fewer sites at the chain's
    end), as G[o1, i1, ..., or, ir]. An operator is a name in st_ops or a d x d matrix."""
    rng = max([1] + [s[-1] - s[0] + 1 for _, s, _ in terms])
    r = min(rng, N - site + 1)
    G = None
    for ops, sites, c in terms:
        if sites[0] != site:
            continue
        mats = []
        for k in range(r):
            s = site + k
            if s in sites:
                o = ops[sites.index(s)]
                mats.append(st_ops[o] if isinstance(o, str) else np.asarray(o))
            else:
                mats.append(st_ops["id"])
        T = mats[0]
        for m in mats[1:]:
            T = np.multiply.outer(T, m)
        G = c * T if G is None else G + c * T
    return G


def gate_from_tensor(G, tau):
    """exp(tau * G, [2, 4, ..]) as a dense operator on the r sites from `site`, kron order
    (site slowest): G[o1, i1, ..., or, ir] -> matrix [(o1..or), (i1..ir)]."""
    r = G.ndim // 2
    d = G.shape[0]
    M = np.transpose(G, list(range(0, 2 * r, 2)) + list(range(1, 2 * r, 2))).reshape(d ** r, d ** r)
    return sla.expm(tau * M)


def trotter_layers(terms, N, dt, order=2, evol="imag", st_ops=OPS):
    """Layers of (site, dense gate) following trotterize, gatelist.jl:75-121."""
    rng = max([1] + [s[-1] - s[0] + 1 for _, s, _ in terms])
    order = 1 if rng == 1 else order
    dt = -1j * dt if evol == "real" else dt
    layers = []
    for i in range(1, rng + 1):
        time = dt / 2 if (i < rng and order == 2) else dt
        layer = []
        site = i
        while site <= N:
            G = local_tensor(terms, st_ops, site, N)
            if G is not None:
                layer.append((site, gate_from_tensor(G, time)))
            site += rng
        layers.append(layer)
    if order == 2:
        for i in range(1, rng):
            layers.append(layers[rng - i - 1])
    return layers


def apply_dense(v, site, U, N, d=2):
    """Apply a k-site dense gate on sites site..site+k-1 of a kron-ordered state vector."""
    k = int(round(np.log(U.shape[0]) / np.log(d)))
    psi = v.reshape((d ** (site - 1), d ** k, d ** (N - site - k + 1)))
    psi = np.einsum("ab,xby->xay", U, psi)
    return psi.reshape(-1)


def tebd_dense(v, layers, N, nsteps):
    """nsteps Trotter steps with renormalisation (tebd.jl:62-78) on a dense vector."""
    for _ in range(nsteps):
        for layer in layers:
            for site, U in layer:
                v = apply_dense(v, site, U, N)
        v = v / np.linalg.norm(v)
    return v


def expect_dense(v, N, op, site):
    """<v| op_site |v> for a one-site operator."""
    w = apply_dense(v, site, op, N)
    return np.vdot(v, w)


def rows_from_layers(layers, d=2):
    """Trotter layers of dense gates (kron order, site slowest) as the rows of a GateList:
    (sites, tensors) with r-site gates G[o1, i1, ..., or, ir] (O[o, i] for one site)."""
    rows = []
    for layer in layers:
        sites, tensors = [], []
        for site, U in layer:
            sites.append(site)
            r = int(round(np.log(U.shape[0]) / np.log(d)))
            perm = [p for k in range(r) for p in (k, r + k)]
            tensors.append(np.transpose(np.asarray(U).reshape((d,) * (2 * r)), perm))
        rows.append((sites, tensors))
    return rows


def applygate_mps(psi, site, gate, direction, cutoff=0.0, maxdim=0, mindim=1):
    """applygate(psi, site, gate, direction; cutoff, maxdim, mindim, error=false),
    src/structures/mps/gatelist.jl:137-171, on an oracle MPS: theta = A1 A2 (:142-145),
    theta'[a,o1,o2,b] = sum G[o1,i1,o2,i2] theta[a,i1,i2,b] (:149-154), split without
    normalisation (:157). A gate on r > 2 sites contracts r sites and splits them with
    replacesites!'s r-site branch (gmps.jl:215-261). A one-site gate O[o, i] acts on the
    site, which becomes the centre, and the centre then moves one site on (replacesites!'s
    one-site branch, gmps.jl:204-213). Returns the (largest) discarded weight."""
    from .mps import movecenter, replacesites2, replacesites_n

    g = np.asarray(gate)
    if g.ndim >= 6:
        r = g.ndim // 2
        theta = psi[site]
        for j in range(1, r):                                  # :142-145
            theta = np.tensordot(theta, psi[site + j], axes=([theta.ndim - 1], [0]))
        theta = np.tensordot(theta, g, axes=(list(range(1, r + 1)), list(range(1, 2 * r, 2))))
        theta = np.moveaxis(theta, 1, -1)                      # (a, o1..or, b), :149-154
        sites, disc = replacesites_n(theta, direction, cutoff, maxdim, mindim)  # :157
        for j in range(r):
            psi[site + j] = sites[j]
        psi.center = site if direction else site + r - 1
        return disc
    if g.ndim == 4:
        theta = np.tensordot(psi[site], psi[site + 1], axes=([2], [0]))
        theta = np.einsum("xiyj,aijb->axyb", g, theta)
        A1, A2, _, disc = replacesites2(theta, direction, False, cutoff, maxdim, mindim)
        psi[site], psi[site + 1] = A1, A2
        psi.center = site if direction else site + 1
        return disc
    if g.ndim == 2:
        psi[site] = np.einsum("st,atb->asb", g, psi[site])
        psi.center = site
        nxt = site + 1 - 2 * int(bool(direction))
        if 0 < nxt <= len(psi):
            movecenter(psi, nxt)
        return 0.0
    raise ValueError("gates act on one or two sites")


def applygates_mps(psi, rows, cutoff=0.0, maxdim=0, mindim=1):
    """applygates!(psi, gates; kwargs...), gatelist.jl:191-223: per row, the direction that
    starts at the end of the row nearer the centre (:201), then for each gate a truncated
    centre move to the gate (:208-214) and the gate (:217). `rows` as rows_from_layers."""
    from .mps import movecenter

    def size(g):
        return np.ndim(g) // 2

    for sites, gates in rows:
        if not sites:
            continue
        first, last = sites[0], sites[-1] + size(gates[-1]) - 1
        direction = not (abs(psi.center - first) < abs(psi.center - last))
        n = len(gates)
        for i in range(1, n + 1):
            k = n + 1 - i if direction else i
            ctr = sites[k - 1] if not direction else sites[k - 1] + size(gates[k - 1]) - 1
            movecenter(psi, ctr, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
            applygate_mps(psi, sites[k - 1], gates[k - 1], direction, cutoff, maxdim, mindim)
