"""TEBD / QJMC references: the Trotter gate sequence of src/structures/mps/gatelist.jl
applied to a dense state vector (exact when the MPS is not truncated), and a dense
quantum-jump trajectory following src/algorithms/mps/qjmc.jl:28-167 with caller-supplied
uniforms."""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from .exact import OPS


def local_tensor(terms, st_ops, site, N, d=2):
    """sitetensor(oplist, st, site), src/structures/mps/oplist.jl:164-183, for terms of range
    <= 2: G[o1,i1,o2,i2] (or G[o,i] for a one-site block at the last site). An operator is
    a name in st_ops or a d x d matrix."""
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
    """exp(tau * G, [2, 4]) as a dense operator on (site, site+1), kron order (site slowest)."""
    if G.ndim == 2:
        return sla.expm(tau * G)
    d = G.shape[0]
    # G[o1,i1,o2,i2] -> matrix [(o1,o2), (i1,i2)] with o1 slowest (kron order)
    M = np.transpose(G, (0, 2, 1, 3)).reshape(d * d, d * d)
    return sla.expm(tau * M)


def trotter_layers(terms, N, dt, order=2, evol="imag", st_ops=OPS):
    """Layers of (site, dense gate) following trotterize, gatelist.jl:75-121 (range <= 2)."""
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
