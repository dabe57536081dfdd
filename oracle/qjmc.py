"""Quantum-jump Monte Carlo references (test infrastructure only): dense state-vector
trajectories that follow src/algorithms/mps/qjmc.jl with caller-supplied uniforms, and the
dense Lindblad master equation that the trajectory average converges to.

Uniform consumption per step, classical branch (qjmc.jl:59-112): one draw that the
classical branch discards (:64), one for the jump test rand() > exp(-sum(rates) dt) (:97),
and, when a jump happens, one for the channel (:99). Non-classical branch (:65-87): one
for the test r > |psi|^2 and, on a jump, one for the channel. Trajectories are rows of U;
each keeps its own cursor, so M trajectories run vectorised with the exact per-trajectory
order of the reference."""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla

from .exact import OPS
from .tebd import trotter_layers


def _mat(o, st_ops):
    return st_ops[o] if isinstance(o, str) else np.asarray(o)


def escape_terms(jump_terms, st_ops=OPS):
    """The escape operators of qjmc_gates, qjmc.jl:11-20: for each jump term
    (ops, sites, c) the term (dag(O) O on each site, sites, c^2)."""
    out = []
    for ops, sites, c in jump_terms:
        mats = [_mat(o, st_ops).conj().T @ _mat(o, st_ops) for o in ops]
        out.append((mats, list(sites), c ** 2))
    return out


def qjmc_layers(H_terms, jump_terms, N, dt, order=2, st_ops=OPS):
    """Heff = -i H - escape / 2 (qjmc.jl:23) trotterized with evol = "imag", i.e. gates
    exp(tau Heff) (gatelist.jl:86-100)."""
    terms = [(o, s, -1j * c) for o, s, c in H_terms]
    terms += [(m, s, -0.5 * c) for m, s, c in escape_terms(jump_terms, st_ops)]
    return trotter_layers(terms, N, dt, order=order, evol="imag", st_ops=st_ops)


def _apply_site_cols(V, site, U, N, d=2):
    """Apply a k-site dense operator on sites site..site+k-1 to every column of V."""
    M = V.shape[1]
    k = int(round(math.log(U.shape[0]) / math.log(d)))
    T = V.reshape((d ** (site - 1), d ** k, d ** (N - site - k + 1), M))
    T = np.einsum("ab,xbym->xaym", U, T)
    return T.reshape(-1, M)


def emission_rates_dense(V, jump_terms, N, st_ops=OPS):
    """qjmc_emission_rates, qjmc.jl:170-220: |c^2 <psi| prod dag(O) O |psi>| per jump
    term, for every column of V (returns M x n_terms)."""
    out = np.zeros((V.shape[1], len(jump_terms)))
    for t, (mats, sites, c2) in enumerate(escape_terms(jump_terms, st_ops)):
        W = V
        for m, s in zip(mats, sites):
            W = _apply_site_cols(W, s, m, N)
        out[:, t] = np.abs(c2 * np.einsum("im,im->m", V.conj(), W))
    return out


def qjmc_dense(v0, H_terms, jump_terms, N, dt, tmax, U, save=0.0, obs=(), classical=True,
               order=2, st_ops=OPS):
    """Run M = U.shape[0] trajectories from the state v0 (qjmc_simulation, qjmc.jl:28-167,
    without truncation). obs: one-site operators (name or matrix, site) measured as
    <psi|O|psi> at time 0 and at every save point (QJMCOperators, :252-260).
    Returns (jumps, jumptimes, times, measurements[M, nsave, nobs]); jump indices are
    1-based, as the reference stores them."""
    U = np.atleast_2d(np.asarray(U, dtype=np.float64))
    M = U.shape[0]
    layers = qjmc_layers(H_terms, jump_terms, N, dt, order=order, st_ops=st_ops)
    save = dt if save == 0 else save
    steps = int(math.ceil(round(tmax / dt, 5)))
    savesteps = int(math.ceil(round(save / dt)))
    V = np.repeat(np.asarray(v0, dtype=np.complex128)[:, None], M, axis=1)
    pos = np.zeros(M, dtype=np.int64)
    rows = np.arange(M)
    jumps = [[] for _ in range(M)]
    jtimes = [[] for _ in range(M)]

    def draw(mask):
        if np.any(pos[mask] >= U.shape[1]):
            raise ValueError("not enough uniforms supplied")
        u = np.full(M, np.nan)
        u[mask] = U[rows[mask], pos[mask]]
        pos[mask] += 1
        return u

    def measure():
        vals = np.zeros((M, len(obs)), dtype=np.complex128)
        for j, (o, s) in enumerate(obs):
            W = _apply_site_cols(V, s, _mat(o, st_ops), N)
            vals[:, j] = np.einsum("im,im->m", V.conj(), W)
        return vals

    times, meas = [0.0], [measure()]
    time = 0.0
    allm = np.ones(M, dtype=bool)
    for i in range(1, steps + 1):
        for layer in layers:
            for site, G in layer:
                V = _apply_site_cols(V, site, G, N)
        if classical:
            draw(allm)  # qjmc.jl:64, unused on this branch
            V = V / np.linalg.norm(V, axis=0)
            rates = emission_rates_dense(V, jump_terms, N, st_ops)
            er = rates.sum(axis=1)
            prob = np.exp(-er * dt)
            jump = draw(allm) > prob
        else:
            r = draw(allm)
            prob = np.linalg.norm(V, axis=0) ** 2
            V = V / np.linalg.norm(V, axis=0)
            jump = r > prob
            rates = emission_rates_dense(V, jump_terms, N, st_ops) if jump.any() else None
            er = rates.sum(axis=1) if rates is not None else None
        if jump.any():
            r = draw(jump)
            for m in np.nonzero(jump)[0]:
                cum = np.cumsum(rates[m]) / er[m]
                hit = np.nonzero(r[m] < cum)[0]
                if hit.size == 0:
                    raise ValueError("no jump channel selected (findfirst returned nothing)")
                idx = int(hit[0])
                ops, sites, _ = jump_terms[idx]
                v = V[:, m:m + 1]
                for o, s in zip(ops, sites):
                    v = _apply_site_cols(v, s, _mat(o, st_ops), N)
                V[:, m] = v[:, 0] / np.linalg.norm(v)
                jumps[m].append(idx + 1)
                jtimes[m].append(time + dt)
        time += dt
        if i % savesteps == 0:
            times.append(time)
            meas.append(measure())
    return jumps, jtimes, np.array(times), np.stack(meas, axis=1)


def lindblad_dense(rho0, H_terms, jump_terms, N, times, st_ops=OPS):
    """rho(t) of d rho/dt = -i[H, rho] + sum_t c_t^2 (L rho L^H - {L^H L, rho}/2) with
    L_t = prod of the term's operators, by the exact propagator of the Liouvillian."""
    from .exact import dense_hamiltonian

    dim = rho0.shape[0]
    H = dense_hamiltonian(N, [(list(o), list(s), c) for o, s, c in H_terms])
    I = np.eye(dim)
    # column-stacking vec: vec(A X B) = (B^T kron A) vec(X)
    Lv = -1j * (np.kron(I, H) - np.kron(H.T, I))
    for ops, sites, c in jump_terms:
        L = dense_hamiltonian(N, [([_mat(o, st_ops) for o in ops], list(sites), 1.0)])
        LdL = L.conj().T @ L
        Lv += c ** 2 * (np.kron(L.conj(), L) - 0.5 * np.kron(I, LdL) - 0.5 * np.kron(LdL.T, I))
    out = []
    r0 = rho0.reshape(-1, order="F")
    for t in times:
        out.append((sla.expm(t * Lv) @ r0).reshape(dim, dim, order="F"))
    return out


def qjmc_mps(states, H_terms, jump_terms, N, dt, tmax, U, save=0.0, obs=(), cutoff=1e-12, maxdim=0,
             mindim=1, order=2, st_ops=OPS):
    """One trajectory of qjmc_simulation, src/algorithms/mps/qjmc.jl:28-167 (classical
    branch), on an MPS: from the product state of the local vectors `states` (productMPS,
    mps.jl:68-76), per step applygates! with truncation (:60), the discarded draw (:64),
    normalize! (:91), the emission rates c^2 <L^H L> (qjmc_emission_rates, :170-220), the jump
    test rand() > exp(-sum(rates) dt) (:97), the channel (:99), and on a jump applyop! at the
    centre 1 followed by movecenter!(psi, N) and the truncated movecenter!(psi, 1) and
    normalize! (:101-107). U is this trajectory's row of uniforms. obs: one-site operators
    (name or matrix, site) measured as <psi|O|psi> at time 0 and at every save point
    (QJMCOperators, :252-260). Returns (jumps, jumptimes, times, measurements[nsave, nobs],
    the final maximum bond dimension)."""
    from .mps import inner_oplist_env, movecenter, normalize, product_mps
    from .tebd import applygates_mps, rows_from_layers

    U = np.asarray(U, dtype=np.float64).reshape(-1)
    rows = rows_from_layers(qjmc_layers(H_terms, jump_terms, N, dt, order=order, st_ops=st_ops))
    esc = escape_terms(jump_terms, st_ops)
    save = dt if save == 0 else save
    steps = int(math.ceil(round(tmax / dt, 5)))
    savesteps = int(math.ceil(round(save / dt)))
    psi = product_mps(states)
    pos = [0]

    def draw():
        if pos[0] >= U.size:
            raise ValueError("not enough uniforms supplied")
        pos[0] += 1
        return float(U[pos[0] - 1])

    obs_terms = [([_mat(o, st_ops)], [s], 1.0) for o, s in obs]

    def measure():
        return inner_oplist_env(psi, obs_terms) if obs_terms else np.zeros(0, complex)

    jumps, jtimes, times, meas = [], [], [0.0], [measure()]
    time = 0.0
    for i in range(1, steps + 1):
        applygates_mps(psi, rows, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
        draw()  # qjmc.jl:64, unused on the classical branch
        normalize(psi)
        rates = np.abs(inner_oplist_env(psi, esc))
        er = rates.sum()
        prob = math.exp(-er * dt)
        if draw() > prob:
            r = draw()
            hit = np.nonzero(r < np.cumsum(rates) / er)[0]
            if hit.size == 0:
                raise ValueError("no jump channel selected (findfirst returned nothing)")
            idx = int(hit[0])
            ops, sites, _ = jump_terms[idx]
            movecenter(psi, 1)
            for o, s in zip(ops, sites):  # applyop!, mps.jl:141-151
                psi[s] = np.einsum("ts,asb->atb", _mat(o, st_ops), psi[s])
            movecenter(psi, N)
            movecenter(psi, 1, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
            normalize(psi)
            jumps.append(idx + 1)
            jtimes.append(time + dt)
        time += dt
        if i % savesteps == 0:
            times.append(time)
            meas.append(measure())
    return jumps, jtimes, np.array(times), np.stack(meas), psi.maxbonddim()
