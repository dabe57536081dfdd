"""Exact references: spin-1/2 operators (src/lattices/spinhalf.jl:1-23), dense
Hamiltonians from operator lists, dense MPO contraction, free-fermion (BdG)
solutions of the open transverse-field Ising chain."""
from __future__ import annotations

import numpy as np

# spinhalf(): Pauli matrices, s+ / s- with up = [1, 0], src/lattices/spinhalf.jl:1-23
OPS = {
    "x": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
    "id": np.eye(2, dtype=np.complex128),
    "pu": np.array([[1, 0], [0, 0]], dtype=np.complex128),
    "pd": np.array([[0, 0], [0, 1]], dtype=np.complex128),
    "n": np.array([[1, 0], [0, 0]], dtype=np.complex128),
    "s+": np.array([[0, 1], [0, 0]], dtype=np.complex128),
    "s-": np.array([[0, 0], [1, 0]], dtype=np.complex128),
}


def dense_hamiltonian(N, terms, d=2):
    """sum_t coeff_t (x)_i O_{t,i} as a dense 2^N matrix, site 1 most significant.
    terms: iterable of (op_names, sites (1-based), coeff)."""
    dim = d ** N
    H = np.zeros((dim, dim), dtype=np.complex128)
    for ops, sites, coeff in terms:
        mats = [OPS["id"]] * N
        for o, s in zip(ops, sites):
            mats = list(mats)
            mats[s - 1] = OPS[o] if isinstance(o, str) else np.asarray(o)
        M = mats[0]
        for m in mats[1:]:
            M = np.kron(M, m)
        H += coeff * M
    return H


def mpo_to_dense(Ws):
    """Contract MPO tensors W[w, s_out, s_in, w'] into a dense operator (site 1 slowest)."""
    cur = Ws[0][0]  # (s_out, s_in, w')
    for W in Ws[1:]:
        # cur[(O),(I), w] * W[w, o, i, w'] -> [(O,o),(I,i), w']
        O, I, _ = cur.shape
        nxt = np.einsum("abw,woiv->aobiv", cur, W)
        cur = nxt.reshape((O * W.shape[1], I * W.shape[2], W.shape[3]))
    return cur[:, :, 0]


def tfim_bdg_ground_energy(J, h):
    """Ground energy of H = -sum_i J_i sz_i sz_{i+1} - sum_i h_i sx_i on an open chain
    (len(h) = N, len(J) = N-1): E0 = -sum_k s_k with s_k the singular values of the
    N x N bidiagonal matrix B (B_ii = h_i, B_{i,i+1} = J_i)."""
    h = np.asarray(h, dtype=float)
    J = np.asarray(J, dtype=float)
    N = h.size
    B = np.diag(h)
    B[np.arange(N - 1), np.arange(1, N)] = J
    s = np.linalg.svd(B, compute_uv=False)
    return -float(s.sum())


def tfim_bdg_first_excited_energy(J, h):
    """First excited energy of the same open chain: one quasiparticle, E1 = E0 + 2 s_min
    (H = sum_k 2 s_k (n_k - 1/2)); checked against ED in the tests."""
    h = np.asarray(h, dtype=float)
    J = np.asarray(J, dtype=float)
    N = h.size
    B = np.diag(h)
    B[np.arange(N - 1), np.arange(1, N)] = J
    s = np.linalg.svd(B, compute_uv=False)
    return -float(s.sum()) + 2.0 * float(s.min())


def tfim_critical_ground_energy(N):
    """Closed form E0(N) = 1 - csc(pi / (4N + 2)) for J = h = 1."""
    return 1.0 - 1.0 / np.sin(np.pi / (4 * N + 2))


def ground_state_dense(H):
    w, v = np.linalg.eigh(H)
    return float(w[0]), v[:, 0]
