"""Pins of the CPU oracle against what the mathematics fixes (the reference ships no
golden vectors: test/runtests.jl is an empty testset). Each check is chosen so that a
plausible restatement mistake (dropped term, wrong index, transposed operand,
missing conjugate) fails it."""
import numpy as np
import pytest

from oracle import exact
from oracle.krylov import eigsolve_ground
from oracle.mps import MPS, entropy, movecenter, random_mps, replacesites2, to_dense, inner_oplist
from oracle.projmps import (ProjMPS, buildleft, buildright, calculate, edgeblock, heff_dense,
                            product_optimal, product_reference)
from oracle.tensors import contract, moveidx, svd, truncation_rank
from oracle.dmrg import dmrg


def randn(rng, shape, cplx=False):
    a = rng.standard_normal(shape)
    return a + 1j * rng.standard_normal(shape) if cplx else a


def herm_env(rng, chi, D, cplx=True):
    X = randn(rng, (chi, D, chi), cplx)
    return 0.5 * (X + np.conj(np.transpose(X, (2, 1, 0))))


# ------------------------------------------------------------------ tensors.jl

def test_contract_index_order():
    x = np.arange(24.0).reshape(2, 3, 4)
    y = np.arange(20.0).reshape(4, 5)
    z = contract(x, y, 3, 1)
    assert z.shape == (2, 3, 5)
    assert np.allclose(z, np.einsum("abc,cd->abd", x, y))
    z2 = contract(y, x, 1, 3)
    assert z2.shape == (5, 2, 3)


def test_moveidx_matches_reference_permutation():
    # moveidx(x, 1, -1) on (a,b,c) gives (b,c,a); moveidx(x, 3, 1) gives (c,a,b)
    x = np.zeros((2, 3, 4))
    assert moveidx(x, 1, -1).shape == (3, 4, 2)
    assert moveidx(x, 3, 1).shape == (4, 2, 3)
    assert moveidx(x, 2, 2).shape == (2, 3, 4)


@pytest.mark.parametrize("cutoff,maxdim,mindim,expect", [
    (0.0, 0, 1, 6),        # no truncation
    (0.0, 3, 1, 3),        # maxdim caps
    (0.0, 10, 1, 6),       # maxdim > n is n
    (1e-3, 0, 1, 3),       # tail weight from index 4 = (1e-4+1e-6+1e-8)/tot <= 1e-3
    (1e-3, 2, 1, 2),
    (1e-3, 0, 5, 5),       # mindim wins
    (1e-20, 0, 1, 6),
    (0.5, 0, 1, 1),
])
def test_truncation_rule(cutoff, maxdim, mindim, expect):
    S = np.array([1.0, 0.5, 0.1, 1e-2, 1e-3, 1e-4])
    k, disc = truncation_rank(S, cutoff, maxdim, mindim)
    assert k == expect
    assert np.isclose(disc, (S[k:] ** 2).sum() / (S ** 2).sum())


def test_truncation_zero_matrix_keeps_one():
    k, _ = truncation_rank(np.zeros(4), 1e-10, 0, 1)
    assert k == 1


def test_svd_tensor_layout(rng):
    x = randn(rng, (3, 2, 5), True)
    U, S, V, _ = svd(x, 1)  # rows (s, b), cols a: U has the new bond first
    assert U.shape == (min(3, 10), 2, 5)
    rec = np.einsum("ksb,k,ka->asb", U, S, V)
    assert np.allclose(rec, x)
    U, S, V, _ = svd(x, 3)
    assert U.shape == (3, 2, 5)
    assert np.allclose(np.einsum("ask,k,kb->asb", U, S, V), x)


# ----------------------------------------------------------------- projmps.jl

@pytest.mark.parametrize("cplx", [False, True])
def test_product_orders_agree(rng, cplx):
    L, R = herm_env(rng, 6, 3, cplx), herm_env(rng, 5, 4, cplx)
    W1, W2 = randn(rng, (3, 2, 2, 2), cplx), randn(rng, (2, 2, 2, 4), cplx)
    x = randn(rng, (6, 2, 2, 5), cplx)
    a = product_reference(L, W1, W2, R, x, 0.7)
    b = product_optimal(L, W1, W2, R, x, 0.7)
    assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(a)


def test_product_brute_force(rng):
    """Element-wise sum over all contracted indices for tiny bonds."""
    L, R = randn(rng, (2, 2, 2), True), randn(rng, (3, 2, 3), True)
    W1, W2 = randn(rng, (2, 2, 2, 3), True), randn(rng, (3, 2, 2, 2), True)
    x = randn(rng, (2, 2, 2, 3), True)
    out = np.zeros((2, 2, 2, 3), complex)
    for a, o1, o2, bp in np.ndindex(out.shape):
        s = 0
        for ap, w, i1, w1, i2, w2, b in np.ndindex(2, 2, 2, 3, 2, 2, 3):
            s += L[a, w, ap] * W1[w, o1, i1, w1] * W2[w1, o2, i2, w2] * R[bp, w2, b] * x[ap, i1, i2, b]
        out[a, o1, o2, bp] = s
    assert np.allclose(product_reference(L, W1, W2, R, x), out, atol=1e-12)


def test_heff_hermitian(rng):
    L, R = herm_env(rng, 4, 3), herm_env(rng, 3, 3)
    # W Hermitian in (s_out, s_in) per channel pair and symmetric in channels
    W = randn(rng, (3, 2, 2, 3), True)
    W = 0.5 * (W + np.conj(np.transpose(W, (0, 2, 1, 3))))
    W = 0.5 * (W + np.transpose(W, (3, 1, 2, 0)))
    H = heff_dense(L, W, W, R)
    # H_eff is Hermitian when L, R, W are; a missing conjugate breaks this
    Lh = 0.5 * (L + np.conj(np.transpose(L, (2, 1, 0))))
    assert np.allclose(L, Lh)
    x, y = randn(rng, 48, True), randn(rng, 48, True)
    assert np.isclose(np.vdot(x, H @ y), np.vdot(H @ x, y))


def test_identity_mpo_left_canonical_gives_identity(rng):
    """Left-canonical A with an identity MPO: L' = identity (D = 1)."""
    M = randn(rng, (6, 5), True)
    Q, _ = np.linalg.qr(M)
    A = Q.reshape((3, 2, 5), order="F")
    W = np.eye(2).reshape(1, 2, 2, 1)
    L = np.eye(3).reshape(3, 1, 3)
    Lp = buildleft(L, A, W)
    assert np.allclose(Lp[:, 0, :], np.eye(5))
    # mirror: right-canonical B gives R' = identity
    Q2, _ = np.linalg.qr(randn(rng, (6, 3), True))
    B = Q2.conj().T.reshape((3, 2, 3), order="F")
    Rp = buildright(np.eye(3).reshape(3, 1, 3), B, W)
    assert np.allclose(Rp[:, 0, :], np.eye(3))


def tfim_mpo(N, J=1.0, h=1.0):
    """H = -sum J sz sz - sum h sx as W[w, s_out, s_in, w'] (D = 3)."""
    X, Z, I = exact.OPS["x"].real, exact.OPS["z"].real, np.eye(2)
    Ws = []
    for i in range(N):
        W = np.zeros((3, 2, 2, 3))
        W[0, :, :, 0] = I
        W[2, :, :, 2] = I
        W[0, :, :, 1] = -J * Z
        W[1, :, :, 2] = Z
        W[0, :, :, 2] = -h * X
        if i == 0:
            W = W[:1]
        if i == N - 1:
            W = W[:, :, :, 2:]
        Ws.append(W)
    return Ws


def test_full_chain_equals_dense_expectation(rng):
    """Contracting the left blocks across the chain reproduces <psi|H|psi> from ED."""
    N = 6
    Ws = tfim_mpo(N, 1.0, 0.7)
    Hd = exact.dense_hamiltonian(N, [(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)] +
                                 [(["x"], [i], -0.7) for i in range(1, N + 1)])
    assert np.allclose(exact.mpo_to_dense(Ws), Hd)
    psi = random_mps(N, 2, 4, rng)
    v = to_dense(psi)
    L = edgeblock()
    for i in range(1, N + 1):
        L = buildleft(L, psi[i], Ws[i - 1])
    assert np.isclose(L.reshape(-1)[0], np.vdot(v, Hd @ v))
    # and the right blocks give the same number
    R = edgeblock()
    for i in range(N, 0, -1):
        R = buildright(R, psi[i], Ws[i - 1])
    assert np.isclose(R.reshape(-1)[0], np.vdot(v, Hd @ v))
    P = ProjMPS(psi.tensors, Ws, center=3)
    assert np.isclose(P.calculate(), np.vdot(v, Hd @ v))


# ------------------------------------------------------------------- krylov

def test_lanczos_ground_state(rng):
    n = 40
    A = randn(rng, (n, n))
    A = A + A.T
    w = np.linalg.eigvalsh(A)
    e, v, nmv = eigsolve_ground(lambda x: A @ x, randn(rng, n), krylovdim=12, maxiter=30)
    assert abs(e - w[0]) < 1e-9
    assert np.linalg.norm(A @ v - e * v) < 1e-5
    # cost model: krylovdim matvecs, then krylovdim-1 per restart
    _, _, nmv2 = eigsolve_ground(lambda x: A @ x, randn(rng, n), krylovdim=3, maxiter=2, tol=0)
    assert nmv2 == 3 + 2


def test_lanczos_invariant_subspace():
    A = np.diag([1.0, 2.0, 3.0, 4.0])
    x0 = np.array([1.0, 0, 0, 0])  # already an eigenvector: breakdown after one step
    e, v, nmv = eigsolve_ground(lambda x: A @ x, x0, krylovdim=3, maxiter=2)
    assert np.isclose(e, 1.0) and nmv == 1 and np.isclose(abs(v[0]), 1.0)


# -------------------------------------------------------------------- mps

def test_replacesites_both_directions(rng):
    th = randn(rng, (3, 2, 2, 4), True)
    for direction in (False, True):
        A1, A2, S, disc = replacesites2(th, direction, False)
        rec = np.tensordot(A1, A2, axes=([2], [0]))
        assert np.allclose(rec, th)
        if direction:  # A2 right-orthonormal
            B = A2.reshape(A2.shape[0], -1, order="F")
            assert np.allclose(B @ B.conj().T, np.eye(B.shape[0]))
        else:  # A1 left-orthonormal
            B = A1.reshape(-1, A1.shape[2], order="F")
            assert np.allclose(B.conj().T @ B, np.eye(B.shape[1]))
        assert np.allclose(np.sort(S)[::-1], np.linalg.svd(th.reshape(6, 8, order="F"), compute_uv=False))


def test_random_mps_is_right_canonical_and_normalised(rng):
    psi = random_mps(8, 2, 8, rng)
    assert psi.center == 1
    assert np.isclose(np.linalg.norm(to_dense(psi)), 1.0)
    for i in range(2, 9):
        B = psi[i].reshape(psi[i].shape[0], -1, order="F")
        assert np.allclose(B @ B.conj().T, np.eye(B.shape[0]))
    assert psi.bonddims() == [2, 4, 8, 8, 8, 4, 2]


def test_entropy_matches_dense(rng):
    psi = random_mps(6, 2, 4, rng)
    v = to_dense(psi).reshape(8, 8)
    s = np.linalg.svd(v, compute_uv=False) ** 2
    assert np.isclose(entropy(psi, 3), -(s * np.log(s)).sum())


def test_inner_oplist_matches_dense(rng):
    psi = random_mps(5, 2, 3, rng)
    v = to_dense(psi)
    terms = [([exact.OPS["z"]], [2], 1.0), ([exact.OPS["x"], exact.OPS["x"]], [3, 4], 0.5)]
    got = inner_oplist(psi, terms)
    Z2 = exact.dense_hamiltonian(5, [(["z"], [2], 1.0)])
    XX = exact.dense_hamiltonian(5, [(["x", "x"], [3, 4], 0.5)])
    assert np.allclose(got, [np.vdot(v, Z2 @ v), np.vdot(v, XX @ v)])


def test_inner_oplist_two_states_matches_dense(rng):
    """inner(st, psi, oplist, phi) with phi != psi (complex, other bond dims): both oracle
    forms against <psi|O|phi> of the dense vectors."""
    from oracle.mps import inner_oplist_env

    psi = random_mps(6, 2, 3, rng)
    phi = random_mps(6, 2, 4, rng)
    for i in range(1, 7):
        phi[i] = phi[i] + 1j * rng.standard_normal(phi[i].shape)
    vpsi, vphi = to_dense(psi), to_dense(phi)
    terms = [([exact.OPS["z"]], [1], 1.0), ([exact.OPS["x"], exact.OPS["y"]], [3, 5], 0.5),
             ([exact.OPS["x"]], [6], -2.0)]
    want = [c * np.vdot(vpsi, exact.dense_hamiltonian(6, [(ops, sites, 1.0)]) @ vphi)
            for ops, sites, c in terms]
    assert np.allclose(inner_oplist(psi, terms, phi), want, atol=1e-12)
    assert np.allclose(inner_oplist_env(psi, terms, phi), want, atol=1e-12)


# --------------------------------------------------------------- closed forms

@pytest.mark.parametrize("N", [3, 8, 10])
def test_tfim_closed_form_vs_ed(N):
    H = exact.dense_hamiltonian(N, [(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)] +
                                [(["x"], [i], -1.0) for i in range(1, N + 1)])
    e0 = np.linalg.eigvalsh(H)[0]
    assert abs(e0 - exact.tfim_critical_ground_energy(N)) < 5e-14 * N
    assert abs(e0 - exact.tfim_bdg_ground_energy(np.ones(N - 1), np.ones(N))) < 5e-14 * N


def test_tfim_bdg_inhomogeneous(rng):
    N = 8
    J, h = rng.uniform(0.8, 1.2, N - 1), rng.uniform(0.5, 1.5, N)
    H = exact.dense_hamiltonian(N, [(["z", "z"], [i, i + 1], -J[i - 1]) for i in range(1, N)] +
                                [(["x"], [i], -h[i - 1]) for i in range(1, N + 1)])
    assert abs(np.linalg.eigvalsh(H)[0] - exact.tfim_bdg_ground_energy(J, h)) < 1e-12


@pytest.mark.parametrize("N,h", [(8, 0.9), (10, 1.0), (9, 1.3)])
def test_tfim_first_excited_energy_vs_ed(N, h):
    terms = ([(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)]
             + [(["x"], [i], -h) for i in range(1, N + 1)])
    ev = np.linalg.eigvalsh(exact.dense_hamiltonian(N, terms))
    assert abs(exact.tfim_bdg_first_excited_energy(np.ones(N - 1), h * np.ones(N)) - ev[1]) < 1e-12


def test_tfim_n100_closed_form_value():
    assert abs(exact.tfim_critical_ground_energy(100) - (-126.961876739680733)) < 1e-9


def test_oracle_dmrg_matches_ed(rng):
    N = 10
    Ws = tfim_mpo(N)
    psi = random_mps(N, 2, 4, rng)
    psi, e = dmrg(psi, Ws, maxdim=32, cutoff=1e-12, maxsweeps=8, krylovdim=4)
    assert abs(e - exact.tfim_critical_ground_energy(N)) < 1e-9


@pytest.mark.parametrize("evol", ["real", "imag"])
def test_applygates_mps_matches_dense(evol):
    """The oracle's MPS applygates! (gatelist.jl:191-223) reproduces the dense Trotter
    evolution when nothing is truncated. Odd N, so every second layer ends in a one-site
    gate; x and z fields and a non-symmetric start, so a transposed gate index or a
    mirrored site order fails."""
    from oracle.mps import normalize, product_mps, to_dense
    from oracle.tebd import applygates_mps, rows_from_layers, tebd_dense, trotter_layers

    N, dt, steps = 7, 0.05, 6
    terms = ([(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)]
             + [(["x"], [i], -0.7) for i in range(1, N + 1)]
             + [(["z"], [i], -0.2) for i in range(1, N + 1)])
    layers = trotter_layers(terms, N, dt, order=2, evol=evol)
    up, dn = np.array([1.0, 0.0]), np.array([0.0, 1.0])
    states = [up, up, dn, up, dn, dn, up]
    psi = product_mps(states)
    rows = rows_from_layers(layers)
    assert any(np.ndim(g) == 2 for _, gs in rows for g in gs)
    for _ in range(steps):
        applygates_mps(psi, rows)
        normalize(psi)
    v0 = states[0]
    for s in states[1:]:
        v0 = np.kron(v0, s)
    v = tebd_dense(v0.astype(complex), layers, N, steps)
    assert np.allclose(to_dense(psi), v, atol=1e-10)
