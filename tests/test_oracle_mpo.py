"""Pins of the oracle's MPO restatements (CPU): the automaton of MPO(st, H) and its
compression against the dense Hamiltonian and the minimal bond dimensions, the products
MPO * MPS and MPO * MPO against dense matrix products, the operator-to-MPO split against
its dense input, and trotterMPO against the ordered product of the exact term exponentials."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import exact
from oracle.exact import OPS, dense_hamiltonian, mpo_to_dense
from oracle.mpo import automaton, combineidxs, compress, mpo, mpo_mpo_product, mpo_mps_product, qo_to_mpo, trotter_mpo
from oracle.mps import random_mps, to_dense


def heisenberg(N):
    t = []
    for i in range(1, N):
        t += [(["s+", "s-"], [i, i + 1], 0.5), (["s-", "s+"], [i, i + 1], 0.5), (["z", "z"], [i, i + 1], 0.25)]
    return t


def tfim(N, g=0.8):
    return [(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)] + [(["x"], [i], -g) for i in range(1, N + 1)]


def mixed(N):
    """Fields, neighbours, a gapped two-site term, a three-site term and a redundant
    duplicate (two terms that the compression merges)."""
    return (tfim(N) + [(["y", "x"], [2, 4], 0.3), (["x", "z", "x"], [1, 2, 3], -0.7),
                       (["z", "z"], [3, 4], 0.2), (["z", "z"], [3, 4], 0.1)])


def bonds(Ws):
    return [w.shape[3] for w in Ws[:-1]]


def test_combineidxs_matches_reference_semantics():
    x = np.arange(2 * 3 * 4 * 5).reshape(2, 3, 4, 5)
    y, (idxs, dims) = combineidxs(x, [1, 3])
    assert y.shape == (3, 5, 8) and dims == (2, 4)
    # the combined index runs over (i1, i3) with i1 fastest
    for a in range(2):
        for c in range(4):
            assert np.array_equal(y[:, :, a + 2 * c], x[a, :, c, :])


@pytest.mark.parametrize("name,terms,N,D", [("heisenberg", heisenberg(6), 6, 5), ("tfim", tfim(7), 7, 3),
                                            ("mixed", mixed(6), 6, None)])
def test_mpo_automaton_and_compression_equal_dense(name, terms, N, D):
    H = dense_hamiltonian(N, terms)
    A = automaton(terms, N)
    assert np.allclose(mpo_to_dense(A), H, atol=1e-12)
    C = mpo(terms, N)
    assert np.allclose(mpo_to_dense(C), H, atol=1e-11)
    if D is not None:  # the textbook automaton is already minimal: the compression keeps it
        assert max(bonds(C)) == D
    assert all(c <= a for c, a in zip(bonds(C), bonds(A)))


def test_compression_removes_redundant_channels():
    """The same two-site term listed twice with split coefficients needs one channel, not two."""
    N = 5
    terms = [(["x", "x"], [2, 3], 0.4), (["x", "x"], [2, 3], 0.6)] + [(["z"], [i], 1.0) for i in range(1, N + 1)]
    A, C = automaton(terms, N), mpo(terms, N)
    assert bonds(A)[1] == 4 and bonds(C)[1] == 3
    assert np.allclose(mpo_to_dense(C), dense_hamiltonian(N, terms), atol=1e-12)


def test_host_automaton_compresses_to_the_same_bonds():
    """The library's own automaton (tnhip.models.mpo_tensors, a different channel layout)
    has the same operator and, after the reference's compression, the same bond dimensions."""
    from tnhip.models import OpList, mpo_tensors, spinhalf

    N = 6
    terms = mixed(N)
    ol = OpList(N)
    for ops, sites, c in terms:
        ol.add(list(ops), list(sites), c)
    W = [np.asarray(w, dtype=complex) for w in mpo_tensors(spinhalf(), ol)]
    assert np.allclose(mpo_to_dense(W), dense_hamiltonian(N, terms), atol=1e-12)
    assert bonds(compress(W)) == bonds(mpo(terms, N))


@pytest.mark.parametrize("cplx", [False, True])
def test_mpo_mps_product_equals_dense(cplx):
    N = 6
    rng = np.random.default_rng(3)
    psi = random_mps(N, 2, 4, rng)
    if cplx:
        psi[2] = psi[2] * (0.3 + 0.8j)
    terms = mixed(N)
    phi = mpo_mps_product(mpo(terms, N), psi)
    assert phi.center == 1
    assert np.allclose(to_dense(phi), dense_hamiltonian(N, terms) @ to_dense(psi), atol=1e-11)
    # right-canonical after the final sweep
    for i in range(2, N + 1):
        B = phi[i].reshape(phi[i].shape[0], -1, order="F")
        assert np.allclose(B @ B.conj().T, np.eye(B.shape[0]), atol=1e-12)


def test_mpo_mps_product_truncated_is_best_approximation():
    """With maxdim the final sweep truncates a right-canonical... left-to-right canonical
    state from the right: the error is bounded by the discarded weights (Eckart-Young per
    bond) and the result has the requested bond dimension."""
    N = 6
    rng = np.random.default_rng(4)
    psi = random_mps(N, 2, 4, rng)
    Ws = mpo(heisenberg(N), N)
    exact_v = dense_hamiltonian(N, heisenberg(N)) @ to_dense(psi)
    phi = mpo_mps_product(Ws, psi, maxdim=4)
    assert max(phi.bonddims()) <= 4
    err = np.linalg.norm(to_dense(phi) - exact_v) / np.linalg.norm(exact_v)
    assert err < 0.5  # a genuine truncation, not a broken product
    full = mpo_mps_product(Ws, psi)
    assert max(full.bonddims()) > 4


def test_mpo_mpo_product_equals_dense():
    N = 5
    A, B = mpo(tfim(N), N), mpo(mixed(N), N)
    P = mpo_mpo_product(A, B)
    assert np.allclose(mpo_to_dense(P), mpo_to_dense(A) @ mpo_to_dense(B), atol=1e-10)


def test_qo_to_mpo_round_trip():
    rng = np.random.default_rng(5)
    n, d = 3, 2
    T = rng.standard_normal((d,) * (2 * n)) + 1j * rng.standard_normal((d,) * (2 * n))
    ts = qo_to_mpo(T)
    # contract the MPO back to T[o1, i1, o2, i2, o3, i3]
    R = ts[0][0]
    for w in ts[1:]:
        R = np.tensordot(R, w, axes=([R.ndim - 1], [0]))
    assert np.allclose(R[..., 0], T, atol=1e-12)


@pytest.mark.parametrize("t", [0.1, -0.05j])
def test_trotter_mpo_is_the_ordered_product_of_term_exponentials(t):
    """trotterMPO multiplies exp(t c_k O_k) onto the running MPO from the left, term by term:
    the dense result is exp(t c_n O_n) ... exp(t c_1 O_1), which differs from exp(t H) at
    O(t^2) (checked too, so that an unordered or mis-scaled product fails)."""
    N = 4
    terms = tfim(N) + [(["y", "y"], [1, 3], 0.5)]
    O = trotter_mpo(terms, N, t)
    ref = np.eye(2 ** N, dtype=complex)
    for ops, sites, c in terms:
        ref = sla.expm(t * dense_hamiltonian(N, [(ops, sites, c)])) @ ref
    assert np.allclose(mpo_to_dense(O), ref, atol=1e-11)
    H = dense_hamiltonian(N, terms)
    assert np.linalg.norm(ref - sla.expm(t * H)) > 1e-4


def test_inner_chains_match_dense():
    """inner(args::GMPS...) (mpo.jl:181-217) against dense algebra: <phi|psi>,
    <phi|O|psi>, <phi|O1 O2|psi>, and an operator before the bra, which the reference moves
    past it as its adjoint: inner(O, phi, psi) = <phi|O^H|psi>, inner(O1, O2, phi, psi) =
    <phi|O1^H O2^H|psi>. Complex states of different bond dimensions, non-Hermitian O."""
    from oracle.exact import mpo_to_dense
    from oracle.mpo import inner
    from oracle.mps import random_mps, to_dense

    rng = np.random.default_rng(41)
    N = 5

    def cmps(chi):
        p = random_mps(N, 2, chi, rng)
        return [t + 1j * rng.standard_normal(t.shape) for t in p.tensors]

    def cmpo(D):
        ts = []
        for i in range(N):
            shape = (1 if i == 0 else D, 2, 2, 1 if i == N - 1 else D)
            ts.append(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        return ts

    from oracle.mps import MPS

    phi, psi = cmps(3), cmps(4)
    O1, O2 = cmpo(2), cmpo(3)
    vphi, vpsi = to_dense(MPS(phi)), to_dense(MPS(psi))
    M1, M2 = mpo_to_dense(O1), mpo_to_dense(O2)
    assert abs(inner(phi, psi) - np.vdot(vphi, vpsi)) <= 1e-10 * abs(np.vdot(vphi, vpsi))
    ref = np.vdot(vphi, M1 @ vpsi)
    assert abs(inner(phi, O1, psi) - ref) <= 1e-10 * abs(ref)
    ref = np.vdot(vphi, M1 @ (M2 @ vpsi))
    assert abs(inner(phi, O1, O2, psi) - ref) <= 1e-10 * abs(ref)
    ref = np.vdot(vphi, M1.conj().T @ vpsi)
    assert abs(inner(O1, phi, psi) - ref) <= 1e-10 * abs(ref)
    ref = np.vdot(vphi, M1.conj().T @ (M2.conj().T @ vpsi))
    assert abs(inner(O1, O2, phi, psi) - ref) <= 1e-10 * abs(ref)
    ref = np.vdot(vphi, M1.conj().T @ (M2 @ vpsi))  # one operator on each side of the bra
    assert abs(inner(O1, phi, O2, psi) - ref) <= 1e-10 * abs(ref)
    with pytest.raises(ValueError):
        inner(phi, O1, psi, O2)  # the last argument must be the ket
    with pytest.raises(ValueError):
        inner(phi, O1)
