"""GPU parity of the MPO paths against the oracle's restatements of
src/structures/mps/mpo.jl and dense operators: MPO(st, H) (the host automaton and the device
compression sweeps), MPO * MPS and MPO * MPO products with and without truncation, and
trotterMPO. Gauges of the SVD factors differ between LAPACK and the device eigensolver, so
parity is stated on dense operators and states (exact to 1e-11 where nothing is truncated),
bond dimensions, and truncated results against the oracle's truncated result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oplist(N, terms):
    from tnhip.models import OpList

    ol = OpList(N)
    for ops, sites, c in terms:
        ol.add(list(ops), list(sites), c)
    return ol


def mixed(N):
    t = [(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)] + [(["x"], [i], -0.8) for i in range(1, N + 1)]
    return t + [(["y", "x"], [2, 4], 0.3), (["x", "z", "x"], [1, 2, 3], -0.7), (["z", "z"], [3, 4], 0.2),
                (["z", "z"], [3, 4], 0.1)]


def heisenberg(N):
    t = []
    for i in range(1, N):
        t += [(["s+", "s-"], [i, i + 1], 0.5), (["s-", "s+"], [i, i + 1], 0.5), (["z", "z"], [i, i + 1], 0.25)]
    return t


@pytest.mark.parametrize("name", ["heisenberg", "mixed"])
def test_mpo_builder_matches_oracle(ctx, name):
    """tnh_mpo_from_oplist: without compression exactly the host automaton
    (tnhip.models.mpo_tensors); with it the oracle's MPO(st, H) bond dimensions and the
    dense Hamiltonian."""
    from tnhip import spinhalf
    from tnhip.models import MPO, mpo_tensors
    from tnhip.mpo import mpo_device
    from oracle.exact import dense_hamiltonian, mpo_to_dense
    from oracle.mpo import mpo

    N = 7
    terms = heisenberg(N) if name == "heisenberg" else mixed(N)
    ol = oplist(N, terms)
    raw = [t.numpy() for t in mpo_device(spinhalf(), ol, cutoff=-1.0)]
    host = mpo_tensors(spinhalf(), ol)
    for a, b in zip(raw, host):
        assert a.shape == b.shape and np.array_equal(a, b)
    M = MPO(spinhalf(), ol)
    ref = mpo(terms, N)
    assert [w.shape[3] for w in M.tensors[:-1]] == [w.shape[3] for w in ref[:-1]]
    H = dense_hamiltonian(N, terms)
    assert np.allclose(mpo_to_dense([np.asarray(w, complex) for w in M.tensors]), H, atol=1e-11)


@pytest.mark.parametrize("cplx", [False, True])
def test_mpo_mps_product_matches_dense_and_oracle(ctx, cplx):
    from tnhip import spinhalf
    from tnhip.models import MPO
    from tnhip.mpo import MPOMPSProduct
    from oracle.exact import dense_hamiltonian
    from oracle.mpo import mpo, mpo_mps_product
    from oracle.mps import MPS, random_mps, to_dense

    N = 7
    rng = np.random.default_rng(6)
    psi = random_mps(N, 2, 6, rng)
    if cplx:
        psi[3] = psi[3] * (0.4 - 0.9j)
    terms = mixed(N)
    M = MPO(spinhalf(), oplist(N, terms))
    phi = MPOMPSProduct(M, psi)
    assert phi.center == 1
    got = to_dense(MPS(phi.to_gmps().tensors))
    exact = dense_hamiltonian(N, terms) @ to_dense(psi)
    assert np.allclose(got, exact, atol=1e-11 * np.linalg.norm(exact))
    ref = mpo_mps_product(mpo(terms, N), psi)
    assert phi.bonddims() == ref.bonddims()
    # truncated product: same bond dimensions and the same state as the oracle's
    # truncation (the kept Schmidt spaces are gauge independent)
    phi_t = MPOMPSProduct(M, psi, maxdim=5)
    ref_t = mpo_mps_product(mpo(terms, N), psi, maxdim=5)
    assert phi_t.bonddims() == ref_t.bonddims()
    a, b = to_dense(MPS(phi_t.to_gmps().tensors)), to_dense(ref_t)
    assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(b)


def test_mpo_mpo_product_and_trotter_mpo(ctx):
    import scipy.linalg as sla
    from tnhip import spinhalf
    from tnhip.models import MPO
    from tnhip.mpo import MPOMPOProduct, trotterMPO
    from oracle.exact import dense_hamiltonian, mpo_to_dense
    from oracle.mpo import trotter_mpo

    N = 5
    st = spinhalf()
    t1 = mixed(N)
    t2 = heisenberg(N)
    A, B = MPO(st, oplist(N, t1)), MPO(st, oplist(N, t2))
    P = [w.numpy() for w in MPOMPOProduct(A, B)]
    assert np.allclose(mpo_to_dense([np.asarray(w, complex) for w in P]),
                       dense_hamiltonian(N, t1) @ dense_hamiltonian(N, t2), atol=1e-10)
    terms = [(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)] + [(["x"], [i], -0.6) for i in range(1, N + 1)]
    terms += [(["y", "y"], [1, 3], 0.5)]
    for t in (0.1, -0.05j):
        O = trotterMPO(st, oplist(N, terms), t)
        ref = np.eye(2 ** N, dtype=complex)
        for ops, sites, c in terms:
            ref = sla.expm(t * dense_hamiltonian(N, [(ops, sites, c)])) @ ref
        assert np.allclose(mpo_to_dense([np.asarray(w, complex) for w in O.tensors]), ref, atol=1e-10)
        Oo = trotter_mpo(terms, N, t)
        assert [w.shape[3] for w in O.tensors[:-1]] == [w.shape[3] for w in Oo[:-1]]


def test_dmrg_with_compressed_mpo(ctx):
    """The compressed MPO (the reference's default) in the device DMRG: same ground energy as
    the closed form of the critical TFIM."""
    from tnhip import dmrg, randomMPS, spinhalf
    from tnhip.models import MPO
    from oracle.exact import tfim_critical_ground_energy

    N = 20
    terms = [(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)] + [(["x"], [i], -1.0) for i in range(1, N + 1)]
    M = MPO(spinhalf(), oplist(N, terms))
    assert max(w.shape[3] for w in M.tensors[:-1]) == 3
    psi = randomMPS(2, N, 8, np.random.default_rng(7))
    psi, e = dmrg(psi, M, maxdim=64, cutoff=1e-12, maxsweeps=12, krylovdim=4, verbose=False)
    assert abs(e - tfim_critical_ground_energy(N)) / N <= 1e-10


def test_inner_chains_match_oracle_and_dense(ctx):
    """inner(args::GMPS...) (mpo.jl:181-217) on the device (tnh_inner_mpo, with several MPOs
    folded by exact MPOMPOProduct): <phi|psi>, <phi|O|psi>, <phi|O1 O2|psi>, and operators
    before the bra (their adjoints), against the oracle's restatement and dense algebra.
    Complex states of different bond dimensions, non-Hermitian complex MPOs; 1e-11 relative."""
    from tnhip import GMPS
    from tnhip.mpo import dot, inner
    from oracle.exact import mpo_to_dense
    from oracle.mpo import inner as oinner
    from oracle.mps import MPS, random_mps, to_dense

    rng = np.random.default_rng(43)
    N = 6

    def cmps(chi):
        p = random_mps(N, 2, chi, rng)
        return [np.asfortranarray(t + 1j * rng.standard_normal(t.shape)) for t in p.tensors]

    def cmpo(D):
        out = []
        for i in range(N):
            shape = (1 if i == 0 else D, 2, 2, 1 if i == N - 1 else D)
            out.append(np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)))
        return out

    phi, psi = cmps(3), cmps(5)
    O1, O2 = cmpo(2), cmpo(3)
    P, Q = GMPS(1, 2, phi, 0), GMPS(1, 2, psi, 0)
    A, B = GMPS(2, 2, O1, 0), GMPS(2, 2, O2, 0)
    vphi, vpsi = to_dense(MPS(phi)), to_dense(MPS(psi))
    M1, M2 = mpo_to_dense(O1), mpo_to_dense(O2)
    cases = [((P, Q), (phi, psi), np.vdot(vphi, vpsi)),
             ((P, A, Q), (phi, O1, psi), np.vdot(vphi, M1 @ vpsi)),
             ((P, A, B, Q), (phi, O1, O2, psi), np.vdot(vphi, M1 @ (M2 @ vpsi))),
             ((A, P, Q), (O1, phi, psi), np.vdot(vphi, M1.conj().T @ vpsi)),
             ((A, B, P, Q), (O1, O2, phi, psi), np.vdot(vphi, M1.conj().T @ (M2.conj().T @ vpsi)))]
    for gargs, oargs, want in cases:
        got = inner(*gargs)
        assert abs(got - want) <= 1e-11 * abs(want), (len(gargs), got, want)
        assert abs(got - oinner(*oargs)) <= 1e-11 * abs(want)
    assert abs(dot(P, Q) - np.vdot(vphi, vpsi)) <= 1e-11 * abs(np.vdot(vphi, vpsi))
