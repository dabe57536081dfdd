"""Pins of oracle functions that the other CPU tests leave open: the tensor exponential
and trace of src/tensors.jl, the second-order Trotter split of gatelist.jl:75-121 (by the
dt^3 scaling of its one-step error against the exact propagator), and the per-cycle Ritz
values of the restarted Lanczos that stands in for KrylovKit's eigsolve (each cycle is the
Rayleigh-Ritz problem of a Krylov space the mathematics fixes)."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import exact
from oracle.krylov import eigsolve_ground
from oracle.tebd import apply_dense, gate_from_tensor, trotter_layers
from oracle.tensors import expm_tensor, trace


def randn(rng, shape, cplx=False):
    a = rng.standard_normal(shape)
    return a + 1j * rng.standard_normal(shape) if cplx else a


def herm(rng, n, cplx=True):
    X = randn(rng, (n, n), cplx)
    return 0.5 * (X + X.conj().T)


# ----------------------------------------------------------------- tensors.jl


def test_expm_tensor_of_a_product_gate_is_the_kron_exponential(rng):
    """G[o1,i1,o2,i2] = a[o1,i1] b[o2,i2]: with outer indices [2, 4] the matrix is
    G[(o1,o2),(i1,i2)] (column-major, o1 fastest) = kron(b, a), whose exponential is taken
    here through its eigendecomposition (an independent route from scipy's Pade)."""
    a, b = herm(rng, 2), herm(rng, 3)
    G = np.einsum("ij,kl->ijkl", a, b)
    E = expm_tensor(G, [2, 4])
    w, V = np.linalg.eigh(np.kron(b, a))
    M = (V * np.exp(w)) @ V.conj().T
    Mt = np.transpose(E, (0, 2, 1, 3)).reshape(6, 6, order="F")
    assert np.allclose(Mt, M, atol=1e-13)


def test_expm_tensor_inverse_and_other_implementation(rng):
    """exp(G) exp(-G) = 1 contracted over the inner/outer pairs, and agreement with the
    kron-order gate builder of the TEBD reference (a transposed index fails one of them)."""
    d = 2
    G = randn(rng, (d, d, d, d), True)
    Ep, Em = expm_tensor(G, [2, 4]), expm_tensor(-G, [2, 4])
    prod = np.einsum("aibj,icjd->abcd", Ep, Em)  # [o1,o2,i1,i2]
    assert np.allclose(prod, np.einsum("ac,bd->abcd", np.eye(d), np.eye(d)), atol=1e-12)
    U = gate_from_tensor(G, 1.0)  # rows (o1, o2) with o1 slowest
    assert np.allclose(np.transpose(Ep, (0, 2, 1, 3)).reshape(d * d, d * d), U, atol=1e-12)


def test_expm_tensor_diagonal_is_elementwise():
    diag = np.array([0.3, -1.2, 2.0, 0.0])
    M = np.diag(diag)  # [(o1,o2),(i1,i2)] column-major
    G = np.transpose(M.reshape(2, 2, 2, 2, order="F"), (0, 2, 1, 3))
    E = expm_tensor(G, [2, 4])
    Em = np.transpose(E, (0, 2, 1, 3)).reshape(4, 4, order="F")
    assert np.allclose(Em, np.diag(np.exp(diag)), atol=1e-14)


def test_host_gate_exponential_matches_oracle(rng):
    """The host-side exp of tnhip.tebd (used to build Trotter gates) against the oracle."""
    from tnhip.tebd import expm_tensor as host_expm

    G = randn(rng, (2, 2, 2, 2), True)
    assert np.allclose(host_expm(0.1 * G, [2, 4]), expm_tensor(0.1 * G, [2, 4]), atol=1e-14)
    g1 = randn(rng, (3, 3), True)
    assert np.allclose(host_expm(g1, [2]), expm_tensor(g1, [2]), atol=1e-13)


def test_trace_pairs_the_given_indices(rng):
    x = randn(rng, (3, 4, 3, 2), True)
    t = trace(x, 1, 3)
    ref = sum(x[i, :, i, :] for i in range(3))
    assert t.shape == (4, 2) and np.allclose(t, ref)
    y = randn(rng, (2, 5, 5))
    assert np.allclose(trace(y, 2, 3), [sum(y[a, i, i] for i in range(5)) for a in range(2)])


# ----------------------------------------------------------- Trotter splitting


def _layers_operator(layers, N, d=2):
    dim = d ** N
    cols = []
    for c in range(dim):
        v = np.zeros(dim, dtype=complex)
        v[c] = 1.0
        for layer in layers:
            for site, U in layer:
                v = apply_dense(v, site, U, N)
        cols.append(v)
    return np.stack(cols, axis=1)


def _chain_terms(N):
    return ([(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)]
            + [(["x"], [i], -0.8) for i in range(1, N + 1)]
            + [(["z"], [i], -0.3 * (-1) ** i) for i in range(1, N + 1)])


@pytest.mark.parametrize("evol", ["imag", "real"])
def test_second_order_trotter_step_error_scales_as_dt_cubed(evol):
    """One step of the layers (odd dt/2, even dt, odd dt/2) against exp(tau H) with
    tau = dt ("imag") or -i dt ("real"): the local error of a symmetric split is O(dt^3), so
    halving dt divides it by 8; a first-order split (or a mis-weighted symmetric one)
    divides it by 4 at most. Odd N, so every second layer ends in a one-site gate."""
    N = 5
    terms = _chain_terms(N)
    H = exact.dense_hamiltonian(N, terms)

    def err(dt, order):
        layers = trotter_layers(terms, N, dt, order=order, evol=evol)
        tau = dt if evol == "imag" else -1j * dt
        return np.linalg.norm(_layers_operator(layers, N) - sla.expm(tau * H), 2)

    r2 = err(0.02, 2) / err(0.01, 2)
    r1 = err(0.02, 1) / err(0.01, 1)
    assert 7.6 < r2 < 8.4, r2
    assert 3.6 < r1 < 4.4, r1
    # and the symmetric split is the far more accurate one at the same dt
    assert err(0.01, 2) < 0.2 * err(0.01, 1)


def test_trotter_layer_structure():
    """Second order: range-2 terms give layers (odd, even, odd) with times dt/2, dt, dt/2,
    the last layer being the same gates as the first (gatelist.jl:112-117)."""
    N = 6
    terms = _chain_terms(N)
    dt = 0.1
    layers = trotter_layers(terms, N, dt, order=2, evol="imag")
    assert len(layers) == 3
    # the even layer ends in a one-site block at site 6 (the range-2 block is cut there)
    assert [s for s, _ in layers[0]] == [1, 3, 5] and [s for s, _ in layers[1]] == [2, 4, 6]
    assert layers[1][-1][1].shape == (2, 2)
    assert layers[2] is layers[0]
    full = trotter_layers(terms, N, 2 * dt, order=1, evol="imag")
    # the odd layer at dt/2 squared is the odd layer of a first-order step at dt
    for (s1, U1), (s2, U2) in zip(layers[0], trotter_layers(terms, N, dt, order=1, evol="imag")[0]):
        assert s1 == s2 and np.allclose(U1 @ U1, U2, atol=1e-13)
    assert len(full) == 2


def _range3_terms(N):
    """Nearest and next-nearest zz, a three-site z x z term and a field: trotterize's site
    range is 3, so its gates act on three sites (fewer at the chain's end)."""
    return ([(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)]
            + [(["z", "z"], [i, i + 2], -0.45) for i in range(1, N - 1)]
            + [(["z", "x", "z"], [i, i + 1, i + 2], 0.3) for i in range(1, N - 1)]
            + [(["x"], [i], -0.8) for i in range(1, N + 1)])


@pytest.mark.parametrize("evol", ["imag", "real"])
def test_second_order_trotter_range3_error_scales_as_dt_cubed(evol):
    """The same order check for a range-3 list: layers (1, 2 at dt/2; 3 at dt; 2, 1 at
    dt/2) form a symmetric split, local error O(dt^3)."""
    N = 6
    terms = _range3_terms(N)
    H = exact.dense_hamiltonian(N, terms)

    def err(dt, order):
        layers = trotter_layers(terms, N, dt, order=order, evol=evol)
        tau = dt if evol == "imag" else -1j * dt
        return np.linalg.norm(_layers_operator(layers, N) - sla.expm(tau * H), 2)

    r2 = err(0.02, 2) / err(0.01, 2)
    r1 = err(0.02, 1) / err(0.01, 1)
    assert 7.6 < r2 < 8.4, r2
    assert 3.6 < r1 < 4.4, r1


def test_trotter_layer_structure_range3():
    N = 7
    layers = trotter_layers(_range3_terms(N), N, 0.1, order=2, evol="imag")
    assert len(layers) == 5
    assert [[s for s, _ in L] for L in layers[:3]] == [[1, 4, 7], [2, 5], [3, 6]]
    assert [U.shape[0] for _, U in layers[0]] == [8, 8, 2]  # site 7: a one-site block
    assert [U.shape[0] for _, U in layers[2]] == [8, 4]     # site 6: cut to two sites
    assert layers[3] is layers[1] and layers[4] is layers[0]


@pytest.mark.parametrize("direction", [False, True])
def test_replacesites_n_untruncated(rng, direction):
    """The r-site branch of replacesites! (gmps.jl:215-261): the sites reproduce theta,
    sweeping right every site but the last is left-orthonormal, sweeping left every site but
    the first right-orthonormal; bonds are the Schmidt ranks of the cuts."""
    from oracle.mps import replacesites_n

    shape = (3, 2, 2, 2, 5)
    theta = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    sites, disc = replacesites_n(theta, direction)
    assert disc < 1e-28
    out = sites[0]
    for t in sites[1:]:
        out = np.tensordot(out, t, axes=([out.ndim - 1], [0]))
    assert np.allclose(out, theta, atol=1e-12)
    for i, t in enumerate(sites):
        cl, d, cr = t.shape
        if not direction and i < len(sites) - 1:
            M = t.reshape(cl * d, cr, order="F")
            assert np.allclose(M.conj().T @ M, np.eye(cr), atol=1e-12)
        if direction and i > 0:
            M = t.reshape(cl, d * cr, order="F")
            assert np.allclose(M @ M.conj().T, np.eye(cl), atol=1e-12)
    for j in range(1, 4):  # Schmidt rank of each cut
        M = theta.reshape(int(np.prod(shape[: j + 1])), -1, order="F")
        assert sites[j - 1].shape[2] == np.linalg.matrix_rank(M)


def test_mps_range3_gates_match_dense():
    """applygates! of range-3 Trotter rows on an MPS (gatelist.jl:137-223 with r-site gates)
    reproduces the dense Trotter evolution when nothing is truncated."""
    from oracle.mps import normalize, product_mps, to_dense
    from oracle.tebd import applygates_mps, rows_from_layers, tebd_dense

    N, dt, steps = 7, 0.05, 5
    terms = _range3_terms(N)
    layers = trotter_layers(terms, N, dt, order=2, evol="real")
    up, dn = np.array([1.0, 0.0]), np.array([0.0, 1.0])
    states = [up, dn, dn, up, dn, up, up]
    psi = product_mps(states)
    rows = rows_from_layers(layers)
    assert any(np.ndim(g) == 6 for _, gs in rows for g in gs)
    for _ in range(steps):
        applygates_mps(psi, rows)
        normalize(psi)
    v0 = states[0]
    for s in states[1:]:
        v0 = np.kron(v0, s)
    v = tebd_dense(v0.astype(complex), layers, N, steps)
    assert np.allclose(to_dense(psi), v, atol=1e-11)


# ------------------------------------------------------------------- Krylov


def _rayleigh_ritz_lowest(A, x, k):
    """Lowest Ritz pair of A on K_k(A, x), from an orthonormal basis of the monomial
    Krylov vectors (QR), independent of the Lanczos recurrences."""
    K = [x / np.linalg.norm(x)]
    for _ in range(k - 1):
        w = A @ K[-1]
        K.append(w / np.linalg.norm(w))
    Q, _ = np.linalg.qr(np.stack(K, axis=1))
    w, V = np.linalg.eigh(Q.conj().T @ A @ Q)
    return w[0], Q @ V[:, 0]


@pytest.mark.parametrize("cplx", [False, True])
def test_lanczos_cycles_are_rayleigh_ritz_on_fixed_krylov_spaces(rng, cplx):
    """Cycle 1 is Rayleigh-Ritz on K_k(A, x0). A restart keeps the Ritz vector y and the
    residual direction; since A y = theta y + b v_k, the next cycle's space is K_k(A, y).
    So the Ritz value after c cycles is Rayleigh-Ritz on K_k(A, y_{c-1}), and the matvec
    count is k + (c - 1)(k - 1)."""
    n, k = 30, 4
    A = herm(rng, n, cplx)
    x0 = randn(rng, n, cplx)
    y = x0
    for cycles in (1, 2, 3):
        e, v, nmv = eigsolve_ground(lambda x: A @ x, x0, krylovdim=k, maxiter=cycles, tol=0.0)
        e_rr, v_rr = _rayleigh_ritz_lowest(A, y, k)
        assert nmv == k + (cycles - 1) * (k - 1)
        assert abs(e - e_rr) <= 1e-11 * np.abs(np.linalg.eigvalsh(A)).max(), (cycles, e, e_rr)
        assert abs(abs(np.vdot(v, v_rr)) - 1) <= 1e-9
        y = v
    # the sequence decreases monotonically towards the ground state
    es = [eigsolve_ground(lambda x: A @ x, x0, k, c, 0.0)[0] for c in (1, 2, 3, 6)]
    assert all(a >= b - 1e-12 for a, b in zip(es, es[1:]))


# ----------------------------------------------------------- Arnoldi (ishermitian = false)


def _rayleigh_ritz_smallest_real(A, x, k):
    """Ritz pair of smallest real part of A on K_k(A, x), from a QR basis of the monomial
    Krylov vectors, independent of the Arnoldi recurrences."""
    K = [x / np.linalg.norm(x)]
    for _ in range(k - 1):
        w = A @ K[-1]
        K.append(w / np.linalg.norm(w))
    Q, _ = np.linalg.qr(np.stack(K, axis=1))
    w, V = np.linalg.eig(Q.conj().T @ A @ Q)
    i = int(np.argmin(w.real))
    v = Q @ V[:, i]
    return w[i], v / np.linalg.norm(v)


def test_arnoldi_cycles_are_rayleigh_ritz_on_fixed_krylov_spaces(rng):
    """As for the Lanczos: cycle 1 is Rayleigh-Ritz on K_k(A, x0) for a non-normal complex A,
    and after a restart on K_k(A, y) of the kept Ritz vector y (A y = lambda y + b v_k)."""
    from oracle.krylov import eigsolve_arnoldi

    n, k = 30, 5
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A += np.diag(np.linspace(-2.0, 2.0, n))
    x0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = x0
    for cycles in (1, 2, 3):
        e, v, nmv = eigsolve_arnoldi(lambda x: A @ x, x0, krylovdim=k, maxiter=cycles, tol=0.0)
        e_rr, v_rr = _rayleigh_ritz_smallest_real(A, y, k)
        assert nmv == k + (cycles - 1) * (k - 1)
        assert abs(e - e_rr) <= 1e-10 * np.abs(A).max(), (cycles, e, e_rr)
        assert abs(abs(np.vdot(v, v_rr)) - 1) <= 1e-8
        y = v


def test_arnoldi_converges_to_dense_smallest_real_part(rng):
    from oracle.krylov import eigsolve_arnoldi

    n = 40
    A = rng.standard_normal((n, n)) / np.sqrt(n) + np.diag(np.linspace(-3.0, 3.0, n))
    A[0, 0] = -6.0  # a well separated smallest real part
    x0 = rng.standard_normal(n)
    e, v, _ = eigsolve_arnoldi(lambda x: A @ x, x0, krylovdim=12, maxiter=40, tol=1e-13)
    ev = np.linalg.eigvals(A)
    want = ev[np.argmin(ev.real)]
    assert abs(e - want) <= 1e-10
    assert np.linalg.norm(A @ v - e * v) <= 1e-9


def hatano_nelson_terms(N, g, delta):
    """Asymmetric hopping e^{+g} / e^{-g} plus zz: non-Hermitian, but with open boundaries
    similar to the Hermitian chain of g = 0, so its spectrum is real and known from ED."""
    return ([(["s+", "s-"], [i, i + 1], -np.exp(g)) for i in range(1, N)]
            + [(["s-", "s+"], [i, i + 1], -np.exp(-g)) for i in range(1, N)]
            + [(["z", "z"], [i, i + 1], delta) for i in range(1, N)])


def test_oracle_nonhermitian_dmrg_reaches_the_similar_hermitian_spectrum():
    """dmrg(...; ishermitian=false) on the Hatano-Nelson chain (N = 8, bonds that hold the
    exact state): the cost is the eigenvalue of smallest real part, which equals the lowest
    eigenvalue of the Hermitian chain it is similar to."""
    from oracle.dmrg import dmrg
    from oracle.mpo import mpo
    from oracle.mps import random_mps

    N, g, delta = 8, 0.25, 0.5
    Hd = exact.dense_hamiltonian(N, hatano_nelson_terms(N, g, delta))
    assert np.abs(Hd - Hd.conj().T).max() > 0.1  # genuinely non-Hermitian
    ev = np.linalg.eigvals(Hd)
    want = ev[np.argmin(ev.real)]
    e_h = np.linalg.eigvalsh(exact.dense_hamiltonian(N, hatano_nelson_terms(N, 0.0, delta)))[0]
    assert abs(want.real - e_h) < 1e-10 and abs(want.imag) < 1e-10
    Ws = mpo(hatano_nelson_terms(N, g, delta), N)
    psi, e = dmrg(random_mps(N, 2, 16, np.random.default_rng(4)), Ws, maxsweeps=10,
                  krylovdim=6, maxdim=16, cutoff=1e-14, ishermitian=False)
    assert abs(e - want) < 1e-9
