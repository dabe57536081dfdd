"""Pins of the oracle's rank-1 ("squared") Hamiltonian terms, ProjMPS(phi, psi; rank=2,
squared=true) of src/structures/mps/projmps.jl, against dense linear algebra.

* calculate() at every centre against <phi|psi> of the dense vectors;
* product() on the two-site window against B^H (w |phi><phi|) B, B the dense isometry that
  embeds the window's tensor into the full state (psi in mixed-canonical form there);
  complex data, so a missing conjugate or a transposed block fails;
* the DMRG driver with the penalty term against exact diagonalisation of H + w |phi><phi|
  (the first excited state of H when phi is its ground state).
"""
import numpy as np
import pytest

from oracle import exact
from oracle.dmrg import dmrg
from oracle.mps import movecenter, random_mps, to_dense
from oracle.projmps import ProjMPSSquared, ProjMPSSum, ProjMPS


def _random_mps(N, d, chi, rng, cplx):
    psi = random_mps(N, d, chi, rng)
    if cplx:
        for i in range(1, N + 1):
            t = psi[i]
            psi[i] = t + 1j * rng.standard_normal(t.shape)
        psi.center = 0  # no longer canonical
    return psi


def _window_dense(psi, site1, theta):
    """Dense state of psi with sites (site1, site1 + 1) replaced by the two-site tensor
    theta, in to_dense's order (site 1 slowest)."""
    N = len(psi)
    left = np.ones((1, 1))
    for i in range(1, site1):
        t = psi[i]
        left = np.einsum("Ia,asb->Isb", left, t).reshape(-1, t.shape[2])
    right = np.ones((1, 1))
    for i in range(N, site1 + 1, -1):
        t = psi[i]
        right = np.einsum("asb,bJ->asJ", t, right).reshape(t.shape[0], -1)
    v = np.einsum("Ia,asrb,bJ->IsrJ", left, theta, right)
    return v.reshape(-1)


@pytest.mark.parametrize("cplx", [False, True])
def test_overlap_blocks_calculate_every_centre(cplx):
    rng = np.random.default_rng(11)
    N, w = 7, 0.37
    phi = _random_mps(N, 2, 3, rng, cplx)
    psi = _random_mps(N, 2, 4, rng, cplx)
    ref = w * np.vdot(to_dense(phi), to_dense(psi))
    P = ProjMPSSquared(phi.tensors, psi.tensors, center=N, coeff=w)
    for c in list(range(N, 0, -1)) + list(range(2, N + 1)):
        P.movecenter(c)
        assert abs(P.calculate() - ref) < 1e-12 * max(1.0, abs(ref))


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("site1", [1, 3, 5])
def test_product_squared_is_projected_rank_one_operator(cplx, site1):
    rng = np.random.default_rng(5 + site1)
    N, w = 6, 1.7
    phi = _random_mps(N, 2, 3, rng, cplx)
    psi = _random_mps(N, 2, 3, rng, cplx)
    movecenter(psi, site1)  # sites < site1 left-, > site1 right-canonical
    theta = np.tensordot(psi[site1], psi[site1 + 1], axes=([2], [0]))
    shape = theta.shape
    n = theta.size
    B = np.zeros((2 ** N, n), dtype=complex)
    for k, idx in enumerate(np.ndindex(shape)):
        E = np.zeros(shape)
        E[idx] = 1.0
        B[:, k] = _window_dense(psi, site1, E)
    assert np.allclose(B.conj().T @ B, np.eye(n), atol=1e-12)  # the window embeds isometrically
    phid = to_dense(phi)
    Heff = w * (B.conj().T @ phid)[:, None] * (phid.conj() @ B)[None, :]
    proj = ProjMPSSquared(phi.tensors, psi.tensors, center=site1, coeff=w)
    A = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    got = proj.product(A, False, 2)
    want = (Heff @ A.reshape(-1)).reshape(shape)
    assert got.shape == shape
    assert np.allclose(got, want, atol=1e-12)
    # the same window reached from the right (direction = true, centre site1 + 1)
    proj.movecenter(site1 + 1)
    assert np.allclose(proj.product(A, True, 2), want, atol=1e-12)


def _tfim(N, h):
    from test_oracle import tfim_mpo

    Ws = tfim_mpo(N, 1.0, h)
    Hd = exact.dense_hamiltonian(N, [(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)] +
                                 [(["x"], [i], -h) for i in range(1, N + 1)])
    return Ws, Hd


def test_dmrg_penalty_term_finds_first_excited_state():
    """dmrg(psi, H, phi0; coeffs=[1, w]) minimises H + w |phi0><phi0|: with phi0 the
    ground state it lands on the first excited state (dmrg.jl:128-154)."""
    N, h, w = 8, 0.9, 10.0
    Ws, Hd = _tfim(N, h)
    ev = np.linalg.eigvalsh(Hd)
    rng = np.random.default_rng(3)
    phi0, e0 = dmrg(random_mps(N, 2, 4, rng), Ws, maxdim=16, cutoff=1e-14, maxsweeps=8,
                    krylovdim=4)
    assert abs(e0 - ev[0]) < 1e-10
    phid = to_dense(phi0)
    Hp = Hd + w * np.outer(phid, phid.conj())
    target = np.linalg.eigvalsh(Hp)[0]
    psi = random_mps(N, 2, 4, rng)
    psi, e1 = dmrg(psi, Ws, maxdim=16, cutoff=1e-14, maxsweeps=12, krylovdim=4,
                   terms=[(phi0.tensors, w)])
    assert abs(e1 - target) < 1e-9
    assert abs(e1 - ev[1]) < 1e-8
    v = to_dense(psi)
    assert abs(np.vdot(phid, v)) < 1e-5
    assert abs(np.vdot(v, Hd @ v) / np.vdot(v, v) - ev[1]) < 1e-8


def test_projmps_sum_adds_terms():
    rng = np.random.default_rng(8)
    N = 5
    Ws, Hd = _tfim(N, 0.6)
    phi = _random_mps(N, 2, 2, rng, True)
    psi = _random_mps(N, 2, 3, rng, True)
    movecenter(psi, 2)
    a = ProjMPS(psi.tensors, Ws, center=2, coeff=0.5)
    b = ProjMPSSquared(phi.tensors, psi.tensors, center=2, coeff=-2.0)
    s = ProjMPSSum([a, b])
    theta = np.tensordot(psi[2], psi[3], axes=([2], [0]))
    assert np.allclose(s.product(theta), a.product(theta) + b.product(theta))
    psid, phid = to_dense(psi), to_dense(phi)
    want = 0.5 * np.vdot(psid, Hd @ psid) - 2.0 * np.vdot(phid, psid)
    assert abs(s.calculate() - want) < 1e-10


# ------------------------------------------------------------ one-site update (nsites = 1)


def _site_window_dense(psi, site, A):
    """Dense state of psi with site `site` replaced by the one-site tensor A."""
    N = len(psi)
    left = np.ones((1, 1))
    for i in range(1, site):
        t = psi[i]
        left = np.einsum("Ia,asb->Isb", left, t).reshape(-1, t.shape[2])
    right = np.ones((1, 1))
    for i in range(N, site, -1):
        t = psi[i]
        right = np.einsum("asb,bJ->asJ", t, right).reshape(t.shape[0], -1)
    return np.einsum("Ia,asb,bJ->IsJ", left, A, right).reshape(-1)


@pytest.mark.parametrize("site", [1, 3, 6])
def test_one_site_products_are_projected_operators(site):
    """product(projV, A, direction, nsites=1) of an MPO term and of a rank-1 term
    (projmps.jl:103-179) against B^H O B for the one-site window's dense isometry B."""
    rng = np.random.default_rng(30 + site)
    N, w = 6, 0.9
    Ws, Hd = _tfim(N, 0.7)
    psi = _random_mps(N, 2, 3, rng, True)
    phi = _random_mps(N, 2, 2, rng, True)
    movecenter(psi, site)
    shape = psi[site].shape
    B = np.zeros((2 ** N, int(np.prod(shape))), dtype=complex)
    for k, idx in enumerate(np.ndindex(shape)):
        E = np.zeros(shape)
        E[idx] = 1.0
        B[:, k] = _site_window_dense(psi, site, E)
    assert np.allclose(B.conj().T @ B, np.eye(B.shape[1]), atol=1e-12)
    A = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    H = ProjMPS(psi.tensors, Ws, center=site, coeff=0.5)
    want = 0.5 * (B.conj().T @ Hd @ B) @ A.reshape(-1)
    assert np.allclose(H.product(A, False, 1), want.reshape(shape), atol=1e-12)
    assert np.allclose(H.product(A, True, 1), want.reshape(shape), atol=1e-12)
    P = ProjMPSSquared(phi.tensors, psi.tensors, center=site, coeff=w)
    phid = to_dense(phi)
    want = w * (B.conj().T @ phid) * (phid.conj() @ B @ A.reshape(-1))
    assert np.allclose(P.product(A, False, 1), want.reshape(shape), atol=1e-12)


def test_one_site_dmrg_reaches_ed_and_excited_state():
    """dmrg(...; nsites=1) (dmrg.jl:35-63 with replacesites!'s one-site branch) from a
    start whose bonds already hold the exact state (N = 8, bond dimension 16) reaches the
    ED ground energy, and with the penalty term the first excited state."""
    N, h, w = 8, 0.9, 10.0
    Ws, Hd = _tfim(N, h)
    ev = np.linalg.eigvalsh(Hd)
    rng = np.random.default_rng(17)
    phi0, e0 = dmrg(random_mps(N, 2, 16, rng), Ws, nsites=1, maxsweeps=10, krylovdim=4)
    assert abs(e0 - ev[0]) < 1e-10
    assert [t.shape[2] for t in phi0.tensors[:-1]] == [2, 4, 8, 16, 8, 4, 2]
    psi, e1 = dmrg(random_mps(N, 2, 16, rng), Ws, nsites=1, maxsweeps=14, krylovdim=4,
                   terms=[(phi0.tensors, w)])
    assert abs(e1 - ev[1]) < 1e-9
