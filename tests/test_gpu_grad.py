"""GPU gradient of E(c) = <psi|H(c)|psi> with respect to the MPO tensors and the term
coefficients, checked against dense expectation values (E and <O_t> exactly), central
finite differences of the device energy (step 1e-5), the device Hellmann-Feynman values
<O_t> from inner(), and, at a converged ground state of the inhomogeneous TFIM, the
derivatives of the free-fermion ground energy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tfim(N, J, h, extra_y=None):
    from tnhip import OpList

    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -J[i - 1])
    for i in range(1, N + 1):
        H.add("x", i, -h[i - 1])
    if extra_y is not None:
        H.add(["y", "x"], [2, 4], extra_y)  # a range-3 term with a gap and a complex operator
    return H


def terms(ol):
    return [(list(o), list(s), c) for o, s, c in zip(ol.ops, ol.sites, ol.coeffs)]


def complex_mps(N, chi, rng):
    from oracle.mps import MPS, movecenter, normalize

    ts = []
    for i in range(N):
        a = 1 if i == 0 else chi
        b = 1 if i == N - 1 else chi
        ts.append(rng.standard_normal((a, 2, b)) + 1j * rng.standard_normal((a, 2, b)))
    psi = MPS(ts, 0)
    movecenter(psi, N)
    movecenter(psi, 1)
    normalize(psi)
    return psi


@pytest.mark.parametrize("block", [1, 3, 8, 100])
def test_grad_matches_dense(ctx, rng, block):
    from tnhip import spinhalf
    from tnhip.models import mpo_tensors
    from tnhip.grad import coupling_gradient, mpo_expectation
    from tnhip.tebd import DeviceMPS
    from oracle.mps import to_dense
    from oracle.exact import dense_hamiltonian

    N = 8
    r = np.random.default_rng(3)
    H = tfim(N, r.uniform(0.8, 1.2, N - 1), r.uniform(0.5, 1.5, N), extra_y=0.3)
    psi = complex_mps(N, 4, rng)
    v = to_dense(psi)
    st = spinhalf()
    d = DeviceMPS(psi.tensors, psi.center)
    E, g = coupling_gradient(st, d, H, block=block)
    Hd = dense_hamiltonian(N, terms(H))
    assert abs(E - np.vdot(v, Hd @ v)) < 1e-12
    ref = [np.vdot(v, dense_hamiltonian(N, [(o, s, 1.0)]) @ v) for o, s, _ in terms(H)]
    assert np.allclose(g, ref, atol=1e-12)
    # E is linear in each W_i: E = sum W_i . dE/dW_i for every site
    W = mpo_tensors(st, H)
    E2, G = mpo_expectation(d, W, block=block, grad=True)
    for i in range(N):
        assert abs(np.sum(W[i] * G[i]) - E2) < 1e-12, i


def test_grad_finite_differences_and_hellmann_feynman(ctx, rng):
    """N = 64, random right-canonical MPS (chi 32): device gradient vs central differences
    of the device energy (E is linear in c, so only rounding, ~eps |E| / step, remains) and
    vs inner()'s <O_t>."""
    from tnhip import spinhalf
    from tnhip.grad import coupling_gradient
    from tnhip.tebd import DeviceMPS, inner
    from oracle.mps import random_mps

    N = 64
    r = np.random.default_rng(1)
    J, h = r.uniform(0.8, 1.2, N - 1), r.uniform(0.5, 1.5, N)
    H = tfim(N, J, h)
    st = spinhalf()
    psi = random_mps(N, 2, 32, rng)
    d = DeviceMPS(psi.tensors, psi.center)
    E, g = coupling_gradient(st, d, H, block=8)
    hf = inner(st, d, H)  # coeff_t * <O_t>
    coeffs = np.array(H.coeffs)
    assert np.allclose(g * coeffs, hf, atol=1e-11)
    step = 1e-5
    for t in [0, 17, N - 2, N - 1 + 5, len(H.coeffs) - 1]:
        Hp, Hm = H * 1.0, H * 1.0
        Hp.coeffs[t] += step
        Hm.coeffs[t] -= step
        Ep, _ = coupling_gradient(st, d, Hp, block=8)
        Em, _ = coupling_gradient(st, d, Hm, block=8)
        fd = (Ep - Em).real / (2 * step)
        # two separate sweeps round differently at ~1e-13 |E|: 1e-13 * 64 / 2e-5 ~ 3e-7 worst
        assert abs(fd - g[t].real) < 1e-6, (t, fd, g[t])


def test_grad_at_ground_state_matches_free_fermions(ctx):
    """dE0/dJ_i = -<zz>, dE0/dh_i = -<x> at the ground state (Hellmann-Feynman): the
    device gradient at a converged DMRG state vs central differences of the exact BdG
    ground energy of the inhomogeneous open TFIM."""
    from tnhip import spinhalf, MPO, randomMPS
    from tnhip.mps import dmrg
    from tnhip.grad import coupling_gradient
    from tnhip.tebd import DeviceMPS
    from oracle.exact import tfim_bdg_ground_energy

    N = 16
    r = np.random.default_rng(1)
    J, h = r.uniform(0.8, 1.2, N - 1), r.uniform(0.5, 1.5, N)
    st = spinhalf()
    H = tfim(N, J, h)
    psi = randomMPS(2, N, 8, rng=np.random.default_rng(4))
    psi, E = dmrg(psi, MPO(st, H), maxsweeps=10, cutoff=1e-14, maxdim=64, verbose=False)
    e0 = tfim_bdg_ground_energy(J, h)
    assert abs(E - e0) < 1e-9 * N
    d = DeviceMPS.from_gmps(psi)
    Eg, g = coupling_gradient(st, d, H, block=8)
    assert abs(Eg.real - e0) < 1e-9 * N
    dJ = -g[: N - 1].real  # coefficient of term i is -J_i
    dh = -g[N - 1:].real
    # <O> at an approximate eigenstate errs linearly in the state error, ~sqrt(dE / gap)
    eps = 1e-5
    for i in range(N - 1):
        Jp, Jm = J.copy(), J.copy()
        Jp[i] += eps
        Jm[i] -= eps
        ref = (tfim_bdg_ground_energy(Jp, h) - tfim_bdg_ground_energy(Jm, h)) / (2 * eps)
        assert abs(dJ[i] - ref) < 1e-5, (i, dJ[i], ref)
    for i in range(N):
        hp, hm = h.copy(), h.copy()
        hp[i] += eps
        hm[i] -= eps
        ref = (tfim_bdg_ground_energy(J, hp) - tfim_bdg_ground_energy(J, hm)) / (2 * eps)
        assert abs(dh[i] - ref) < 1e-5, (i, dh[i], ref)


def test_energy_function_autograd(ctx, rng):
    import torch
    from tnhip import spinhalf
    from tnhip.grad import coupling_gradient, energy_function
    from tnhip.tebd import DeviceMPS

    N = 10
    H = tfim(N, np.ones(N - 1), 0.7 * np.ones(N))
    st = spinhalf()
    psi = complex_mps(N, 4, rng)
    d = DeviceMPS(psi.tensors, psi.center)
    f = energy_function(st, d, H)
    c = torch.tensor(np.real(H.coeffs), dtype=torch.float64, requires_grad=True)
    E = f(c)
    E.backward()
    E_ref, g = coupling_gradient(st, d, H)
    assert abs(E.item() - E_ref.real) < 1e-12
    assert np.allclose(c.grad.numpy(), g.real, atol=1e-12)
