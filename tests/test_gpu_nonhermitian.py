"""GPU parity of dmrg(...; ishermitian=false) (src/algorithms/mps/dmrg.jl:6, :51-53: the local
problem solved by Arnoldi for the eigenvalue of smallest real part, on a complex state).

The non-Hermitian case is the Hatano-Nelson chain (asymmetric hopping e^{+-g} plus zz), whose
spectrum with open boundaries equals that of the Hermitian chain of g = 0: the run is checked
pass by pass against the oracle's Arnoldi DMRG on the same start (1e-9 per site) and its
result against ED (1e-10 per site). On a Hermitian H the Arnoldi path must reach the same
ground energy as the Lanczos (the critical TFIM closed form, 1e-10 per site)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hatano_nelson(N, g, delta, dense=True):
    """The MPO and, for an ED-sized chain, the dense matrix (None otherwise: at N = 16 it
    would be a 2^16 x 2^16 complex matrix of 69 GB)."""
    from tnhip import MPO, spinhalf
    from tnhip.models import OpList
    from oracle import exact

    H = OpList(N)
    for i in range(1, N):
        H.add(["s+", "s-"], [i, i + 1], -np.exp(g))
        H.add(["s-", "s+"], [i, i + 1], -np.exp(-g))
        H.add(["z", "z"], [i, i + 1], delta)
    if not dense:
        return MPO(spinhalf(), H), None
    Hd = exact.dense_hamiltonian(N, [(o, s, c) for o, s, c in zip(H.ops, H.sites, H.coeffs)])
    return MPO(spinhalf(), H), Hd


def test_nonhermitian_dmrg_vs_oracle_and_ed(ctx):
    from tnhip import GMPS, dmrg
    from oracle.dmrg import dmrg as odmrg
    from oracle.mps import MPS, random_mps

    N, g, delta = 8, 0.25, 0.5
    M, Hd = _hatano_nelson(N, g, delta)
    ev = np.linalg.eigvals(Hd)
    want = ev[np.argmin(ev.real)]
    start = random_mps(N, 2, 16, np.random.default_rng(4))
    kw = dict(maxsweeps=8, krylovdim=6, maxdim=16, cutoff=1e-14, ishermitian=False)
    hist, ohist = [], []
    psi, e = dmrg(GMPS(1, 2, [t.copy() for t in start.tensors], 1), M, verbose=False, history=hist, **kw)
    _, e_ref = odmrg(MPS([t.copy() for t in start.tensors], 1), M.tensors, history=ohist, **kw)
    assert isinstance(e, complex)
    assert abs(e - want) / N <= 1e-10
    assert len(hist) == len(ohist)
    for h, o in zip(hist, ohist):
        assert abs(h["energy"] - o[1]) / N <= 1e-9
    assert np.iscomplexobj(psi.tensors[0])


def test_nonhermitian_path_on_a_hermitian_h(ctx):
    from tnhip import MPO, dmrg, randomMPS, spinhalf
    from oracle.exact import tfim_critical_ground_energy
    from test_gpu_svd_dmrg import tfim_H

    N = 12
    rng = np.random.default_rng(3)
    psi, e = dmrg(randomMPS(2, N, 4, rng), MPO(spinhalf(), tfim_H(N)), maxdim=32, cutoff=1e-12,
                  maxsweeps=8, krylovdim=4, ishermitian=False, verbose=False)
    assert abs(e.real - tfim_critical_ground_energy(N)) / N <= 1e-10
    assert abs(e.imag) / N <= 1e-10


def test_nonhermitian_dmrg_with_truncation_n16(ctx):
    """N = 16 with the bond dimension capped at 48 (truncation active): the GPU's Arnoldi DMRG
    against the oracle's on the same start, pass by pass (1e-9 per site; measured 1.4e-10:
    the truncation of a non-normal problem's right eigenvector amplifies the rounding
    differences of the two codes), and against the Hermitian partner chain's ground energy
    from the GPU's Lanczos DMRG (2e-9 per site; the oracle's own distance is 2.2e-10)."""
    from tnhip import GMPS, dmrg
    from oracle.dmrg import dmrg as odmrg
    from oracle.mps import MPS, random_mps

    N, g, delta = 16, 0.1, 0.5
    M, _ = _hatano_nelson(N, g, delta, dense=False)
    Mh, _ = _hatano_nelson(N, 0.0, delta, dense=False)
    start = random_mps(N, 2, 8, np.random.default_rng(2))
    kw = dict(maxdim=48, cutoff=1e-12, maxsweeps=12, krylovdim=4)
    hist, ohist = [], []
    _, e = dmrg(GMPS(1, 2, [t.copy() for t in start.tensors], 1), M, ishermitian=False,
                verbose=False, history=hist, **kw)
    _, e_ref = odmrg(MPS([t.copy() for t in start.tensors], 1), M.tensors, ishermitian=False,
                     history=ohist, **kw)
    for h, o in zip(hist, ohist):
        assert abs(h["energy"] - o[1]) / N <= 1e-9
    _, eh = dmrg(GMPS(1, 2, random_mps(N, 2, 8, np.random.default_rng(1)).tensors, 1), Mh,
                 verbose=False, **kw)
    assert abs(e - eh) / N <= 2e-9
