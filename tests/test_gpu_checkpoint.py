"""Checkpoint and resume of the device DMRG: dmrg(..., checkpoint=path) writes psi after every
pass in the reference's HDF5 layout; reading it back and calling dmrg again continues the
run to the same ground energy (critical TFIM, N = 16, closed form E0 = 1 - csc(pi / 66))."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_dmrg_checkpoint_and_resume(ctx, tmp_path):
    from tnhip import MPO, dmrg, randomMPS, spinhalf
    from tnhip.checkpoint import read_mps
    from tnhip.models import OpList
    from oracle.exact import tfim_critical_ground_energy

    N = 16
    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        H.add("x", i, -1.0)
    M = MPO(spinhalf(), H)
    path = tmp_path / "run.h5"
    psi = randomMPS(2, N, 4, np.random.default_rng(9))
    psi, e1 = dmrg(psi, M, maxsweeps=3, cutoff=1e-12, maxdim=64, verbose=False, checkpoint=path)
    saved = read_mps(path)
    assert saved.center == psi.center and len(saved) == N
    for a, b in zip(saved.tensors, psi.tensors):
        assert np.allclose(a, b, atol=0, rtol=0)
    psi2, e2 = dmrg(saved, M, maxsweeps=12, cutoff=1e-12, maxdim=64, verbose=False)
    assert abs(e2 - tfim_critical_ground_energy(N)) / N <= 1e-10
    assert e2 <= e1 + 1e-12
