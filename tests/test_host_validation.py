"""Host-side argument checks of the DMRG options and of inner(args...), which the Python
mirror makes before any device call (so they run without a GPU): the reference's error
messages for a missing Hamiltonian, a term that is neither an MPS nor an MPO, a state that
is not rank 1, an unsupported nsites, and inner arguments without a bra-ket structure."""
import numpy as np
import pytest


def _mps(N, chi=2):
    ts = [np.ones((1 if i == 0 else chi, 2, 1 if i == N - 1 else chi)) for i in range(N)]
    from tnhip import GMPS

    return GMPS(1, 2, ts, 1)


def _mpo(N):
    from tnhip import GMPS

    return GMPS(2, 2, [np.ones((1, 2, 2, 1)) for _ in range(N)], 0)


def test_dmrg_arguments():
    from tnhip import GMPS, dmrg

    N = 4
    psi, H = _mps(N), _mpo(N)
    with pytest.raises(ValueError, match="nsites"):
        dmrg(psi, H, nsites=3, verbose=False)
    with pytest.raises(ValueError, match="atleast one"):
        dmrg(psi, verbose=False)
    with pytest.raises(ValueError, match="rank 1"):
        dmrg(H, H, verbose=False)
    bad = GMPS(3, 2, [np.ones((1, 2, 2, 2, 1)) for _ in range(N)], 0)
    with pytest.raises(ValueError, match="MPOs \\(rank 2\\) or MPSs \\(rank 1\\)"):
        dmrg(psi, bad, verbose=False)
    with pytest.raises(ValueError, match="same properties"):
        dmrg(psi, _mpo(N + 1), verbose=False)


def test_inner_arguments():
    from tnhip.mpo import inner

    N = 4
    phi, psi, O = _mps(N), _mps(N, 3), _mpo(N)
    with pytest.raises(ValueError, match="atleast 2"):
        inner(phi)
    with pytest.raises(ValueError, match="braket"):
        inner(phi, O)
    with pytest.raises(ValueError, match="final GMPS"):
        inner(phi, psi, O)
    with pytest.raises(ValueError, match="same length"):
        inner(phi, _mps(N + 1))
