"""GPU parity of rank-1 ("squared") Hamiltonian terms in the two-site DMRG,
dmrg(psi, H, phi; coeffs) (src/algorithms/mps/dmrg.jl:128-154, which builds
ProjMPS(phi, psi; rank=2, squared=true) for each rank-1 argument; product, project and
calculate of src/structures/mps/projmps.jl:135-216), through tnh_dmrg_add_projector.

Checks: the state's calculate() against the dense <psi|H|psi> + w <phi|psi> (1e-12); the
penalty run H + w |phi0><phi0| against the oracle's DMRG on the same start (1e-10 per site)
and against exact diagonalisation of that operator, which with phi0 the ground state is the
first excited state of H (1e-10 per site); a phase-rotated complex phi0; and a Hamiltonian
made of a rank-1 term alone, whose ground state is phi itself with energy -|phi|^2.
Also the one-site update, dmrg(...; nsites=1): every pass against the oracle's one-site
DMRG (1e-10 per site), and ED with bonds that hold the exact state."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tfim(N, h):
    from tnhip import MPO, spinhalf
    from tnhip.models import OpList
    from oracle import exact

    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        H.add("x", i, -h)
    Hd = exact.dense_hamiltonian(N, [(o, s, c) for o, s, c in zip(H.ops, H.sites, H.coeffs)])
    return MPO(spinhalf(), H), Hd


def _complex_mps(N, chi, rng):
    from oracle.mps import MPS, movecenter, random_mps

    psi = random_mps(N, 2, chi, rng)
    ts = [t + 1j * rng.standard_normal(t.shape) for t in psi.tensors]
    out = MPS(ts, 0)
    movecenter(out, 1)
    return out


@pytest.mark.parametrize("cplx", [False, True])
def test_calculate_with_rank1_term_matches_dense(ctx, cplx):
    from tnhip import GMPS
    from tnhip.mps import DMRGState
    from oracle.mps import random_mps, to_dense, MPS

    N, h, w = 8, 0.8, 0.65
    M, Hd = _tfim(N, h)
    rng = np.random.default_rng(21)
    if cplx:
        phi, psi0 = _complex_mps(N, 3, rng), _complex_mps(N, 4, rng)
    else:
        phi, psi0 = random_mps(N, 2, 3, rng), random_mps(N, 2, 4, rng)
    st = DMRGState(GMPS(1, 2, [t.copy() for t in psi0.tensors], 1),
                   [M, GMPS(1, 2, phi.tensors, 1)], coeffs=[1.0, w])
    try:
        e = st.energy()
        psid = to_dense(MPS(st.tensors()))
    finally:
        st.close()
    phid = to_dense(phi)
    want = np.vdot(psid, Hd @ psid).real + w * np.vdot(phid, psid).real
    assert abs(e - want) <= 1e-12 * max(1.0, abs(want))


def _ground(M, N, seed):
    from oracle.dmrg import dmrg as odmrg
    from oracle.mps import random_mps

    phi0, _ = odmrg(random_mps(N, 2, 4, np.random.default_rng(seed)), M.tensors, maxdim=32,
                    cutoff=1e-14, maxsweeps=8, krylovdim=4)
    return phi0


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_dmrg_penalty_term_vs_oracle_and_ed(ctx, precision):
    from tnhip import GMPS, dmrg
    from oracle.dmrg import dmrg as odmrg
    from oracle.mps import MPS, random_mps, to_dense

    N, h, w = 10, 0.9, 10.0
    M, Hd = _tfim(N, h)
    ev = np.linalg.eigvalsh(Hd)
    phi0 = _ground(M, N, 1)
    phid = to_dense(phi0)
    target = np.linalg.eigvalsh(Hd + w * np.outer(phid, phid.conj()))[0]
    assert abs(target - ev[1]) < 1e-9  # phi0 is the ground state: the target is E_1
    start = random_mps(N, 2, 4, np.random.default_rng(2))
    kw = dict(maxdim=32, cutoff=1e-12, maxsweeps=10, krylovdim=4)
    psi = GMPS(1, 2, [t.copy() for t in start.tensors], 1)
    psi, e = dmrg(psi, M, GMPS(1, 2, phi0.tensors, 1), coeffs=[1.0, w], precision=precision,
                  verbose=False, **kw)
    assert abs(e - target) / N <= 1e-10
    v = to_dense(MPS(psi.tensors))
    assert abs(np.vdot(phid, v)) <= 1e-6 * np.linalg.norm(v)
    if precision == "f64":
        ref = MPS([t.copy() for t in start.tensors], 1)
        _, e_ref = odmrg(ref, M.tensors, terms=[(phi0.tensors, w)], **kw)
        assert abs(e - e_ref) / N <= 1e-10


def test_dmrg_penalty_term_complex_phase(ctx):
    """phi0 times a phase is the same projector: a complex phi and a complex start give the
    same excited state, on the GPU and in the oracle."""
    from tnhip import GMPS, dmrg
    from oracle.dmrg import dmrg as odmrg
    from oracle.mps import MPS, to_dense

    N, h, w = 8, 0.7, 6.0
    M, Hd = _tfim(N, h)
    phi0 = _ground(M, N, 4)
    phic = [t.astype(np.complex128) for t in phi0.tensors]
    phic[0] = phic[0] * np.exp(0.6j)
    phid = to_dense(MPS(phic))
    target = np.linalg.eigvalsh(Hd + w * np.outer(phid, phid.conj()))[0]
    start = _complex_mps(N, 4, np.random.default_rng(9))
    kw = dict(maxdim=16, cutoff=1e-12, maxsweeps=10, krylovdim=4)
    psi = GMPS(1, 2, [t.copy() for t in start.tensors], 1)
    psi, e = dmrg(psi, M, GMPS(1, 2, phic, 1), coeffs=[1.0, w], verbose=False, **kw)
    _, e_ref = odmrg(MPS([t.copy() for t in start.tensors], 1), M.tensors,
                     terms=[(phic, w)], **kw)
    assert abs(e - target) / N <= 1e-10
    assert abs(e - e_ref) / N <= 1e-10
    assert np.iscomplexobj(psi.tensors[0])


def test_dmrg_rank1_term_alone(ctx):
    """H = -|phi><phi| alone (no MPO term): the ground state is phi, energy -<phi|phi>."""
    from tnhip import GMPS, dmrg
    from oracle.mps import MPS, random_mps, to_dense

    N = 9
    rng = np.random.default_rng(13)
    phi = _complex_mps(N, 3, rng)
    phid = to_dense(phi)
    nrm2 = np.vdot(phid, phid).real
    start = random_mps(N, 2, 2, rng)
    psi = GMPS(1, 2, [t.copy() for t in start.tensors], 1)
    psi, e = dmrg(psi, GMPS(1, 2, phi.tensors, 1), coeffs=[-1.0], maxdim=8, cutoff=1e-14,
                  maxsweeps=4, krylovdim=3, verbose=False)
    assert abs(e + nrm2) <= 1e-12 * nrm2
    v = to_dense(MPS(psi.tensors))
    assert abs(abs(np.vdot(phid, v)) / np.sqrt(nrm2 * np.vdot(v, v).real) - 1) <= 1e-10


def test_dmrg_one_site_matches_oracle_per_pass(ctx):
    """dmrg(...; nsites=1) (dmrg.jl:35-63, replacesites!'s one-site branch gmps.jl:204-213)
    from a start whose bonds (6) cannot hold the exact state: every pass's energy against
    the oracle's one-site DMRG on the same start, 1e-10 per site; the bonds never change."""
    from tnhip import GMPS, dmrg
    from oracle.dmrg import dmrg as odmrg
    from oracle.mps import MPS, random_mps

    N = 10
    M, _ = _tfim(N, 0.9)
    start = random_mps(N, 2, 6, np.random.default_rng(5))
    kw = dict(nsites=1, maxsweeps=6, krylovdim=4)
    hist, ohist = [], []
    psi = GMPS(1, 2, [t.copy() for t in start.tensors], 1)
    psi, e = dmrg(psi, M, verbose=False, history=hist, **kw)
    _, e_ref = odmrg(MPS([t.copy() for t in start.tensors], 1), M.tensors, history=ohist, **kw)
    assert len(hist) == len(ohist) == 6
    for h, o in zip(hist, ohist):
        assert abs(h["energy"] - o[1]) / N <= 1e-10
    assert [t.shape for t in psi.tensors] == [t.shape for t in start.tensors]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_dmrg_one_site_full_bonds_reach_ed(ctx, precision):
    """With bonds that hold the exact state (N = 8, bond dimension 16) the one-site update
    reaches the ED ground energy, in FP64 and with FP32 passes refined in FP64; with the
    ground state as a penalty term it reaches the first excited state."""
    from tnhip import GMPS, dmrg
    from oracle.mps import random_mps

    N, w = 8, 10.0
    M, Hd = _tfim(N, 0.9)
    ev = np.linalg.eigvalsh(Hd)
    rng = np.random.default_rng(6)
    psi = GMPS(1, 2, random_mps(N, 2, 16, rng).tensors, 1)
    psi, e0 = dmrg(psi, M, nsites=1, maxsweeps=10, krylovdim=4, precision=precision,
                   verbose=False)
    assert abs(e0 - ev[0]) / N <= 1e-10
    phi0 = GMPS(1, 2, [t.copy() for t in psi.tensors], psi.center)
    psi1 = GMPS(1, 2, random_mps(N, 2, 16, rng).tensors, 1)
    psi1, e1 = dmrg(psi1, M, phi0, coeffs=[1.0, w], nsites=1, maxsweeps=14, krylovdim=4,
                    precision=precision, verbose=False)
    assert abs(e1 - ev[1]) / N <= 1e-10
