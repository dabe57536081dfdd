"""GPU parity at the full sizes of the five benchmark workloads, each against the FP64 oracle,
LAPACK, closed forms or free fermions on the same seeded inputs:

  Heisenberg N = 32 DMRG        (chi0 = 8 seed 0, chi_max 64, cutoff 1e-10, 8 passes, krylovdim 4)
  H_eff matvec + 2chi SVD       (chi in {256, 512, 1024, 2048}, FP64 / FP32 / split-BF16x3)
  critical TFIM N = 100 DMRG    (chi0 = 16, cutoff 1e-12, FP64 / FP32 / split-BF16x3 + FP64 passes)
  inhomogeneous TFIM gradient   (N = 64, converged ground state held at bond dimension 512)

Tolerances: energies |dE|/N <= 1e-10 (FP64 target of the workloads); matvec relative
2-norm error 1e-13 (FP64), 2e-6 sqrt(K / 64) (FP32, K = D chi) and 3e-5 (split-BF16x3), as
in test_gpu_contract.py; Gram-path singular values |s^2 - s_ref^2| <= 1e-13 sqrt(n) s_max^2
against LAPACK; observables and entropies 1e-8 between converged GPU and CPU states."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def heisenberg_oplist(N):
    from tnhip.models import OpList

    H = OpList(N)
    for i in range(1, N):
        H.add(["s+", "s-"], [i, i + 1], 0.5)
        H.add(["s-", "s+"], [i, i + 1], 0.5)
        H.add(["z", "z"], [i, i + 1], 0.25)
    return H


def tfim_oplist(N, J, h):
    from tnhip.models import OpList

    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -J[i - 1])
    for i in range(1, N + 1):
        H.add("x", i, -h[i - 1])
    return H


def bond_entropies(tensors):
    from oracle.mps import MPS, entropy

    psi = MPS([t.copy() for t in tensors], 0)
    return np.array([entropy(psi, i) for i in range(1, len(tensors))])


def zz_correlators(tensors):
    from oracle.exact import OPS
    from oracle.mps import MPS, inner_oplist

    N = len(tensors)
    terms = [([OPS["z"], OPS["z"]], [i, i + 1], 1.0) for i in range(1, N)]
    return np.real(inner_oplist(MPS(tensors), terms))


# ------------------------------------------------------------ Heisenberg N = 32


def test_heisenberg_n32_matches_oracle_dmrg(ctx):
    """Same initial tensors (oracle randomMPS, N(0,1) from seed 0, chi 8), same sweep
    schedule: energies, final discarded weights, bond entropies and <z_i z_i+1>."""
    from tnhip import GMPS, MPO, dmrg, spinhalf
    from oracle.dmrg import dmrg as odmrg
    from oracle.mps import random_mps

    N = 32
    M = MPO(spinhalf(), heisenberg_oplist(N))
    ref = random_mps(N, 2, 8, np.random.default_rng(0))
    psi = GMPS(1, 2, [t.copy() for t in ref.tensors], 1)
    hist, ohist = [], []
    psi, e = dmrg(psi, M, maxdim=64, cutoff=1e-10, maxsweeps=8, krylovdim=4, verbose=False, history=hist)
    ref, e_ref = odmrg(ref, M.tensors, maxdim=64, cutoff=1e-10, maxsweeps=8, krylovdim=4, history=ohist)
    print(f"N=32 Heisenberg: GPU {e:.14f} oracle {e_ref:.14f}")
    assert len(hist) == len(ohist) == 8
    assert abs(e - e_ref) / N <= 1e-10
    # per-pass energies follow the same trajectory (the passes are not yet converged)
    for h, (_, c, _, _) in zip(hist, ohist):
        assert abs(h["energy"] - c) / N <= 1e-9
    assert abs(hist[-1]["discarded"] - ohist[-1][3]) <= 1e-11
    assert np.max(np.abs(bond_entropies(psi.tensors) - bond_entropies(ref.tensors))) <= 1e-8
    assert np.max(np.abs(zz_correlators(psi.tensors) - zz_correlators(ref.tensors))) <= 1e-8
    # the bond dimensions agree except where a singular value sits at the cutoff
    gdims = [t.shape[2] for t in psi.tensors[:-1]]
    assert max(abs(a - b) for a, b in zip(gdims, ref.bonddims())) <= 1


# ------------------------------------------------- matvec and 2chi x 2chi SVD

_MATVEC_TOL = {"f64": 1e-13, "f32": 2e-6, "bf16x3": 3e-5}


@pytest.mark.parametrize("chi", [256, 512, 1024, 2048])
def test_heff_matvec_and_split_full_size(ctx, chi):
    """The synthetic matvec workload at every bond dimension, in every precision, against
    one FP64 oracle matvec (optimal order, numpy/OpenBLAS), then the truncated SVD of the
    2chi x 2chi matrix [(a, s1), (s2, b)] of the FP64 result to rank chi against LAPACK."""
    from tnhip import Tensor, ops as T
    from oracle.projmps import product_optimal
    from tnhip import MPO, spinhalf
    from oracle.tensors import truncation_rank

    print(f"chi={chi}: building inputs", flush=True)
    rng = np.random.default_rng(0)
    D = 5
    # the automaton bulk tensor (5, 2, 2, 5); the compressed MPO of a 4-site chain has D = 4
    # at its first bond, where one site spans only d^2 = 4 operators
    W = np.asfortranarray(MPO(spinhalf(), heisenberg_oplist(4), compress=False).tensors[1])
    assert W.shape == (D, 2, 2, D) and not np.iscomplexobj(W)

    def env():
        X = rng.standard_normal((chi, D, chi))
        return np.asfortranarray(0.5 * (X + X.transpose(2, 1, 0)))

    L, R = env(), env()
    x = np.asfortranarray(rng.standard_normal((chi, 2, 2, chi)))
    x /= np.linalg.norm(x)
    ref = product_optimal(L, W, W, R, x)
    dW, dx = Tensor.from_numpy(W), Tensor.from_numpy(x)
    d64 = (Tensor.from_numpy(L), Tensor.from_numpy(R))
    d32 = (Tensor.from_numpy(L, np.float32), Tensor.from_numpy(R, np.float32))
    out64 = None
    for prec in ("f64", "f32", "bf16x3"):
        dL, dR = d64 if prec == "f64" else d32
        out = T.heff_apply(dL, dW, dW, dR, dx, precision=prec)
        err = rel(out.numpy(), ref)
        tol = _MATVEC_TOL[prec] * (max(1.0, np.sqrt(D * chi / 64.0)) if prec == "f32" else 1.0)
        print(f"chi={chi} {prec}: matvec rel err {err:.3e} (tol {tol:.1e})", flush=True)
        assert err <= tol
        if prec == "f64":
            out64 = out
    del d32, ref
    # 2chi x 2chi split of the normalised FP64 result, truncated to chi
    Mh = out64.numpy().reshape(2 * chi, 2 * chi, order="F")
    Mh = np.asfortranarray(Mh / np.linalg.norm(Mh))
    U, S, Vt, disc, fb = T.svd_trunc(Tensor.from_numpy(Mh), 0.0, chi, 1)
    s_ref = np.linalg.svd(Mh, compute_uv=False)
    k_ref, disc_ref = truncation_rank(s_ref, 0.0, chi, 1)
    s = S.numpy()
    n = 2 * chi
    print(f"chi={chi}: svd k={s.size} fallback={fb} max|ds^2|/s0^2="
          f"{np.max(np.abs(s ** 2 - s_ref[:s.size] ** 2)) / s_ref[0] ** 2:.2e} disc={disc:.6e}", flush=True)
    assert not fb and s.size == k_ref == chi
    assert np.max(np.abs(s ** 2 - s_ref[:chi] ** 2)) <= 1e-13 * np.sqrt(n) * s_ref[0] ** 2
    assert abs(disc - disc_ref) <= 1e-12
    u, vt = U.numpy(), Vt.numpy()
    assert np.abs(u.T @ u - np.eye(chi)).max() <= 1e-12
    # the rank-chi reconstruction misses exactly the discarded weight (Eckart-Young)
    err2 = np.linalg.norm(Mh - (u * s) @ vt) ** 2
    assert abs(err2 - disc_ref) <= 1e-10


# --------------------------------------------------------- critical TFIM N = 100


def test_tfim_n100_one_site_passes_after_two_site_growth(ctx):
    """The config-3 workload finished by one-site passes (dmrg(...; nsites=1), dmrg.jl:35-63
    with replacesites!'s one-site branch): the two-site run grows the bonds, then four
    one-site passes keep every bond, do not raise the energy by more than rounding, and stay
    within 1e-10 per site of the closed form."""
    from tnhip import GMPS, MPO, dmrg, spinhalf
    from oracle.exact import tfim_critical_ground_energy
    from oracle.mps import random_mps

    N = 100
    E0 = tfim_critical_ground_energy(N)
    M = MPO(spinhalf(), tfim_oplist(N, np.ones(N - 1), np.ones(N)))
    psi, e2 = dmrg(GMPS(1, 2, random_mps(N, 2, 16, np.random.default_rng(0)).tensors, 1), M,
                   maxdim=2048, cutoff=1e-12, maxsweeps=24, krylovdim=3, verbose=False)
    bonds = [t.shape for t in psi.tensors]
    hist = []
    psi, e1 = dmrg(psi, M, nsites=1, maxsweeps=4, minsweeps=4, krylovdim=3, verbose=False,
                   history=hist)
    print(f"N=100 TFIM two-site E={e2:.14f}, one-site E={e1:.14f}", flush=True)
    assert [t.shape for t in psi.tensors] == bonds
    assert e1 <= e2 + 1e-11 * N
    assert abs(e1 - E0) / N <= 1e-10


def test_tfim_n100_first_excited_state_vs_free_fermions(ctx):
    """The config-3 workload with a rank-1 penalty (dmrg(psi, H, phi0; coeffs = [1, 10]),
    dmrg.jl:144-145): the converged N = 100 ground state as phi0, then H + 10 |phi0><phi0|
    from a fresh start, against the free-fermion first excited energy E1 = E0 + 2 s_min (the
    closed form's one-quasiparticle state, pinned to ED at small N); the two states are
    orthogonal (inner(phi0, psi1) on the device, mpo.jl:181-217)."""
    from tnhip import GMPS, MPO, dmrg, spinhalf
    from tnhip.mpo import inner
    from oracle.exact import tfim_bdg_first_excited_energy, tfim_critical_ground_energy
    from oracle.mps import random_mps

    N = 100
    E0 = tfim_critical_ground_energy(N)
    E1 = tfim_bdg_first_excited_energy(np.ones(N - 1), np.ones(N))
    M = MPO(spinhalf(), tfim_oplist(N, np.ones(N - 1), np.ones(N)))
    kw = dict(maxdim=2048, cutoff=1e-12, maxsweeps=24, krylovdim=3, verbose=False)
    psi0, e0 = dmrg(GMPS(1, 2, random_mps(N, 2, 16, np.random.default_rng(0)).tensors, 1), M, **kw)
    psi1, e1 = dmrg(GMPS(1, 2, random_mps(N, 2, 16, np.random.default_rng(1)).tensors, 1), M, psi0,
                    coeffs=[1.0, 10.0], **kw)
    print(f"N=100 TFIM E0={e0:.14f} |dE0|/N={abs(e0 - E0) / N:.2e} E1={e1:.14f} "
          f"|dE1|/N={abs(e1 - E1) / N:.2e} chi={psi1.maxbonddim()}", flush=True)
    assert abs(e0 - E0) / N <= 1e-10
    assert abs(e1 - E1) / N <= 1e-10
    assert abs(inner(psi0, psi1)) <= 1e-5


def test_tfim_n100_precisions_vs_closed_form_and_cpu(ctx):
    """The N = 100 critical TFIM from one random chi = 16 start (seed 0), cutoff 1e-12,
    in FP64, FP32-MFMA and split-BF16x3 (each finishing with FP64 passes), against the
    closed form E0 = 1 - csc(pi / 402), against each other, and against the FP64 CPU oracle
    DMRG on the same initial tensors (energy, middle-bond entropy, <x_50>)."""
    from tnhip import GMPS, MPO, dmrg, spinhalf
    from oracle.dmrg import dmrg as odmrg
    from oracle.exact import OPS, tfim_critical_ground_energy
    from oracle.mps import MPS, inner_oplist, random_mps

    N = 100
    E0 = tfim_critical_ground_energy(N)
    M = MPO(spinhalf(), tfim_oplist(N, np.ones(N - 1), np.ones(N)))
    start = random_mps(N, 2, 16, np.random.default_rng(0))
    kw = dict(maxdim=2048, cutoff=1e-12, maxsweeps=24, krylovdim=3)
    res = {}
    for prec in ("f64", "f32", "bf16x3"):
        psi = GMPS(1, 2, [t.copy() for t in start.tensors], 1)
        hist = []
        psi, e = dmrg(psi, M, verbose=False, precision=prec, history=hist, **kw)
        res[prec] = (psi, e)
        print(f"N=100 TFIM {prec}: E={e:.14f} |dE|/N={abs(e - E0) / N:.2e} passes={len(hist)} "
              f"chi={psi.maxbonddim()}", flush=True)
        assert abs(e - E0) / N <= 1e-10
        assert hist[-1]["precision"] == "f64"
    for prec in ("f32", "bf16x3"):
        assert abs(res[prec][1] - res["f64"][1]) / N <= 1e-10
    ref, e_cpu = odmrg(start.copy(), M.tensors, **kw)
    print(f"N=100 TFIM CPU oracle: E={e_cpu:.14f}", flush=True)
    assert abs(res["f64"][1] - e_cpu) / N <= 1e-10
    g = res["f64"][0]
    s_gpu = bond_entropies(g.tensors)[N // 2 - 1]
    s_cpu = bond_entropies(ref.tensors)[N // 2 - 1]
    assert abs(s_gpu - s_cpu) <= 1e-8
    x50 = [([OPS["x"]], [50], 1.0)]
    assert abs(inner_oplist(MPS(g.tensors), x50)[0] - inner_oplist(MPS(ref.tensors), x50)[0]) <= 1e-8


# ------------------------------------------------ gradient at a converged state


def test_gradient_inhomogeneous_tfim_n64_chi512(ctx):
    """E(J, h) = <psi|H|psi> for the N = 64 inhomogeneous TFIM (J_i ~ U(0.8, 1.2),
    h_i ~ U(0.5, 1.5) from seed 1) at its DMRG ground state held at bond dimension 512
    (converged at cutoff 1e-14, then six passes with mindim = maxdim = 512): the device
    gradient (checkpoints every 8 sites) against central differences of the device energy
    (step 1e-5) and against Hellmann-Feynman from free fermions, dE0/dJ_i and dE0/dh_i by
    central differences of the exact BdG ground energy."""
    from tnhip import MPO, randomMPS, spinhalf
    from tnhip.grad import coupling_gradient
    from tnhip.mps import dmrg
    from tnhip.tebd import DeviceMPS
    from oracle.exact import tfim_bdg_ground_energy

    N = 64
    r = np.random.default_rng(1)
    J, h = r.uniform(0.8, 1.2, N - 1), r.uniform(0.5, 1.5, N)
    st = spinhalf()
    H = tfim_oplist(N, J, h)
    W = MPO(st, H)
    psi = randomMPS(2, N, 16, rng=np.random.default_rng(2))
    # the disordered chain converges slowly with the default 3-vector Lanczos (1e-5 per site
    # after 30 passes); a 10-vector Krylov space with 6 restarts converges in ~5 passes
    kw = dict(krylovdim=10, kryloviter=6, verbose=False)
    psi, E = dmrg(psi, W, maxsweeps=12, cutoff=1e-14, maxdim=128, **kw)
    # a split of a (chi, 2, 2, chi) theta keeps at most 2 chi values: from chi ~ 30 the bond
    # reaches 512 after five passes at mindim = maxdim = 512, cutoff 0
    psi, E = dmrg(psi, W, maxsweeps=6, cutoff=0.0, maxdim=512, mindim=512, **kw)
    e0 = tfim_bdg_ground_energy(J, h)
    print(f"N=64 gradient state: chi={psi.maxbonddim()} |E-E0|/N={abs(E - e0) / N:.2e}", flush=True)
    assert psi.maxbonddim() == 512
    assert abs(E - e0) / N <= 1e-10
    d = DeviceMPS.from_gmps(psi)
    Eg, g = coupling_gradient(st, d, H, block=8)
    assert abs(Eg.real - e0) / N <= 1e-10
    step = 1e-5
    for t in [0, 31, N - 2, N - 1, N - 1 + 40, 2 * N - 2]:
        Hp, Hm = H * 1.0, H * 1.0
        Hp.coeffs[t] += step
        Hm.coeffs[t] -= step
        Ep, _ = coupling_gradient(st, d, Hp, block=8)
        Em, _ = coupling_gradient(st, d, Hm, block=8)
        fd = (Ep - Em).real / (2 * step)
        assert abs(fd - g[t].real) < 1e-6, (t, fd, g[t])
    dJ, dh = -g[: N - 1].real, -g[N - 1:].real  # term coefficients are -J_i and -h_i
    eps = 1e-5
    errs = []
    for i in range(N - 1):
        Jp, Jm = J.copy(), J.copy()
        Jp[i] += eps
        Jm[i] -= eps
        errs.append(dJ[i] - (tfim_bdg_ground_energy(Jp, h) - tfim_bdg_ground_energy(Jm, h)) / (2 * eps))
    for i in range(N):
        hp, hm = h.copy(), h.copy()
        hp[i] += eps
        hm[i] -= eps
        errs.append(dh[i] - (tfim_bdg_ground_energy(J, hp) - tfim_bdg_ground_energy(J, hm)) / (2 * eps))
    print(f"N=64 Hellmann-Feynman max error {np.max(np.abs(errs)):.2e}", flush=True)
    # <O> at an approximate eigenstate errs linearly in the state error, ~sqrt(dE / gap)
    assert np.max(np.abs(errs)) < 1e-5


# ------------------------------------------------------------ QJMC N = 40


def _qjmc_model(N, g, gamma):
    from tnhip.models import OpList

    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        H.add("x", i, -g)
    J = OpList(N)
    for i in range(1, N + 1):
        J.add("s-", i, gamma ** 0.5)
    zs = OpList(N)
    for i in range(1, N + 1):
        zs.add("z", i)
    return H, J, zs


def _terms(ol):
    return [(list(o), list(s), c) for o, s, c in zip(ol.ops, ol.sites, ol.coeffs)]


def test_qjmc_n40_workload_vs_oracle_trajectories(ctx):
    """The batched-QJMC workload itself: 512 trajectories of the dissipative TFIM (N = 40,
    g = 0.5, s-_i at rate 0.1, dt = 0.01, 200 second-order Trotter steps, cutoff 1e-8,
    chi_max 256, uniforms from seed 0) on the GPU (engine "auto"); four of them (the three
    with the most jumps and one without) rerun by the CPU oracle's MPS trajectory on the
    same uniforms: identical jump records and times, <sz_i> at every step to 1e-8. The
    trajectory average <sz_i(t)> is bounded by the oracle spread (a sanity check of the
    bookkeeping of all 512 rows: mean of |<sz>| <= 1)."""
    from tnhip import spinhalf
    from tnhip.qjmc import qjmc_trajectories
    from oracle.qjmc import qjmc_mps

    N, g, gamma, dt, steps = 40, 0.5, 0.1, 0.01, 200
    H, J, zs = _qjmc_model(N, g, gamma)
    U = np.random.default_rng(0).random((512, 4 * steps + 16))
    out = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, steps * dt, dt, U, observables=zs, cutoff=1e-8,
                            maxdim=256)
    meas = out["measurements"].real
    assert meas.shape == (512, steps + 1, N)
    assert np.all(np.abs(meas) <= 1 + 1e-9)
    nj = np.array(out["njumps"])
    print(f"QJMC N=40: engine {out['engine']}, {nj.sum()} jumps, max bond {max(out['maxbonddim'])}", flush=True)
    order = np.argsort(-nj, kind="stable")
    pick = list(order[:3]) + [int(np.nonzero(nj == 0)[0][0])] if (nj == 0).any() else list(order[:4])
    up = np.array([1.0, 0.0])
    for t in pick:
        j, jt, times, ms, chi = qjmc_mps([up] * N, _terms(H), _terms(J), N, dt, steps * dt, U[t],
                                         obs=[("z", i) for i in range(1, N + 1)], cutoff=1e-8, maxdim=256)
        print(f"  trajectory {t}: jumps {j} max bond {chi}", flush=True)
        assert out["jumps"][t] == j, (t, out["jumps"][t], j)
        assert np.allclose(out["jumptimes"][t], jt, atol=1e-12)
        assert np.abs(meas[t] - ms.real).max() <= 1e-8
        assert out["maxbonddim"][t] == chi


def test_qjmc_overflowing_rows_vs_oracle(ctx):
    """The regime above the narrow batched slots: dissipative TFIM at g = 1 (N = 10,
    gamma = 0.1, cutoff 1e-12, 70 steps of dt = 0.05), where most trajectories reach bond
    dimension 24-32 and continue on the wide slots (one 1024-thread workgroup per
    trajectory, divide-and-conquer splits) while the others complete in the narrow ones.
    Every row, narrow or wide, against the oracle MPS trajectory on its
    uniforms: identical jump records, bond dimensions, <sz_i> to 1e-8."""
    from tnhip import spinhalf
    from tnhip.qjmc import qjmc_trajectories
    from oracle.qjmc import qjmc_mps

    N, g, gamma, dt, steps = 10, 1.0, 0.1, 0.05, 70
    H, J, zs = _qjmc_model(N, g, gamma)
    U = np.random.default_rng(31).random((8, 4 * steps + 8))
    out = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, steps * dt, dt, U, observables=zs, cutoff=1e-12,
                            nworkers=4)
    assert out["engine"] == "batched"
    chis = out["maxbonddim"]
    assert min(chis) <= 16 < max(chis)
    up = np.array([1.0, 0.0])
    for t in range(len(U)):
        j, jt, times, ms, chi = qjmc_mps([up] * N, _terms(H), _terms(J), N, dt, steps * dt, U[t],
                                         obs=[("z", i) for i in range(1, N + 1)], cutoff=1e-12)
        assert out["jumps"][t] == j, t
        assert out["maxbonddim"][t] == chi, (t, out["maxbonddim"][t], chi)
        assert np.abs(out["measurements"][t].real - ms.real).max() <= 1e-8
