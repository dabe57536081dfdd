"""GPU parity of quantum-jump Monte Carlo (qjmc_simulation, qjmc.jl:28-167) against the
dense oracle trajectory on the same supplied uniforms. Without truncation (cutoff 0) the MPS
trajectory is the dense one up to the SVD's accuracy, so the jump records must agree
exactly and <sz_i> at every save point to 1e-8."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def model(N, g, gamma):
    from tnhip import OpList

    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        H.add("x", i, -g)
    J = OpList(N)
    for i in range(1, N + 1):
        J.add("s-", i, gamma ** 0.5)
    return H, J


def terms(ol):
    return [(list(o), list(s), c) for o, s, c in zip(ol.ops, ol.sites, ol.coeffs)]


def zlist(N):
    from tnhip import OpList

    zs = OpList(N)
    for i in range(1, N + 1):
        zs.add("z", i)
    return zs


@pytest.mark.parametrize("classical", [True, False])
def test_qjmc_trajectories_match_dense(ctx, classical):
    from tnhip import productMPS, spinhalf
    from tnhip.qjmc import QJMCOperators, QJMCActivity, qjmc_simulation
    from oracle.qjmc import qjmc_dense

    N, g, gamma, dt, tmax = 6, 0.5, 0.5, 0.05, 1.0
    H, J = model(N, g, gamma)
    U = np.random.default_rng(11).random((3, 200))
    v0 = np.zeros(2 ** N, complex)
    v0[0] = 1.0
    jr, tr, times, meas = qjmc_dense(v0, terms(H), terms(J), N, dt, tmax, U, save=0.25,
                                     obs=[("z", i) for i in range(1, N + 1)], classical=classical)
    assert sum(len(j) for j in jr) >= 2  # the case exercises jumps
    for m in range(U.shape[0]):
        st = spinhalf()
        psi = productMPS(st, ["up"] * N)
        it = iter(U[m])
        obs, act = QJMCOperators(zlist(N), st), QJMCActivity()
        jumps, jtimes = qjmc_simulation(st, psi, H, J, tmax, dt, [obs, act], save=0.25, cutoff=0.0,
                                        classical=classical, verbose=False, rand=lambda: float(next(it)))
        assert jumps == jr[m], (m, jumps, jr[m])
        assert np.allclose(jtimes, tr[m], atol=1e-12)
        assert act.jumps == len(jumps)
        got = np.array(obs.measurements).real
        assert got.shape == meas[m].shape
        assert np.allclose(got, meas[m].real, atol=1e-8)


@pytest.mark.parametrize("engine", ["workers", "batched"])
def test_native_runner_matches_dense(ctx, engine):
    """tnh_qjmc_run (C++ step loop, two worker threads with their own streams) and
    tnh_qjmc_run_batched (all trajectories per launch) against the dense oracle on the same
    uniforms: identical jump records, <sz_i> to 1e-8."""
    from tnhip import spinhalf
    from tnhip.qjmc import qjmc_trajectories
    from oracle.qjmc import qjmc_dense

    N, g, gamma, dt, tmax = 6, 0.5, 0.5, 0.05, 1.0
    H, J = model(N, g, gamma)
    U = np.random.default_rng(12).random((5, 200))
    v0 = np.zeros(2 ** N, complex)
    v0[0] = 1.0
    jr, tr, times, meas = qjmc_dense(v0, terms(H), terms(J), N, dt, tmax, U, save=0.25,
                                     obs=[("z", i) for i in range(1, N + 1)])
    assert sum(len(j) for j in jr) >= 2
    out = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, tmax, dt, U, observables=zlist(N), save=0.25,
                            cutoff=0.0, nworkers=2, engine=engine)
    assert out["engine"] == engine
    assert out["jumps"] == jr
    for m in range(U.shape[0]):
        assert np.allclose(out["jumptimes"][m], tr[m], atol=1e-12)
    assert np.allclose(out["times"], times)
    assert np.allclose(out["measurements"].real, meas.real, atol=1e-8)


def test_batched_matches_workers_truncated(ctx):
    """The dissipative TFIM with truncation (cutoff 1e-8, maxdim 256), 24 trajectories of
    60 steps from all spins up: the batched engine against the per-trajectory workers on
    the same uniforms. Both run the same splits (Gram eigensolve, truncation rule, one
    Newton-Schulz step) in different summation orders, so jump records must agree exactly
    and measurements to 1e-9."""
    from tnhip import spinhalf
    from tnhip.qjmc import qjmc_trajectories

    N, g, gamma, dt, steps = 10, 0.5, 0.1, 0.01, 60
    H, J = model(N, g, gamma)
    U = np.random.default_rng(21).random((24, 4 * steps + 8))
    kw = dict(observables=zlist(N), cutoff=1e-8, maxdim=256)
    a = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, steps * dt, dt, U, engine="batched", **kw)
    b = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, steps * dt, dt, U, engine="workers", nworkers=4, **kw)
    assert sum(a["njumps"]) >= 3  # the case exercises jumps
    assert a["jumps"] == b["jumps"]
    assert a["jumptimes"] == b["jumptimes"]
    assert a["maxbonddim"] == b["maxbonddim"]
    assert np.abs(a["measurements"] - b["measurements"]).max() <= 1e-9


def test_batched_wide_slots_without_truncation(ctx):
    """Without truncation the bond dimension of a 10-site chain reaches 32, above the narrow
    batched slots (16 for spin-1/2): those trajectories continue inside the batched call on
    the wide slots (bond dimension 64, the divide-and-conquer split per trajectory), so the
    batched engine completes them itself and matches the per-trajectory workers. Untruncated
    splits resolve singular values near sqrt(eps) s_max only to that level (the Gram path's
    limit), and the two engines split the small Gram matrices (side <= 8) with different
    eigensolvers, so measurements agree to 1e-9 (observed ~3e-12 after 12 steps)."""
    from tnhip import spinhalf
    from tnhip.qjmc import qjmc_trajectories

    N, dt, steps = 10, 0.05, 12
    H, J = model(N, 1.5, 1e-4)
    U = np.random.default_rng(22).random((3, 4 * steps + 8))
    kw = dict(observables=zlist(N), cutoff=0.0)
    a = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, steps * dt, dt, U, engine="batched", **kw)
    b = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, steps * dt, dt, U, engine="workers", **kw)
    assert a["engine"] == "batched"
    assert max(a["maxbonddim"]) == 32
    assert a["jumps"] == b["jumps"]
    assert np.abs(a["measurements"] - b["measurements"]).max() <= 1e-9


def test_batched_partial_overflow(ctx):
    """Dissipative TFIM (g = 1, gamma = 0.1, cutoff 1e-12), 8 trajectories of 70 steps: most
    trajectories reach bond dimension 24-32, above the narrow slots (16 for spin-1/2), while
    a few stay at 16 and complete there. The others continue on the wide slots in the same
    call (from the start of the step they outgrew the narrow ones in), and every trajectory matches the workers' own run: jump records exactly, bond
    dimensions, measurements to 1e-8. At cutoff 1e-12 the kept singular values reach ~1e-6
    s_max, whose vectors the Gram path resolves only to ~eps s_max^2 / s_k^2; the engines
    form their Gram matrices in different orders, so after 70 steps their states differ at
    ~1e-9 (the bound the oracle comparison of this problem uses is 1e-8)."""
    from tnhip import spinhalf
    from tnhip.qjmc import qjmc_trajectories

    N, g, gamma, dt, steps = 10, 1.0, 0.1, 0.05, 70
    H, J = model(N, g, gamma)
    U = np.random.default_rng(31).random((8, 4 * steps + 8))
    kw = dict(observables=zlist(N), cutoff=1e-12, nworkers=4)
    a = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, steps * dt, dt, U, engine="auto", **kw)
    b = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, steps * dt, dt, U, engine="workers", **kw)
    fit = [t for t in range(len(U)) if b["maxbonddim"][t] <= 16]
    assert 0 < len(fit) < len(U)  # some trajectories fit the narrow slots, some do not
    assert a["engine"] == "batched"
    assert a["maxbonddim"] == b["maxbonddim"]
    assert a["jumps"] == b["jumps"]
    assert a["jumptimes"] == b["jumptimes"]
    assert np.abs(a["measurements"] - b["measurements"]).max() <= 1e-8


def test_batched_wide_slots_second_step(ctx):
    """Weakly dissipative critical TFIM, N = 14 (g = 1, gamma = 0.02, cutoff 1e-10), 4
    trajectories of 84 steps of dt = 0.05: the central bond passes 16 near t = 2 and 64 near
    t = 3.7 (up to 90-128 without jumps), so trajectories go narrow slots -> wide slots of 64
    -> wide slots of 128 (Gram sides up to 256), each continued from the start of the step
    it outgrew the smaller slots in. Against the per-trajectory workers: jump records, bond dimensions, measurements
    to 1e-8 (the cross-engine bound of the truncated runs above)."""
    from tnhip import spinhalf
    from tnhip.qjmc import qjmc_trajectories

    N, g, gamma, dt, steps = 14, 1.0, 0.02, 0.05, 84
    H, J = model(N, g, gamma)
    U = np.random.default_rng(41).random((4, 4 * steps + 8))
    kw = dict(observables=zlist(N), cutoff=1e-10, nworkers=4)
    a = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, steps * dt, dt, U, engine="batched", **kw)
    b = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, steps * dt, dt, U, engine="workers", **kw)
    print(f"max bonds {a['maxbonddim']}, jumps {a['njumps']}", flush=True)
    assert max(a["maxbonddim"]) > 64
    assert a["engine"] == "batched"
    assert a["maxbonddim"] == b["maxbonddim"]
    assert a["jumps"] == b["jumps"]
    assert np.abs(a["measurements"] - b["measurements"]).max() <= 1e-8


def _random_sitetypes(d, rng):
    from tnhip.models import Sitetypes

    st = Sitetypes(d)
    st.add_state("up", np.eye(d)[0])
    st.add_op("id", np.eye(d), "id")
    for name in ("a", "b", "h"):
        x = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        st.add_op(name, (x + x.conj().T) / (2 * d), name)
    low = np.diag(np.ones(d - 1), -1)
    st.add_op("l", low, "l+")
    st.add_op("l+", low.T, "l")
    return st


def test_batched_beyond_wide_slots_reruns_on_workers(ctx):
    """d = 8 (a random Hermitian chain of 6 sites, strongly coupled, maxdim 100): the narrow
    slots hold bond dimension 4 and the wide ones 64 (Gram side 512), so trajectories whose
    bonds grow past 64 are marked: engine="batched" raises, engine="auto" reruns them on the
    per-trajectory workers and matches the workers' own run."""
    from tnhip import OpList, TnhError
    from tnhip.qjmc import qjmc_trajectories

    rng = np.random.default_rng(5)
    d, N, dt, steps = 8, 6, 0.1, 3
    st = _random_sitetypes(d, rng)
    H = OpList(N)
    for i in range(1, N):
        H.add(["a", "b"], [i, i + 1], 20.0)
        H.add(["b", "a"], [i, i + 1], 15.0)
    for i in range(1, N + 1):
        H.add("h", i, 0.5)
    J = OpList(N)
    J.add("l", 1, 0.3)
    obs = OpList(N)
    for i in range(1, N + 1):
        obs.add("h", i)
    U = np.random.default_rng(51).random((2, 4 * steps + 8))
    kw = dict(observables=obs, cutoff=1e-12, maxdim=100, nworkers=2)
    with pytest.raises(TnhError, match="bond dimension above"):
        qjmc_trajectories(st, ["up"] * N, H, J, steps * dt, dt, U, engine="batched", **kw)
    a = qjmc_trajectories(st, ["up"] * N, H, J, steps * dt, dt, U, engine="auto", **kw)
    b = qjmc_trajectories(st, ["up"] * N, H, J, steps * dt, dt, U, engine="workers", **kw)
    assert a["engine"] == "batched+workers"
    assert max(a["maxbonddim"]) > 64
    assert a["jumps"] == b["jumps"]
    assert a["maxbonddim"] == b["maxbonddim"]
    assert np.abs(a["measurements"] - b["measurements"]).max() <= 1e-8


def test_batched_matches_workers_spin_one(ctx):
    """d = 3 (spin-1: slot capacity 10, Gram matrices up to 30 x 30): spin-1 XXZ-like chain
    with a transverse field and S- decay, truncated to maxdim 8. The batched engine against
    the per-trajectory workers (the library's general operators) on the same uniforms:
    identical jump records and bond dimensions, measurements to 1e-9. The dense oracle is
    spin-1/2 only; the spin-1/2 engines are pinned to it above."""
    from tnhip import OpList
    from tnhip.models import Sitetypes
    from tnhip.qjmc import qjmc_trajectories

    s2 = 2 ** -0.5
    st = Sitetypes(3)
    st.add_state("up", [1, 0, 0])
    st.add_op("id", np.eye(3), "id")
    st.add_op("sz", np.diag([1.0, 0.0, -1.0]), "sz")
    st.add_op("sx", s2 * np.array([[0, 1, 0], [1, 0, 1], [0, 1, 0]]), "sx")
    st.add_op("s+", 2 ** 0.5 * np.array([[0, 1, 0], [0, 0, 1], [0, 0, 0]]), "s-")
    st.add_op("s-", 2 ** 0.5 * np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), "s+")
    N, dt, steps = 6, 0.02, 40
    H = OpList(N)
    for i in range(1, N):
        H.add(["sz", "sz"], [i, i + 1], 1.0)
        H.add(["sx", "sx"], [i, i + 1], 0.5)
    for i in range(1, N + 1):
        H.add("sx", i, 0.7)
    J = OpList(N)
    for i in range(1, N + 1):
        J.add("s-", i, 0.4)
    obs = OpList(N)
    for i in range(1, N + 1):
        obs.add("sz", i)
    U = np.random.default_rng(31).random((12, 4 * steps + 8))
    kw = dict(observables=obs, cutoff=1e-10, maxdim=8)
    a = qjmc_trajectories(st, ["up"] * N, H, J, steps * dt, dt, U, engine="batched", **kw)
    b = qjmc_trajectories(st, ["up"] * N, H, J, steps * dt, dt, U, engine="workers", nworkers=3, **kw)
    assert a["engine"] == "batched"
    assert sum(a["njumps"]) >= 3
    assert max(a["maxbonddim"]) >= 4  # entangled, truncated states
    assert a["jumps"] == b["jumps"]
    assert a["maxbonddim"] == b["maxbonddim"]
    assert np.abs(a["measurements"] - b["measurements"]).max() <= 1e-9


def test_native_runner_rejects_short_uniforms(ctx):
    from tnhip import spinhalf, TnhError
    from tnhip.qjmc import qjmc_trajectories

    N = 4
    H, J = model(N, 0.5, 0.5)
    with pytest.raises(TnhError):
        qjmc_trajectories(spinhalf(), ["up"] * N, H, J, 1.0, 0.05, np.full((1, 3), 0.5), cutoff=0.0)


def test_emission_rates_match_dense(ctx, rng):
    from tnhip import spinhalf
    from tnhip.qjmc import qjmc_emission_rates
    from tnhip.tebd import DeviceMPS
    from oracle.mps import random_mps, to_dense, normalize
    from oracle.qjmc import emission_rates_dense

    N = 5
    _, J = model(N, 0.5, 0.3)
    J.add(["s-", "s-"], [2, 3], 0.7)
    psi = random_mps(N, 2, 4, rng)
    normalize(psi)
    v = to_dense(psi)
    ref = emission_rates_dense(v[:, None], terms(J), N)[0]
    got = qjmc_emission_rates(spinhalf(), DeviceMPS(psi.tensors, psi.center), J)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-14)
