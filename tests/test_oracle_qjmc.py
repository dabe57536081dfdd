"""Pins of the QJMC oracle (CPU): the trajectory average against the exact Lindblad
evolution, and the uniform-consumption order of qjmc.jl:59-112 on a case whose jump
time is fixed by the uniforms."""
import numpy as np
import pytest

from oracle.qjmc import emission_rates_dense, escape_terms, lindblad_dense, qjmc_dense
from oracle.exact import OPS, dense_hamiltonian


def dissipative_tfim(N, g, gamma):
    H = [(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)] + [(["x"], [i], -g) for i in range(1, N + 1)]
    J = [(["s-"], [i], gamma ** 0.5) for i in range(1, N + 1)]
    return H, J


def test_escape_operator_is_projector_up():
    (mats, sites, c2), = escape_terms([(["s-"], [2], 0.3)])
    assert np.allclose(mats[0], OPS["pu"]) and sites == [2] and abs(c2 - 0.09) < 1e-15


def test_trajectory_average_matches_lindblad():
    """M = 4000 trajectories, N = 3: <sz_i>(t) within 4 standard errors plus the O(dt)
    bias of jumps taken at the end of each step (dt = 0.01, gamma = 0.5: ~gamma dt / 2)."""
    N, g, gamma = 3, 0.5, 0.5
    H, J = dissipative_tfim(N, g, gamma)
    v0 = np.zeros(2 ** N, complex)
    v0[0] = 1.0
    M, dt = 4000, 0.01
    U = np.random.default_rng(5).random((M, 400))
    _, _, times, meas = qjmc_dense(v0, H, J, N, dt, 1.0, U, save=0.25, obs=[("z", i) for i in range(1, N + 1)])
    rhos = lindblad_dense(np.outer(v0, v0.conj()), H, J, N, times)
    Z = [dense_hamiltonian(N, [(["z"], [i], 1.0)]) for i in range(1, N + 1)]
    for k in range(1, len(times)):
        exact = np.array([np.trace(z @ rhos[k]).real for z in Z])
        avg = meas[:, k, :].real.mean(0)
        se = meas[:, k, :].real.std(0) / np.sqrt(M)
        assert np.all(np.abs(avg - exact) <= 4 * se + gamma * dt), (times[k], avg, exact)
    # decay is visible: the field alone would keep <sz> near 1 - O(g^2 t^2)
    assert exact.max() < 0.2


def test_uniform_order_fixes_the_jump_step():
    """No Hamiltonian, one site decaying at rate gamma from |up>: after normalisation the
    state stays |up>, so the rate is gamma and the jump probability per step is
    1 - exp(-gamma dt). Each quiet step draws two uniforms (qjmc.jl:64, :97); the jump step
    draws a third (:99); after the jump the rate is 0 and nothing more happens."""
    N, gamma, dt, k = 1, 0.7, 0.1, 4
    J = [(["s-"], [1], gamma ** 0.5)]
    v0 = np.array([1.0, 0.0], complex)
    row = [0.5, 0.0] * (k - 1) + [0.5, 0.999999, 0.3] + [0.5, 0.999999] * 6
    jumps, jt, times, meas = qjmc_dense(v0, [], J, N, dt, 1.0, np.array([row]), obs=[("z", 1)])
    assert jumps[0] == [1] and abs(jt[0][0] - k * dt) < 1e-12
    z = meas[0, :, 0].real
    assert np.allclose(z[:k], 1.0) and np.allclose(z[k:], -1.0)
    # the same with the jump test uniform just below exp(-gamma dt): no jump
    row2 = [0.5, np.exp(-gamma * dt) - 1e-9] * 10
    jumps2, *_ = qjmc_dense(v0, [], J, N, dt, 1.0, np.array([row2]))
    assert jumps2[0] == []


def test_rates_two_site_jump():
    """A two-site jump operator s-_1 s-_2 has rate c^2 <pu_1 pu_2>."""
    rng = np.random.default_rng(2)
    N = 3
    v = rng.standard_normal(8) + 1j * rng.standard_normal(8)
    v /= np.linalg.norm(v)
    r = emission_rates_dense(v[:, None], [(["s-", "s-"], [1, 2], 0.5)], N)
    P = dense_hamiltonian(N, [(["pu", "pu"], [1, 2], 1.0)])
    assert abs(r[0, 0] - 0.25 * np.vdot(v, P @ v).real) < 1e-14


def test_inner_oplist_env_matches_brute_force():
    """The environment form of inner(st, psi, oplist, psi) against the brute-force sweep
    and the dense state, with one-site, neighbouring and gapped two-site terms."""
    from oracle.exact import OPS, dense_hamiltonian
    from oracle.mps import inner_oplist, inner_oplist_env, random_mps, to_dense

    rng = np.random.default_rng(5)
    psi = random_mps(6, 2, 4, rng)
    psi[3] = psi[3] * (1 + 0.5j)  # complex, unnormalised
    terms = [([OPS["z"]], [2], 1.0), ([OPS["x"], OPS["y"]], [3, 4], 0.5), ([OPS["s+"], OPS["s-"]], [1, 5], -2.0)]
    a, b = inner_oplist_env(psi, terms), inner_oplist(psi, terms)
    assert np.allclose(a, b, atol=1e-13)
    v = to_dense(psi)
    for (mats, sites, c), x in zip(terms, a):
        O = dense_hamiltonian(6, [(mats, sites, c)])
        assert abs(x - np.vdot(v, O @ v)) < 1e-12


@pytest.mark.parametrize("seed", [3, 4])
def test_mps_trajectory_matches_dense(seed):
    """The MPS restatement of qjmc_simulation (oracle.qjmc.qjmc_mps) against the dense
    trajectory on the same uniforms, without truncation: identical jump records and times,
    <sz_i> at every save point to 1e-10. N = 6, dissipative TFIM with strong decay so that
    the case contains jumps."""
    from oracle.qjmc import qjmc_dense, qjmc_mps

    N, g, gamma, dt, tmax = 6, 0.5, 0.5, 0.05, 1.0
    H = [(["z", "z"], [i, i + 1], -1.0) for i in range(1, N)] + [(["x"], [i], -g) for i in range(1, N + 1)]
    J = [(["s-"], [i], gamma ** 0.5) for i in range(1, N + 1)]
    U = np.random.default_rng(seed).random((2, 200))
    v0 = np.zeros(2 ** N, complex)
    v0[0] = 1.0
    obs = [("z", i) for i in range(1, N + 1)]
    jr, tr, times, meas = qjmc_dense(v0, H, J, N, dt, tmax, U, save=0.25, obs=obs)
    assert sum(len(j) for j in jr) >= 1
    up = np.array([1.0, 0.0])
    for m in range(U.shape[0]):
        j, t, tm, ms, chi = qjmc_mps([up] * N, H, J, N, dt, tmax, U[m], save=0.25, obs=obs, cutoff=0.0)
        assert j == jr[m]
        assert np.allclose(t, tr[m], atol=1e-12) and np.allclose(tm, times)
        assert np.abs(ms - meas[m]).max() <= 1e-10
