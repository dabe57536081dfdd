"""Host logic of qjmc_trajectories(engine="auto") without a GPU: the batched call marks the
trajectories that outgrew its slots (maxbond = -1) and only those are rerun on the
per-trajectory runner, each with its own row of uniforms, and merged back in place."""
import ctypes as C

import numpy as np


class FakeCtx:
    """Stands in for a Context: writes through the C-ABI argument pointers what the two
    entry points would, with values that identify the uniforms row they came from."""

    def __init__(self, over):
        self.over = set(over)
        self.calls = []

    def call(self, fn, *args):
        self.calls.append(fn)
        prob = args[0]._obj
        ntraj, U, nu = args[1], args[2], args[3]
        if fn == "tnh_qjmc_run_batched":
            meas, nj, jumps, jsteps, maxjumps, maxb = args[4:10]
        else:
            meas, nj, jumps, jsteps, maxjumps, maxb = args[5:11]
        nsave = prob.nsteps // prob.savesteps + 1
        nobs = max(1, prob.nobs)
        Ua = np.ctypeslib.as_array(U, shape=(ntraj, nu))
        M = np.ctypeslib.as_array(meas, shape=(ntraj, nsave, nobs, 2))
        J = np.ctypeslib.as_array(jumps, shape=(ntraj, maxjumps))
        S = np.ctypeslib.as_array(jsteps, shape=(ntraj, maxjumps))
        NJ = np.ctypeslib.as_array(nj, shape=(ntraj,))
        B = np.ctypeslib.as_array(maxb, shape=(ntraj,))
        for t in range(ntraj):
            if fn == "tnh_qjmc_run_batched":
                M[t] = Ua[t, 0]
                NJ[t] = 0
                B[t] = -1 if t in self.over else 5
            else:  # the per-trajectory runner: one jump, at step = 1 + 10 * uniform
                M[t] = Ua[t, 0] + 100.0
                NJ[t] = 1
                J[t, 0] = 1
                S[t, 0] = 1 + int(10 * Ua[t, 0])
                B[t] = 20 + t


def test_auto_engine_reruns_only_marked_trajectories():
    from tnhip import OpList, spinhalf
    from tnhip.qjmc import qjmc_trajectories

    N, ntraj = 4, 6
    H, J, obs = OpList(N), OpList(N), OpList(N)
    for i in range(1, N):
        H.add(["x", "x"], [i, i + 1], 1.0)
    for i in range(1, N + 1):
        J.add("s-", i, 0.3)
        obs.add("z", i)
    U = np.linspace(0.05, 0.9, ntraj)[:, None] * np.ones((ntraj, 40))
    over = [1, 4]
    ctx = FakeCtx(over)
    r = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, 0.2, 0.05, U, observables=obs, engine="auto", ctx=ctx)
    assert ctx.calls == ["tnh_qjmc_run_batched", "tnh_qjmc_run"]
    assert r["engine"] == "batched+workers"
    for t in range(ntraj):
        m = r["measurements"][t].real  # coefficient 1: the values written above
        if t in over:
            k = over.index(t)  # row k of the rerun
            assert np.allclose(m, U[t, 0] + 100.0)
            assert r["maxbonddim"][t] == 20 + k
            assert r["jumps"][t] == [1] and r["njumps"][t] == 1
            assert np.isclose(r["jumptimes"][t][0], (1 + int(10 * U[t, 0])) * 0.05)
        else:
            assert np.allclose(m, U[t, 0])
            assert r["maxbonddim"][t] == 5 and r["jumps"][t] == []


def test_auto_engine_without_overflow_makes_one_call():
    from tnhip import OpList, spinhalf
    from tnhip.qjmc import qjmc_trajectories

    N = 3
    H, J = OpList(N), OpList(N)
    H.add(["x", "x"], [1, 2], 1.0)
    J.add("s-", 2, 0.3)
    ctx = FakeCtx([])
    r = qjmc_trajectories(spinhalf(), ["up"] * N, H, J, 0.1, 0.05, np.full((2, 20), 0.5), engine="auto",
                          ctx=ctx)
    assert ctx.calls == ["tnh_qjmc_run_batched"] and r["engine"] == "batched"
