"""Quantum-jump Monte Carlo on the GPU: qjmc_gates, qjmc_simulation, qjmc_emission_rates
and the QJMC observers of src/algorithms/mps/qjmc.jl.

Per trajectory the MPS stays on the device: the Trotter gates of exp(dt (-iH - sum L^H L / 2))
run through tnh_applygate2 / tnh_apply_site_op, the emission rates c^2 <psi|L^H L|psi> come
from tnh_inner_oplist, and a jump is applyop! followed by the two centre sweeps of
qjmc.jl:103-107. Random numbers come from a caller-supplied source (`rand`, a callable
returning one uniform in [0, 1)), consumed in the order of the reference's rand() calls so
that a trajectory is reproducible from its uniforms.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .models import GMPS, OpList, Sitetypes
from .tebd import DeviceMPS, applygates, entropy, inner, trotterize


def qjmc_gates(st: Sitetypes, H: OpList, jumpops: OpList, dt, order=2, evol="imag"):
    """qjmc_gates(st, H, jumpops, dt; kwargs), qjmc.jl:9-26: the escape operators
    c^2 dag(O) O and the gates of Heff = -i H - escape / 2. Returns (Heff, gates)."""
    escapeops = OpList(jumpops.length)
    for ops, sites, c in zip(jumpops.ops, jumpops.sites, jumpops.coeffs):
        names = [st.opprod([st.dag(o), o]) for o in ops]
        escapeops.add(names, sites, c ** 2)
    Heff = (-1j) * H + (-0.5) * escapeops
    gates = trotterize(st, Heff, dt, order=order, evol=evol)
    return Heff, gates


def _escape_oplist(st: Sitetypes, jumpops: OpList) -> OpList:
    ol = OpList(jumpops.length)
    for ops, sites, c in zip(jumpops.ops, jumpops.sites, jumpops.coeffs):
        ol.add([st.opprod([st.dag(o), o]) for o in ops], sites, c ** 2)
    return ol


def qjmc_emission_rates(st: Sitetypes, psi: DeviceMPS, jumpops: OpList) -> np.ndarray:
    """qjmc_emission_rates(st, psi, jumpops), qjmc.jl:170-220: |c^2 <psi|dag(O) O|psi>| for
    every jump term (psi normalised), computed on the GPU."""
    return np.abs(inner(st, psi, _escape_oplist(st, jumpops)))


def _applyop(st: Sitetypes, psi: DeviceMPS, ops, sites):
    """applyop!(st, psi, ops, sites), mps.jl:141-151 (coefficient 1)."""
    for o, s in zip(ops, sites):
        psi.apply_site_op(s, st.op(o))


def qjmc_simulation(st: Sitetypes, psi, H: OpList, jumpops: OpList, tmax, dt, observers=(), save=0.0,
                    cutoff=1e-12, mindim=1, maxdim=0, classical=True, verbose=True, rand=None,
                    order=2, evol="imag"):
    """qjmc_simulation(st, psi, H, jumpops, tmax, dt, observers; kwargs), qjmc.jl:28-167.
    psi: a GMPS (updated in place at the end) or a DeviceMPS. Returns (jumps, jumptimes)
    with 1-based jump indices, as the reference stores them."""
    if rand is None:
        gen = np.random.default_rng()
        rand = gen.random
    host = psi if isinstance(psi, GMPS) else None
    dpsi = DeviceMPS.from_gmps(psi) if host is not None else psi
    save = dt if save == 0 else save
    steps = int(math.ceil(round(tmax / dt, 5)))
    savesteps = int(math.ceil(round(save / dt)))
    jumps, jumptimes = [], []
    _, gates = qjmc_gates(st, H, jumpops, dt, order=order, evol=evol)
    N = len(dpsi)
    time = 0.0
    for ob in observers:
        ob.measure(time, dpsi, jumps, jumptimes)

    def jump_to(rates, er):
        r = rand()
        cum = np.cumsum(rates) / er
        hit = np.nonzero(r < cum)[0]
        if hit.size == 0:
            raise ValueError("no jump channel selected")
        idx = int(hit[0])
        dpsi.movecenter(1)
        _applyop(st, dpsi, jumpops.ops[idx], jumpops.sites[idx])
        dpsi.movecenter(N)
        dpsi.movecenter(1, cutoff, maxdim, mindim)
        dpsi.normalize()
        jumps.append(idx + 1)
        jumptimes.append(time + dt)

    for i in range(1, steps + 1):
        applygates(dpsi, gates, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
        r = rand()
        if not classical:
            prob = dpsi.norm() ** 2
            dpsi.normalize()
            if r > prob:
                rates = qjmc_emission_rates(st, dpsi, jumpops)
                jump_to(rates, rates.sum())
        else:
            dpsi.normalize()
            rates = qjmc_emission_rates(st, dpsi, jumpops)
            er = rates.sum()
            prob = math.exp(-er * dt)
            if rand() > prob:
                jump_to(rates, er)
        time += dt
        if i % savesteps == 0:
            if verbose:
                print(f"time={time:.5f}, jumps={len(jumps)}, maxbonddim={dpsi.maxbonddim()}, "
                      f"activity={len(jumps) / time:.5f} ")
            for ob in observers:
                ob.measure(time, dpsi, jumps, jumptimes)
    if host is not None:
        g = dpsi.to_gmps()
        host.tensors, host.center = g.tensors, g.center
    return jumps, jumptimes


def _cplx_flat(m):
    """Column-major interleaved complex doubles of a matrix or tensor."""
    a = np.asarray(m, dtype=np.complex128).reshape(-1, order="F")
    return np.ascontiguousarray(a).view(np.float64)


def qjmc_trajectories(st: Sitetypes, init_names, H: OpList, jumpops: OpList, tmax, dt, uniforms,
                      observables: OpList | None = None, save=0.0, cutoff=1e-12, mindim=1, maxdim=0,
                      order=2, nworkers=4, maxjumps=256, engine="auto", ctx=None):
    """Many trajectories of qjmc_simulation (qjmc.jl:28-167, classical branch) run natively by
    tnh_qjmc_run: the step loop in C++, every tensor operation on the GPU, `nworkers` threads
    each with its own stream. All trajectories start from productMPS(st, init_names);
    trajectory t consumes uniforms[t] in the reference's rand() order. Jump operators must
    be single-site (multi-site ones run through qjmc_simulation). engine: "batched" advances
    all trajectories together (tnh_qjmc_run_batched: one workgroup per trajectory in every
    launch, narrow slots of bond dimension min(16, 32 / d) and, for the trajectories that
    outgrow them, wide slots of up to 512 / d); "workers" runs them one by one on `nworkers`
    streams (tnh_qjmc_run, no bond limit, d <= 8); "auto" runs the batched engine and reruns
    on the workers only the trajectories whose bond dimension outgrew the wide slots too
    (each with its own uniforms; "engine" then reads "batched+workers"). Returns a dict with
    `jumps` and `jumptimes` (lists per trajectory, 1-based indices, as the reference stores
    them), `measurements` (ntraj x nsave x nobs, coefficient * <O>), `times`, `maxbonddim`."""
    from ._lib import QjmcProblem, TnhError
    from .core import Context

    ctx = ctx or Context.default()
    N, d = jumpops.length, st.dim
    if len(init_names) != N or H.length != N:
        raise ValueError("OpList and MPS have different lengths.")
    for sites in jumpops.sites:
        if len(sites) != 1:
            raise NotImplementedError("multi-site jump operators: use qjmc_simulation")
    obs = observables if observables is not None else OpList(N)
    for sites in obs.sites:
        if len(sites) != 1:
            raise NotImplementedError("multi-site observables: use qjmc_simulation with QJMCOperators")
    _, gates = qjmc_gates(st, H, jumpops, dt, order=order)
    layer_count = [len(s) for s in gates.sites]
    gsite, grange, goff, gdata, off = [], [], [], [], 0
    for sites, gts in zip(gates.sites, gates.gates):
        for s, g in zip(sites, gts):
            gsite.append(int(s))
            grange.append(g.ndim // 2)
            goff.append(off)
            flat = _cplx_flat(g)
            gdata.append(flat)
            off += flat.size // 2
    gdata = np.concatenate(gdata) if gdata else np.zeros(2)
    jsite = [s[0] for s in jumpops.sites]
    jcoeff = np.array([float(np.real(c)) for c in jumpops.coeffs])
    if any(abs(np.imag(c)) > 0 for c in jumpops.coeffs):
        raise ValueError("jump coefficients must be real")
    jop = np.concatenate([_cplx_flat(st.op(o[0])) for o in jumpops.ops]) if jumpops.ops else np.zeros(2)
    osite = [s[0] for s in obs.sites]
    oop = np.concatenate([_cplx_flat(st.op(o[0])) for o in obs.ops]) if obs.ops else np.zeros(2)
    init = np.concatenate([_cplx_flat(st.state(n)) for n in init_names])
    save = dt if save == 0 else save
    nsteps = int(math.ceil(round(tmax / dt, 5)))
    savesteps = int(math.ceil(round(save / dt)))
    U = np.ascontiguousarray(np.atleast_2d(np.asarray(uniforms, dtype=np.float64)))
    ntraj, nu = U.shape

    def iarr(x):
        return (C.c_int * max(1, len(x)))(*x)

    keep = [np.ascontiguousarray(a) for a in (gdata, jcoeff, jop, oop, init)]
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    p = QjmcProblem()
    p.N, p.d, p.nlayers = N, d, len(layer_count)
    p.layer_count, p.gate_site, p.gate_range = iarr(layer_count), iarr(gsite), iarr(grange)
    p.gate_offset = (C.c_int64 * max(1, len(goff)))(*goff)
    p.gate_data = dp(keep[0])
    p.njump, p.jump_site, p.jump_coeff, p.jump_op = len(jsite), iarr(jsite), dp(keep[1]), dp(keep[2])
    p.nobs, p.obs_site, p.obs_op = len(osite), iarr(osite), dp(keep[3])
    p.init_state = dp(keep[4])
    p.dt, p.nsteps, p.savesteps = float(dt), nsteps, savesteps
    p.cutoff, p.maxdim, p.mindim = float(cutoff), int(maxdim), int(mindim)
    nsave = nsteps // savesteps + 1
    meas = np.zeros((ntraj, nsave, max(1, len(osite)), 2))
    njumps = np.zeros(ntraj, dtype=np.int32)
    jumps = np.zeros((ntraj, maxjumps), dtype=np.int32)
    jsteps = np.zeros((ntraj, maxjumps), dtype=np.int32)
    maxb = np.zeros(ntraj, dtype=np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    if engine not in ("auto", "batched", "workers"):
        raise ValueError("engine must be 'auto', 'batched' or 'workers'")
    used = "workers"
    if engine in ("auto", "batched"):
        ctx.call("tnh_qjmc_run_batched", C.byref(p), ntraj, dp(U), nu, dp(meas), ip(njumps), ip(jumps),
                 ip(jsteps), int(maxjumps), ip(maxb))
        used = "batched"
        over = np.flatnonzero(maxb < 0)  # trajectories that outgrew the batched slots
        if over.size:
            if engine == "batched":
                raise TnhError(-7, f"qjmc batched: {over.size} trajectories need a bond dimension above "
                                   f"the slot capacity (first: {int(over[0])}); use engine='auto' or "
                                   "'workers'", "tnh_qjmc_run_batched")
            # rerun only those on the per-trajectory runner, each with its own uniforms
            Us = np.ascontiguousarray(U[over])
            m2 = np.zeros((over.size,) + meas.shape[1:])
            nj2 = np.zeros(over.size, dtype=np.int32)
            j2 = np.zeros((over.size, maxjumps), dtype=np.int32)
            s2 = np.zeros((over.size, maxjumps), dtype=np.int32)
            b2 = np.zeros(over.size, dtype=np.int32)
            ctx.call("tnh_qjmc_run", C.byref(p), int(over.size), dp(Us), nu, int(nworkers), dp(m2), ip(nj2),
                     ip(j2), ip(s2), int(maxjumps), ip(b2))
            meas[over], njumps[over], jumps[over], jsteps[over], maxb[over] = m2, nj2, j2, s2, b2
            used = "batched+workers"
    if used == "workers":
        ctx.call("tnh_qjmc_run", C.byref(p), ntraj, dp(U), nu, int(nworkers), dp(meas), ip(njumps), ip(jumps),
                 ip(jsteps), int(maxjumps), ip(maxb))
    coeffs = np.array(obs.coeffs, dtype=complex) if obs.coeffs else np.zeros(0)
    vals = meas[..., 0] + 1j * meas[..., 1]
    vals = vals[:, :, : len(osite)] * coeffs[None, None, :]
    nj = njumps.tolist()
    return {
        "jumps": [jumps[t, : min(nj[t], maxjumps)].tolist() for t in range(ntraj)],
        "jumptimes": [(jsteps[t, : min(nj[t], maxjumps)] * dt).tolist() for t in range(ntraj)],
        "njumps": nj,
        "measurements": vals,
        "times": np.arange(nsave) * savesteps * dt,
        "maxbonddim": maxb.tolist(),
        "engine": used,
    }


class QJMCOperators:
    """QJMCOperators(oplist, st), qjmc.jl:238-260: inner(st, psi, oplist, psi) at every
    save point."""

    def __init__(self, oplist: OpList, st: Sitetypes):
        self.oplist, self.st = oplist, st
        self.times, self.measurements = [], []

    def measure(self, time, psi, jumps, jumptimes):
        self.times.append(time)
        self.measurements.append(inner(self.st, psi, self.oplist))


class QJMCActivity:
    """QJMCActivity(), qjmc.jl:264-285: time and number of jumps so far."""

    def __init__(self):
        self.time, self.jumps = 0.0, 0

    def measure(self, time, psi, jumps, jumptimes):
        self.time, self.jumps = time, len(jumps)


class QJMCEntropy:
    """QJMCEntropy(), qjmc.jl:293-310: the entanglement entropy of every bond."""

    def __init__(self):
        self.times, self.measurements = [], []

    def measure(self, time, psi, jumps, jumptimes):
        self.times.append(time)
        self.measurements.append([entropy(psi, i) for i in range(1, len(psi))])
