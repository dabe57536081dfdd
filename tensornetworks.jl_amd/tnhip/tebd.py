"""TEBD on the GPU: trotterize, applygate(s), tebd and its observers, inner, entropy.

Mirrors src/structures/mps/gatelist.jl, src/algorithms/mps/tebd.jl,
src/structures/mps/mps.jl:87-151 and gmps.jl:184-189. The MPS stays on the device
(DeviceMPS); gates are built once on the host (matrix exponentials of the local terms,
as the reference's setup does) and applied by tnh_applygate2 / tnh_apply_site_op.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import scipy.linalg as sla

from ._lib import lib
from .core import Context, Tensor
from .models import GMPS, OpList, Sitetypes


# ------------------------------------------------------------------ device MPS

class DeviceMPS:
    """A rank-1 GMPS whose site tensors live on the GPU (1-based indexing, as GMPS)."""

    def __init__(self, tensors, center=0, dim=None, ctx: Context | None = None):
        self.ctx = ctx or Context.default()
        self.sites = [t if isinstance(t, Tensor) else Tensor.from_numpy(t, ctx=self.ctx) for t in tensors]
        self.center = center
        self.dim = dim if dim is not None else self.sites[0].shape[1]
        self.rank = 1

    @classmethod
    def from_gmps(cls, psi: GMPS, ctx=None):
        return cls(list(psi.tensors), psi.center, psi.dim, ctx)

    def to_gmps(self) -> GMPS:
        return GMPS(1, self.dim, [t.numpy() for t in self.sites], self.center)

    def copy(self) -> "DeviceMPS":
        return DeviceMPS([t.astype(t.dtype) for t in self.sites], self.center, self.dim, self.ctx)

    def __len__(self):
        return len(self.sites)

    def __getitem__(self, i):
        return self.sites[i - 1]

    def bonddims(self):
        return [t.shape[2] for t in self.sites[:-1]]

    def maxbonddim(self):
        """maxbonddim(psi), src/structures/mps/abstractmps.jl:75-81."""
        return max([0] + self.bonddims())

    def _handles(self):
        return (C.c_void_p * len(self.sites))(*[t.h.value for t in self.sites])

    def _adopt(self, arr):
        for i, t in enumerate(self.sites):
            if t.h.value != arr[i]:
                t.h = None  # freed by the library
                self.sites[i] = Tensor(self.ctx, C.c_void_p(arr[i]))

    def movecenter(self, idx, cutoff=0.0, maxdim=0, mindim=1):
        """movecenter!(psi, idx; cutoff, maxdim, mindim), gmps.jl:90-112."""
        arr = self._handles()
        try:
            self.ctx.call("tnh_mps_movecenter", len(self), arr, int(self.center), int(idx), float(cutoff),
                          int(maxdim), int(mindim))
        finally:
            self._adopt(arr)  # moved sites are replaced in place, also before a failure
        self.center = idx

    def norm(self):
        """norm(psi), gmps.jl:29-37."""
        if self.center == 0:
            self.movecenter(1)
        v = C.c_double()
        self.ctx.call("tnh_buf_norm", self[self.center].h, C.byref(v))
        return v.value

    def normalize(self):
        """normalize!(psi), gmps.jl:46-51."""
        n = self.norm()
        self.ctx.call("tnh_buf_scale", self[self.center].h, 1.0 / n, 0.0)

    def apply_site_op(self, site, op):
        """A <- op A at one site (applyop!, mps.jl:141-151)."""
        G = Tensor.from_numpy(np.asarray(op), ctx=self.ctx)
        h = C.c_void_p(self.sites[site - 1].h.value)
        self.ctx.call("tnh_apply_site_op", C.byref(h), G.h)
        self.sites[site - 1].h = None
        self.sites[site - 1] = Tensor(self.ctx, h)


# ------------------------------------------------------------------ gates

class GateList:
    """GateList(length, sites, gates), src/structures/mps/gatelist.jl:8-13."""

    def __init__(self, length):
        self.length = length
        self.sites: list[list[int]] = []
        self.gates: list[list[np.ndarray]] = []
        self._dev: list[list[Tensor]] = []

    def add(self, sites, gates):
        """add!(gl, sites, gates), gatelist.jl:40-62."""
        order = np.argsort(sites, kind="stable")
        sites = [sites[i] for i in order]
        gates = [gates[i] for i in order]
        if len(sites) != len(gates):
            raise ValueError("The site and gate list must be the same length.")
        for i in range(len(sites) - 1):
            if sites[i] + gatesize(gates[i]) - 1 >= sites[i + 1]:
                raise ValueError("Gates in one sequence cannot be shared by sites.")
        if sites and sites[-1] + gatesize(gates[-1]) - 1 > self.length:
            raise ValueError("The bond gates cannot exceed the length of the lattice.")
        self.sites.append(sites)
        self.gates.append(gates)
        self._dev.append([None] * len(gates))

    def device_gate(self, row, i, ctx):
        if self._dev[row][i] is None:
            self._dev[row][i] = Tensor.from_numpy(self.gates[row][i], ctx=ctx)
        return self._dev[row][i]


def gatesize(gate):
    """gatesize(gate), gatelist.jl:28-32."""
    if gate.ndim % 2 == 1:
        raise ValueError("The gate must have even number of dimensions.")
    return gate.ndim // 2


def expm_tensor(x, outeridxs):
    """exp(x, outeridxs), src/tensors.jl:319-338: matrix exponential with the listed
    ("in") indices as columns."""
    n = x.ndim
    inner = [i for i in range(1, n + 1) if i not in outeridxs]
    perm = [i - 1 for i in inner] + [i - 1 for i in outeridxs]
    y = np.transpose(x, perm)
    ish, osh = y.shape[: len(inner)], y.shape[len(inner):]
    M = y.reshape((int(np.prod(ish)), int(np.prod(osh))), order="F")
    E = sla.expm(M).reshape(ish + osh, order="F")
    return np.asfortranarray(np.transpose(E, np.argsort(perm)))


def trotterize(st: Sitetypes, ops: OpList, dt, order=2, evol="imag") -> GateList:
    """trotterize(st, ops, dt; order, evol), src/structures/mps/gatelist.jl:75-121."""
    gl = GateList(ops.length)
    rng = ops.siterange()
    order = 1 if rng == 1 else order
    if order < 0 or order > 2:
        raise ValueError("Only trotter order 1 and 2 are supported.")
    dt = -1j * dt if evol == "real" else dt
    for i in range(1, rng + 1):
        time = dt / 2 if (i < rng and order == 2) else dt
        gates, sites = [], []
        site = i
        while site <= ops.length:
            g = ops.sitetensor(st, site)
            if g is not None:
                g = expm_tensor(time * g, [2 * k for k in range(1, gatesize(g) + 1)])
                gates.append(g)
                sites.append(site)
            site += rng
        gl.add(sites, gates)
    if order == 2:
        for i in range(1, rng):
            gl.add(gl.sites[rng - i - 1], gl.gates[rng - i - 1])
    return gl


def applygate(psi: DeviceMPS, site, gate_dev: Tensor, gate_np, direction, cutoff=0.0, maxdim=0, mindim=1):
    """applygate(psi, site, gate, direction; cutoff, maxdim, mindim), gatelist.jl:137-171
    (error=false). Two-site gates run tnh_applygate2, wider ones tnh_applygate_sites; one-site
    gates tnh_apply_site_op followed by replacesites!'s one-site branch (gmps.jl:204-213)."""
    rng = gatesize(gate_np)
    if rng == 2:
        h1 = C.c_void_p(psi[site].h.value)
        h2 = C.c_void_p(psi[site + 1].h.value)
        k = C.c_int()
        disc = C.c_double()
        try:
            psi.ctx.call("tnh_applygate2", C.byref(h1), C.byref(h2), gate_dev.h, int(bool(direction)),
                         float(cutoff), int(maxdim), int(mindim), C.byref(k), C.byref(disc))
        finally:
            for j, h in ((site, h1), (site + 1, h2)):  # replaced buffers, also before a failure
                if psi.sites[j - 1].h.value != h.value:
                    psi.sites[j - 1].h = None
                    psi.sites[j - 1] = Tensor(psi.ctx, h)
        psi.center = site if direction else site + 1
        return disc.value
    if rng == 1:
        h = C.c_void_p(psi[site].h.value)
        psi.ctx.call("tnh_apply_site_op", C.byref(h), gate_dev.h)
        psi.sites[site - 1].h = None
        psi.sites[site - 1] = Tensor(psi.ctx, h)
        psi.center = site
        nxt = site + 1 - 2 * int(bool(direction))
        if 0 < nxt <= len(psi):
            psi.movecenter(nxt)
        return 0.0
    # rng >= 3: the r-site branch of replacesites! (gmps.jl:215-257) on the device
    arr = (C.c_void_p * rng)(*[psi[site + j].h.value for j in range(rng)])
    ks = (C.c_int * (rng - 1))()
    disc = C.c_double()
    try:
        psi.ctx.call("tnh_applygate_sites", rng, arr, gate_dev.h, int(bool(direction)), float(cutoff),
                     int(maxdim), int(mindim), ks, C.byref(disc))
    finally:
        for j in range(rng):  # replaced buffers, also before a failure
            if psi.sites[site + j - 1].h.value != arr[j]:
                psi.sites[site + j - 1].h = None
                psi.sites[site + j - 1] = Tensor(psi.ctx, C.c_void_p(arr[j]))
    psi.center = site if direction else site + rng - 1
    return disc.value


def applygates(psi: DeviceMPS, gates: GateList, cutoff=0.0, maxdim=0, mindim=1):
    """applygates!(psi, gates; kwargs), gatelist.jl:191-227."""
    for row in range(len(gates.gates)):
        sites, gts = gates.sites[row], gates.gates[row]
        if not sites:
            continue
        firstsite = sites[0]
        lastsite = sites[-1] + gatesize(gts[-1]) - 1
        direction = not (abs(psi.center - firstsite) < abs(psi.center - lastsite))
        n = len(gts)
        for i in range(1, n + 1):
            g = n + 1 - i if direction else i
            ctr = sites[g - 1] if not direction else sites[g - 1] + gatesize(gts[g - 1]) - 1
            psi.movecenter(ctr, cutoff, maxdim, mindim)
            applygate(psi, sites[g - 1], gates.device_gate(row, g - 1, psi.ctx), gts[g - 1], direction,
                      cutoff, maxdim, mindim)


# ------------------------------------------------------------------ observables

def _oplist_payload(st: Sitetypes, oplist: OpList):
    firsts, ranges, mats = [], [], []
    for ops, sites in zip(oplist.ops, oplist.sites):
        s0, r = sites[0], sites[-1] - sites[0] + 1
        firsts.append(s0)
        ranges.append(r)
        for s in range(s0, s0 + r):
            m = st.op(ops[sites.index(s)]) if s in sites else st.op("id")
            m = np.asarray(m, dtype=np.complex128)
            mats.append(np.asfortranarray(m).reshape(-1, order="F"))
    flat = np.concatenate(mats).view(np.float64) if mats else np.zeros(0)
    return firsts, ranges, np.ascontiguousarray(flat)


def inner(st: Sitetypes, psi: DeviceMPS, oplist: OpList, phi: DeviceMPS | None = None) -> np.ndarray:
    """inner(st, psi, oplist, phi), src/structures/mps/mps.jl:87-134: coeff * <psi|O|phi>
    for every term (phi defaults to psi), computed on the GPU."""
    n = len(oplist.ops)
    firsts, ranges, flat = _oplist_payload(st, oplist)
    out = np.zeros(2 * max(n, 1))
    hs = psi._handles()
    args = (n, (C.c_int * max(n, 1))(*firsts), (C.c_int * max(n, 1))(*ranges),
            flat.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)))
    if phi is None or phi is psi:
        psi.ctx.call("tnh_inner_oplist", len(psi), hs, *args)
    else:
        if len(phi) != len(psi):
            raise ValueError("GMPS must share the same length.")
        psi.ctx.call("tnh_inner_oplist_pair", len(psi), hs, phi._handles(), *args)
    vals = out[0::2][:n] + 1j * out[1::2][:n]
    return np.array([c * v for c, v in zip(oplist.coeffs, vals)])


def entropy(psi: DeviceMPS, site):
    """entropy(psi, site), gmps.jl:184-189: -sum s^2 ln s^2 across (site, site + 1)."""
    from .ops import svd_trunc

    psi.movecenter(site)
    A = psi[site]
    cl, d, cr = A.shape
    M = Tensor.wrap(A.device_ptr(), (cl * d, cr), A.dtype, owner=A, ctx=psi.ctx)
    _, S, _, _, _ = svd_trunc(M, 0.0, 0, 1, want=(False, True, False))
    s2 = S.numpy() ** 2
    s2 = s2[s2 > 0]
    return float(-(s2 * np.log(s2)).sum())


# ------------------------------------------------------------------ tebd

class TEBDNorm:
    """TEBDNorm(tol), tebd.jl:114-148."""

    def __init__(self, tol=1e-10):
        self.times, self.measurements, self.tol = [], [], tol

    def measure(self, time, psi, norm, energy):
        self.times.append(time)
        self.measurements.append(norm)

    def checkdone(self):
        if len(self.times) < 3:
            return False
        E1 = (self.measurements[-1] - self.measurements[-2]) / (self.times[-1] - self.times[-2])
        E2 = (self.measurements[-2] - self.measurements[-3]) / (self.times[-2] - self.times[-3])
        if abs(E2 + E1) < 1e-8:
            return abs(E2 - E1) < self.tol
        return abs((E2 - E1) / (0.5 * (E2 + E1))) < self.tol


class TEBDEnergy:
    """TEBDEnergy(tol), tebd.jl:150-182."""

    def __init__(self, tol=1e-10):
        self.times, self.measurements, self.tol = [], [], tol

    def measure(self, time, psi, norm, energy):
        self.times.append(time)
        self.measurements.append(energy)

    def checkdone(self):
        if len(self.times) < 3:
            return False
        E1, E2 = self.measurements[-1], self.measurements[-2]
        if 0.5 * abs(E2 + E1) < 1e-8:
            return abs(E2 - E1) < self.tol
        return abs(2 * (E2 - E1) / (E2 + E1)) < self.tol


class TEBDOperators:
    """TEBDOperators(st, oplist), tebd.jl:184-219."""

    def __init__(self, st, oplist):
        self.st, self.oplist = st, oplist
        self.times, self.measurements = [], []

    def measure(self, time, psi, norm, energy):
        self.times.append(time)
        self.measurements.append(inner(self.st, psi, self.oplist))

    def checkdone(self):
        return False


def tebd(st: Sitetypes, psi, H: OpList, dt, tmax, save, observers=(), cutoff=1e-12, maxdim=0, mindim=1,
         evol="imag", order=2, norm=0.0, verbose=True):
    """tebd(st, psi, H, dt, tmax, save, observers; kwargs), src/algorithms/mps/tebd.jl:3-100
    (without projectors). psi may be a GMPS (copied to the device and updated in place at
    the end) or a DeviceMPS. Returns (psi, energy)."""
    if st.dim != (psi.dim if hasattr(psi, "dim") else psi[1].shape[1]):
        raise ValueError("Sitetypes and MPS physical dimensions do not match.")
    if len(psi) != H.length:
        raise ValueError("OpList and MPS have different lengths.")
    if evol not in ("imag", "real"):
        raise ValueError("evol must be real of imag.")
    if order not in (1, 2):
        raise ValueError("Only first or second order trotter supported.")
    host = psi if isinstance(psi, GMPS) else None
    dpsi = DeviceMPS.from_gmps(psi) if host is not None else psi
    gates = trotterize(st, H, dt, order=order, evol=evol)
    nsteps = int(round(tmax / dt))
    save = dt if save < dt else save
    nsave = int(round(save / dt))
    normal = norm
    energy = float(np.real(inner(st, dpsi, H).sum()))
    for ob in observers:
        ob.measure(0.0, dpsi, normal, energy)
    converged = False
    step = 0
    while not converged:
        applygates(dpsi, gates, cutoff=cutoff, maxdim=maxdim, mindim=mindim)
        psinorm = math.log(dpsi.norm())
        normal += psinorm
        dpsi.normalize()
        step += 1
        if step >= nsteps:
            converged = True
        if step % nsave == 0:
            if verbose:
                print(f"time={step * dt:.4f}, energy={energy:.12f}, maxbonddim={dpsi.maxbonddim()} ")
            energy = float(np.real(inner(st, dpsi, H).sum()))
            for ob in observers:
                ob.measure(step * dt, dpsi, normal, energy)
                converged = converged or ob.checkdone()
    if host is not None:
        g = dpsi.to_gmps()
        host.tensors, host.center = g.tensors, g.center
        return host, energy
    return dpsi, energy
