"""MPS operations of the reference, executed on the GPU through the C-ABI:
random/product states, centre moves with truncation, the two-site split and DMRG.

Mirrors the names and keyword arguments of src/structures/mps/{gmps,mps}.jl and
src/algorithms/mps/dmrg.jl. Site tensors travel as Fortran-ordered arrays."""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import DmrgParams
from .core import Context, Tensor, precision_code
from .models import GMPS


def _upload_all(ts, ctx):
    return [Tensor.from_numpy(t, ctx=ctx) for t in ts]


def _handles(ts):
    arr = (C.c_void_p * len(ts))(*[t.h.value for t in ts])
    return arr


def movecenter(psi: GMPS, idx: int, cutoff=0.0, maxdim=0, mindim=1, ctx: Context | None = None):
    """movecenter!(psi, idx; cutoff, maxdim, mindim), src/structures/mps/gmps.jl:90-112."""
    ctx = ctx or Context.default()
    dev = _upload_all(psi.tensors, ctx)
    arr = _handles(dev)
    try:
        ctx.call("tnh_mps_movecenter", len(psi), arr, int(psi.center), int(idx), float(cutoff),
                 int(maxdim), int(mindim))
    finally:
        # the library replaces the handles in place as it moves (also when it fails part-way,
        # the array then holds the moved and the untouched sites): take ownership from it
        out = [Tensor(ctx, C.c_void_p(arr[i])) for i in range(len(psi))]
        for t in dev:
            t.h = None  # freed or reused by the library
    psi.tensors = [t.numpy() for t in out]
    psi.center = idx
    return psi


def normalize(psi: GMPS):
    """normalize!(psi), gmps.jl:46-51 (host arithmetic on one tensor)."""
    if psi.center == 0:
        movecenter(psi, 1)
    A = psi[psi.center]
    psi[psi.center] = A / np.sqrt(np.vdot(A, A).real)
    return psi


def randomMPS(dim: int, length: int, bonddim: int, rng=None) -> GMPS:
    """randomMPS(dim, length, bonddim) -> randomGMPS(1, ...), src/structures/mps/mps.jl:32-34,
    gmps.jl:276-291: N(0,1) sites, centre moved to N then to 1 (GPU SVDs), first tensor
    redrawn, normalised. `rng` is a numpy Generator (Julia's RNG stream is not available)."""
    rng = rng if rng is not None else np.random.default_rng()
    ts = []
    for i in range(1, length + 1):
        D1 = 1 if i == 1 else bonddim
        D2 = 1 if i == length else bonddim
        ts.append(rng.standard_normal((D1, dim, D2)))
    psi = GMPS(1, dim, ts, 0)
    movecenter(psi, length)
    movecenter(psi, 1)
    psi[1] = rng.standard_normal((1, dim, min(dim, bonddim)))
    normalize(psi)
    return psi


def replacesites(theta, direction: bool, normalize_: bool = True, cutoff=0.0, maxdim=0, mindim=1):
    """replacesites!(psi, theta, site, direction, normalize; ...), gmps.jl:199-267 on a host
    theta of two sites, returning (A1, A2, schmidt values, discarded weight); for three to six
    sites (the r-site branch, :215-257) (list of site tensors, None, largest discarded
    weight)."""
    from . import ops

    t = Tensor.from_numpy(theta)
    n = np.ndim(theta) - 2
    if n == 2:
        A1, A2, s, disc = ops.replacesites2(t, direction, normalize_, cutoff, maxdim, mindim)
        return A1.numpy(), A2.numpy(), s, disc
    if n < 2:
        raise ValueError("replacesites: theta must hold at least two sites")
    ctx = t.ctx
    arr = (C.c_void_p * n)()
    ks = (C.c_int * (n - 1))()
    disc = C.c_double()
    try:
        ctx.call("tnh_replacesites_sites", n, t.h, int(bool(direction)), int(bool(normalize_)),
                 float(cutoff), int(maxdim), int(mindim), arr, ks, C.byref(disc))
    finally:
        out = [Tensor(ctx, C.c_void_p(arr[i])) for i in range(n) if arr[i]]
    return [o.numpy() for o in out], None, disc.value


class DMRGState:
    """Device-resident dmrg state (tnh_dmrg): psi, the Hamiltonian terms and their blocks.

    Hs mixes MPOs (rank 2: ProjMPS(psi, H, psi)) and MPSs (rank 1: the term
    coeff |phi><phi|, ProjMPS(phi, psi; squared=true)), as dmrg(psi, Hs...) does
    (dmrg.jl:139-149); the rank-1 terms are summed after the MPO terms."""

    def __init__(self, psi: GMPS, Hs, coeffs=None, ctx: Context | None = None, complex_state=False):
        self.ctx = ctx or Context.default()
        N = len(psi)
        Hs = list(Hs)
        self.N = N
        coeffs = list(coeffs) if coeffs is not None else [1.0] * len(Hs)
        if len(coeffs) != len(Hs):
            raise ValueError("coeffs must have one entry per Hamiltonian term")
        mpos = [(H, c) for H, c in zip(Hs, coeffs) if H.rank == 2]
        phis = [(H, c) for H, c in zip(Hs, coeffs) if H.rank == 1]
        if len(mpos) + len(phis) != len(Hs):
            raise ValueError("Hamiltonian must be composed of MPOs (rank 2) or MPSs (rank 1).")
        # a complex rank-1 term makes the state complex (the library keeps one dtype)
        cplx = complex_state or any(np.iscomplexobj(t) for H, _ in phis for t in H.tensors)
        pts = [np.asarray(t, dtype=np.complex128) for t in psi.tensors] if cplx else psi.tensors
        dpsi = _upload_all(pts, self.ctx)
        dW = [Tensor.from_numpy(w, ctx=self.ctx) for H, _ in mpos for w in H.tensors]
        carr = (C.c_double * max(1, len(mpos)))(*[c for _, c in mpos])
        h = C.c_void_p()
        self.ctx.call("tnh_dmrg_create", N, _handles(dpsi), len(mpos),
                      _handles(dW) if dW else None, carr, C.byref(h))
        self.h = h
        try:
            for H, c in phis:
                dphi = _upload_all(H.tensors, self.ctx)
                self.ctx.check(_lib.lib.tnh_dmrg_add_projector(h, _handles(dphi), float(c)),
                               "tnh_dmrg_add_projector")
        except Exception:
            self.close()
            raise

    def sweep(self, direction: bool, krylovdim=3, maxiter=2, lanczos_tol=1e-14, cutoff=1e-12,
              maxdim=1000, mindim=1, precision="f64", nsites=2, nonhermitian=False):
        """One pass; returns (last Ritz value, max bond dim, max discarded weight). With
        nonhermitian the Ritz value (smallest real part) is returned as a complex number."""
        p = DmrgParams(int(krylovdim), int(maxiter), float(lanczos_tol), float(cutoff), int(maxdim),
                       int(mindim), precision_code(precision), int(nsites), int(bool(nonhermitian)))
        e = C.c_double()
        D = C.c_int()
        disc = C.c_double()
        self.ctx.check(_lib.lib.tnh_dmrg_sweep(self.h, int(bool(direction)), C.byref(p), C.byref(e),
                                               C.byref(D), C.byref(disc)), "tnh_dmrg_sweep")
        if nonhermitian:
            re, im = C.c_double(), C.c_double()
            self.ctx.check(_lib.lib.tnh_dmrg_last_ritz(self.h, C.byref(re), C.byref(im)),
                           "tnh_dmrg_last_ritz")
            return complex(re.value, im.value), D.value, disc.value
        return e.value, D.value, disc.value

    def energy(self):
        from ._lib import lib

        e = C.c_double()
        self.ctx.check(lib.tnh_dmrg_energy(self.h, 0, C.byref(e)), "tnh_dmrg_energy")
        return e.value

    def bonddims(self):
        from ._lib import lib

        dims = (C.c_int64 * (self.N + 1))()
        lib.tnh_dmrg_bonddims(self.h, dims)
        return list(dims)

    def center(self):
        from ._lib import lib

        c = C.c_int()
        lib.tnh_dmrg_center(self.h, C.byref(c))
        return c.value

    def site(self, i) -> np.ndarray:
        from ._lib import lib

        b = C.c_void_p()
        self.ctx.check(lib.tnh_dmrg_site(self.h, int(i), C.byref(b)), "tnh_dmrg_site")
        t = Tensor(self.ctx, b)
        try:
            return t.numpy()
        finally:
            t.h = None  # borrowed

    def tensors(self):
        return [self.site(i) for i in range(1, self.N + 1)]

    def close(self):
        from ._lib import lib

        if self.h:
            lib.tnh_dmrg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dmrg(psi: GMPS, *Hs: GMPS, nsites=2, krylovdim=3, kryloviter=2, ishermitian=True, minsweeps=1,
         maxsweeps=1000, tol=1e-10, tolgrad=1e-5, numconverges=4, verbose=True, cutoff=1e-12,
         maxdim=1000, mindim=1, coeffs=None, precision="f64", refine_sweeps=2, refine_tol=None,
         history=None, state_out=None, checkpoint=None):
    """dmrg(psi, H::GMPS...; kwargs...), src/algorithms/mps/dmrg.jl:128-154 and :1-99.

    psi is updated in place and (psi, cost) returned, as in the reference. The whole bond
    step runs on the GPU (tnh_dmrg_sweep). ishermitian=False solves each local problem by
    Arnoldi for the eigenvalue of smallest real part (KrylovKit's eigsolve with
    ishermitian=false) on a complex copy of psi; the returned cost is then complex.
    Additional keywords of this implementation:
      precision      "f64" | "f32" | "bf16x3": arithmetic of the matvecs and environment
                     updates; Krylov vectors, the Gram/SVD and reductions stay FP64.
      refine_sweeps  with a lower precision, switch to FP64 for the final passes: once
                     the relative energy change falls below refine_tol (default: just above
                     the precision's noise floor, measured on the critical TFIM at N = 100:
                     1e-5 for "f32", whose passes stall at ~1e-6, and 1e-4 for "bf16x3") or
                     when only refine_sweeps passes remain before maxsweeps; at least
                     refine_sweeps FP64 passes are run before stopping, except that
                     maxsweeps always ends the run (with maxsweeps <= refine_sweeps every
                     pass is FP64).
      checkpoint     a path: after every pass psi is written there in the reference's
                     HDF5 layout (HDF5.write of GMPS, gmps.jl:359-370, group "psi"); a run
                     resumes by reading it back (tnhip.checkpoint.read_mps) and calling dmrg.
    """
    if nsites not in (1, 2):
        raise ValueError("nsites must be 1 or 2")
    if psi.rank != 1:
        raise ValueError("Psi must be a GMPS of rank 1 (vector).")
    if len(Hs) == 0:
        raise ValueError("You must provide atleast one MPS/MPO for the Hamiltonian.")
    for H in Hs:
        if len(H) != len(psi) or H.dim != psi.dim:
            raise ValueError("GMPS must share the same properties.")
        if H.rank not in (1, 2):
            raise ValueError("Hamiltonian must be composed of MPOs (rank 2) or MPSs (rank 1).")
    st = DMRGState(psi, Hs, coeffs, complex_state=not ishermitian)
    if state_out is not None:
        state_out.append(st)
    cost = st.energy()                                    # :22
    lastcost = cost
    D = psi.maxbonddim()
    lastD = D
    grad = 0.0
    direction = False
    converged = False
    convergedsweeps = convergedgrad = 0
    sweeps = 0
    prec = precision
    fp64_passes = 0
    switch_tol = refine_tol if refine_tol is not None else {"f64": 0.0, "f32": 1e-5, "bf16x3": 1e-4}[precision]

    def diff(x, y):
        return abs(x - y) if abs(x) < 1e-10 else abs((x - y) / x)

    while not converged:
        if prec != "f64" and maxsweeps and sweeps >= maxsweeps - refine_sweeps:
            prec = "f64"  # the last refine_sweeps passes before maxsweeps (all, if fewer)
        t_sweep = time.perf_counter()
        pass_prec = prec
        cost, D, disc = st.sweep(direction, krylovdim, kryloviter, 1e-14, cutoff, maxdim, mindim, prec,
                                 nsites, not ishermitian)
        t_sweep = time.perf_counter() - t_sweep  # the sweep returns host scalars: synchronised
        if prec == "f64":
            fp64_passes += 1
        direction = not direction
        sweeps += 1                                       # :66
        if prec != "f64" and sweeps >= 2 and diff(cost, lastcost) < switch_tol:
            prec = "f64"                                  # reached the FP32 floor
        if sweeps >= minsweeps:                           # :69-86
            dc = diff(cost, lastcost)
            convergedsweeps = convergedsweeps + 1 if (dc < tol and lastD == D) else 0
            if (dc + grad) != 0 and abs((dc - grad) / (dc + grad)) < tolgrad and lastD == D:
                convergedgrad += 1
            else:
                convergedgrad = 0
            if max(convergedsweeps, convergedgrad) >= numconverges:
                converged = True
            if sweeps >= maxsweeps and maxsweeps != 0:
                converged = True
            # finish with FP64 passes, but never beyond maxsweeps (dmrg.jl:80-82 stops there)
            if (converged and precision != "f64" and fp64_passes < refine_sweeps
                    and not (maxsweeps and sweeps >= maxsweeps)):
                converged = False
                prec = "f64"
        grad = abs(diff(cost, lastcost))
        lastcost = cost
        lastD = D
        if history is not None:
            history.append({"sweep": sweeps, "energy": cost, "maxbonddim": D, "discarded": disc,
                            "precision": pass_prec, "seconds": t_sweep})
        if checkpoint is not None:
            from .checkpoint import write_mps

            write_mps(checkpoint, GMPS(1, psi.dim, st.tensors(), st.center()))
        if verbose:
            print(f"Sweep={sweeps}, energy={np.real(cost):.12f}, maxbonddim={D} ")
    psi.tensors = st.tensors()
    psi.center = st.center()
    if state_out is None:
        st.close()
    return psi, cost
