"""Site types, operator lists and MPO construction: the host-side setup of the
reference (src/sitetypes.jl, src/lattices/spinhalf.jl, src/structures/mps/oplist.jl,
src/structures/mps/mpo.jl:323-459). These run once per problem on the host, as the
reference's Julia setup code does; every tensor they produce is then handed to the
GPU library."""
from __future__ import annotations

import numpy as np


class Sitetypes:
    """Sitetypes(dim, statenames, states, opnames, ops, opdags), src/sitetypes.jl:1-18."""

    def __init__(self, dim: int):
        self.dim = dim
        self.states: dict[str, np.ndarray] = {}
        self.ops: dict[str, np.ndarray] = {}
        self.dags: dict[str, str] = {}
        self.temp = 0

    def add_state(self, name, vec):
        self.states[name] = np.asarray(vec, dtype=np.complex128)

    def add_op(self, name, mat, dag=None):
        self.ops[name] = np.asarray(mat, dtype=np.complex128)
        self.dags[name] = dag if dag is not None else name

    def state(self, name):
        """state(st, name), src/sitetypes.jl:26-37."""
        if name not in self.states:
            raise KeyError(f"The state {name} is undefined.")
        return self.states[name].copy()

    def op(self, name):
        """op(st, name), src/sitetypes.jl:45-55."""
        if isinstance(name, np.ndarray):
            return name
        if name not in self.ops:
            raise KeyError(f"The operator {name} is undefined.")
        return self.ops[name].copy()

    def dag(self, name):
        """dag(st, name), src/sitetypes.jl:62-72."""
        return self.dags[name]

    def opprod(self, names):
        """opprod(st, names), src/sitetypes.jl:144-166: name of the product of operators,
        registering it (and its adjoint) when new."""
        prod = self.op(names[0])
        for nm in names[1:]:
            prod = prod @ self.op(nm)
        for nm, m in self.ops.items():
            if np.abs(prod - m).sum() < 1e-12:
                return nm
        name, dagname = f"temp{self.temp}", f"temp{self.temp + 1}"
        self.add_op(name, prod, dagname)
        self.add_op(dagname, prod.conj().T, name)
        self.temp += 2
        return name


def spinhalf() -> Sitetypes:
    """spinhalf(), src/lattices/spinhalf.jl:1-23 (Pauli matrices; up = [1, 0])."""
    st = Sitetypes(2)
    st.add_state("up", [1, 0])
    st.add_state("dn", [0, 1])
    st.add_state("s", [0.5 ** 0.5, 0.5 ** 0.5])
    st.add_state("as", [0.5 ** 0.5, -(0.5 ** 0.5)])
    st.add_op("x", [[0, 1], [1, 0]], "x")
    st.add_op("y", [[0, -1j], [1j, 0]], "y")
    st.add_op("z", [[1, 0], [0, -1]], "z")
    st.add_op("id", [[1, 0], [0, 1]], "id")
    st.add_op("pu", [[1, 0], [0, 0]], "pu")
    st.add_op("pd", [[0, 0], [0, 1]], "pd")
    st.add_op("n", [[1, 0], [0, 0]], "n")
    st.add_op("s+", [[0, 1], [0, 0]], "s-")
    st.add_op("s-", [[0, 0], [1, 0]], "s+")
    return st


class OpList:
    """OpList(length): list of (ops, sites, coeff) terms, src/structures/mps/oplist.jl:1-59."""

    def __init__(self, length: int):
        self.length = length
        self.ops: list[list[str]] = []
        self.sites: list[list[int]] = []
        self.coeffs: list[complex] = []

    def __len__(self):
        return self.length

    def add(self, ops, sites, coeff=1.0):
        """add!(oplist, ops, sites, coeff), oplist.jl:33-59 (1-based sites)."""
        if isinstance(ops, str):
            ops, sites = [ops], [sites]
        ops, sites = list(ops), [int(s) for s in sites]
        if len(ops) != len(sites):
            raise ValueError("The lists must be the same length.")
        order = np.argsort(sites, kind="stable")
        sites = [sites[i] for i in order]
        ops = [ops[i] for i in order]
        last = 0
        for s in sites:
            if s < 0 or s > self.length:
                raise ValueError(f"The sites must be between 1 and {self.length}.")
            if sites.count(s) > 1:
                raise ValueError("There are two or more operators on the same site.")
            if s <= last:
                raise ValueError("The site list must be ordered.")
            last = s
        self.ops.append(ops)
        self.sites.append(sites)
        self.coeffs.append(coeff)
        return self

    def __add__(self, other):
        """add(oplist1, oplist2), oplist.jl:67-77."""
        o = OpList(max(self.length, other.length))
        o.ops, o.sites, o.coeffs = self.ops + other.ops, self.sites + other.sites, self.coeffs + other.coeffs
        return o

    def __mul__(self, x):
        o = OpList(self.length)
        o.ops, o.sites, o.coeffs = list(self.ops), list(self.sites), [c * x for c in self.coeffs]
        return o

    __rmul__ = __mul__

    def siterange(self):
        """siterange(oplist), oplist.jl:102-109."""
        return max([1] + [s[-1] - s[0] + 1 for s in self.sites])

    def siteindexs(self, site):
        """siteindexs(oplist, site), oplist.jl:117-125 (0-based term indices)."""
        return [i for i, s in enumerate(self.sites) if s[0] == site]

    def totensor(self, st: Sitetypes, idx: int):
        """totensor(oplist, st, idx), oplist.jl:134-156: G[o1, i1, o2, i2, ...]."""
        ops, sites = self.ops[idx], self.sites[idx]
        rng = min(self.siterange(), self.length - sites[0] + 1)
        mats = []
        k = 0
        for s in range(rng):
            if sites[0] + s in sites:
                mats.append(st.op(ops[k]))
                k += 1
            else:
                mats.append(st.op("id"))
        G = mats[0]
        for m in mats[1:]:
            G = np.multiply.outer(G, m)
        return self.coeffs[idx] * G

    def sitetensor(self, st: Sitetypes, site: int):
        """sitetensor(oplist, st, idx), oplist.jl:164-183 (None when no term starts here)."""
        idxs = self.siteindexs(site)
        if not idxs:
            return None
        ten = self.totensor(st, idxs[0])
        for i in idxs[1:]:
            ten = ten + self.totensor(st, i)
        return ten


def _realify(arrs):
    if all(np.abs(a.imag).max(initial=0.0) == 0.0 for a in arrs):
        return [np.asfortranarray(a.real) for a in arrs]
    return [np.asfortranarray(a) for a in arrs]


def mpo_tensors(st: Sitetypes, H: OpList, slots: list | None = None):
    """Finite-state-automaton MPO of an operator list, after MPO(st, H),
    src/structures/mps/mpo.jl:323-440: W[w, s_out, s_in, w'] with channel 0 = "nothing
    placed yet", the last channel = "term complete", one fresh channel per bond a
    multi-site term crosses (operators in between, identities in gaps), and single-site
    terms in the [0, last] corner. The reference then compresses by SVD with cutoff
    1e-15 (:442-457); that changes only the gauge of W for the Hamiltonians used here,
    so it is not applied (energies and the operator are identical: see the tests).
    Returns real tensors when every coefficient and operator is real. If `slots` is a list,
    it receives, per term, (site, row, col, op) such that the term's coefficient c enters
    only as W_site[row, :, :, col] += c * op (1-based site): dW/dc for the chain rule."""
    N, d = H.length, st.dim
    nchan = [0] * (N + 1)  # channels crossing bond b (between sites b and b+1), 1-based b
    plan = []
    for ops, sites, coeff in zip(H.ops, H.sites, H.coeffs):
        if len(sites) == 1:
            plan.append((ops, sites, coeff, None))
            continue
        chans = {}
        for b in range(sites[0], sites[-1]):
            nchan[b] += 1
            chans[b] = nchan[b]
        plan.append((ops, sites, coeff, chans))
    dims = [1] + [2 + nchan[b] for b in range(1, N)] + [1]  # dims[b] = bond b (0..N)
    I = st.op("id")
    Ws = []
    for i in range(1, N + 1):
        Dl, Dr = dims[i - 1], dims[i]
        W = np.zeros((Dl, d, d, Dr), dtype=np.complex128)
        last_l, last_r = Dl - 1, Dr - 1
        if i > 1:
            W[last_l, :, :, last_r] = I  # finished terms pass through
        if i < N:
            W[0, :, :, 0] = I  # nothing placed yet
        Ws.append(W)
    for ops, sites, coeff, chans in plan:
        if chans is None:
            i = sites[0]
            W = Ws[i - 1]
            W[0, :, :, W.shape[3] - 1] += coeff * st.op(ops[0])
            if slots is not None:
                slots.append((i, 0, W.shape[3] - 1, st.op(ops[0])))
            continue
        for i in range(sites[0], sites[-1] + 1):
            W = Ws[i - 1]
            opm = st.op(ops[sites.index(i)]) if i in sites else I
            row = 0 if i == sites[0] else chans[i - 1]
            col = W.shape[3] - 1 if i == sites[-1] else chans[i]
            W[row, :, :, col] += (coeff if i == sites[0] else 1.0) * opm
            if slots is not None and i == sites[0]:
                slots.append((i, row, col, opm))
    return _realify(Ws)


class GMPS:
    """Host container mirroring GMPS(rank, dim, tensors, center),
    src/structures/mps/gmps.jl:8-21. Tensors are Fortran-ordered numpy arrays."""

    def __init__(self, rank: int, dim: int, tensors, center: int = 0):
        self.rank = rank
        self.dim = dim
        self.tensors = [np.asfortranarray(t) for t in tensors]
        self.center = center

    def __len__(self):
        return len(self.tensors)

    def __getitem__(self, i):  # 1-based, as in Julia
        return self.tensors[i - 1]

    def __setitem__(self, i, v):
        self.tensors[i - 1] = np.asfortranarray(v)

    def bonddim(self, site):
        """bonddim(psi, site), src/structures/mps/abstractmps.jl:64-67."""
        if site < 1 or site > len(self):
            return None
        return self.tensors[site].shape[0]

    def maxbonddim(self):
        """maxbonddim(psi), abstractmps.jl:75-81."""
        return max([0] + [self.tensors[i].shape[0] for i in range(1, len(self))])


def MPO(st: Sitetypes, H: OpList, cutoff=1e-15, maxdim=0, mindim=1, compress=True) -> GMPS:
    """MPO(st, H; cutoff=1e-15, maxdim=0, mindim=1), src/structures/mps/mpo.jl:323-459: the
    automaton MPO followed by the reference's SVD compression sweeps, which run on the GPU
    (tnh_mpo_from_oplist). compress=False returns the host-built automaton alone
    (mpo_tensors), as CPU-only callers need."""
    if not compress:
        return GMPS(2, st.dim, mpo_tensors(st, H), 0)
    from .mpo import mpo_device

    return GMPS(2, st.dim, [t.numpy() for t in mpo_device(st, H, cutoff, maxdim, mindim)], 0)


def productMPS(st: Sitetypes, names) -> GMPS:
    """productMPS(st, names), src/structures/mps/mps.jl:68-76."""
    ts = [st.state(n).reshape((1, st.dim, 1)) for n in names]
    return GMPS(1, st.dim, _realify(ts), 0)
