"""GPU parity of TEBD (trotterize + applygates + tebd) and of the device observables
(inner, entropy) against dense state-vector references.

Without truncation (cutoff 0, unlimited maxdim) TEBD applies the same Trotter gates
exactly, so the MPS must reproduce the dense evolution up to the SVD's accuracy. The
FP64 Gram eigensolver (M M^H) fixes the kept subspace to ~eps * s_max^2 / s, so when a
split keeps min(m, n) < m columns, directions with singular values near the noise floor
sqrt(eps) * s_max are chosen only to that level; observed errors are ~1e-10 per local
observable, hence 1e-9 on each <Z_i> and 1e-8 relative on the 20-term energy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tfim_oplist(N, J=1.0, h=1.0, g=0.0):
    from tnhip import OpList

    H = OpList(N)
    for i in range(1, N + 1):
        H.add("x", i, h)
        if g:
            H.add("z", i, g)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], J)
    return H


def terms_of(H):
    return [(o, s, c) for o, s, c in zip(H.ops, H.sites, H.coeffs)]


def test_inner_matches_dense(ctx, rng):
    from tnhip import GMPS, OpList, spinhalf
    from tnhip.tebd import DeviceMPS, inner
    from oracle.mps import MPS, to_dense
    from oracle import exact

    N = 6
    ts = [rng.standard_normal((1 if i == 0 else 3, 2, 1 if i == N - 1 else 3)) +
          1j * rng.standard_normal((1 if i == 0 else 3, 2, 1 if i == N - 1 else 3)) for i in range(N)]
    v = to_dense(MPS(ts))
    st = spinhalf()
    ol = OpList(N)
    ol.add("z", 2, 1.0)
    ol.add(["x", "y"], [3, 5], 0.5)  # range-3 term with an identity gap
    ol.add(["z", "z"], [5, 6], 2.0)
    got = inner(st, DeviceMPS(ts), ol)
    ref = [np.vdot(v, exact.dense_hamiltonian(N, [(o, s, c)]) @ v) for o, s, c in terms_of(ol)]
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("cplx_phi", [False, True])
def test_inner_two_states_matches_dense(ctx, rng, cplx_phi):
    """inner(st, psi, oplist, phi) with phi != psi (mps.jl:87-134): <psi|O_t|phi> per term,
    psi complex and phi real or complex with other bond dimensions, terms at both chain
    ends and one with an identity gap."""
    from tnhip import OpList, spinhalf
    from tnhip.tebd import DeviceMPS, inner
    from oracle.mps import MPS, to_dense
    from oracle import exact

    N = 7

    def rand(chi, cplx):
        out = []
        for i in range(N):
            shape = (1 if i == 0 else chi, 2, 1 if i == N - 1 else chi)
            t = rng.standard_normal(shape)
            out.append(t + 1j * rng.standard_normal(shape) if cplx else t)
        return out

    ps, ph = rand(3, True), rand(5, cplx_phi)
    vpsi, vphi = to_dense(MPS(ps)), to_dense(MPS(ph))
    ol = OpList(N)
    ol.add("z", 1, 1.0)
    ol.add(["x", "y"], [3, 5], 0.5)  # range-3 term with an identity gap
    ol.add(["s+", "s-"], [6, 7], 2.0)
    ol.add("x", 4, -1.5)
    got = inner(spinhalf(), DeviceMPS(ps), ol, DeviceMPS(ph))
    ref = [np.vdot(vpsi, exact.dense_hamiltonian(N, [(o, s, c)]) @ vphi)
           for o, s, c in terms_of(ol)]
    scale = np.linalg.norm(vpsi) * np.linalg.norm(vphi)
    assert np.allclose(got, ref, rtol=0, atol=1e-12 * scale)


def _window_mps(rng, N, chi, site):
    from oracle.mps import movecenter, random_mps

    psi = random_mps(N, 2, chi, rng)
    for i in range(1, N + 1):
        psi[i] = psi[i] + 1j * rng.standard_normal(psi[i].shape)
    psi.center = 0
    movecenter(psi, site)
    return psi


@pytest.mark.parametrize("direction", [False, True])
@pytest.mark.parametrize("r", [2, 3, 4])
def test_applygate_sites_matches_oracle(ctx, rng, r, direction):
    """tnh_applygate_sites (gatelist.jl:137-171 for an r-site gate, then replacesites!'s
    r-site branch, gmps.jl:215-257) against the oracle's applygate_mps with truncation at
    every split (maxdim 3): same bonds, same state (the kept subspaces do not depend on the
    SVD gauge), and the canonical form the sweep direction leaves."""
    import ctypes as C
    from tnhip import Tensor
    from tnhip.tebd import DeviceMPS
    from oracle.mps import MPS, to_dense
    from oracle.tebd import applygate_mps

    N, site = 7, 2
    ref = _window_mps(rng, N, 4, site if not direction else site + r - 1)
    shape = (2,) * (2 * r)
    G = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    dev = DeviceMPS([t.copy() for t in ref.tensors], ref.center)
    arr = (C.c_void_p * r)(*[dev[site + j].h.value for j in range(r)])
    ks = (C.c_int * (r - 1))()
    disc = C.c_double()
    Gd = Tensor.from_numpy(np.asfortranarray(G), ctx=dev.ctx)
    try:
        ctx.call("tnh_applygate_sites", r, arr, Gd.h, int(direction), 1e-14, 3, 1, ks, C.byref(disc))
    finally:
        for j in range(r):  # the window's handles are replaced in place
            if dev.sites[site - 1 + j].h.value != arr[j]:
                dev.sites[site - 1 + j].h = None
                dev.sites[site - 1 + j] = Tensor(dev.ctx, C.c_void_p(arr[j]))
    d_ref = applygate_mps(ref, site, G, direction, cutoff=1e-14, maxdim=3, mindim=1)
    got = [t.numpy() for t in dev.sites]
    assert [t.shape for t in got] == [t.shape for t in ref.tensors]
    assert list(ks) == [ref.tensors[site - 1 + j].shape[2] for j in range(r - 1)]
    v, w = to_dense(MPS(got)), to_dense(ref)
    assert np.allclose(v, w, atol=1e-10 * np.linalg.norm(w))
    assert abs(disc.value - d_ref) <= 1e-10 * max(1.0, d_ref)
    for j in range(r):
        t = got[site - 1 + j]
        cl, d, cr = t.shape
        if not direction and j < r - 1:
            M = t.reshape(cl * d, cr, order="F")
            assert np.allclose(M.conj().T @ M, np.eye(cr), atol=1e-12)
        if direction and j > 0:
            M = t.reshape(cl, d * cr, order="F")
            assert np.allclose(M @ M.conj().T, np.eye(cl), atol=1e-12)


def range3_oplist(N):
    from tnhip import OpList

    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N - 1):
        H.add(["z", "z"], [i, i + 2], -0.45)
        H.add(["z", "x", "z"], [i, i + 1, i + 2], 0.3)
    for i in range(1, N + 1):
        H.add("x", i, -0.8)
    return H


@pytest.mark.parametrize("evol", ["imag", "real"])
def test_tebd_range3_exact_without_truncation(ctx, evol):
    """TEBD with three-site gates (next-nearest zz and a z x z term: trotterize's range is
    3) reproduces the dense Trotter evolution when nothing is truncated."""
    from tnhip import productMPS, spinhalf, OpList
    from tnhip.tebd import tebd, TEBDOperators
    from oracle.tebd import trotter_layers, tebd_dense, expect_dense
    from oracle.exact import OPS, dense_hamiltonian

    N, dt, steps = 8, 0.05, 6
    st = spinhalf()
    H = range3_oplist(N)
    names = ["up", "dn", "dn", "up", "dn", "up", "up", "dn"]
    zs = OpList(N)
    for i in range(1, N + 1):
        zs.add("z", i)
    obs = TEBDOperators(st, zs)
    psi, energy = tebd(st, productMPS(st, names), H, dt, dt * steps, dt * steps, [obs], cutoff=0.0,
                       maxdim=0, evol=evol, verbose=False)
    layers = trotter_layers(terms_of(H), N, dt, order=2, evol=evol)
    assert len(layers) == 5
    v0 = np.zeros(2 ** N, complex)
    v0[int("".join("0" if n == "up" else "1" for n in names), 2)] = 1.0
    v = tebd_dense(v0, layers, N, steps)
    ref = np.array([expect_dense(v, N, OPS["z"], i) for i in range(1, N + 1)])
    assert np.allclose(obs.measurements[-1], ref, atol=1e-9)
    e_ref = np.vdot(v, dense_hamiltonian(N, terms_of(H)) @ v).real
    assert abs(energy - e_ref) <= 1e-8 * max(1, abs(e_ref))


def test_tebd_range3_truncated_matches_oracle(ctx):
    """Real-time TEBD with three-site gates and the bond dimension capped at 4: device
    against the oracle's applygates_mps on the same rows, step by step."""
    from tnhip import productMPS, spinhalf, OpList
    from tnhip.tebd import TEBDOperators, tebd
    from oracle.exact import OPS
    from oracle.mps import normalize, product_mps
    from oracle.tebd import applygates_mps, rows_from_layers, trotter_layers

    N, dt, steps = 10, 0.05, 8
    trunc = dict(cutoff=1e-10, maxdim=4, mindim=1)
    st = spinhalf()
    H = range3_oplist(N)
    names = ["up" if i % 2 else "dn" for i in range(1, N + 1)]
    zs = OpList(N)
    for i in range(1, N + 1):
        zs.add("z", i)
    obs = TEBDOperators(st, zs)
    psi, _ = tebd(st, productMPS(st, names), H, dt, dt * steps, dt, [obs], evol="real",
                  verbose=False, **trunc)
    rows = rows_from_layers(trotter_layers(terms_of(H), N, dt, order=2, evol="real"))
    up, dn = np.array([1.0, 0.0]), np.array([0.0, 1.0])
    ref = product_mps([up if n == "up" else dn for n in names])
    zref = []
    for _ in range(steps):
        applygates_mps(ref, rows, **trunc)
        normalize(ref)
        zref.append([_expect_z(ref, i, OPS["z"]) for i in range(1, N + 1)])
    assert [np.asarray(t).shape[2] for t in psi.tensors[:-1]] == ref.bonddims()
    assert max(ref.bonddims()) == 4
    got = np.real(np.array(obs.measurements[1:]))
    assert np.allclose(got, np.array(zref), atol=1e-9)


def test_trotterize_layers():
    from tnhip import spinhalf
    from tnhip.tebd import trotterize

    N = 5
    gl = trotterize(spinhalf(), tfim_oplist(N), 0.1, order=2)
    assert gl.sites == [[1, 3, 5], [2, 4], [1, 3, 5]]
    assert [g.ndim for g in gl.gates[0]] == [4, 4, 2]  # site 5: one-site gate (oplist.jl:138)


@pytest.mark.parametrize("evol", ["imag", "real"])
def test_tebd_exact_without_truncation(ctx, evol):
    from tnhip import productMPS, spinhalf
    from tnhip.tebd import DeviceMPS, tebd, TEBDOperators, trotterize
    from tnhip import OpList
    from oracle.tebd import trotter_layers, tebd_dense, expect_dense
    from oracle.exact import OPS

    N, dt, steps = 7, 0.05, 8
    st = spinhalf()
    H = tfim_oplist(N, J=1.0, h=0.7, g=0.2)
    Hrun = H
    names = ["up" if i % 2 else "dn" for i in range(1, N + 1)]
    psi = productMPS(st, names)
    zs = OpList(N)
    for i in range(1, N + 1):
        zs.add("z", i)
    obs = TEBDOperators(st, zs)
    psi, energy = tebd(st, psi, Hrun, dt, dt * steps, dt * steps, [obs], cutoff=0.0, maxdim=0,
                       evol=evol, verbose=False)
    layers = trotter_layers(terms_of(H), N, dt, order=2, evol=evol)
    v0 = np.zeros(2 ** N, complex)
    idx = int("".join("0" if n == "up" else "1" for n in names), 2)
    v0[idx] = 1.0
    v = tebd_dense(v0, layers, N, steps)
    ref = np.array([expect_dense(v, N, OPS["z"], i) for i in range(1, N + 1)])
    assert np.allclose(obs.measurements[-1], ref, atol=1e-9)
    from oracle.exact import dense_hamiltonian
    e_ref = np.vdot(v, dense_hamiltonian(N, terms_of(H)) @ v).real
    assert abs(energy - e_ref) <= 1e-8 * max(1, abs(e_ref))


def test_tebd_imaginary_time_ground_state(ctx):
    """Imaginary-time TEBD with truncation approaches the TFIM ground energy. The
    reference's trotterize applies exp(+dt H) for evol = "imag" (gatelist.jl:86-100), so
    the ground state of H_phys is reached by evolving with H = -H_phys; the returned
    energy is <H> = -E0."""
    from tnhip import productMPS, spinhalf
    from tnhip.tebd import tebd
    from oracle.exact import tfim_bdg_ground_energy

    N = 12
    st = spinhalf()
    minus_H = tfim_oplist(N, J=1.0, h=1.2)  # -(-ZZ - 1.2 X)
    psi = productMPS(st, ["up"] * N)
    psi, e = tebd(st, psi, minus_H, 0.05, 4.0, 1.0, [], cutoff=1e-12, maxdim=32, verbose=False)
    e0 = tfim_bdg_ground_energy(np.ones(N - 1), 1.2 * np.ones(N))
    # second-order Trotter error at dt = 0.05 ~ 1e-3 relative
    assert abs(-e - e0) / abs(e0) < 5e-3


def test_entropy_matches_dense(ctx, rng):
    from tnhip.tebd import DeviceMPS, entropy
    from oracle.mps import MPS, to_dense, random_mps

    psi = random_mps(6, 2, 4, rng)
    v = to_dense(psi).reshape(8, 8)
    s = np.linalg.svd(v, compute_uv=False) ** 2
    ref = -(s[s > 0] * np.log(s[s > 0])).sum()
    d = DeviceMPS(psi.tensors, psi.center)
    assert abs(entropy(d, 3) - ref) < 1e-10


def test_applyop_and_norm(ctx, rng):
    from tnhip.tebd import DeviceMPS
    from oracle.mps import MPS, to_dense, random_mps
    from oracle.exact import OPS
    from oracle.tebd import apply_dense

    psi = random_mps(5, 2, 4, rng)
    d = DeviceMPS(psi.tensors, psi.center)
    d.apply_site_op(3, OPS["s-"])
    v = apply_dense(to_dense(psi), 3, OPS["s-"], 5)
    w = to_dense(MPS([t.numpy() for t in d.sites]))
    assert np.allclose(w, v, atol=1e-13)
    # applyop! leaves the centre flag alone; re-canonicalise as qjmc.jl:103-106 does
    d.movecenter(5)
    d.movecenter(1)
    assert abs(d.norm() - np.linalg.norm(v)) < 1e-12


@pytest.mark.parametrize("trunc", [dict(cutoff=1e-10, maxdim=3, mindim=1), dict(cutoff=1e-4, maxdim=0, mindim=1)])
def test_tebd_truncated_matches_oracle(ctx, trunc):
    """Real-time TEBD with the truncation active (bond dimension capped at 3, or a cutoff of
    1e-4 that discards weight at every step), N = 10 TFIM quench from a Neel state: the
    device tebd against the oracle's applygates_mps (gatelist.jl:191-223) plus the same
    normalisation per step, on the same Trotter rows. A truncated step keeps the top Schmidt
    subspace, which does not depend on the SVD's gauge, so the bond dimensions agree exactly
    and the local <sz_i> and the norm history to rounding."""
    from tnhip import productMPS, spinhalf
    from tnhip import OpList
    from tnhip.tebd import TEBDOperators, tebd
    from oracle.exact import OPS
    from oracle.mps import normalize, product_mps
    from oracle.tebd import applygates_mps, rows_from_layers, trotter_layers

    N, dt, steps = 10, 0.05, 12
    st = spinhalf()
    H = tfim_oplist(N, J=1.0, h=0.9, g=0.3)
    names = ["up" if i % 2 else "dn" for i in range(1, N + 1)]
    zs = OpList(N)
    for i in range(1, N + 1):
        zs.add("z", i)
    obs = TEBDOperators(st, zs)
    psi, _ = tebd(st, productMPS(st, names), H, dt, dt * steps, dt, [obs], evol="real", verbose=False, **trunc)
    rows = rows_from_layers(trotter_layers(terms_of(H), N, dt, order=2, evol="real"))
    up, dn = np.array([1.0, 0.0]), np.array([0.0, 1.0])
    ref = product_mps([up if n == "up" else dn for n in names])
    zref = []
    for _ in range(steps):
        applygates_mps(ref, rows, **trunc)
        normalize(ref)
        zref.append([_expect_z(ref, i, OPS["z"]) for i in range(1, N + 1)])
    assert [np.asarray(t).shape[2] for t in psi.tensors[:-1]] == ref.bonddims()
    assert max(ref.bonddims()) > 1
    got = np.real(np.array(obs.measurements[1:]))
    assert got.shape == (steps, N)
    assert np.allclose(got, np.array(zref), atol=1e-9)


def _expect_z(psi, site, op):
    """<psi|op_site|psi> of a normalised oracle MPS by a full contraction."""
    from oracle.mps import to_dense

    v = to_dense(psi)
    N = len(psi)
    v = v.reshape([2] * N)
    w = np.tensordot(op, v, axes=([1], [site - 1]))
    w = np.moveaxis(w, 0, site - 1)
    return float(np.real(np.vdot(v.ravel(), w.ravel())))


def test_random_mps_matches_oracle_construction(ctx):
    """randomMPS(d, N, chi) (gmps.jl:276-291) against the oracle's random_mps on the same
    generator: same draws in the same order, the centre moved to N and back to 1, site 1
    redrawn, normalised. The two SVD implementations may choose opposite signs for the bond-1
    basis vectors (the only gauge that reaches the state, since site 1 is redrawn after the
    move), so the device state must equal one of the oracle state's sign variants of site 1's
    columns to 1e-12; every site 2..N is right-canonical and the bond dimensions agree."""
    import itertools
    from tnhip import randomMPS
    from oracle.mps import MPS, random_mps, to_dense

    N, d, chi = 8, 2, 4
    psi = randomMPS(d, N, chi, rng=np.random.default_rng(11))
    ref = random_mps(N, d, chi, np.random.default_rng(11))
    assert [np.asarray(t).shape[2] for t in psi.tensors[:-1]] == ref.bonddims() and psi.center == ref.center == 1
    got = to_dense(MPS([np.asarray(t) for t in psi.tensors], 1))
    assert abs(np.linalg.norm(got) - 1.0) < 1e-12
    best = np.inf
    k = ref[1].shape[2]
    for signs in itertools.product([1.0, -1.0], repeat=k):
        var = MPS([t.copy() for t in ref.tensors], 1)
        var[1] = var[1] * np.array(signs)[None, None, :]
        best = min(best, np.linalg.norm(to_dense(var) - got))
    assert best < 1e-12
    for t in psi.tensors[1:]:
        B = np.asarray(t).reshape(t.shape[0], -1, order="F")
        assert np.allclose(B @ B.conj().T, np.eye(B.shape[0]), atol=1e-12)
