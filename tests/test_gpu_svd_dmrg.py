"""GPU parity of the truncated SVD (Gram + Jacobi, and the one-sided Jacobi fallback),
the two-site split, the Lanczos eigensolver, centre moves and full DMRG sweeps
against the FP64 oracle, LAPACK, dense diagonalisation and closed forms.

Tolerances: singular values 1e-12 relative to s_max (Gram path: entries above
1e-6 s_max, the squared-condition limit of a Gram eigensolve); reconstructions and
orthonormality 1e-12; Lanczos energy vs the oracle's identical algorithm 1e-11;
DMRG energies |dE|/N <= 1e-10 at FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def randn(rng, shape, cplx=False):
    a = rng.standard_normal(shape)
    if cplx:
        a = a + 1j * rng.standard_normal(shape)
    return np.asfortranarray(a)


def decaying(rng, m, n, cplx=False, decay=0.5):
    """m x n matrix with singular values decay**i (well separated from any cutoff)."""
    k = min(m, n)
    U, _ = np.linalg.qr(randn(rng, (m, k), cplx))
    V, _ = np.linalg.qr(randn(rng, (n, k), cplx))
    s = decay ** np.arange(k)
    return np.asfortranarray((U * s) @ V.conj().T), s


@pytest.mark.parametrize("fallback", [False, True])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(1, 1), (5, 3), (3, 5), (40, 64), (97, 130), (130, 97), (256, 256)])
def test_svd_full_rank(ctx, rng, fallback, cplx, shape):
    from tnhip import Tensor, ops as T

    M = randn(rng, shape, cplx)
    U, S, Vt, disc, fb = T.svd_trunc(Tensor.from_numpy(M), 0.0, 0, 1, force_fallback=fallback)
    assert fb == fallback
    u, s, vt = U.numpy(), S.numpy(), Vt.numpy()
    ref = np.linalg.svd(M, compute_uv=False)
    assert s.shape == ref.shape
    if fallback:  # one-sided Jacobi on M: absolute accuracy eps * s_max per value
        assert np.max(np.abs(s - ref)) <= 1e-12 * ref[0]
    else:  # Gram path: s^2 carries the eigensolver's absolute error ~ eps sqrt(n) s_max^2
        assert np.max(np.abs(s ** 2 - ref ** 2)) <= 1e-13 * np.sqrt(max(shape)) * ref[0] ** 2
    assert np.linalg.norm((u * s) @ vt - M) <= 1e-12 * np.linalg.norm(M)
    k = s.size
    # the eigenvector (basis) factor is orthonormal to working precision; on the Gram
    # path the other factor is formed as diag(1/s) U^H M and carries eps * cond^2
    basis_left = shape[0] <= shape[1]
    kappa2 = 1.0 if fallback else (ref[0] / ref[-1]) ** 2
    tu = 1e-12 if (fallback or basis_left) else 1e-13 * kappa2
    tv = 1e-12 if (fallback or not basis_left) else 1e-13 * kappa2
    assert np.abs(u.conj().T @ u - np.eye(k)).max() <= tu
    assert np.abs(vt @ vt.conj().T - np.eye(k)).max() <= tv
    assert disc == 0.0


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(20, 24), (100, 150)])
@pytest.mark.parametrize("scale", [1e100, 1e-100])
def test_svd_extreme_norm(ctx, rng, cplx, shape, scale):
    """Gram entries near 1e200 (1e-200): the Jacobi rotation test multiplies two diagonal
    entries, which would overflow (underflow) without the exact power-of-two rescaling of
    G. Shapes cover the single-workgroup eigensolve (p = 20) and the block path (p = 100)."""
    from tnhip import Tensor, ops as T

    M = np.asfortranarray(randn(rng, shape, cplx) * scale)
    U, S, Vt, disc, fb = T.svd_trunc(Tensor.from_numpy(M), 0.0, 0, 1)
    assert not fb
    u, s, vt = U.numpy(), S.numpy(), Vt.numpy()
    ref = np.linalg.svd(M, compute_uv=False)
    assert np.max(np.abs((s / ref[0]) ** 2 - (ref / ref[0]) ** 2)) <= 1e-13 * np.sqrt(max(shape))
    assert np.linalg.norm((u * s) @ vt - M) <= 1e-12 * np.linalg.norm(M)
    assert np.abs(u.conj().T @ u - np.eye(s.size)).max() <= 1e-12


@pytest.mark.parametrize("fallback", [False, True])
@pytest.mark.parametrize("cutoff,maxdim,mindim", [(1e-10, 0, 1), (1e-4, 0, 1), (0.0, 7, 1), (1e-6, 5, 1),
                                                  (1e-2, 0, 9), (1e-14, 0, 1)])
def test_svd_truncation_rule(ctx, rng, fallback, cutoff, maxdim, mindim):
    from tnhip import Tensor, ops as T
    from oracle.tensors import truncation_rank

    M, s_true = decaying(rng, 64, 48, decay=0.55)
    U, S, Vt, disc, fb = T.svd_trunc(Tensor.from_numpy(M), cutoff, maxdim, mindim, force_fallback=fallback)
    k_ref, disc_ref = truncation_rank(np.linalg.svd(M, compute_uv=False), cutoff, maxdim, mindim)
    assert S.shape[0] == k_ref
    # Gram path: each s^2 / sum s^2 carries an absolute error ~ eps sqrt(n)
    assert abs(disc - disc_ref) <= 1e-9 * disc_ref + (1e-15 if fallback else 1e-13)


@pytest.mark.parametrize("direction", [False, True])
@pytest.mark.parametrize("cplx", [False, True])
def test_replacesites_vs_oracle(ctx, rng, direction, cplx):
    from tnhip import Tensor, ops as T
    from oracle.mps import replacesites2

    th = randn(rng, (12, 2, 2, 9), cplx)
    th /= np.linalg.norm(th)
    A1, A2, s, disc = T.replacesites2(Tensor.from_numpy(th), direction, True, 1e-8, 0, 1)
    B1, B2, s_ref, disc_ref = replacesites2(th, direction, True, 1e-8, 0, 1)
    a1, a2 = A1.numpy(), A2.numpy()
    assert a1.shape == B1.shape and a2.shape == B2.shape
    assert np.max(np.abs(s - s_ref)) <= 1e-12
    assert abs(disc - disc_ref) <= 1e-12
    rec = np.tensordot(a1, a2, axes=([2], [0]))
    assert np.linalg.norm(rec - np.tensordot(B1, B2, axes=([2], [0]))) <= 1e-12
    if direction:
        B = a2.reshape(a2.shape[0], -1, order="F")
        assert np.allclose(B @ B.conj().T, np.eye(B.shape[0]), atol=1e-12)
        assert abs(np.linalg.norm(a1) - 1) <= 1e-13
    else:
        B = a1.reshape(-1, a1.shape[2], order="F")
        assert np.allclose(B.conj().T @ B, np.eye(B.shape[1]), atol=1e-12)
        assert abs(np.linalg.norm(a2) - 1) <= 1e-13


@pytest.mark.parametrize("direction", [False, True])
@pytest.mark.parametrize("n", [3, 4])
def test_replacesites_multi_site_vs_oracle(ctx, rng, n, direction):
    """replacesites! of a theta of n sites (gmps.jl:215-257, tnh_replacesites_sites) with
    truncation (maxdim 5) and normalize = true, against the oracle's restatement: same bonds
    and discarded weight, the same truncated state, the canonical form of the sweep and a
    unit-norm centre."""
    from tnhip.mps import replacesites
    from oracle.mps import replacesites_n

    shape = (3,) + (2,) * n + (4,)
    th = randn(rng, shape, True)
    th /= np.linalg.norm(th)
    sites, _, disc = replacesites(th, direction, True, 1e-12, 5, 1)
    ref, disc_ref = replacesites_n(th, direction, 1e-12, 5, 1)
    assert [t.shape for t in sites] == [t.shape for t in ref]
    assert max(t.shape[2] for t in sites[:-1]) == 5
    assert abs(disc - disc_ref) <= 1e-12

    def chain(ts):
        out = ts[0]
        for t in ts[1:]:
            out = np.tensordot(out, t, axes=([out.ndim - 1], [0]))
        return out

    c = ref[0 if direction else -1]
    ref[0 if direction else -1] = c / np.linalg.norm(c)  # normalize!(psi) on the new centre
    assert np.linalg.norm(chain(sites) - chain(ref)) <= 1e-11
    assert abs(np.linalg.norm(sites[0 if direction else -1]) - 1) <= 1e-13
    for i, t in enumerate(sites):
        cl, d, cr = t.shape
        if not direction and i < n - 1:
            B = t.reshape(cl * d, cr, order="F")
            assert np.allclose(B.conj().T @ B, np.eye(cr), atol=1e-12)
        if direction and i > 0:
            B = t.reshape(cl, d * cr, order="F")
            assert np.allclose(B @ B.conj().T, np.eye(cl), atol=1e-12)


def test_split_rank_deficient_keeps_orthonormal_factor(ctx, rng):
    """theta of rank 3, cutoff 1e-10 (k_cut = 3) and mindim 6 (tensors.jl:215: k = 6):
    the kept orthonormal factor stays orthonormal."""
    from tnhip import Tensor, ops as T

    a = randn(rng, (8, 3)) @ randn(rng, (3, 8))
    th = np.asfortranarray(a.reshape(4, 2, 2, 4, order="F"))
    A1, A2, s, disc = T.replacesites2(Tensor.from_numpy(th), False, False, 1e-10, 0, 6)
    B = A1.numpy().reshape(8, -1, order="F")
    assert B.shape[1] == 6
    assert np.allclose(B.T @ B, np.eye(6), atol=1e-10)
    assert np.allclose(np.tensordot(A1.numpy(), A2.numpy(), axes=([2], [0])), th, atol=1e-12)


def tfim_w():
    """Bulk TFIM tensor: every (w, w') block is a real symmetric 2x2 matrix, so H_eff is
    Hermitian whenever L and R are Hermitian per MPO index."""
    from tnhip.models import OpList, MPO, spinhalf

    H = OpList(4)
    for i in range(1, 4):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, 5):
        H.add("x", i, -0.9)
    return MPO(spinhalf(), H, compress=False).tensors[1]


@pytest.mark.parametrize("cplx", [False, True])
def test_eigsolve_matches_oracle_lanczos(ctx, rng, cplx):
    from tnhip import Tensor, ops as T
    from oracle.krylov import eigsolve_ground
    from oracle.projmps import product_optimal, heff_dense

    W = tfim_w()
    chi = 6
    X = randn(rng, (chi, 3, chi), cplx)
    L = np.asfortranarray(0.5 * (X + np.conj(np.transpose(X, (2, 1, 0)))))
    Y = randn(rng, (chi, 3, chi), cplx)
    R = np.asfortranarray(0.5 * (Y + np.conj(np.transpose(Y, (2, 1, 0)))))
    x0 = randn(rng, (chi, 2, 2, chi), cplx)
    e_ref, v_ref, nmv_ref = eigsolve_ground(lambda x: product_optimal(L, W, W, R, x), x0, 4, 3, 1e-14)
    th = Tensor.from_numpy(x0)
    dL, dW, dR = Tensor.from_numpy(L), Tensor.from_numpy(W), Tensor.from_numpy(R)
    e, nmv = T.eigsolve_ground(dL, dW, dW, dR, th, 4, 3, 1e-14)
    assert nmv == nmv_ref
    assert abs(e - e_ref) <= 1e-11 * abs(e_ref)
    v = th.numpy()
    assert abs(abs(np.vdot(v, v_ref)) - 1) <= 1e-9
    # with a large Krylov space it converges to the dense ground state
    H = heff_dense(L, W, W, R)
    assert np.allclose(H, H.conj().T)
    th2 = Tensor.from_numpy(x0)
    e2, _ = T.eigsolve_ground(dL, dW, dW, dR, th2, 30, 30, 1e-12)
    assert abs(e2 - np.linalg.eigvalsh(H)[0]) <= 1e-9 * abs(e2)


def test_eigsolve_invariant_subspace(ctx):
    """theta0 an exact eigenvector: breakdown after one matvec, eigenpair returned."""
    from tnhip import Tensor, ops as T

    W = np.zeros((1, 2, 2, 1))
    W[0, :, :, 0] = np.diag([1.0, 2.0])
    x0 = np.zeros((1, 2, 2, 1), order="F")
    x0[0, 0, 0, 0] = 1.0
    th = Tensor.from_numpy(x0)
    dW = Tensor.from_numpy(W)
    e, nmv = T.eigsolve_ground(None, dW, dW, None, th, 3, 2)
    assert nmv == 1 and abs(e - 1.0) < 1e-14


def test_movecenter_preserves_state(ctx, rng):
    from tnhip import GMPS, movecenter
    from oracle.mps import MPS, to_dense

    ts = [rng.standard_normal((1 if i == 0 else 4, 2, 1 if i == 5 else 4)) for i in range(6)]
    v0 = to_dense(MPS(ts))
    psi = GMPS(1, 2, ts, 0)
    movecenter(psi, 3)
    assert psi.center == 3
    v1 = to_dense(MPS(psi.tensors))
    assert np.allclose(v1, v0, atol=1e-12 * np.linalg.norm(v0))
    for i in (1, 2):  # left-orthonormal
        B = psi[i].reshape(-1, psi[i].shape[2], order="F")
        assert np.allclose(B.T @ B, np.eye(B.shape[1]), atol=1e-12)
    for i in (4, 5, 6):  # right-orthonormal
        B = psi[i].reshape(psi[i].shape[0], -1, order="F")
        assert np.allclose(B @ B.T, np.eye(B.shape[0]), atol=1e-12)
    assert [t.shape[2] for t in psi.tensors[:-1]] == [2, 4, 4, 4, 2]


def tfim_H(N, J=1.0, g=1.0):
    from tnhip.models import OpList

    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -J)
    for i in range(1, N + 1):
        H.add("x", i, -g)
    return H


def test_dmrg_tfim_closed_form(ctx, rng):
    from tnhip import MPO, dmrg, randomMPS, spinhalf
    from oracle.exact import tfim_critical_ground_energy

    N = 24
    psi = randomMPS(2, N, 4, rng)
    psi, e = dmrg(psi, MPO(spinhalf(), tfim_H(N)), maxdim=64, cutoff=1e-12, maxsweeps=10, krylovdim=4,
                  verbose=False)
    assert abs(e - tfim_critical_ground_energy(N)) / N <= 1e-10


def test_dmrg_heisenberg_matches_oracle_and_ed(ctx, rng):
    """Config-1 style run (N=12 so that ED applies), GPU vs the oracle's DMRG."""
    from tnhip import MPO, dmrg, spinhalf, GMPS
    from tnhip.models import OpList
    from oracle import exact
    from oracle.dmrg import dmrg as odmrg
    from oracle.mps import MPS, random_mps

    N = 12
    H = OpList(N)
    for i in range(1, N):
        for o in ("s+", "s-"):
            H.add([o, "s-" if o == "s+" else "s+"], [i, i + 1], 0.5)
        H.add(["z", "z"], [i, i + 1], 0.25)
    M = MPO(spinhalf(), H)
    ref_psi = random_mps(N, 2, 8, np.random.default_rng(0))
    psi = GMPS(1, 2, [t.copy() for t in ref_psi.tensors], 1)
    psi, e = dmrg(psi, M, maxdim=64, cutoff=1e-10, maxsweeps=8, krylovdim=4, verbose=False)
    _, e_ref = odmrg(ref_psi, M.tensors, maxdim=64, cutoff=1e-10, maxsweeps=8, krylovdim=4)
    Hd = exact.dense_hamiltonian(N, [(o, s, c) for o, s, c in zip(H.ops, H.sites, H.coeffs)])
    e_ed = np.linalg.eigvalsh(Hd)[0]
    assert abs(e - e_ed) / N <= 1e-10
    assert abs(e - e_ref) / N <= 1e-10
    from oracle.mps import to_dense

    v = exact.ground_state_dense(Hd)[1]
    assert abs(abs(np.vdot(to_dense(MPS(psi.tensors)), v)) - 1) <= 1e-8


def test_dmrg_complex_mpo_pauli_y(ctx, rng):
    """The automaton MPO with Pauli y is complex: the C128 path gives the same energy."""
    from tnhip import MPO, dmrg, randomMPS, spinhalf
    from tnhip.models import OpList
    from oracle import exact

    N = 10
    H = OpList(N)
    for i in range(1, N):
        for o in ("x", "y", "z"):
            H.add([o, o], [i, i + 1], 0.25)
    M = MPO(spinhalf(), H)
    assert np.iscomplexobj(M.tensors[3])
    psi = randomMPS(2, N, 4, rng)
    psi, e = dmrg(psi, M, maxdim=64, cutoff=1e-12, maxsweeps=8, krylovdim=4, verbose=False)
    Hd = exact.dense_hamiltonian(N, [(o, s, c) for o, s, c in zip(H.ops, H.sites, H.coeffs)])
    assert abs(e - np.linalg.eigvalsh(Hd)[0]) / N <= 1e-10


@pytest.mark.parametrize("precision", ["f32", "bf16x3"])
def test_dmrg_mixed_precision_refined(ctx, rng, precision):
    """Lower-precision sweeps then FP64 refinement reach |dE|/N <= 1e-10."""
    from tnhip import MPO, dmrg, randomMPS, spinhalf
    from oracle.exact import tfim_critical_ground_energy

    N = 20
    psi = randomMPS(2, N, 4, rng)
    hist = []
    psi, e = dmrg(psi, MPO(spinhalf(), tfim_H(N)), maxdim=64, cutoff=1e-12, maxsweeps=10, krylovdim=4,
                  verbose=False, precision=precision, refine_sweeps=2, history=hist)
    assert any(h["precision"] != "f64" for h in hist) and hist[-1]["precision"] == "f64"
    assert abs(e - tfim_critical_ground_energy(N)) / N <= 1e-10


def test_dmrg_f32_pass_gram_floor_refined(ctx):
    """An FP32 pass at bond dimension 300 (Gram sides 600, block Jacobi: the FP32 pass stops
    its eigensolves at 1e-9 max G_ii instead of the FP64 rounding floor) followed by one FP64
    pass lands on the energy of two FP64 passes from the same state (a copy of it) to
    1e-10 per site: the critical TFIM at N = 20, bonds forced to 300 from a state converged
    at chi 32."""
    from tnhip import MPO, dmrg, randomMPS, spinhalf
    from tnhip.models import GMPS
    from tnhip.mps import DMRGState

    N, chi = 20, 300
    M = MPO(spinhalf(), tfim_H(N))
    psi = randomMPS(2, N, 8, np.random.default_rng(3))
    psi, _ = dmrg(psi, M, maxdim=32, cutoff=1e-12, maxsweeps=6, minsweeps=6, tol=0.0, verbose=False)
    st = DMRGState(psi, [M])
    for dr in (False, True) * 3:  # grow the bonds to chi (centre back at site 1)
        st.sweep(dr, 3, 2, 1e-14, 0.0, chi, chi, "f64")
    assert max(st.bonddims()) == chi
    start = GMPS(1, 2, st.tensors(), st.center())
    st.sweep(False, 3, 2, 1e-14, 0.0, chi, chi, "f32")
    e_mixed, _, _ = st.sweep(True, 3, 2, 1e-14, 0.0, chi, chi, "f64")
    st.close()
    twin = DMRGState(start, [M])
    twin.sweep(False, 3, 2, 1e-14, 0.0, chi, chi, "f64")
    e_f64, _, _ = twin.sweep(True, 3, 2, 1e-14, 0.0, chi, chi, "f64")
    twin.close()
    print(f"refined {e_mixed:.13f} vs two FP64 passes {e_f64:.13f}", flush=True)
    assert abs(e_mixed - e_f64) / N <= 1e-10


def test_dmrg_two_terms_sum(ctx, rng):
    """ProjMPSSum: two MPO terms with coefficients (projmpssum.jl:63-73)."""
    from tnhip import MPO, dmrg, randomMPS, spinhalf
    from tnhip.models import OpList
    from oracle.exact import tfim_bdg_ground_energy

    N = 16
    st = spinhalf()
    Hzz, Hx = OpList(N), OpList(N)
    for i in range(1, N):
        Hzz.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        Hx.add("x", i, -1.0)
    psi = randomMPS(2, N, 4, rng)
    # field 1.3: gapped paramagnet, so 10 passes of the default 4-vector Lanczos converge
    psi, e = dmrg(psi, MPO(st, Hzz), MPO(st, Hx), coeffs=[1.0, 1.3], maxdim=64, cutoff=1e-12,
                  maxsweeps=10, krylovdim=4, verbose=False)
    assert abs(e - tfim_bdg_ground_energy(np.ones(N - 1), 1.3 * np.ones(N))) / N <= 1e-10


@pytest.mark.parametrize("p", [512, 2048])
@pytest.mark.parametrize("cutoff", [1e-10, 1e-12])
@pytest.mark.parametrize("decay", [0.9, 0.8])
def test_svd_cutoff_rank_matches_lapack_large(ctx, p, cutoff, decay):
    """Cutoff-driven truncation at Gram sides 512 and 2048: a p x p matrix with singular
    values decay^i (random orthogonal factors), whose discarded-weight tail crosses the
    cutoff among values of relative size ~cutoff, where the Gram eigensolve carries an
    absolute error ~eps sqrt(p) s_max^2. The kept rank equals the truncation rule applied
    to LAPACK's singular values (tensors.jl:202-215) and the discarded weight agrees."""
    from tnhip import Tensor, ops as T
    from oracle.tensors import truncation_rank

    rng = np.random.default_rng(p + int(100 * decay))
    U, _ = np.linalg.qr(rng.standard_normal((p, p)))
    V, _ = np.linalg.qr(rng.standard_normal((p, p)))
    M = np.asfortranarray((U * decay ** np.arange(p)) @ V.T)
    s_ref = np.linalg.svd(M, compute_uv=False)
    k_ref, disc_ref = truncation_rank(s_ref, cutoff, 0, 1)
    _, S, _, disc, fb = T.svd_trunc(Tensor.from_numpy(M), cutoff, 0, 1, want=(False, True, False))
    assert not fb
    assert S.shape[0] == k_ref, (S.shape[0], k_ref)
    # each s^2 / sum s^2 of the Gram path carries an absolute error ~ eps sqrt(p)
    assert abs(disc - disc_ref) <= 1e-13
