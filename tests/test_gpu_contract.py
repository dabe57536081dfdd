"""GPU parity of the MFMA GEMM and of the two-site contractions against the FP64 oracle.

Tolerances (relative 2-norm error against FP64):
  f64     <= 1e-13
  f32     <= 2e-6 * sqrt(K / 64) (inputs rounded to FP32, FP32 accumulation)
  bf16x3  <= 3e-5 (hi*hi + hi*lo + lo*hi split products, FP32 accumulation)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


TOL = {"f64": 1e-13, "f32": 2e-6, "bf16x3": 3e-5}


def tol_for(prec, k):
    return TOL[prec] * (max(1.0, np.sqrt(k / 64.0)) if prec == "f32" else 1.0)


def randn(rng, shape, cplx=False):
    a = rng.standard_normal(shape)
    if cplx:
        a = a + 1j * rng.standard_normal(shape)
    return np.asfortranarray(a)


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
@pytest.mark.parametrize("ops", ["NN", "NT", "TN"])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (7, 5, 3), (64, 64, 16), (130, 97, 33), (300, 260, 129)])
def test_gemm_real(ctx, rng, prec, ops, mnk):
    from tnhip import Tensor, ops as T

    M, N, K = mnk
    A = randn(rng, (K, M) if ops[0] == "T" else (M, K))
    B = randn(rng, (N, K) if ops[1] == "T" else (K, N))
    C0 = randn(rng, (M, N))
    dt = np.float64 if prec == "f64" else np.float32
    dA, dB, dC = Tensor.from_numpy(A, dt), Tensor.from_numpy(B, dt), Tensor.from_numpy(C0, dt)
    T.gemm(ops[0], ops[1], M, N, K, 0.75, dA, A.shape[0], dB, B.shape[0], -0.5, dC, M, prec)
    opA = A.T if ops[0] == "T" else A
    opB = B.T if ops[1] == "T" else B
    if prec != "f64":
        opA, opB, C0r = opA.astype(np.float32).astype(np.float64), opB.astype(np.float32).astype(np.float64), C0.astype(np.float32).astype(np.float64)
    else:
        C0r = C0
    ref = 0.75 * opA @ opB - 0.5 * C0r
    assert rel(dC.numpy(np.float64), ref) <= tol_for(prec, K) * 2


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("C", "N"), ("N", "C"), ("N", "T")])
def test_gemm_complex(ctx, rng, opA, opB):
    from tnhip import Tensor, ops as T

    M, N, K = 70, 45, 37
    A = randn(rng, (K, M) if opA != "N" else (M, K), True)
    B = randn(rng, (N, K) if opB != "N" else (K, N), True)
    dA, dB = Tensor.from_numpy(A), Tensor.from_numpy(B)
    dC = Tensor.from_numpy(np.zeros((M, N), complex, order="F"))
    T.gemm(opA, opB, M, N, K, 1.0, dA, A.shape[0], dB, B.shape[0], 0.0, dC, M, "f64")
    f = {"N": lambda x: x, "T": lambda x: x.T, "C": lambda x: x.conj().T}
    assert rel(dC.numpy(), f[opA](A) @ f[opB](B)) <= 1e-13


def heisenberg_w(D=5):
    """Real XXX MPO bulk tensor W[w, s_out, s_in, w'] with S.S = 1/4 sz sz + 1/2 (s+ s- + s- s+)."""
    from oracle.exact import OPS

    sp, sm, sz, I = (OPS[k].real for k in ("s+", "s-", "z", "id"))
    W = np.zeros((5, 2, 2, 5))
    W[0, :, :, 0] = I
    W[4, :, :, 4] = I
    W[0, :, :, 1] = 0.5 * sp
    W[0, :, :, 2] = 0.5 * sm
    W[0, :, :, 3] = 0.25 * sz
    W[1, :, :, 4] = sm
    W[2, :, :, 4] = sp
    W[3, :, :, 4] = sz
    return W


def random_env(rng, chi, D, cplx=False):
    """(chi, D, chi) blocks, Hermitian per MPO index."""
    X = randn(rng, (chi, D, chi), cplx)
    return np.asfortranarray(0.5 * (X + np.conj(np.transpose(X, (2, 1, 0)))))


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
@pytest.mark.parametrize("chi", [1, 3, 16, 40, 130])
def test_heff_real_vs_oracle(ctx, rng, prec, chi):
    from tnhip import Tensor, ops as T
    from oracle.projmps import product_reference

    W = heisenberg_w()
    L, R = random_env(rng, chi, 5), random_env(rng, chi + 1, 5)
    x = randn(rng, (chi, 2, 2, chi + 1))
    ref = product_reference(L, W, W, R, x, coeff=0.7)
    dt = np.float64 if prec == "f64" else np.float32
    dL, dR = Tensor.from_numpy(L, dt), Tensor.from_numpy(R, dt)
    dW = Tensor.from_numpy(W)
    dx = Tensor.from_numpy(x)
    out = T.heff_apply(dL, dW, dW, dR, dx, coeff=0.7, precision=prec)
    assert rel(out.numpy(), ref) <= tol_for(prec, 5 * chi) * 4


def test_heff_complex_vs_oracle(ctx, rng):
    from tnhip import Tensor, ops as T
    from oracle.projmps import product_reference

    chi = 24
    D = 3
    W1 = randn(rng, (D, 2, 2, 4), True)
    W2 = randn(rng, (4, 2, 2, D), True)
    L, R = random_env(rng, chi, D, True), random_env(rng, chi - 5, D, True)
    x = randn(rng, (chi, 2, 2, chi - 5), True)
    ref = product_reference(L, W1, W2, R, x, coeff=0.3 - 0.2j)
    out = T.heff_apply(*(Tensor.from_numpy(a) for a in (L, W1, W2, R, x)), coeff=0.3 - 0.2j)
    assert rel(out.numpy(), ref) <= 1e-13


@pytest.mark.parametrize("cplx", [False, True])
def test_heff_f32_large_theta_vs_oracle(ctx, rng, cplx):
    """FP32 matvec on a theta of >= 2^20 elements, where theta is first rounded to an FP32
    copy (chi = 512: 4 chi^2 = 2^20), real and complex, against the FP64 oracle."""
    from tnhip import Tensor, ops as T
    from oracle.projmps import product_optimal

    chi = 512
    W = heisenberg_w()
    L, R = random_env(rng, chi, 5, cplx), random_env(rng, chi, 5, cplx)
    x = randn(rng, (chi, 2, 2, chi), cplx)
    ref = product_optimal(L, W, W, R, x, coeff=0.7)
    cdt = np.complex64 if cplx else np.float32
    dL, dR = Tensor.from_numpy(L, cdt), Tensor.from_numpy(R, cdt)
    out = T.heff_apply(dL, Tensor.from_numpy(W), Tensor.from_numpy(W), dR, Tensor.from_numpy(x), coeff=0.7,
                       precision="f32")
    assert rel(out.numpy(), ref) <= tol_for("f32", 5 * chi) * 4


def test_heff_edge_blocks(ctx, rng):
    """L = R = None stands for the ones(1,1,1) block at the chain ends."""
    from tnhip import Tensor, ops as T
    from oracle.projmps import product_reference, edgeblock

    W1 = heisenberg_w()[:1]
    W2 = heisenberg_w()[:, :, :, 4:]
    x = randn(rng, (1, 2, 2, 1))
    ref = product_reference(edgeblock(), W1, W2, edgeblock(), x)
    out = T.heff_apply(None, Tensor.from_numpy(W1), Tensor.from_numpy(W2), None, Tensor.from_numpy(x))
    assert rel(out.numpy(), ref) <= 1e-14


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_env_updates_vs_oracle(ctx, rng, cplx, prec):
    from tnhip import Tensor, ops as T
    from oracle.projmps import buildleft, buildright

    ca, cb, D = 37, 29, 5
    W = heisenberg_w() if not cplx else randn(rng, (D, 2, 2, D), True)
    A = randn(rng, (ca, 2, cb), cplx)
    L = random_env(rng, ca, D, cplx)
    R = random_env(rng, cb, D, cplx)
    dA, dW = Tensor.from_numpy(A), Tensor.from_numpy(W)
    Lp = T.env_left(Tensor.from_numpy(L), dA, dW, precision=prec)
    Rp = T.env_right(Tensor.from_numpy(R), dA, dW, precision=prec)
    t = 1e-13 if prec == "f64" else 1e-5
    assert rel(Lp.numpy(), buildleft(L, A, W)) <= t
    assert rel(Rp.numpy(), buildright(R, A, W)) <= t


def test_calculate_and_theta(ctx, rng):
    from tnhip import Tensor, ops as T
    from oracle.projmps import calculate

    ca, cb = 12, 9
    W = heisenberg_w()
    A1 = randn(rng, (ca, 2, cb))
    A2 = randn(rng, (cb, 2, 7))
    L, R = random_env(rng, ca, 5), random_env(rng, cb, 5)
    e = T.calculate(Tensor.from_numpy(L), Tensor.from_numpy(A1), Tensor.from_numpy(W), Tensor.from_numpy(R))
    assert abs(e - calculate(L, A1, W, R)) <= 1e-12 * abs(calculate(L, A1, W, R))
    th = T.two_site_theta(Tensor.from_numpy(A1), Tensor.from_numpy(A2)).numpy()
    assert rel(th, np.tensordot(A1, A2, axes=([2], [0]))) <= 1e-14
