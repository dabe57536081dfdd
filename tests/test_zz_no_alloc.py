"""tnh_heff_apply_noalloc (the matvec whose temporaries come from the library's static
per-GPU scratch) against tnh_heff_apply on the same inputs, through the C-ABI: bitwise the
same result (same kernels and summation order), the oracle's product within the FP64
matvec tolerance, two calls in a row with the second smaller, every precision, complex
operands, edge blocks, and the clear rejection of shapes beyond the scratch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def operands(rng, chi_l, chi_r, D, cplx=False):
    from test_gpu_contract import heisenberg_w, random_env, randn

    W = heisenberg_w() if D == 5 else randn(rng, (D, 2, 2, D), cplx)
    return (random_env(rng, chi_l, D, cplx), W, random_env(rng, chi_r, D, cplx),
            randn(rng, (chi_l, 2, 2, chi_r), cplx))


def both(L, W, R, x, prec="f64", coeff=1.0):
    from tnhip import Tensor, ops as T

    dt = None if prec == "f64" else (np.complex64 if np.iscomplexobj(L) else np.float32)
    dL = Tensor.from_numpy(L, dt) if L is not None else None
    dR = Tensor.from_numpy(R, dt) if R is not None else None
    dW, dx = Tensor.from_numpy(W), Tensor.from_numpy(x)
    ref = T.heff_apply(dL, dW, dW, dR, dx, coeff=coeff, precision=prec).numpy()
    out = Tensor.empty(x.shape, x.dtype)
    got = T.heff_apply_noalloc(dL, dW, dW, dR, dx, out, coeff=coeff, precision=prec).numpy()
    return got, ref


def test_noalloc_two_calls_second_smaller(ctx):
    from oracle.projmps import product_reference

    rng = np.random.default_rng(11)
    L, W, R, x = operands(rng, 200, 190, 5)
    got, ref = both(L, W, R, x, coeff=0.7)
    assert np.array_equal(got, ref)
    o = product_reference(L, W, W, R, x, coeff=0.7)
    assert np.linalg.norm(got - o) <= 1e-13 * np.linalg.norm(o)
    L, W, R, x = operands(rng, 37, 29, 5)
    got, ref = both(L, W, R, x)
    assert np.array_equal(got, ref)
    o = product_reference(L, W, W, R, x)
    assert np.linalg.norm(got - o) <= 1e-13 * np.linalg.norm(o)


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
def test_noalloc_lower_precisions(ctx, prec):
    rng = np.random.default_rng(12)
    for chi in (160, 48):  # larger first, then smaller: the shared scratch is reused
        L, W, R, x = operands(rng, chi, chi, 5)
        got, ref = both(L, W, R, x, prec=prec)
        assert np.array_equal(got, ref)


def test_noalloc_complex_and_edges(ctx):
    from oracle.projmps import edgeblock, product_reference

    rng = np.random.default_rng(13)
    L, W, R, x = operands(rng, 24, 19, 3, cplx=True)
    got, ref = both(L, W, R, x, coeff=0.3 - 0.2j)
    assert np.array_equal(got, ref)
    o = product_reference(L, W, W, R, x, coeff=0.3 - 0.2j)
    assert np.linalg.norm(got - o) <= 1e-13 * np.linalg.norm(o)
    W1 = W[:1]
    x1 = np.asfortranarray(rng.standard_normal((1, 2, 2, 1)))
    from tnhip import Tensor, ops as T

    out = Tensor.empty(x1.shape, np.complex128)
    got = T.heff_apply_noalloc(None, Tensor.from_numpy(W1), Tensor.from_numpy(W[:, :, :, :1]), None,
                               Tensor.from_numpy(x1), out).numpy()
    o = product_reference(edgeblock(), W1, W[:, :, :, :1], edgeblock(), x1)
    assert np.linalg.norm(got - o) <= 1e-14 * max(np.linalg.norm(o), 1e-300)


def test_noalloc_rejects_shapes_beyond_the_scratch(ctx):
    from tnhip import Tensor, TnhError, ops as T

    chi = 1024  # FP64: T1 and T2 alone are 2 x 168 MB
    W = np.zeros((5, 2, 2, 5))
    dL = Tensor.empty((chi, 5, chi))
    dx = Tensor.empty((chi, 2, 2, chi))
    out = Tensor.empty((chi, 2, 2, chi))
    with pytest.raises(TnhError) as e:
        T.heff_apply_noalloc(dL, Tensor.from_numpy(W), Tensor.from_numpy(W), dL, dx, out)
    assert e.value.status == -4 and "static scratch" in e.value.message
    # the context and the entry point still work after the rejection
    rng = np.random.default_rng(14)
    L, W, R, x = operands(rng, 16, 16, 5)
    got, ref = both(L, W, R, x)
    assert np.array_equal(got, ref)
