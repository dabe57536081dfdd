"""Error paths of the C-ABI on the GPU: non-finite input reaches the caller as a status
(TNH_ERR_NUMERIC or TNH_ERR_NOCONV, mirroring the reference's LAPACK/KrylovKit exceptions),
the Gram path hands a non-finite problem to the one-sided fallback before giving up, and
the context stays usable after a failed call. Every kernel involved loops a bounded number
of sweeps, and a NaN comparison never rotates, so these inputs end normally."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NUMERIC, NOCONV = -5, -6


def poisoned(rng, shape, value, cplx=False, where=(0, 0)):
    a = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    a = np.asfortranarray(a)
    a[where] = value
    return a


@pytest.mark.parametrize("value", [np.nan, np.inf])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(20, 30), (100, 150)])
def test_svd_non_finite_input_is_an_error(ctx, rng, value, cplx, shape):
    """p <= 64 (one-workgroup eigensolve) and p = 100 (block Jacobi): both the Gram path and
    the one-sided fallback see the non-finite entry; the call fails with a status."""
    from tnhip import Tensor, TnhError, ops as T

    M = Tensor.from_numpy(poisoned(rng, shape, value, cplx, where=(3, 7)))
    with pytest.raises(TnhError) as e:
        T.svd_trunc(M, 1e-12, 0, 1)
    assert e.value.status in (NUMERIC, NOCONV)
    with pytest.raises(TnhError) as e:
        T.svd_trunc(M, 1e-12, 0, 1, force_fallback=True)
    assert e.value.status in (NUMERIC, NOCONV)
    # the context still works
    good = np.asfortranarray(rng.standard_normal(shape))
    U, S, Vt, disc, fb = T.svd_trunc(Tensor.from_numpy(good), 0.0, 0, 1)
    assert np.allclose(S.numpy(), np.linalg.svd(good, compute_uv=False), atol=1e-12)


def test_replacesites_nan_theta_is_an_error(ctx, rng):
    from tnhip import Tensor, TnhError, ops as T

    th = Tensor.from_numpy(poisoned(rng, (6, 2, 2, 5), np.nan, where=(1, 0, 1, 2)))
    with pytest.raises(TnhError) as e:
        T.replacesites2(th, False, True, 1e-10, 0, 1)
    assert e.value.status in (NUMERIC, NOCONV)


@pytest.mark.parametrize("where", ["theta", "L"])
def test_eigsolve_non_finite_is_numeric(ctx, rng, where):
    from tnhip import Tensor, TnhError, ops as T
    from test_gpu_svd_dmrg import tfim_w

    chi = 6
    W = tfim_w()
    X = rng.standard_normal((chi, 3, chi))
    L = np.asfortranarray(0.5 * (X + X.transpose(2, 1, 0)))
    R = L.copy(order="F")
    x0 = np.asfortranarray(rng.standard_normal((chi, 2, 2, chi)))
    if where == "theta":
        x0[2, 1, 0, 3] = np.nan
    else:
        L[1, 2, 4] = np.nan
    dL, dW, dR = Tensor.from_numpy(L), Tensor.from_numpy(W), Tensor.from_numpy(R)
    with pytest.raises(TnhError) as e:
        T.eigsolve_ground(dL, dW, dW, dR, Tensor.from_numpy(x0), 4, 2, 1e-14)
    assert e.value.status == NUMERIC


def test_dmrg_with_nan_coupling_raises(ctx, rng):
    """A NaN coupling in the MPO: the first bond's eigensolve fails and dmrg raises."""
    from tnhip import MPO, TnhError, dmrg, randomMPS, spinhalf
    from tnhip.models import OpList

    N = 8
    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        H.add("x", i, np.nan if i == 4 else -1.0)
    psi = randomMPS(2, N, 4, rng)
    with pytest.raises(TnhError) as e:
        dmrg(psi, MPO(spinhalf(), H), maxdim=16, maxsweeps=2, verbose=False)
    assert e.value.status in (NUMERIC, NOCONV)
