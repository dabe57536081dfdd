"""Error paths of the C-ABI on the GPU: non-finite input reaches the caller as a status
(TNH_ERR_NUMERIC or TNH_ERR_NOCONV, mirroring the reference's LAPACK/KrylovKit exceptions),
the Gram path hands a non-finite problem to the one-sided fallback before giving up, and
the context stays usable after a failed call. Every kernel involved loops a bounded number
of sweeps, and a NaN comparison never rotates, so these inputs end normally. Argument errors
of the rank-1 terms, the one-site update, wide gates, multi-site splits and two-state inner
are refused on the host before any device work."""
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


ARG, SHAPE, UNSUPPORTED = -1, -3, -7


def test_rank1_term_and_nsites_arguments_are_refused(ctx, rng):
    """tnh_dmrg_add_projector refuses a complex phi on a real state and a phi whose bonds do
    not chain; a sweep refuses nsites = 3; all before any device work, and the state keeps
    working after each refusal."""
    from tnhip import GMPS, MPO, Tensor, TnhError, spinhalf
    from tnhip._lib import lib
    from tnhip.mps import DMRGState, _handles
    from test_gpu_svd_dmrg import tfim_H
    from oracle.mps import random_mps

    N = 6
    st = DMRGState(GMPS(1, 2, random_mps(N, 2, 3, rng).tensors, 1), [MPO(spinhalf(), tfim_H(N))])
    try:
        phi = random_mps(N, 2, 2, rng)
        cphi = [Tensor.from_numpy(np.asfortranarray(t + 1j * rng.standard_normal(t.shape)))
                for t in phi.tensors]
        with pytest.raises(TnhError) as e:
            st.ctx.check(lib.tnh_dmrg_add_projector(st.h, _handles(cphi), 1.0), "add_projector")
        assert e.value.status == ARG
        bad = [Tensor.from_numpy(t) for t in phi.tensors]
        bad[2] = Tensor.from_numpy(np.ones((5, 2, 2)))
        with pytest.raises(TnhError) as e:
            st.ctx.check(lib.tnh_dmrg_add_projector(st.h, _handles(bad), 1.0), "add_projector")
        assert e.value.status == SHAPE
        with pytest.raises(TnhError) as e:
            st.sweep(False, nsites=3)
        assert e.value.status == ARG
        e1, D, _ = st.sweep(False, 3, 2, 1e-14, 1e-12, 8, 1)
        assert np.isfinite(e1) and D <= 8
    finally:
        st.close()


def test_wide_gate_and_split_arguments_are_refused(ctx, rng):
    """A gate with d^r > 64 (d = 3, four sites) is UNSUPPORTED, replacesites of one site and
    inner between states of different physical dimensions are refused; the context stays
    usable."""
    import ctypes as C
    from tnhip import Tensor, TnhError, OpList, spinhalf
    from tnhip.tebd import DeviceMPS, inner

    N, r = 5, 4
    ts = [np.asfortranarray(rng.standard_normal((1 if i == 0 else 2, 3, 1 if i == N - 1 else 2)))
          for i in range(N)]
    dev = DeviceMPS(ts, 0)
    G = Tensor.from_numpy(np.asfortranarray(rng.standard_normal((3,) * (2 * r))))
    arr = (C.c_void_p * r)(*[dev[1 + j].h.value for j in range(r)])
    ks = (C.c_int * (r - 1))()
    disc = C.c_double()
    with pytest.raises(TnhError) as e:
        ctx.call("tnh_applygate_sites", r, arr, G.h, 0, 0.0, 0, 1, ks, C.byref(disc))
    assert e.value.status == UNSUPPORTED
    assert [arr[j] for j in range(r)] == [dev[1 + j].h.value for j in range(r)]  # untouched
    th = Tensor.from_numpy(np.asfortranarray(rng.standard_normal((2, 3, 2))))
    outs = (C.c_void_p * 1)()
    with pytest.raises(TnhError) as e:
        ctx.call("tnh_replacesites_sites", 1, th.h, 0, 0, 0.0, 0, 1, outs, None, C.byref(disc))
    assert e.value.status == ARG
    two = DeviceMPS([np.asfortranarray(rng.standard_normal((1 if i == 0 else 2, 2, 1 if i == N - 1 else 2)))
                     for i in range(N)], 0)
    ol = OpList(N)
    ol.add("z", 2, 1.0)
    with pytest.raises(TnhError) as e:
        inner(spinhalf(), two, ol, dev)
    assert e.value.status == SHAPE
    assert np.isfinite(inner(spinhalf(), two, ol)[0])
