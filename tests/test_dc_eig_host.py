"""The small-matrix eigensolver of the split (tensornetworks.jl_amd/csrc/dc_eig.hpp:
Householder tridiagonalisation, divide and conquer with deflation and Gu-Eisenstat
eigenvectors) compiled for the host from the same header the GPU kernel instantiates, with a
sequential execution policy, and checked against LAPACK (numpy.linalg.eigh): eigenvalues to
1e-13 sqrt(p) of the largest, residual and orthogonality at rounding level. The cases are the
ones the algorithm's branches are there for: real and complex, p from 1 to the 512 limit,
exact and near degeneracies (deflation by rotation), clusters at the noise floor and exactly
zero eigenvalues (deflation by small coupling), graded spectra, zero and extreme scales,
tridiagonal inputs (Wilkinson, discrete Laplacian, nearly diagonal)."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent
SRC = ROOT / "host" / "dc_eig_host.cpp"


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    out = tmp_path_factory.mktemp("dc") / "dc_eig_host.so"
    subprocess.run([gxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-o", str(out), str(SRC)], check=True)
    L = C.CDLL(str(out))
    dp = C.POINTER(C.c_double)
    L.dc_eig_host.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp]
    L.dc_eig_host.restype = C.c_int
    return L


def eig(lib, G):
    p = G.shape[0]
    cplx = np.iscomplexobj(G)
    dp = C.POINTER(C.c_double)
    Gr = np.asfortranarray(G.real.copy())
    Gi = np.asfortranarray(G.imag.copy() if cplx else np.zeros((p, p)))
    lam = np.zeros(p)
    Vr = np.zeros((p, p), order="F")
    Vi = np.zeros((p, p), order="F")
    rc = lib.dc_eig_host(p, int(cplx), Gr.ctypes.data_as(dp), Gi.ctypes.data_as(dp), lam.ctypes.data_as(dp),
                         Vr.ctypes.data_as(dp), Vi.ctypes.data_as(dp))
    return rc, lam, (Vr + 1j * Vi if cplx else Vr)


def check(lib, G):
    rc, lam, V = eig(lib, G)
    p = G.shape[0]
    assert rc == 0
    assert np.all(np.diff(lam) >= 0)
    ref = np.linalg.eigvalsh(G)
    sc = max(np.abs(ref).max(), 1e-300)
    assert np.abs(lam - ref).max() / sc <= 1e-13 * np.sqrt(p)
    assert np.linalg.norm((G / sc) @ V - V * (lam / sc)) <= 1e-13 * p
    assert np.linalg.norm(V.conj().T @ V - np.eye(p)) <= 1e-12 * np.sqrt(p)


def unitary(rng, p, cplx):
    A = rng.standard_normal((p, p)) + (1j * rng.standard_normal((p, p)) if cplx else 0)
    return np.linalg.qr(A)[0]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("p", [1, 2, 3, 5, 9, 33, 64])
def test_random_gram(lib, p, cplx):
    rng = np.random.default_rng(p)
    A = rng.standard_normal((p, p)) + (1j * rng.standard_normal((p, p)) if cplx else 0)
    check(lib, A @ A.conj().T)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("p", [100, 256, 300, 512])
def test_structured_spectra(lib, p, cplx):
    rng = np.random.default_rng(p + 7)
    Uq = unitary(rng, p, cplx)
    # a decaying part above a cluster at the noise floor
    check(lib, (Uq * np.r_[np.exp(-np.arange(p // 4)), np.full(p - p // 4, 1e-17)]) @ Uq.conj().T)
    # four-fold exact degeneracies
    check(lib, (Uq * np.repeat(np.exp(-0.5 * np.arange(p // 4 + 1)), 4)[:p]) @ Uq.conj().T)
    # rank p / 2 (exact zeros), the shape of a two-site split's Gram matrix
    M = (Uq[:, : p // 2] * np.exp(-0.3 * np.arange(p // 2))) @ rng.standard_normal((p // 2, p))
    check(lib, M @ M.conj().T)


def test_special_matrices(lib):
    for G in (np.eye(7), np.diag(np.arange(9.0)), np.zeros((6, 6)), np.diag(np.r_[np.ones(5), np.zeros(5)]),
              1e200 * np.eye(5) + 1e199, 1e-200 * (np.eye(5) + 0.1)):
        check(lib, G)
    n = 41
    check(lib, np.diag(np.abs(np.arange(n) - 20.0)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1))
    check(lib, np.diag(np.full(60, 2.0)) - np.diag(np.ones(59), 1) - np.diag(np.ones(59), -1))
    check(lib, np.eye(60) + np.diag(np.full(59, 1e-12), 1) + np.diag(np.full(59, 1e-12), -1))


def test_dmrg_split_gram(lib):
    """The Gram matrix of a two-site theta of the critical Ising chain after a few oracle
    DMRG passes (bond dimension 40: Gram side 80), whose spectrum spans many decades."""
    from oracle.dmrg import dmrg
    from oracle.mps import random_mps
    from tnhip.models import OpList, mpo_tensors, spinhalf

    N = 16
    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0)
    for i in range(1, N + 1):
        H.add("x", i, -1.0)
    psi = random_mps(N, 2, 40, np.random.default_rng(3))
    psi, _ = dmrg(psi, [np.asarray(w, float) for w in mpo_tensors(spinhalf(), H)], maxsweeps=3, minsweeps=3,
                  tol=0.0, cutoff=0.0, maxdim=40)
    th = np.tensordot(psi[8], psi[9], axes=([2], [0]))
    M = th.reshape(th.shape[0] * th.shape[1], -1, order="F")
    check(lib, M @ M.conj().T)


@pytest.mark.parametrize("value", [np.nan, np.inf])
@pytest.mark.parametrize("p", [9, 20, 150])
def test_non_finite_input_is_reported(lib, p, value):
    """A non-finite entry is reported through the status (the caller then fails the split),
    and every gather stays inside its matrix: the merge sorts order NaN last, so their ranks
    are permutations whatever the values (the host copy of the workgroup state starts with
    arbitrary contents, as LDS does)."""
    rng = np.random.default_rng(p)
    A = rng.standard_normal((p, p + 10))
    A[3, 7] = value
    with np.errstate(all="ignore"):
        G = A @ A.T
    rc, _, _ = eig(lib, G)
    assert rc == 1
