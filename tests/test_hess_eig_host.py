"""The small non-Hermitian eigensolver of the Arnoldi path (tensornetworks.jl_amd/csrc/
hess_eig.hpp: implicit single-shift QR on an upper Hessenberg matrix, inverse iteration for
one eigenvector) compiled for the host and checked against LAPACK (numpy.linalg.eigvals):
every eigenvalue to 1e-12 of ||H|| (matched as multisets), the eigenvector of the smallest
real part with residual ||H y - lam y|| <= 1e-12 ||H||. Cases: real and complex, m from 1
to 31, Hessenberg matrices with complex-conjugate pairs, a defective-looking Jordan-like
block, a zero subdiagonal (already split), a diagonal matrix and a zero matrix."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent
SRC = ROOT / "host" / "hess_eig_host.cpp"


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    out = tmp_path_factory.mktemp("hess") / "hess_eig_host.so"
    subprocess.run([gxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-o", str(out), str(SRC)], check=True)
    L = C.CDLL(str(out))
    dp = C.POINTER(C.c_double)
    L.hess_eig_host.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, C.POINTER(C.c_int)]
    L.hess_eig_host.restype = C.c_int
    return L


def run(lib, H):
    m = H.shape[0]
    dp = C.POINTER(C.c_double)
    hr = np.asfortranarray(H.real.copy())
    hi = np.asfortranarray(np.imag(H).copy())
    lr, li, yr, yi = (np.zeros(m) for _ in range(4))
    w = C.c_int()
    rc = lib.hess_eig_host(m, hr.ctypes.data_as(dp), hi.ctypes.data_as(dp), lr.ctypes.data_as(dp),
                           li.ctypes.data_as(dp), yr.ctypes.data_as(dp), yi.ctypes.data_as(dp),
                           C.byref(w))
    assert rc == 0
    return lr + 1j * li, yr + 1j * yi, w.value


def check(lib, H):
    H = np.triu(H, -1)
    lam, y, w = run(lib, H)
    ref = np.linalg.eigvals(H)
    scale = max(1.0, np.abs(H).max())
    # match as multisets (greedy nearest)
    left = list(ref)
    for l in lam:
        j = int(np.argmin([abs(l - r) for r in left]))
        assert abs(l - left[j]) <= 1e-10 * scale, (l, left[j])
        left.pop(j)
    assert lam[w].real <= ref.real.min() + 1e-10 * scale
    assert abs(np.linalg.norm(y) - 1) <= 1e-12
    assert np.linalg.norm(H @ y - lam[w] * y) <= 1e-10 * scale


@pytest.mark.parametrize("m", [1, 2, 3, 5, 8, 16, 31])
@pytest.mark.parametrize("cplx", [False, True])
def test_random_hessenberg(lib, m, cplx):
    rng = np.random.default_rng(m + 100 * cplx)
    for _ in range(5):
        H = rng.standard_normal((m, m)) + (1j * rng.standard_normal((m, m)) if cplx else 0)
        check(lib, H)


def test_special_structures(lib):
    rng = np.random.default_rng(7)
    m = 12
    J = np.diag(np.full(m, 0.3)) + np.diag(np.full(m - 1, 1.0), -1) + 1e-3 * np.triu(rng.standard_normal((m, m)))
    check(lib, J)
    H = np.triu(rng.standard_normal((m, m)), -1)
    H[6, 5] = 0.0  # already split
    check(lib, H)
    check(lib, np.diag(rng.standard_normal(m)))
    check(lib, np.zeros((m, m)))
    # an Arnoldi projection of a non-symmetric matrix (conjugate pairs, clustered values)
    A = rng.standard_normal((60, 60)) / np.sqrt(60) + np.diag(np.linspace(-1, 1, 60))
    V = np.zeros((60, 21))
    Hh = np.zeros((21, 20))
    v = rng.standard_normal(60)
    V[:, 0] = v / np.linalg.norm(v)
    for j in range(20):
        w = A @ V[:, j]
        for _ in range(2):
            c = V[:, : j + 1].T @ w
            w -= V[:, : j + 1] @ c
            Hh[: j + 1, j] += c
        Hh[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / Hh[j + 1, j]
    check(lib, Hh[:20, :20])
