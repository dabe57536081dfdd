"""Restarted Lanczos for the lowest eigenpair, standing in for KrylovKit's
eigsolve(Heff, A0, 1, :SR; maxiter, krylovdim, ishermitian=true, tol) call at
src/algorithms/mps/dmrg.jl:51-53.

KrylovKit.jl is a third-party dependency absent from the reference tree (no pinned
version: Project.toml lists UUIDs only, Manifest.toml is git-ignored). Its published
algorithm for Hermitian problems is a Lanczos factorisation with full
re-orthogonalisation, restarted Krylov-Schur style keeping the wanted Ritz vector
(keep = div(3*krylovdim + 2*converged, 5) = 1 vector for krylovdim <= 3 and one
wanted pair). This restatement keeps exactly one Ritz vector per restart, so one
cycle costs krylovdim matvecs first and krylovdim - 1 after each restart.

Pins (tests/test_oracle_pins.py, tests/test_oracle.py): cycle c returns the lowest
Rayleigh-Ritz pair of H on K_k(H, y_{c-1}) (y_0 = x0, y_c the Ritz vector of cycle c),
computed independently from a QR basis of the monomial Krylov vectors, with
k + (c - 1)(k - 1) matvecs; the converged pair is pinned by dense diagonalisation.
Parity unpinned only for KrylovKit itself: that its version keeps exactly one vector per
restart for these parameters cannot be checked without running it.

eigsolve_arnoldi is the non-Hermitian counterpart (ishermitian = false, the eigenvalue of
smallest real part): pinned the same way (Rayleigh-Ritz on K_k(A, y_{c-1}) cycle by cycle,
the converged pair against dense eigenvalues, the DMRG built on it against ED of the
Hatano-Nelson chain), with the same unpinned caveat about KrylovKit's own restart.
"""
from __future__ import annotations

import numpy as np


def eigsolve_ground(matvec, x0, krylovdim=3, maxiter=2, tol=1e-14):
    """Return (energy, normalised Ritz vector, number of matvecs)."""
    shape = x0.shape
    x = x0.reshape(-1, order="F").astype(np.result_type(x0, np.float64), copy=True)
    n = x.size
    kdim = max(1, min(krylovdim, n))
    V = np.zeros((kdim + 1, n), dtype=x.dtype)
    H = np.zeros((kdim + 1, kdim), dtype=x.dtype)
    V[0] = x / np.linalg.norm(x)
    start = 0
    nmv = 0

    def mv(v):
        return matvec(v.reshape(shape, order="F")).reshape(-1, order="F")

    theta, y, m = 0.0, None, 0
    for it in range(maxiter):
        m = kdim
        breakdown = False
        for j in range(start, kdim):
            w = mv(V[j])
            nmv += 1
            if np.iscomplexobj(w) and not np.iscomplexobj(V):  # complex operator, real start
                V, H = V.astype(np.complex128), H.astype(np.complex128)
            for _ in range(2):  # classical Gram-Schmidt, twice
                c = V[: j + 1].conj() @ w
                H[: j + 1, j] += c
                w = w - c @ V[: j + 1]
            beta = np.linalg.norm(w)
            H[j + 1, j] = beta
            scale = max(1.0, np.abs(H[: j + 2, : j + 1]).max())
            if beta <= 1e-14 * scale:
                m = j + 1
                breakdown = True
                break
            V[j + 1] = w / beta
        Hm = H[:m, :m]
        Hm = 0.5 * (Hm + Hm.conj().T)
        evals, evecs = np.linalg.eigh(Hm)
        theta, y = float(evals[0]), evecs[:, 0]
        b = H[m, :m] @ y  # residual coefficient on V[m]
        res = abs(b)
        if breakdown or res <= tol or it == maxiter - 1:
            break
        # restart: keep the Ritz vector and the residual direction
        ritz = y @ V[:m]
        V2 = np.zeros_like(V)
        V2[0] = ritz / np.linalg.norm(ritz)
        V2[1] = V[m]
        V = V2
        H = np.zeros_like(H)
        H[0, 0] = theta
        H[1, 0] = b
        start = 1
    ritz = y @ V[:m]
    ritz = ritz / np.linalg.norm(ritz)
    return theta, ritz.reshape(shape, order="F"), nmv


def eigsolve_arnoldi(matvec, x0, krylovdim=3, maxiter=2, tol=1e-14):
    """The non-Hermitian counterpart, eigsolve(Heff, A0, 1, :SR; ishermitian=false, ...) of
    dmrg.jl:51-53: Arnoldi with the same two-pass Gram-Schmidt and breakdown test, the Ritz
    value of smallest real part of the Hessenberg projection (scipy.linalg.eig) and the same
    one-vector restart (A u = lambda u + b v_m, so the next projection starts with
    [lambda; b] in column 0). KrylovKit's non-Hermitian eigsolve restarts from a Schur form;
    with one wanted value and one kept vector that is the same subspace. Returns (complex Ritz
    value, normalised Ritz vector, number of matvecs). Pinned like the Lanczos: cycle c is
    Rayleigh-Ritz on K_k(A, y_{c-1}), and the converged pair against dense eigenvalues."""
    import scipy.linalg as sla

    shape = x0.shape
    x = x0.reshape(-1, order="F").astype(np.complex128, copy=True)
    n = x.size
    kdim = max(1, min(krylovdim, n))
    V = np.zeros((kdim + 1, n), dtype=np.complex128)
    H = np.zeros((kdim + 1, kdim), dtype=np.complex128)
    V[0] = x / np.linalg.norm(x)
    start = 0
    nmv = 0

    def mv(v):
        return matvec(v.reshape(shape, order="F")).reshape(-1, order="F")

    lam, y, m = 0.0, None, 0
    for it in range(maxiter):
        m = kdim
        breakdown = False
        for j in range(start, kdim):
            w = mv(V[j]).astype(np.complex128)
            nmv += 1
            for _ in range(2):
                c = V[: j + 1].conj() @ w
                H[: j + 1, j] += c
                w = w - c @ V[: j + 1]
            beta = np.linalg.norm(w)
            H[j + 1, j] = beta
            scale = max(1.0, np.abs(H[: j + 2, : j + 1]).max())
            if beta <= 1e-14 * scale:
                m = j + 1
                breakdown = True
                break
            V[j + 1] = w / beta
        evals, evecs = sla.eig(H[:m, :m])
        i = int(np.argmin(evals.real))                            # :SR
        lam, y = complex(evals[i]), evecs[:, i] / np.linalg.norm(evecs[:, i])
        b = H[m, :m] @ y
        if breakdown or abs(b) <= tol or it == maxiter - 1:
            break
        ritz = y @ V[:m]
        V2 = np.zeros_like(V)
        V2[0] = ritz / np.linalg.norm(ritz)
        V2[1] = V[m]
        V = V2
        H = np.zeros_like(H)
        H[0, 0] = lam
        H[1, 0] = b
        start = 1
    ritz = y @ V[:m]
    ritz = ritz / np.linalg.norm(ritz)
    return lam, ritz.reshape(shape, order="F"), nmv
