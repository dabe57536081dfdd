"""Tensor primitives of src/tensors.jl restated in numpy (column-major semantics).

Indices are 1-based in the arguments, as in the reference.
"""
from __future__ import annotations

import numpy as np


def _as_list(i):
    return [i] if isinstance(i, (int, np.integer)) else list(i)


def contract(x, y, idx1, idx2, conjx=False, conjy=False):
    """contract(x, y, idxs1, idxs2, conjx, conjy), src/tensors.jl:9-18.

    Result indices: the open indices of x in order, then the open indices of y.
    """
    i1, i2 = _as_list(idx1), _as_list(idx2)
    if len(i1) != len(i2):
        raise ValueError("The length of contracting indexs differ.")
    x = np.conj(x) if conjx else x
    y = np.conj(y) if conjy else y
    return np.tensordot(x, y, axes=([i - 1 for i in i1], [i - 1 for i in i2]))


def moveidx(x, currentidx, newidx):
    """moveidx(x, currentidx, newidx), src/tensors.jl:135-160 (newidx = -1 means last)."""
    n = x.ndim
    newidx = n if newidx == -1 else newidx
    return np.moveaxis(x, currentidx - 1, newidx - 1).copy()


def trace(x, idx1, idx2):
    """trace(x, idx1, idx2), src/tensors.jl:75-78."""
    return np.trace(x, axis1=idx1 - 1, axis2=idx2 - 1)


def truncation_rank(S, cutoff=0.0, maxdim=0, mindim=1):
    """Number of singular values kept: the rule of src/tensors.jl:202-215.

    maxdim = 0 means unlimited; with cutoff != 0 the kept rank is the last i whose
    tail weight sum_{j>=i} s_j^2 / sum s^2 exceeds cutoff (1 if none). The
    `findfirst(S == 0)` test at :205-206 compares a vector with a scalar and never
    fires, so it is not restated. Returns (k, discarded_weight).
    """
    S = np.asarray(S, dtype=np.float64)
    sz = S.shape[0]
    mindim = min(mindim, sz)
    maxdim = sz if (maxdim == 0 or maxdim > sz) else maxdim
    maxdim = 1 if maxdim == 0 else maxdim
    S2 = S ** 2
    total = S2.sum()
    if cutoff != 0:
        with np.errstate(invalid="ignore", divide="ignore"):
            S2cum = np.cumsum(S2[::-1])[::-1] / total
        above = np.nonzero(S2cum > cutoff)[0]
        idxs = int(above[-1]) + 1 if above.size else 1
        maxdim = min(maxdim, idxs)
    k = max(maxdim, mindim)
    disc = float(S2[k:].sum() / total) if total > 0 else 0.0
    return k, disc


def svd_matrix(M, cutoff=0.0, maxdim=0, mindim=1):
    """Truncated thin SVD of a matrix (LAPACK gesdd, gesvd if it raises): the
    decomposition of src/tensors.jl:189-220. Returns U (m,k), S (k,), Vt (k,n), disc."""
    try:
        U, S, Vt = np.linalg.svd(M, full_matrices=False)
    except np.linalg.LinAlgError:
        import scipy.linalg as sl

        U, S, Vt = sl.svd(M, full_matrices=False, lapack_driver="gesvd")
    k, disc = truncation_rank(S, cutoff, maxdim, mindim)
    return U[:, :k], S[:k], Vt[:k, :], disc


def svd(x, idx, cutoff=0.0, maxdim=0, mindim=1):
    """svd(x, idx; cutoff, maxdim, mindim), src/tensors.jl:168-227.

    Rows are all indices except idx (in order, first fastest), columns are idx.
    U is returned with the new bond at position idx; S as a vector; V is (k, dim_idx).
    """
    n = x.ndim
    idx = n if idx == -1 else idx
    y = np.moveaxis(x, idx - 1, -1)
    other = y.shape[:-1]
    M = y.reshape((-1, y.shape[-1]), order="F")
    U, S, Vt, disc = svd_matrix(M, cutoff, maxdim, mindim)
    U = U.reshape(other + (U.shape[1],), order="F")
    U = np.moveaxis(U, -1, idx - 1)
    return U, S, Vt, disc


def expm_tensor(x, outeridxs):
    """exp(x, outeridxs), src/tensors.jl:319-338, for a gate tensor G[o1,i1,o2,i2]
    with outer ("in") indices [2, 4]: the matrix exponential of
    G[(o1,o2),(i1,i2)] reshaped back."""
    import scipy.linalg as sl

    inner = [i for i in range(1, x.ndim + 1) if i not in outeridxs]
    perm = [i - 1 for i in inner] + [i - 1 for i in outeridxs]
    y = np.transpose(x, perm)
    ish = y.shape[: len(inner)]
    osh = y.shape[len(inner):]
    M = y.reshape((int(np.prod(ish)), int(np.prod(osh))), order="F")
    E = sl.expm(M)
    E = E.reshape(ish + osh, order="F")
    return np.transpose(E, np.argsort(perm))
