// Eigenpairs of a small complex upper Hessenberg matrix: the Arnoldi projection of the
// non-Hermitian eigsolve (KrylovKit's eigsolve with ishermitian = false, called from
// src/algorithms/mps/dmrg.jl:51-53). Host code, no HIP: the matrix is at most 32 x 32.
//
//   eigenvalues  all eigenvalues by the implicit single-shift QR algorithm on the active
//                block (Givens bulge chase, Wilkinson shift from the trailing 2 x 2 block,
//                an exceptional shift after 10 and 20 stalled iterations, deflation when a
//                subdiagonal entry falls below eps times its diagonal neighbours);
//   eigenvector  the eigenvector of one eigenvalue by inverse iteration on H - mu I (Gaussian
//                elimination with partial pivoting, mu nudged off the eigenvalue by a few ulps
//                of ||H|| so the system is solvable), unit 2-norm.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace tnh {
namespace hess {

using cd = std::complex<double>;

// H: m x m, column-major (H[i + j m]), upper Hessenberg (entries below the subdiagonal are
// ignored). lam receives the m eigenvalues. Returns false if an eigenvalue did not converge
// within 30 m iterations.
inline bool eigenvalues(int m, std::vector<cd> H, std::vector<cd> &lam) {
  lam.assign(m, cd(0, 0));
  auto at = [&](int i, int j) -> cd & { return H[size_t(i) + size_t(j) * m]; };
  const double eps = 2.220446049250313e-16;
  int hi = m - 1, iter = 0, total = 0;
  while (hi >= 0) {
    if (hi == 0) {
      lam[0] = at(0, 0);
      break;
    }
    // deflation: the lowest l with a negligible H[l][l-1] in the active block
    int l = hi;
    for (; l > 0; --l) {
      const double s = std::abs(at(l - 1, l - 1)) + std::abs(at(l, l));
      if (std::abs(at(l, l - 1)) <= eps * (s > 0 ? s : 1.0)) {
        at(l, l - 1) = 0;
        break;
      }
    }
    if (l == hi) {  // a 1 x 1 block has split off
      lam[hi] = at(hi, hi);
      --hi;
      iter = 0;
      continue;
    }
    if (++total > 30 * m) return false;
    // shift: the eigenvalue of the trailing 2 x 2 block nearer its last diagonal entry
    cd mu;
    if (iter == 10 || iter == 20) {
      mu = at(hi, hi) + cd(std::abs(at(hi, hi - 1)) * (iter == 10 ? 0.75 : -0.4375), 0);
    } else {
      const cd a = at(hi - 1, hi - 1), b = at(hi - 1, hi), c = at(hi, hi - 1), d = at(hi, hi);
      const cd half = 0.5 * (a - d);
      const cd disc = std::sqrt(half * half + b * c);
      const cd m1 = 0.5 * (a + d) + disc, m2 = 0.5 * (a + d) - disc;
      mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
    }
    ++iter;
    // one implicit shifted QR step on rows / columns l..hi
    for (int k = l; k < hi; ++k) {
      cd x, y;
      if (k == l) {
        x = at(l, l) - mu;
        y = at(l + 1, l);
      } else {
        x = at(k, k - 1);
        y = at(k + 1, k - 1);
      }
      const double nx = std::abs(x), ny = std::abs(y);
      const double r = std::hypot(nx, ny);
      if (r == 0.0) continue;
      // G = [[c, s], [-conj(s), c]] with G [x; y] = [r'; 0]: c = |x| / r, s = sign(x) conj(y) / r
      const cd sx = nx > 0 ? x / nx : cd(1, 0);
      const double c = nx / r;
      const cd s = sx * std::conj(y) / r;
      // rows k, k+1 (columns from max(l, k - 1) to hi)
      for (int j = std::max(l, k - 1); j <= hi; ++j) {
        const cd u = at(k, j), v = at(k + 1, j);
        at(k, j) = c * u + s * v;
        at(k + 1, j) = -std::conj(s) * u + c * v;
      }
      if (k > l) at(k + 1, k - 1) = 0;
      // columns k, k+1 (rows from l to min(k + 2, hi))
      for (int i = l; i <= std::min(k + 2, hi); ++i) {
        const cd u = at(i, k), v = at(i, k + 1);
        at(i, k) = c * u + std::conj(s) * v;
        at(i, k + 1) = -s * u + c * v;
      }
    }
  }
  return true;
}

// Unit-norm eigenvector y of the m x m matrix H (column-major) for the eigenvalue mu.
inline void eigenvector(int m, const std::vector<cd> &H, cd mu, std::vector<cd> &y) {
  double hn = 0;
  for (const cd &h : H) hn = std::max(hn, std::abs(h));
  const double eps = 2.220446049250313e-16;
  const cd shift = mu + cd(8 * eps * (hn > 0 ? hn : 1.0), 0);
  y.assign(m, cd(1.0 / std::sqrt(double(m)), 0));
  std::vector<cd> A(size_t(m) * m), b(m);
  for (int it = 0; it < 3; ++it) {
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) A[i + size_t(j) * m] = H[i + size_t(j) * m] - (i == j ? shift : cd(0, 0));
    b = y;
    // Gaussian elimination with partial pivoting
    for (int k = 0; k < m; ++k) {
      int piv = k;
      for (int i = k + 1; i < m; ++i)
        if (std::abs(A[i + size_t(k) * m]) > std::abs(A[piv + size_t(k) * m])) piv = i;
      if (piv != k) {
        for (int j = 0; j < m; ++j) std::swap(A[k + size_t(j) * m], A[piv + size_t(j) * m]);
        std::swap(b[k], b[piv]);
      }
      cd dkk = A[k + size_t(k) * m];
      if (std::abs(dkk) == 0.0) dkk = A[k + size_t(k) * m] = cd(eps * (hn > 0 ? hn : 1.0), 0);
      for (int i = k + 1; i < m; ++i) {
        const cd f = A[i + size_t(k) * m] / dkk;
        if (f == cd(0, 0)) continue;
        for (int j = k; j < m; ++j) A[i + size_t(j) * m] -= f * A[k + size_t(j) * m];
        b[i] -= f * b[k];
      }
    }
    for (int i = m - 1; i >= 0; --i) {
      cd s = b[i];
      for (int j = i + 1; j < m; ++j) s -= A[i + size_t(j) * m] * y[j];
      y[i] = s / A[i + size_t(i) * m];
    }
    double nr = 0;
    for (const cd &v : y) nr += std::norm(v);
    nr = std::sqrt(nr);
    for (cd &v : y) v /= nr;
  }
}

}  // namespace hess
}  // namespace tnh
