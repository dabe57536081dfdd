// Host entry point over csrc/hess_eig.hpp for tests/test_hess_eig_host.py: all eigenvalues of
// an m x m upper Hessenberg matrix (planar real / imaginary parts, column-major) and the
// unit-norm eigenvector of the one with the smallest real part.
#include "../../tensornetworks.jl_amd/csrc/hess_eig.hpp"

extern "C" int hess_eig_host(int m, const double *hr, const double *hi, double *lr, double *li,
                             double *yr, double *yi, int *which) {
  using tnh::hess::cd;
  std::vector<cd> H(size_t(m) * m);
  for (size_t e = 0; e < H.size(); ++e) H[e] = cd(hr[e], hi[e]);
  std::vector<cd> lam, y;
  if (!tnh::hess::eigenvalues(m, H, lam)) return 1;
  int w = 0;
  for (int i = 0; i < m; ++i) {
    lr[i] = lam[i].real();
    li[i] = lam[i].imag();
    if (lam[i].real() < lam[w].real()) w = i;
  }
  tnh::hess::eigenvector(m, H, lam[w], y);
  for (int i = 0; i < m; ++i) {
    yr[i] = y[i].real();
    yi[i] = y[i].imag();
  }
  *which = w;
  return 0;
}
