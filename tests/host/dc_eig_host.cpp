// Host build of the small-matrix eigensolver (tensornetworks.jl_amd/csrc/dc_eig.hpp): the
// same algorithm the device kernel instantiates, run with a sequential execution policy, plus
// the back-transformation by the Householder vectors. Built by tests/test_dc_eig_host.py.
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../tensornetworks.jl_amd/csrc/dc_eig.hpp"

namespace {
struct HostEx {
  template <class F> void par(int n, F &&f) const {
    for (int t = 0; t < n; ++t) f(t);
  }
  void sync() const {}
  void mark(int) const {}
  template <class F> double sum(int n, F &&f, double *) const {
    double s = 0;
    for (int t = 0; t < n; ++t) s += f(t);
    return s;
  }
  template <class F> double vmax(int n, F &&f, double *) const {
    double s = 0;
    for (int t = 0; t < n; ++t) s = std::fmax(s, f(t));
    return s;
  }
  int gsize(int) const { return 1; }
  int nthreads() const { return 1024; }
  double gsum(double x, int) const { return x; }
  void gsum4(double &, double &, double &, double &, int) const {}
};
}  // namespace

extern "C" int dc_eig_host(int p, int cplx, double *Gr, double *Gi, double *lam, double *Vr, double *Vi) {
  using namespace tnh::dc;
  if (p < 1 || p > kDcPMax) return -1;
  auto *sh = new DcShared();
  // LDS starts with arbitrary contents on the device: so does this copy
  std::memset(static_cast<void *>(sh), 0x5a, sizeof(DcShared));
  std::vector<double> taur(p), taui(p), Q(size_t(p) * p), Q2(size_t(p) * p), U(size_t(p) * p), out(p + 1);
  HostEx ex;
  if (cplx)
    dc_eig<true>(ex, *sh, Gr, Gi, p, taur.data(), taui.data(), Q.data(), Q2.data(), U.data(), out.data());
  else
    dc_eig<false>(ex, *sh, Gr, nullptr, p, taur.data(), taui.data(), Q.data(), Q2.data(), U.data(), out.data());
  delete sh;
  for (int i = 0; i < p; ++i) lam[i] = out[i];
  // V = H_0 H_1 ... H_{p-2} Q
  for (size_t e = 0; e < size_t(p) * p; ++e) {
    Vr[e] = Q[e];
    if (cplx) Vi[e] = 0.0;
  }
  for (int k = p - 2; k >= 0; --k) {
    const double tr = taur[k], ti = taui[k];
    if (tr == 0.0 && ti == 0.0) continue;
    for (int c = 0; c < p; ++c) {
      double sr = 0, si = 0;  // s = v^H V[:, c]
      for (int i = k + 1; i < p; ++i) {
        const double vr = Gr[i + size_t(k) * p], vi = cplx ? Gi[i + size_t(k) * p] : 0.0;
        const double xr = Vr[i + size_t(c) * p], xi = cplx ? Vi[i + size_t(c) * p] : 0.0;
        sr += vr * xr + vi * xi;
        si += vr * xi - vi * xr;
      }
      const double fr = tr * sr - ti * si, fi = tr * si + ti * sr;  // tau s
      for (int i = k + 1; i < p; ++i) {
        const double vr = Gr[i + size_t(k) * p], vi = cplx ? Gi[i + size_t(k) * p] : 0.0;
        Vr[i + size_t(c) * p] -= vr * fr - vi * fi;
        if (cplx) Vi[i + size_t(c) * p] -= vr * fi + vi * fr;
      }
    }
  }
  return out[p] == 1.0 ? 0 : 1;
}
