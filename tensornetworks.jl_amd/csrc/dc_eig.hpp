// Eigen-decomposition of a small Hermitian matrix (p <= kDcPMax) by one workgroup:
// Householder reduction to a real symmetric tridiagonal matrix T, then bottom-up divide and
// conquer on T (Cuppen's tearing; deflation of negligible coupling components and of close
// diagonal pairs by Givens rotations; every secular-equation root by a safeguarded two-pole
// rational iteration in coordinates shifted to the nearer pole; eigenvectors from the
// Gu-Eisenstat recomputed coupling vector, so they are orthogonal to working precision).
// The Householder vectors are left in the matrix for the back-transformation.
//
// A Jacobi eigensolver needs ~p dependent steps per sweep and 11-17 sweeps; this needs p
// dependent Householder steps and log2(p) merge levels of parallel work, which is what a
// single latency-bound workgroup wants.
//
// The algorithm is written once against an execution policy `Ex` that supplies `par(n, f)`
// (f(t) for t in [0, n), spread over the workgroup), `sync()` (a workgroup barrier) and
// `sum(n, f, scratch)` (a workgroup-wide sum returned to every thread). The device policy
// (in the .hip file) maps these to threads and barriers; a host policy that runs the loops
// sequentially compiles the same code for tests on the CPU.
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define TNH_HD __host__ __device__
#else
#define TNH_HD
#endif

namespace tnh {
namespace dc {

constexpr int kDcPMax = 512;   // largest matrix side
constexpr int kDcThreads = 1024;  // workgroup size the partial-sum layout is sized for
constexpr double kEps = 2.220446049250313e-16;

// Workgroup-shared state (LDS on the device). Per merge m (start b, split bm, end b2) the
// per-position arrays hold the merge's entries at [b, b2).
struct DcShared {
  double d[kDcPMax], e[kDcPMax], lam[kDcPMax];
  double zt[kDcPMax], Dt[kDcPMax];  // coupling vector and diagonal, in column order
  double Ds[kDcPMax], zs[kDcPMax];  // the same, sorted ascending within each merge
  double dk[kDcPMax], zk[kDcPMax];  // the kept (not deflated) entries, compacted
  double zh[kDcPMax], tau[kDcPMax];  // recomputed coupling vector; root offsets
  double rc[kDcPMax], rs[kDcPMax];  // Givens rotations of the deflation, per merge
  double vr[kDcPMax], vi[kDcPMax], yr[kDcPMax], yi[kDcPMax];  // Householder step vectors
  double part[2 * kDcThreads];  // matrix-vector partial sums (real, imaginary)
  double red[64];
  double mrho[kDcPMax / 2];
  int src[kDcPMax];   // sorted position -> column it came from
  int kpos[kDcPMax];  // kept compact index -> sorted position (relative to the merge start)
  int dpos[kDcPMax];  // deflated compact index -> sorted position
  int org[kDcPMax];   // root -> compact index of the pole its offset is measured from
  int fsrc[kDcPMax];  // final position -> item (root q < K, or K + deflated index)
  int rj[kDcPMax], ri[kDcPMax];
  int kflag[kDcPMax];  // sorted position kept (1) or deflated (0)
  int mK[kDcPMax / 2], mrot[kDcPMax / 2];
  int status;
};

TNH_HD inline bool finite(double v) { return fabs(v) <= 1.7976931348623157e308; }
// Ordering key of the merge sorts: NaN after everything, so that with ties broken by position
// every rank below is a permutation whatever the values (the gathers index with it).
TNH_HD inline double order_key(double v) { return v == v ? v : INFINITY; }

// Merge geometry at a level whose blocks have size s (the last may be shorter).
struct MergeAt {
  int m, b, bm, b2;
  TNH_HD bool pass() const { return bm >= b2; }
};
TNH_HD inline MergeAt merge_of(int j, int s, int p) {
  MergeAt g;
  g.m = j / (2 * s);
  g.b = 2 * s * g.m;
  g.bm = g.b + s < p ? g.b + s : p;
  g.b2 = g.b + 2 * s < p ? g.b + 2 * s : p;
  return g;
}

// Root j (0-based) of 1 + rho sum_k z2[k] / (dl[k] - lambda) = 0 for dl ascending, rho > 0,
// sum z2 = 1: lambda in (dl[j], dl[j + 1]), the last in (dl[K - 1], dl[K - 1] + rho). Returns
// lambda as dl[*org] + tau with org the nearer pole, so that lambda - dl[i] is formed as
// (dl[org] - dl[i]) + tau without cancellation. Each iteration fits the two poles around the
// root (psi, the terms up to j, and phi, the rest, each matched in value and slope) and takes
// the root of that model; a step leaving the current bracket is replaced by bisection.
TNH_HD inline double secular_root(const double *dl, const double *zk, double rho, int j, int K, int *org,
                                  int *iters) {
  double lo, hi;
  int o;
  if (j < K - 1) {
    const double gap = dl[j + 1] - dl[j], mid = 0.5 * gap;
    double f = 1.0;
    for (int k = 0; k < K; ++k) f += rho * zk[k] * zk[k] / ((dl[k] - dl[j]) - mid);
    if (f >= 0) {  // root in the left half: measured from dl[j]
      o = j;
      lo = 0.0;
      hi = mid;
    } else {  // right half: measured from dl[j + 1]
      o = j + 1;
      lo = -mid;
      hi = 0.0;
    }
  } else {
    o = K - 1;
    lo = 0.0;
    hi = rho;
  }
  *org = o;
  const double dor = dl[o];
  double t = 0.5 * (lo + hi);
  int it = 0;
  for (; it < 96; ++it) {
    double psi = 0, dpsi = 0, phi = 0, dphi = 0;
    for (int k = 0; k < K; ++k) {
      const double df = (dl[k] - dor) - t;
      const double q = rho * zk[k] * zk[k] / df;
      if (k <= j) {
        psi += q;
        dpsi += q / df;
      } else {
        phi += q;
        dphi += q / df;
      }
    }
    const double f = 1.0 + psi + phi;
    // f is zero to within the rounding of its own evaluation: t is a root
    if (fabs(f) <= 8.0 * kEps * (1.0 + fabs(psi) + fabs(phi))) break;
    if (f > 0)
      hi = t < hi ? t : hi;
    else
      lo = t > lo ? t : lo;
    double xn;
    if (j < K - 1) {
      const double a = (dl[j] - dor) - t, bb = (dl[j + 1] - dor) - t;
      const double s = dpsi * a * a, S = dphi * bb * bb;
      const double c = 1.0 + psi - s / a + phi - S / bb;
      const double dj = dl[j] - dor, dj1 = dl[j + 1] - dor;
      // c (dj - x)(dj1 - x) + s (dj1 - x) + S (dj - x) = 0
      const double A2 = c, B2 = -(c * (dj + dj1) + s + S), C2 = c * dj * dj1 + s * dj1 + S * dj;
      xn = 0.5 * (lo + hi);
      if (A2 == 0.0) {
        if (B2 != 0.0) xn = -C2 / B2;
      } else {
        const double disc = B2 * B2 - 4.0 * A2 * C2;
        const double sq = sqrt(disc > 0 ? disc : 0.0);
        const double r1 = (-B2 - copysign(sq, B2)) / (2.0 * A2);
        const double r2 = r1 != 0.0 ? C2 / (A2 * r1) : r1;
        if (lo < r1 && r1 < hi)
          xn = r1;
        else if (lo < r2 && r2 < hi)
          xn = r2;
      }
    } else {
      const double dj = dl[K - 1] - dor, a = dj - t;
      const double s = dpsi * a * a;
      const double c = 1.0 + psi - s / a + phi;
      xn = c != 0.0 ? dj + s / c : 0.5 * (lo + hi);
    }
    if (!(lo < xn && xn < hi)) xn = 0.5 * (lo + hi);
    const double step = fabs(xn - t);
    t = xn;
    if (step <= 4.0 * kEps * fabs(t) || hi - lo <= 4.0 * kEps * (fabs(lo) > fabs(hi) ? fabs(lo) : fabs(hi))) break;
  }
  *iters = it;
  return t;
}

// The reduction and divide-and-conquer on the device-resident p x p Hermitian matrix
// (Gr, Gi planar, leading dimension p; Gi unused when !CPLX). On return: the Householder
// vectors in the columns of G below the subdiagonal (v_k[0] = 1 stored), their factors in
// taur/taui[0 .. p-2], the eigenvalues (ascending) in out[0 .. p), the eigenvectors of T in
// Q (p x p, column-major, matching order), and out[p] = 1 on success (finite values,
// converged roots), 0 otherwise. Q2 and U are p x p work matrices.
template <bool CPLX, class Ex>
TNH_HD void dc_eig(const Ex &ex, DcShared &sh, double *Gr, double *Gi, int p, double *taur, double *taui,
                   double *Q, double *Q2, double *U, double *out) {
  const int64_t P = p;
  // exact Hermitian symmetrisation and a power-of-two scaling to max |G_ii| in [1, 2)
  ex.par(p * p, [&](int e) {
    const int i = e % p, j = e / p;
    if (i < j) {
      const double ar = 0.5 * (Gr[i + j * P] + Gr[j + i * P]);
      Gr[i + j * P] = ar;
      Gr[j + i * P] = ar;
      if (CPLX) {
        const double ai = 0.5 * (Gi[i + j * P] - Gi[j + i * P]);
        Gi[i + j * P] = ai;
        Gi[j + i * P] = -ai;
      }
    } else if (i == j && CPLX) {
      Gi[i + i * P] = 0.0;
    }
  });
  ex.sync();
  double dm = 0;
  for (int i = 0; i < p; ++i) dm = fmax(dm, fabs(Gr[i + i * P]));
  const int ex2 = (dm > 0 && dm < 1e300) ? ilogb(dm) : 0;
  if (ex2 != 0) {
    const double sc = ldexp(1.0, -ex2);
    ex.sync();  // every thread has read the diagonal
    ex.par(p * p, [&](int e) {
      Gr[e] *= sc;
      if (CPLX) Gi[e] *= sc;
    });
    ex.sync();
  }
  // ---- Householder reduction: Q_h^H G Q_h = T, real off-diagonal (beta_k). Five
  // barriers per step: the matrix-vector partial sums; y = B v with v^H y; the rank-2 update,
  // which also sums the squared tail of the next step's column.
  double xn2 = ex.sum(p > 2 ? p - 2 : 0, [&](int i) {
    const double a = Gr[2 + i], b = CPLX ? Gi[2 + i] : 0.0;
    return a * a + b * b;
  }, sh.red);
  for (int k = 0; k + 1 < p; ++k) {
    const int n = p - k - 1;
    double *col_r = Gr + k * P + (k + 1), *col_i = CPLX ? Gi + k * P + (k + 1) : nullptr;
    const double ar = col_r[0], ai = CPLX ? col_i[0] : 0.0;
    const double dkk = Gr[k + k * P];
    double tr = 0, ti = 0, beta = ar, sr = 0, si = 0;
    if (xn2 != 0.0 || ai != 0.0) {
      beta = -copysign(sqrt(ar * ar + ai * ai + xn2), ar);
      tr = (beta - ar) / beta;
      ti = -ai / beta;
      // v = x / (alpha - beta), v_0 = 1
      const double ur = ar - beta, ui = ai, den = ur * ur + ui * ui;
      sr = ur / den;
      si = -ui / den;
    }
    ex.par(1, [&](int) {
      sh.d[k] = dkk;
      sh.e[k] = beta;
      taur[k] = tr;
      taui[k] = ti;
    });
    double *Br = Gr + (k + 1) * P + (k + 1), *Bi = CPLX ? Gi + (k + 1) * P + (k + 1) : nullptr;
    if (tr == 0.0 && ti == 0.0) {  // nothing to reflect: the next column's tail norm
      xn2 = ex.sum(n > 2 ? n - 2 : 0, [&](int i) {
        const double a = Br[2 + i], b = CPLX ? Bi[2 + i] : 0.0;
        return a * a + b * b;
      }, sh.red);
      continue;
    }
    auto vget = [&](int j, double &a, double &b) {
      if (j == 0) {
        a = 1.0;
        b = 0.0;
      } else {
        const double xr = col_r[j], xi = CPLX ? col_i[j] : 0.0;
        a = xr * sr - xi * si;
        b = xr * si + xi * sr;
      }
    };
    // y = B v, B = G[k+1.., k+1..]: nc column chunks per row, partial sums in LDS; v is
    // formed on the fly from the column (still x here)
    const int nc = kDcThreads / n > 0 ? kDcThreads / n : 1;
    ex.par(n * nc, [&](int e) {
      const int i = e % n, c = e / n;
      double sr2 = 0, si2 = 0;
      for (int j = c; j < n; j += nc) {
        double vrj, vij;
        vget(j, vrj, vij);
        const double br = Br[i + j * P];
        if (CPLX) {
          const double bi = Bi[i + j * P];
          sr2 += br * vrj - bi * vij;
          si2 += br * vij + bi * vrj;
        } else {
          sr2 += br * vrj;
        }
      }
      sh.part[e] = sr2;
      if (CPLX) sh.part[kDcThreads + e] = si2;
    });
    ex.sync();
    // y from the partials, v into LDS and into the column (for the back-transformation),
    // and v^H y (real for Hermitian B)
    const double vy = ex.sum(n, [&](int i) {
      double a = 0, b = 0;
      for (int c = 0; c < nc; ++c) {
        a += sh.part[i + c * n];
        if (CPLX) b += sh.part[kDcThreads + i + c * n];
      }
      sh.yr[i] = a;
      sh.yi[i] = b;
      double va, vb;
      vget(i, va, vb);
      sh.vr[i] = va;
      sh.vi[i] = vb;
      col_r[i] = va;
      if (CPLX) col_i[i] = vb;
      return va * a + vb * b;
    }, sh.red);
    const double h = 0.5 * (tr * tr + ti * ti) * vy;
    // B -= w v^H + v w^H with w = tau y - h v; the updated column 0 of B is the next
    // step's x, whose tail (rows 2..) is summed here
    xn2 = ex.sum(n * n, [&](int e) {
      const int i = e % n, j = e / n;
      const double wir = tr * sh.yr[i] - ti * sh.yi[i] - h * sh.vr[i];
      const double wjr = tr * sh.yr[j] - ti * sh.yi[j] - h * sh.vr[j];
      double nr, ni = 0.0;
      if (CPLX) {
        const double wii = tr * sh.yi[i] + ti * sh.yr[i] - h * sh.vi[i];
        const double wji = tr * sh.yi[j] + ti * sh.yr[j] - h * sh.vi[j];
        // w_i conj(v_j) + v_i conj(w_j)
        const double re = wir * sh.vr[j] + wii * sh.vi[j] + sh.vr[i] * wjr + sh.vi[i] * wji;
        const double im = wii * sh.vr[j] - wir * sh.vi[j] + sh.vi[i] * wjr - sh.vr[i] * wji;
        nr = Br[i + j * P] - re;
        ni = Bi[i + j * P] - im;
        Br[i + j * P] = nr;
        Bi[i + j * P] = ni;
      } else {
        nr = Br[i + j * P] - (wir * sh.vr[j] + sh.vr[i] * wjr);
        Br[i + j * P] = nr;
      }
      return (j == 0 && i >= 2) ? nr * nr + ni * ni : 0.0;
    }, sh.red);
  }
  ex.par(1, [&](int) {
    sh.d[p - 1] = Gr[(p - 1) * (P + 1)];
    sh.status = 1;
  });
  ex.sync();
  // a non-finite input leaves a non-finite tridiagonal matrix: report it and stop
  const double bad = ex.sum(p, [&](int i) {
    return (finite(sh.d[i]) && (i + 1 >= p || finite(sh.e[i]))) ? 0.0 : 1.0;
  }, sh.red);
  if (bad != 0.0) {
    ex.par(p + 1, [&](int i) { out[i] = i < p ? NAN : 0.0; });
    return;
  }
  // ---- divide and conquer on T = tridiag(e, d, e): leaves of size 1 with every
  // off-diagonal torn (d_i -= |e_{i-1}| + |e_i|), merged pairwise level by level
  ex.par(p, [&](int i) {
    double v = sh.d[i];
    if (i > 0) v -= fabs(sh.e[i - 1]);
    if (i + 1 < p) v -= fabs(sh.e[i]);
    sh.lam[i] = v;
  });
  ex.par(p * p, [&](int e) { Q[e] = (e % p == e / p) ? 1.0 : 0.0; });
  ex.sync();
  for (int s = 1; s < p; s *= 2) {
    const int nm = (p + 2 * s - 1) / (2 * s);
    // coupling vector z = Q^T u, u = e_{bm-1} + sign(e) e_{bm}: the last row of the left
    // block's eigenvectors and the first row of the right block's
    ex.par(p, [&](int j) {
      const MergeAt g = merge_of(j, s, p);
      if (g.pass()) return;
      const double sg = sh.e[g.bm - 1] >= 0 ? 1.0 : -1.0;
      sh.zt[j] = j < g.bm ? Q[(g.bm - 1) + j * P] : sg * Q[g.bm + j * P];
      sh.Dt[j] = sh.lam[j];
    });
    ex.sync();
    // merge the two ascending halves (stable: the left half first among equals)
    ex.par(p, [&](int j) {
      const MergeAt g = merge_of(j, s, p);
      if (g.pass()) return;
      int rank;
      const double kj = order_key(sh.Dt[j]);
      if (j < g.bm) {
        int c = 0;
        for (int t = g.bm; t < g.b2; ++t) c += order_key(sh.Dt[t]) < kj;
        rank = (j - g.b) + c;
      } else {
        int c = 0;
        for (int t = g.b; t < g.bm; ++t) c += order_key(sh.Dt[t]) <= kj;
        rank = (j - g.bm) + c;
      }
      sh.Ds[g.b + rank] = sh.Dt[j];
      sh.zs[g.b + rank] = sh.zt[j];
      sh.src[g.b + rank] = j;
    });
    ex.sync();
    // deflation, one thread per merge: normalise z (rho absorbs |z|^2); a component with
    // rho |z_i| <= tol leaves (D_i, e_i) as an eigenpair; two poles closer than tol / |cs|
    // are rotated so that one of them has z = 0 and leaves too
    ex.par(nm, [&](int m) {
      const int b = 2 * s * m, bm = b + s < p ? b + s : p, b2 = b + 2 * s < p ? b + 2 * s : p;
      if (bm >= b2) return;
      const int n = b2 - b;
      double zn2 = 0, dmax = 0;
      for (int t = 0; t < n; ++t) {
        zn2 += sh.zs[b + t] * sh.zs[b + t];
        dmax = fmax(dmax, fabs(sh.Ds[b + t]));
      }
      double rho = fabs(sh.e[bm - 1]);
      int K = 0, nd = 0, nr = 0;
      for (int t = 0; t < n; ++t) sh.kflag[b + t] = 0;
      if (zn2 > 0 && rho > 0) {
        const double zn = sqrt(zn2);
        for (int t = 0; t < n; ++t) sh.zs[b + t] /= zn;
        rho *= zn2;
        const double tol = 8.0 * kEps * fmax(dmax, rho);
        int last = -1;  // sorted position of the last kept entry
        for (int t = 0; t < n; ++t) {
          const double zt = sh.zs[b + t];
          if (rho * fabs(zt) <= tol) continue;
          if (last >= 0) {
            const double zj = sh.zs[b + last], r = sqrt(zj * zj + zt * zt);
            const double c = zt / r, sn = zj / r;
            const double Dj = sh.Ds[b + last], Di = sh.Ds[b + t];
            if (fabs((Di - Dj) * c * sn) <= tol) {
              sh.zs[b + last] = 0.0;
              sh.zs[b + t] = r;
              sh.Ds[b + last] = Dj * c * c + Di * sn * sn;
              sh.Ds[b + t] = Dj * sn * sn + Di * c * c;
              sh.rj[b + nr] = last;
              sh.ri[b + nr] = t;
              sh.rc[b + nr] = c;
              sh.rs[b + nr] = sn;
              ++nr;
              sh.kflag[b + last] = 0;
            }
          }
          sh.kflag[b + t] = 1;
          last = t;
        }
      }
      // kept entries compacted in ascending sorted position (so ascending D), and the
      // deflated positions likewise
      for (int t = 0; t < n; ++t) {
        if (sh.kflag[b + t]) {
          sh.kpos[b + K] = t;
          sh.dk[b + K] = sh.Ds[b + t];
          sh.zk[b + K] = sh.zs[b + t];
          ++K;
        } else {
          sh.dpos[b + nd++] = t;
        }
      }
      sh.mK[m] = K;
      sh.mrot[m] = nr;
      sh.mrho[m] = rho;
    });
    ex.sync();
    // working copy of the merge block's eigenvectors in sorted order, then the rotations
    ex.par(p * p, [&](int e) {
      const int r = e % p, c = e / p;
      const MergeAt g = merge_of(c, s, p);
      if (g.pass() || r < g.b || r >= g.b2) return;
      const int sc = sh.src[c];
      if (sc < g.b || sc >= g.b2) {  // not reachable: the rank above is a permutation
        sh.status = 0;
        return;
      }
      Q2[e] = Q[r + sc * P];
    });
    ex.sync();
    ex.par(p, [&](int r) {
      const MergeAt g = merge_of(r, s, p);
      if (g.pass()) return;
      const int nr = sh.mrot[g.m];
      for (int q = 0; q < nr; ++q) {
        const int64_t cj = r + (g.b + sh.rj[g.b + q]) * P, ci = r + (g.b + sh.ri[g.b + q]) * P;
        const double c = sh.rc[g.b + q], sn = sh.rs[g.b + q];
        const double a = Q2[cj], bq = Q2[ci];
        Q2[cj] = c * a - sn * bq;
        Q2[ci] = sn * a + c * bq;
      }
    });
    // roots of the secular equations (one thread per root)
    ex.par(p, [&](int a) {
      const MergeAt g = merge_of(a, s, p);
      if (g.pass()) return;
      const int j = a - g.b, K = sh.mK[g.m];
      if (j >= K) return;
      int o = 0, its = 0;
      const double t = secular_root(sh.dk + g.b, sh.zk + g.b, sh.mrho[g.m], j, K, &o, &its);
      sh.org[a] = o;
      sh.tau[a] = t;
      if (!(t == t) || its >= 96) sh.status = 0;
    });
    ex.sync();
    // Gu-Eisenstat: the coupling vector for which the computed roots are exact
    ex.par(p, [&](int a) {
      const MergeAt g = merge_of(a, s, p);
      if (g.pass()) return;
      const int i = a - g.b, K = sh.mK[g.m];
      if (i >= K) return;
      const double *dl = sh.dk + g.b;
      const double di = dl[i];
      double prod = (dl[sh.org[a]] - di) + sh.tau[a];
      for (int k = 0; k < K; ++k) {
        if (k == i) continue;
        prod *= ((dl[sh.org[g.b + k]] - di) + sh.tau[g.b + k]) / (dl[k] - di);
      }
      sh.zh[a] = copysign(sqrt(fabs(prod) / sh.mrho[g.m]), sh.zk[a]);
    });
    ex.sync();
    // eigenvectors of diag(dk) + rho zh zh^T: u_k(i) = zh_i / (dk_i - lambda_k), normalised
    ex.par(p * p, [&](int e) {
      const int r = e % p, c = e / p;
      const MergeAt g = merge_of(c, s, p);
      if (g.pass()) return;
      const int i = r - g.b, k = c - g.b, K = sh.mK[g.m];
      if (i < 0 || i >= K || k >= K) return;
      const double *dl = sh.dk + g.b;
      U[e] = sh.zh[r] / ((dl[i] - dl[sh.org[c]]) - sh.tau[c]);
    });
    ex.sync();
    ex.par(p, [&](int c) {
      const MergeAt g = merge_of(c, s, p);
      if (g.pass()) return;
      const int k = c - g.b, K = sh.mK[g.m];
      if (k >= K) return;
      double nn = 0;
      for (int i = 0; i < K; ++i) nn += U[(g.b + i) + c * P] * U[(g.b + i) + c * P];
      const double f = 1.0 / sqrt(nn);
      for (int i = 0; i < K; ++i) U[(g.b + i) + c * P] *= f;
    });
    // the merged block's eigenvalues in ascending order: the K roots and the deflated poles
    ex.par(p, [&](int a) {
      const MergeAt g = merge_of(a, s, p);
      if (g.pass()) return;
      const int x = a - g.b, n = g.b2 - g.b, K = sh.mK[g.m];
      const double *dl = sh.dk + g.b;
      auto val = [&](int y) {
        return y < K ? dl[sh.org[g.b + y]] + sh.tau[g.b + y] : sh.Ds[g.b + sh.dpos[g.b + y - K]];
      };
      const double vx = val(x), kx = order_key(vx);
      int rank = 0;
      for (int y = 0; y < n; ++y) {
        const double ky = order_key(val(y));
        rank += (ky < kx) || (ky == kx && y < x);
      }
      sh.fsrc[g.b + rank] = x;
      sh.Dt[g.b + rank] = vx;
    });
    ex.sync();
    ex.par(p, [&](int a) {
      const MergeAt g = merge_of(a, s, p);
      if (g.pass()) return;
      sh.lam[a] = sh.Dt[a];
    });
    // new eigenvectors of the merged block: Q2 (sorted, rotated) times U on the kept columns
    ex.par(p * p, [&](int e) {
      const int r = e % p, c = e / p;
      const MergeAt g = merge_of(c, s, p);
      if (g.pass() || r < g.b || r >= g.b2) return;
      const int x = sh.fsrc[c], K = sh.mK[g.m];
      if (x < 0 || x >= g.b2 - g.b) {  // not reachable: the rank above is a permutation
        sh.status = 0;
        return;
      }
      double v;
      if (x < K) {
        v = 0.0;
        for (int q = 0; q < K; ++q) v += Q2[r + (g.b + sh.kpos[g.b + q]) * P] * U[(g.b + q) + (g.b + x) * P];
      } else {
        v = Q2[r + (g.b + sh.dpos[g.b + x - K]) * P];
      }
      if (!finite(v)) sh.status = 0;
      Q[e] = v;
    });
    ex.sync();
  }
  ex.par(p, [&](int i) {
    const double v = ldexp(sh.lam[i], ex2);
    out[i] = v;
    if (!finite(v)) sh.status = 0;
  });
  ex.sync();
  ex.par(1, [&](int) { out[p] = sh.status ? 1.0 : 0.0; });
}

}  // namespace dc
}  // namespace tnh
