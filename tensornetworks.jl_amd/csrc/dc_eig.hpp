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
// (f(t) for t in [0, n), spread over the workgroup), `sync()` (a workgroup barrier),
// `sum(n, f, scratch)` / `vmax(n, f, scratch)` (a workgroup-wide sum or maximum returned to
// every thread, one barrier: consecutive calls must alternate their scratch buffers),
// `gsize(p)`, `gsum(x, G)` and `gsum4(a, b, c, d, G)` (groups of G consecutive threads that
// share a secular root, and sums over the group), `nthreads()` (the workgroup size, at most kDcThreads; small matrices
// run on fewer threads, where a barrier costs less), and `mark(id)`, a phase boundary that a
// timing policy can record. The
// device policy (DcDeviceEx, at the end of this header in HIP builds) maps these to threads,
// barriers and lane shuffles; a host
// policy that runs the loops sequentially (G = 1) compiles the same code for tests on the
// CPU. The merge products are the one place with a device-only path (matrix-core tiles);
// the host runs the same sums as plain loops.
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define TNH_HD __host__ __device__
#define TNH_INLINE __attribute__((always_inline))
#else
#define TNH_HD
#define TNH_INLINE
#endif

namespace tnh {
namespace dc {

constexpr int kDcPMax = 512;      // largest matrix side
constexpr int kDcThreads = 1024;  // largest workgroup; the partial-sum layout is sized for it
constexpr int kDcLdsMat = 13800;  // doubles of LDS for the active matrix of the reduction
constexpr double kEps = 2.220446049250313e-16;

// Workgroup-shared state (LDS on the device). The reduction and the divide and conquer use
// disjoint arrays, which share storage. Per merge m (start b, split bm, end b2) the
// per-position arrays of the divide and conquer hold the merge's entries at [b, b2).
struct DcShared {
  double d[kDcPMax], e[kDcPMax];  // the tridiagonal matrix
  double red[2][64];
  int status;
  union {
    struct {
      double vr[kDcPMax], vi[kDcPMax], yr[kDcPMax], yi[kDcPMax];  // Householder step vectors
      double part[2 * kDcThreads];  // matrix-vector partial sums (real, imaginary)
      double tr[kDcPMax], ti[kDcPMax];  // reflector factors of the LDS-resident steps
      double A[kDcLdsMat];              // the active matrix once it fits
    } h;
    struct {
      double lam[kDcPMax];
      double zt[kDcPMax], Dt[kDcPMax];  // coupling vector and diagonal, in column order
      double Ds[kDcPMax], zs[kDcPMax];  // the same, sorted ascending within each merge
      double dk[kDcPMax], zk[kDcPMax];  // the kept (not deflated) entries, compacted
      double zh[kDcPMax], tau[kDcPMax];  // recomputed coupling vector; root offsets
      double rc[kDcPMax], rs[kDcPMax];  // Givens rotations of the deflation, per merge
      double cn[kDcPMax];               // reciprocal norms of the secular eigenvectors
      double mrho[kDcPMax / 2];
      int src[kDcPMax];   // sorted position -> column it came from
      int kpos[kDcPMax];  // kept compact index -> sorted position (relative to the merge start)
      int dpos[kDcPMax];  // deflated compact index -> sorted position
      int org[kDcPMax];   // root -> compact index of the pole its offset is measured from
      int finv[kDcPMax];  // item (root q < K, or K + deflated index) -> final position
      int rj[kDcPMax], ri[kDcPMax];
      int kflag[kDcPMax];  // sorted position kept (1) or deflated (0)
      int mK[kDcPMax / 2], mrot[kDcPMax / 2];
    } c;
  } u;
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
// the root of that model; a step leaving the current bracket is replaced by bisection. The G
// threads of a group (lane gl) each sum every G-th term and combine with ex.gsum; the sums
// come out bitwise equal in every lane, so the group iterates in lockstep.
template <class Ex>
TNH_HD double secular_root(const Ex &ex, const double *dl, const double *zk, double rho, int j, int K, int gl,
                           int G, int *org, int *iters) {
  double lo, hi;
  int o;
  if (j < K - 1) {
    const double gap = dl[j + 1] - dl[j], mid = 0.5 * gap;
    double f = 0.0;
    for (int k = gl; k < K; k += G) f += rho * zk[k] * zk[k] / ((dl[k] - dl[j]) - mid);
    f = 1.0 + ex.gsum(f, G);
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
    for (int k = gl; k < K; k += G) {
      const double r = 1.0 / ((dl[k] - dor) - t);
      const double q = rho * zk[k] * zk[k] * r;
      if (k <= j) {
        psi += q;
        dpsi += q * r;
      } else {
        phi += q;
        dphi += q * r;
      }
    }
    ex.gsum4(psi, dpsi, phi, dphi, G);
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

#if defined(__HIP_DEVICE_COMPILE__)
typedef double dc_v4d __attribute__((ext_vector_type(4)));
#endif

// The merged blocks' eigenvectors for the K kept items of every merge at level s:
// Q[(b + r) + finv[b + x] p] = sum_{q < K} Q2[(b + r) + (b + kpos[b + q]) p] U[(b + q) + (b + x) p]
// for r < n (the merge size), x < K. On the device each wave computes 32 x 32 tiles with
// v_mfma_f64_16x16x4f64 (lane (g = lane / 16, rr = lane % 16) supplies A[rr][k_g] and
// B[k_g][rr]; accumulator register i holds C[g + 4 i][rr]), operands straight from memory;
// tiles past the merge size or past K are skipped whole.
TNH_HD inline void merge_product(DcShared &sh, int s, int p, const double *Q2, const double *U, double *Q) {
  const int64_t P = p;
  const int nm = (p + 2 * s - 1) / (2 * s);
  auto &c = sh.u.c;
#if defined(__HIP_DEVICE_COMPILE__)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, rr = lane & 15;
  const int tpd = (2 * s + 31) / 32, tpm = tpd * tpd;
  for (int tile = wave; tile < nm * tpm; tile += int(blockDim.x >> 6)) {
    const int m = tile / tpm, loc = tile % tpm;
    const int b = 2 * s * m, bm = b + s < p ? b + s : p, b2 = b + 2 * s < p ? b + 2 * s : p;
    if (bm >= b2) continue;
    const int n = b2 - b, K = c.mK[m];
    const int row0 = 32 * (loc % tpd), col0 = 32 * (loc / tpd);
    if (row0 >= n || col0 >= K) continue;
    dc_v4d acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) acc[x][y] = dc_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
    for (int q0 = 0; q0 < K; q0 += 4) {
      const int q = q0 + g;
      const bool qok = q < K;
      const int64_t colA = qok ? (b + c.kpos[b + q]) * P : 0;
      double fa[2], fb[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const int r = row0 + 16 * x + rr;
        fa[x] = (qok && r < n) ? Q2[(b + r) + colA] : 0.0;
      }
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        const int cc = col0 + 16 * y + rr;
        fb[y] = (qok && cc < K) ? U[(b + q) + (b + cc) * P] : 0.0;
      }
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[x], fb[y], acc[x][y], 0, 0, 0);
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = row0 + 16 * x + g + 4 * i, cc = col0 + 16 * y + rr;
          if (r < n && cc < K) {
            const int fp = c.finv[b + cc];
            const double v = acc[x][y][i];
            if (!finite(v) || fp < b || fp >= b2) {
              sh.status = 0;
            } else {
              Q[(b + r) + fp * P] = v;
            }
          }
        }
  }
#else
  for (int m = 0; m < nm; ++m) {
    const int b = 2 * s * m, bm = b + s < p ? b + s : p, b2 = b + 2 * s < p ? b + 2 * s : p;
    if (bm >= b2) continue;
    const int n = b2 - b, K = c.mK[m];
    for (int x = 0; x < K; ++x) {
      const int fp = c.finv[b + x];
      if (fp < b || fp >= b2) {  // not reachable: the final rank is a permutation
        sh.status = 0;
        continue;
      }
      for (int r = 0; r < n; ++r) {
        double v = 0.0;
        for (int q = 0; q < K; ++q) v += Q2[(b + r) + (b + c.kpos[b + q]) * P] * U[(b + q) + (b + x) * P];
        if (!finite(v)) sh.status = 0;
        Q[(b + r) + fp * P] = v;
      }
    }
  }
#endif
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
  int rcount = 0;  // alternates the reduction scratch buffers
  auto red = [&]() { return sh.red[(rcount++) & 1]; };
  ex.mark(0);
  // exact Hermitian symmetrisation and a power-of-two scaling to max |G_ii| in [1, 2)
  ex.par(p * p, [&](int e) {
    const int i = e % p, j = e / p;
    if (i < j) {
      const double a = 0.5 * (Gr[i + j * P] + Gr[j + i * P]);
      Gr[i + j * P] = a;
      Gr[j + i * P] = a;
      if (CPLX) {
        const double b = 0.5 * (Gi[i + j * P] - Gi[j + i * P]);
        Gi[i + j * P] = b;
        Gi[j + i * P] = -b;
      }
    } else if (i == j && CPLX) {
      Gi[i + i * P] = 0.0;
    }
  });
  ex.sync();
  const double dm = ex.vmax(p, [&](int i) { return fabs(Gr[i + i * P]); }, red());
  const int ex2 = (dm > 0 && dm < 1e300) ? ilogb(dm) : 0;
  if (ex2 != 0) {
    const double sc = ldexp(1.0, -ex2);
    ex.par(p * p, [&](int e) {
      Gr[e] *= sc;
      if (CPLX) Gi[e] *= sc;
    });
    ex.sync();
  }
  ex.mark(1);
  // ---- Householder reduction: Q_h^H G Q_h = T, real off-diagonal (beta_k). Three barriers
  // per step: the matrix-vector partial sums; y = B v with v^H y; the rank-2 update, which
  // also sums the squared tail of the next step's column. The active matrix A_k = G[k.., k..]
  // (element (i, j) at ar[i + j * lda]) is read from G until it fits in LDS, then copied
  // there; from then on the reflectors and their factors also stay in LDS (v_k in the dead
  // column 0 of A_k) until the end, so no barrier of those steps waits for a global store.
  // The two phases are separate instances of the same step, each with pointers of one
  // address space.
  double xn2 = ex.sum(p > 2 ? p - 2 : 0, [&](int i) {
    const double a = Gr[2 + i], b = CPLX ? Gi[2 + i] : 0.0;
    return a * a + b * b;
  }, red());
  // one step on A_k at (ar, ai) with leading dimension lda; v_k goes to (vcr, vci) (n entries)
  // and (tau_k) to (*tpr, *tpi)
  auto step = [&](int k, double *ar, double *ai, int lda, double *vcr, double *vci, double *tpr,
                  double *tpi) TNH_INLINE {
    const int n = p - k - 1;
    const double *col_r = ar + 1, *col_i = CPLX ? ai + 1 : nullptr;  // column 0 of A_k, rows 1..
    const double alr = col_r[0], ali = CPLX ? col_i[0] : 0.0;
    const double dkk = ar[0];
    double tr = 0, ti = 0, beta = alr, sr = 0, si = 0;
    if (xn2 != 0.0 || ali != 0.0) {
      beta = -copysign(sqrt(alr * alr + ali * ali + xn2), alr);
      tr = (beta - alr) / beta;
      ti = -ali / beta;
      // v = x / (alpha - beta), v_0 = 1
      const double ur = alr - beta, ui = ali, den = ur * ur + ui * ui;
      sr = ur / den;
      si = -ui / den;
    }
    ex.par(1, [&](int) {
      sh.d[k] = dkk;
      sh.e[k] = beta;
      *tpr = tr;
      *tpi = ti;
    });
    double *Br = ar + 1 + lda, *Bi = CPLX ? ai + 1 + lda : nullptr;
    if (tr == 0.0 && ti == 0.0) {  // nothing to reflect (x is zero below row 0): v = e_0
      ex.par(n, [&](int i) {
        vcr[i] = i == 0 ? 1.0 : 0.0;
        if (CPLX) vci[i] = 0.0;
      });
      xn2 = ex.sum(n > 2 ? n - 2 : 0, [&](int i) {
        const double a = Br[2 + i], b = CPLX ? Bi[2 + i] : 0.0;
        return a * a + b * b;
      }, red());
      return;
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
    // y = B v: nc (<= 16) column chunks per row, partial sums in LDS; v formed on the fly.
    // Each thread loads kMvU columns' entries (and x entries) before using any of them, so
    // that many loads are in flight while A_k streams from L2.
    int nc = ex.nthreads() / n > 0 ? ex.nthreads() / n : 1;
    if (nc > 16) nc = 16;
    constexpr int kMvU = 4;
    ex.par(n * nc, [&](int e) {
      const int i = e % n, cc = e / n;
      double sr2 = 0, si2 = 0;
      for (int j0 = cc; j0 < n; j0 += kMvU * nc) {
        double br[kMvU], bi[kMvU], xr[kMvU], xi[kMvU];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
        for (int u = 0; u < kMvU; ++u) {
          const int j = j0 + u * nc, jj = j < n ? j : j0;
          br[u] = Br[i + jj * lda];
          bi[u] = CPLX ? Bi[i + jj * lda] : 0.0;
          xr[u] = col_r[jj];
          xi[u] = CPLX ? col_i[jj] : 0.0;
        }
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
        for (int u = 0; u < kMvU; ++u) {
          const int j = j0 + u * nc;
          if (j >= n) break;
          double vrj = 1.0, vij = 0.0;  // v_0 = 1; v_j = x_j s
          if (j != 0) {
            vrj = xr[u] * sr - xi[u] * si;
            vij = xr[u] * si + xi[u] * sr;
          }
          if (CPLX) {
            sr2 += br[u] * vrj - bi[u] * vij;
            si2 += br[u] * vij + bi[u] * vrj;
          } else {
            sr2 += br[u] * vrj;
          }
        }
      }
      sh.u.h.part[e] = sr2;
      if (CPLX) sh.u.h.part[kDcThreads + e] = si2;
    });
    ex.sync();
    // y from the partials, v into LDS (and into its slot), and v^H y (real for Hermitian B)
    const double vy = ex.sum(n, [&](int i) {
      double a = 0, b = 0;
      for (int cc = 0; cc < nc; ++cc) {
        a += sh.u.h.part[i + cc * n];
        if (CPLX) b += sh.u.h.part[kDcThreads + i + cc * n];
      }
      sh.u.h.yr[i] = a;
      sh.u.h.yi[i] = b;
      double va, vb;
      vget(i, va, vb);
      sh.u.h.vr[i] = va;
      sh.u.h.vi[i] = vb;
      vcr[i] = va;
      if (CPLX) vci[i] = vb;
      return va * a + vb * b;
    }, red());
    const double h = 0.5 * (tr * tr + ti * ti) * vy;
    // B -= w v^H + v w^H with w = tau y - h v; the updated column 0 of B is the next
    // step's x, whose tail (rows 2..) is summed here. Thread (i, c0) updates row i of the
    // columns c0, c0 + nu, ...: no index division per entry, the row's w_i and v_i held in
    // registers, and kUpU entries loaded before any of them is stored, so that many loads
    // are in flight while A_k streams from L2.
    int nu = ex.nthreads() / n > 0 ? ex.nthreads() / n : 1;
    constexpr int kUpU = 1;
    xn2 = ex.sum(n * nu, [&](int e) {
      const int i = e % n, c0 = e / n;
      const double *vr = sh.u.h.vr, *vi = sh.u.h.vi, *yr = sh.u.h.yr, *yi = sh.u.h.yi;
      const double vri = vr[i], vii = CPLX ? vi[i] : 0.0;
      const double wir = tr * yr[i] - ti * yi[i] - h * vri;
      const double wii = CPLX ? tr * yi[i] + ti * yr[i] - h * vii : 0.0;
      double acc = 0.0;
      for (int j0 = c0; j0 < n; j0 += kUpU * nu) {
        double br[kUpU], bi[kUpU];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
        for (int u = 0; u < kUpU; ++u) {
          const int j = j0 + u * nu, jj = j < n ? j : j0;
          br[u] = Br[i + jj * lda];
          bi[u] = CPLX ? Bi[i + jj * lda] : 0.0;
        }
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
        for (int u = 0; u < kUpU; ++u) {
          const int j = j0 + u * nu;
          if (j >= n) break;
          const double wjr = tr * yr[j] - ti * yi[j] - h * vr[j];
          double nr, ni = 0.0;
          if (CPLX) {
            const double wji = tr * yi[j] + ti * yr[j] - h * vi[j];
            // w_i conj(v_j) + v_i conj(w_j)
            const double re = wir * vr[j] + wii * vi[j] + vri * wjr + vii * wji;
            const double im = wii * vr[j] - wir * vi[j] + vii * wjr - vri * wji;
            nr = br[u] - re;
            ni = bi[u] - im;
            Br[i + j * lda] = nr;
            Bi[i + j * lda] = ni;
          } else {
            nr = br[u] - (wir * vr[j] + vri * wjr);
            Br[i + j * lda] = nr;
          }
          if (j == 0 && i >= 2) acc += nr * nr + ni * ni;
        }
      }
      return acc;
    }, red());
  };
  const int planes = CPLX ? 2 : 1;
  int k = 0;
  for (; k + 1 < p; ++k) {  // G-resident steps; v_k into G's column, tau_k into taur/taui
    const int n = p - k - 1;
    if ((n + 1) * (n + 1) * planes <= kDcLdsMat) break;
    step(k, Gr + k * (P + 1), CPLX ? Gi + k * (P + 1) : nullptr, p, Gr + k * P + (k + 1),
         CPLX ? Gi + k * P + (k + 1) : nullptr, taur + k, taui + k);
  }
  double dlast;
  if (k + 1 < p) {
    const int k0 = k, m1 = p - k0;
    double *Lr = sh.u.h.A, *Li = CPLX ? sh.u.h.A + m1 * m1 : nullptr;
    ex.par(m1 * m1, [&](int e) {
      const int i = e % m1, j = e / m1;
      Lr[e] = Gr[(k0 + i) + (k0 + j) * P];
      if (CPLX) Li[e] = Gi[(k0 + i) + (k0 + j) * P];
    });
    ex.sync();
    for (; k + 1 < p; ++k) {  // LDS-resident steps; v_k into A_k's column 0 (rows 1..)
      const int o = (k - k0) * (m1 + 1);
      step(k, Lr + o, CPLX ? Li + o : nullptr, m1, Lr + o + 1, CPLX ? Li + o + 1 : nullptr, sh.u.h.tr + k,
           sh.u.h.ti + k);
    }
    dlast = Lr[(m1 - 1) * (m1 + 1)];
    ex.sync();
    // the LDS reflectors and factors to G's columns and taur/taui
    ex.par(m1 * m1, [&](int e) {
      const int i = e % m1, j = e / m1;  // v_{k0 + j}, entry i (rows k0 + j + 1 + i)
      if (j + 1 >= m1 || i >= m1 - 1 - j) return;
      const int o = j * (m1 + 1) + 1 + i;
      Gr[(k0 + j + 1 + i) + (k0 + j) * P] = Lr[o];
      if (CPLX) Gi[(k0 + j + 1 + i) + (k0 + j) * P] = Li[o];
    });
    ex.par(m1 - 1, [&](int j) {
      taur[k0 + j] = sh.u.h.tr[k0 + j];
      taui[k0 + j] = sh.u.h.ti[k0 + j];
    });
  } else {
    dlast = Gr[(p - 1) * (P + 1)];
  }
  ex.mark(2);
  ex.par(1, [&](int) {
    sh.d[p - 1] = dlast;
    sh.status = 1;
  });
  ex.sync();
  // a non-finite input leaves a non-finite tridiagonal matrix: report it and stop
  const double bad = ex.sum(p, [&](int i) {
    return (finite(sh.d[i]) && (i + 1 >= p || finite(sh.e[i]))) ? 0.0 : 1.0;
  }, red());
  if (bad != 0.0) {
    ex.par(p + 1, [&](int i) { out[i] = i < p ? NAN : 0.0; });
    return;
  }
  // ---- divide and conquer on T = tridiag(e, d, e): leaves of size 1 with every
  // off-diagonal torn (d_i -= |e_{i-1}| + |e_i|), merged pairwise level by level
  auto &c = sh.u.c;
  ex.par(p, [&](int i) {
    double v = sh.d[i];
    if (i > 0) v -= fabs(sh.e[i - 1]);
    if (i + 1 < p) v -= fabs(sh.e[i]);
    c.lam[i] = v;
  });
  ex.par(p * p, [&](int e) { Q[e] = (e % p == e / p) ? 1.0 : 0.0; });
  ex.sync();
  const int G = ex.gsize(p);
  int level = 0;
  for (int s = 1; s < p; s *= 2, ++level) {
    const int nm = (p + 2 * s - 1) / (2 * s);
    ex.mark(8 + 4 * level);
    // coupling vector z = Q^T u, u = e_{bm-1} + sign(e) e_{bm}: the last row of the left
    // block's eigenvectors and the first row of the right block's
    ex.par(p, [&](int j) {
      const MergeAt g = merge_of(j, s, p);
      if (g.pass()) return;
      const double sg = sh.e[g.bm - 1] >= 0 ? 1.0 : -1.0;
      c.zt[j] = j < g.bm ? Q[(g.bm - 1) + j * P] : sg * Q[g.bm + j * P];
      c.Dt[j] = c.lam[j];
    });
    ex.sync();
    // merge the two ascending halves (stable: the left half first among equals)
    ex.par(p, [&](int j) {
      const MergeAt g = merge_of(j, s, p);
      if (g.pass()) return;
      int rank;
      const double kj = order_key(c.Dt[j]);
      if (j < g.bm) {
        int cnt = 0;
        for (int t = g.bm; t < g.b2; ++t) cnt += order_key(c.Dt[t]) < kj;
        rank = (j - g.b) + cnt;
      } else {
        int cnt = 0;
        for (int t = g.b; t < g.bm; ++t) cnt += order_key(c.Dt[t]) <= kj;
        rank = (j - g.bm) + cnt;
      }
      c.Ds[g.b + rank] = c.Dt[j];
      c.zs[g.b + rank] = c.zt[j];
      c.src[g.b + rank] = j;
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
        zn2 += c.zs[b + t] * c.zs[b + t];
        dmax = fmax(dmax, fabs(c.Ds[b + t]));
      }
      double rho = fabs(sh.e[bm - 1]);
      int K = 0, nd = 0, nr = 0;
      for (int t = 0; t < n; ++t) c.kflag[b + t] = 0;
      if (zn2 > 0 && rho > 0) {
        const double zn = sqrt(zn2);
        for (int t = 0; t < n; ++t) c.zs[b + t] /= zn;
        rho *= zn2;
        const double tol = 8.0 * kEps * fmax(dmax, rho);
        int last = -1;  // sorted position of the last kept entry
        for (int t = 0; t < n; ++t) {
          const double zt = c.zs[b + t];
          if (rho * fabs(zt) <= tol) continue;
          if (last >= 0) {
            const double zj = c.zs[b + last], r = sqrt(zj * zj + zt * zt);
            const double cs = zt / r, sn = zj / r;
            const double Dj = c.Ds[b + last], Di = c.Ds[b + t];
            if (fabs((Di - Dj) * cs * sn) <= tol) {
              c.zs[b + last] = 0.0;
              c.zs[b + t] = r;
              c.Ds[b + last] = Dj * cs * cs + Di * sn * sn;
              c.Ds[b + t] = Dj * sn * sn + Di * cs * cs;
              c.rj[b + nr] = last;
              c.ri[b + nr] = t;
              c.rc[b + nr] = cs;
              c.rs[b + nr] = sn;
              ++nr;
              c.kflag[b + last] = 0;
            }
          }
          c.kflag[b + t] = 1;
          last = t;
        }
      }
      // kept entries compacted in ascending sorted position (so ascending D), and the
      // deflated positions likewise
      for (int t = 0; t < n; ++t) {
        if (c.kflag[b + t]) {
          c.kpos[b + K] = t;
          c.dk[b + K] = c.Ds[b + t];
          c.zk[b + K] = c.zs[b + t];
          ++K;
        } else {
          c.dpos[b + nd++] = t;
        }
      }
      c.mK[m] = K;
      c.mrot[m] = nr;
      c.mrho[m] = rho;
    });
    ex.sync();
    // working copy of the merge block's eigenvectors in sorted order, then the rotations
    ex.par(p * p, [&](int e) {
      const int r = e % p, col = e / p;
      const MergeAt g = merge_of(col, s, p);
      if (g.pass() || r < g.b || r >= g.b2) return;
      const int sc = c.src[col];
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
      const int nr = c.mrot[g.m];
      for (int q = 0; q < nr; ++q) {
        const int64_t cj = r + (g.b + c.rj[g.b + q]) * P, ci = r + (g.b + c.ri[g.b + q]) * P;
        const double cs = c.rc[g.b + q], sn = c.rs[g.b + q];
        const double a = Q2[cj], bq = Q2[ci];
        Q2[cj] = cs * a - sn * bq;
        Q2[ci] = sn * a + cs * bq;
      }
    });
    ex.mark(9 + 4 * level);
    // roots of the secular equations: a group of G consecutive threads per root
    ex.par(p * G, [&](int t) {
      const int a = t / G, gl = t % G;
      const MergeAt g = merge_of(a, s, p);
      if (g.pass()) return;
      const int j = a - g.b, K = c.mK[g.m];
      if (j >= K) return;
      int o = 0, its = 0;
      const double tt = secular_root(ex, c.dk + g.b, c.zk + g.b, c.mrho[g.m], j, K, gl, G, &o, &its);
      if (gl == 0) {
        c.org[a] = o;
        c.tau[a] = tt;
        if (!finite(tt) || its >= 96) sh.status = 0;
      }
    });
    ex.sync();
    // Gu-Eisenstat: the coupling vector for which the computed roots are exact
    ex.par(p, [&](int a) {
      const MergeAt g = merge_of(a, s, p);
      if (g.pass()) return;
      const int i = a - g.b, K = c.mK[g.m];
      if (i >= K) return;
      const double *dl = c.dk + g.b;
      const double di = dl[i];
      double prod = (dl[c.org[a]] - di) + c.tau[a];
      for (int k = 0; k < K; ++k) {
        if (k == i) continue;
        prod *= ((dl[c.org[g.b + k]] - di) + c.tau[g.b + k]) / (dl[k] - di);
      }
      c.zh[a] = copysign(sqrt(fabs(prod) / c.mrho[g.m]), c.zk[a]);
    });
    ex.sync();
    ex.mark(10 + 4 * level);
    // eigenvectors of diag(dk) + rho zh zh^T: u_k(i) = zh_i / (dk_i - lambda_k), normalised
    ex.par(p, [&](int col) {
      const MergeAt g = merge_of(col, s, p);
      if (g.pass()) return;
      const int k = col - g.b, K = c.mK[g.m];
      if (k >= K) return;
      const double *dl = c.dk + g.b;
      const double lo = dl[c.org[col]], tk = c.tau[col];
      double nn = 0;
      for (int i = 0; i < K; ++i) {
        const double v = c.zh[g.b + i] / ((dl[i] - lo) - tk);
        nn += v * v;
      }
      c.cn[col] = 1.0 / sqrt(nn);
    });
    ex.sync();
    ex.par(p * p, [&](int e) {
      const int r = e % p, col = e / p;
      const MergeAt g = merge_of(col, s, p);
      if (g.pass()) return;
      const int i = r - g.b, k = col - g.b, K = c.mK[g.m];
      if (i < 0 || i >= K || k >= K) return;
      const double *dl = c.dk + g.b;
      U[e] = c.zh[r] / ((dl[i] - dl[c.org[col]]) - c.tau[col]) * c.cn[col];
    });
    // the merged block's eigenvalues in ascending order: the K roots and the deflated poles
    ex.par(p, [&](int a) {
      const MergeAt g = merge_of(a, s, p);
      if (g.pass()) return;
      const int x = a - g.b, n = g.b2 - g.b, K = c.mK[g.m];
      const double *dl = c.dk + g.b;
      auto val = [&](int y) {
        return y < K ? dl[c.org[g.b + y]] + c.tau[g.b + y] : c.Ds[g.b + c.dpos[g.b + y - K]];
      };
      const double vx = val(x), kx = order_key(vx);
      int rank = 0;
      for (int y = 0; y < n; ++y) {
        const double ky = order_key(val(y));
        rank += (ky < kx) || (ky == kx && y < x);
      }
      c.finv[g.b + x] = g.b + rank;
      c.Dt[g.b + rank] = vx;
    });
    ex.sync();
    ex.mark(11 + 4 * level);
    // new eigenvectors of the merged block: Q2 (sorted, rotated) times U on the kept items;
    // the deflated items' columns are copied
    ex.par(p, [&](int a) {
      const MergeAt g = merge_of(a, s, p);
      if (g.pass()) return;
      c.lam[a] = c.Dt[a];
    });
    merge_product(sh, s, p, Q2, U, Q);
    ex.par(p * p, [&](int e) {
      const int r = e % p, col = e / p;  // col - b: the item's index within its merge
      const MergeAt g = merge_of(col, s, p);
      if (g.pass() || r < g.b || r >= g.b2) return;
      const int x = col - g.b, K = c.mK[g.m];
      if (x < K) return;
      const int fp = c.finv[col];
      if (fp < g.b || fp >= g.b2) {  // not reachable: the final rank is a permutation
        sh.status = 0;
        return;
      }
      const double v = Q2[r + (g.b + c.dpos[g.b + x - K]) * P];
      if (!finite(v)) sh.status = 0;
      Q[r + fp * P] = v;
    });
    ex.sync();
  }
  ex.par(p, [&](int i) {
    const double v = ldexp(c.lam[i], ex2);
    out[i] = v;
    if (!finite(v)) sh.status = 0;
  });
  ex.sync();
  ex.par(1, [&](int) { out[p] = sh.status ? 1.0 : 0.0; });
  ex.mark(63);
}

#ifdef __HIPCC__
// Execution policy of dc_eig on one workgroup of NT threads. The reductions take one barrier
// (the caller alternates two scratch buffers, so a fast thread's next write cannot overtake a
// slow thread's read); secular-root groups are G consecutive lanes of one wave.
template <int NT>
struct DcDeviceEx {
  template <class F> __device__ void par(int n, F &&f) const {
#pragma unroll 4
    for (int t = threadIdx.x; t < n; t += NT) f(t);
  }
  __device__ void sync() const { __syncthreads(); }
  __device__ void mark(int) const {}
  __device__ int nthreads() const { return NT; }
  template <class F> __device__ double sum(int n, F &&f, double *red) const {
    double s = 0;
#pragma unroll 4
    for (int t = threadIdx.x; t < n; t += NT) s += f(t);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (NT == 64) return s;
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    double r = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) r += red[w];
    return r;
  }
  template <class F> __device__ double vmax(int n, F &&f, double *red) const {
    double s = 0;
    for (int t = threadIdx.x; t < n; t += NT) s = fmax(s, f(t));
    for (int o = 32; o > 0; o >>= 1) s = fmax(s, __shfl_xor(s, o, 64));
    if (NT == 64) return s;
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    double r = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) r = fmax(r, red[w]);
    return r;
  }
  // lanes per secular root: the largest power of two with p G <= NT, at most a wave
  __device__ int gsize(int p) const {
    int g = 1;
    while (g < 64 && 2 * g * p <= NT) g *= 2;
    return g;
  }
  __device__ double gsum(double x, int G) const {
    for (int o = G >> 1; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
  }
  // four sums over the group, their shuffle chains interleaved
  __device__ void gsum4(double &a, double &b, double &c, double &d, int G) const {
    for (int o = G >> 1; o > 0; o >>= 1) {
      const double ta = __shfl_xor(a, o, 64), tb = __shfl_xor(b, o, 64);
      const double tc = __shfl_xor(c, o, 64), td = __shfl_xor(d, o, 64);
      a += ta;
      b += tb;
      c += tc;
      d += td;
    }
  }
};

#endif  // __HIPCC__

}  // namespace dc
}  // namespace tnh
