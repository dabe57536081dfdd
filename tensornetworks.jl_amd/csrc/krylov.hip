// Restarted Lanczos for the lowest eigenpair of H_eff (the KrylovKit eigsolve call at
// src/algorithms/mps/dmrg.jl:50-53): Krylov-Schur restart keeping one Ritz vector,
// full re-orthogonalisation by classical Gram-Schmidt applied twice.
//
// The Krylov basis lives in HBM in FP64 (planar complex when theta is complex); the
// matvecs run in the requested precision. The projected matrix (at most 32 x 32) is
// diagonalised on the host.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "tnh_internal.hpp"

namespace tnh {

void heff_apply(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2,
                const tnh_buf *R, const tnh_buf *x, tnh_buf *out, double cr, double ci, int prec);

namespace {

constexpr int kMaxVec = 32;
constexpr int kDotBlocks = 256;

struct Coef {
  double re[kMaxVec];
  double im[kMaxVec];
};

__device__ inline double wsum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// part[(i * gridDim.x + blk) * 2 + {0,1}] = partial <v_i, w>
__global__ __launch_bounds__(256) void mdot_kernel(const double *__restrict__ V, int64_t vstride,
                                                   const double *__restrict__ w, int64_t n,
                                                   bool cplx, double *__restrict__ part) {
  const int i = blockIdx.y;
  const double *v = V + i * vstride;
  double sr = 0, si = 0;
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    const double vr = v[k], wr = w[k];
    if (cplx) {
      const double vi = v[n + k], wi = w[n + k];
      sr += vr * wr + vi * wi;
      si += vr * wi - vi * wr;
    } else {
      sr += vr * wr;
    }
  }
  __shared__ double red[2][4];
  sr = wsum(sr);
  si = wsum(si);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sr;
    red[1][threadIdx.x >> 6] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[(int64_t(i) * gridDim.x + blockIdx.x) * 2] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    part[(int64_t(i) * gridDim.x + blockIdx.x) * 2 + 1] =
        red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

// out[2 i + {0,1}] = <v_i, w> of a short vector (n <= kOneBlockDot), one 1024-thread workgroup
// per vector and no partials pass: each thread loads four elements of v_i and w before summing
// any, so the workgroup keeps 4096 loads in flight
constexpr int64_t kOneBlockDot = 16384;
__global__ __launch_bounds__(1024) void mdot1_kernel(const double *__restrict__ V, int64_t vstride,
                                                     const double *__restrict__ w, int64_t n,
                                                     bool cplx, double *__restrict__ out) {
  const int i = blockIdx.x;
  const double *v = V + i * vstride;
  double sr = 0, si = 0;
  for (int64_t k0 = threadIdx.x; k0 < n; k0 += 4 * 1024) {
    double vr[4], wr[4], vi[4], wi[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = k0 + u * 1024;
      const bool in = k < n;
      vr[u] = in ? v[k] : 0.0;
      wr[u] = in ? w[k] : 0.0;
      vi[u] = (in && cplx) ? v[n + k] : 0.0;
      wi[u] = (in && cplx) ? w[n + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sr += vr[u] * wr[u] + vi[u] * wi[u];
      si += vr[u] * wi[u] - vi[u] * wr[u];
    }
  }
  __shared__ double red[2][16];
  sr = wsum(sr);
  si = wsum(si);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sr;
    red[1][threadIdx.x >> 6] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int q = 0; q < 16; ++q) {
      a += red[0][q];
      b += red[1][q];
    }
    out[2 * i] = a;
    out[2 * i + 1] = b;
  }
}

// out[2 i + {0,1}] = sum of the nblk partials of vector i (one wave per vector)
__global__ void mdot_final_kernel(const double *__restrict__ part, int nblk, int nv,
                                  double *__restrict__ out) {
  const int i = blockIdx.x;
  if (i >= nv) return;
  double sr = 0, si = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    sr += part[(int64_t(i) * nblk + b) * 2];
    si += part[(int64_t(i) * nblk + b) * 2 + 1];
  }
  sr = wsum(sr);
  si = wsum(si);
  if (threadIdx.x == 0) {
    out[2 * i] = sr;
    out[2 * i + 1] = si;
  }
}

// w -= sum_i c_i v_i with the coefficients read from device memory (c[2 i], c[2 i + 1] = re,
// im, as the dot kernels write them), so the Gram-Schmidt pass needs no host round trip
__global__ __launch_bounds__(256) void vsub_dev_kernel(double *__restrict__ w,
                                                       const double *__restrict__ V, int64_t vstride,
                                                       int nv, const double *__restrict__ c, int64_t n,
                                                       bool cplx) {
  __shared__ double cr[kMaxVec], ci[kMaxVec];
  if (threadIdx.x < nv) {
    cr[threadIdx.x] = c[2 * threadIdx.x];
    ci[threadIdx.x] = cplx ? c[2 * threadIdx.x + 1] : 0.0;
  }
  __syncthreads();
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    double ar = w[k], ai = cplx ? w[n + k] : 0.0;
    for (int i = 0; i < nv; ++i) {
      const double vr = V[i * vstride + k];
      if (cplx) {
        const double vi = V[i * vstride + n + k];
        ar -= cr[i] * vr - ci[i] * vi;
        ai -= cr[i] * vi + ci[i] * vr;
      } else {
        ar -= cr[i] * vr;
      }
    }
    w[k] = ar;
    if (cplx) w[n + k] = ai;
  }
}

// dst = src / sqrt(nsq[0]) over m values (dst = 0 when nsq[0] is not positive and finite:
// a breakdown the host detects afterwards, whose vector is then never used)
__global__ __launch_bounds__(256) void normalize_copy_kernel(double *__restrict__ dst,
                                                             const double *__restrict__ src,
                                                             const double *__restrict__ nsq, int64_t m) {
  const double q = nsq[0];
  const double inv = (q > 0 && isfinite(q)) ? 1.0 / sqrt(q) : 0.0;
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < m;
       k += int64_t(gridDim.x) * blockDim.x)
    dst[k] = src[k] * inv;
}

// out = scale * sum_i y_i v_i
__global__ __launch_bounds__(256) void lincomb_kernel(double *__restrict__ out,
                                                      const double *__restrict__ V,
                                                      int64_t vstride, int nv, Coef y, int64_t n,
                                                      bool cplx) {
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    double ar = 0, ai = 0;
    for (int i = 0; i < nv; ++i) {
      const double vr = V[i * vstride + k];
      if (cplx) {
        const double vi = V[i * vstride + n + k];
        ar += y.re[i] * vr - y.im[i] * vi;
        ai += y.re[i] * vi + y.im[i] * vr;
      } else {
        ar += y.re[i] * vr;
      }
    }
    out[k] = ar;
    if (cplx) out[n + k] = ai;
  }
}

int grid1(tnh_ctx *ctx, int64_t n) {
  int64_t g = (n + 255) / 256;
  return int(std::max<int64_t>(1, std::min<int64_t>(g, int64_t(ctx->cu_count) * 4)));
}

}  // namespace

double eigsolve_ground_op(tnh_ctx *ctx, const Matvec &matvec, tnh_buf *theta, int kdim,
                          int maxiter, double tol, int *nmv_out);

// Symmetric eigenproblem of a small dense matrix (column-major n x n) by cyclic
// Jacobi rotations; eigenvalues ascending in w, eigenvectors in the columns of Z.
void jacobi_eig_sym_host(int n, std::vector<double> A, std::vector<double> &w,
                         std::vector<double> &Z) {
  Z.assign(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) Z[i + size_t(i) * n] = 1.0;
  auto a = [&](int i, int j) -> double & { return A[i + size_t(j) * n]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += a(i, j) * a(i, j);
        if (i != j) off += a(i, j) * a(i, j);
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (std::abs(apq) < 1e-300) continue;
        const double tau = (a(q, q) - a(p, p)) / (2 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1 + tau * tau));
        const double c = 1 / std::sqrt(1 + t * t), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double zkp = Z[k + size_t(p) * n], zkq = Z[k + size_t(q) * n];
          Z[k + size_t(p) * n] = c * zkp - s * zkq;
          Z[k + size_t(q) * n] = s * zkp + c * zkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  std::vector<double> Zs(size_t(n) * n);
  w.resize(n);
  for (int j = 0; j < n; ++j) {
    w[j] = a(idx[j], idx[j]);
    for (int k = 0; k < n; ++k) Zs[k + size_t(j) * n] = Z[k + size_t(idx[j]) * n];
  }
  Z.swap(Zs);
}

// Lowest eigenpair of a small Hermitian matrix H (m x m, column-major).
static void herm_lowest(int m, const std::vector<std::complex<double>> &H, double *e0,
                        std::vector<std::complex<double>> &y) {
  // real symmetric embedding [[Re, -Im], [Im, Re]]: every eigenvalue appears twice
  const int n = 2 * m;
  std::vector<double> B(size_t(n) * n, 0.0), w, Z;
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      const std::complex<double> h = 0.5 * (H[i + size_t(j) * m] + std::conj(H[j + size_t(i) * m]));
      B[i + size_t(j) * n] = h.real();
      B[(i + m) + size_t(j + m) * n] = h.real();
      B[(i + m) + size_t(j) * n] = h.imag();
      B[i + size_t(j + m) * n] = -h.imag();
    }
  jacobi_eig_sym_host(n, B, w, Z);
  *e0 = w[0];
  y.resize(m);
  double nrm = 0;
  for (int i = 0; i < m; ++i) {
    y[i] = std::complex<double>(Z[i], Z[i + m]);
    nrm += std::norm(y[i]);
  }
  nrm = std::sqrt(nrm);
  for (auto &v : y) v /= nrm;
}

double eigsolve_ground(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2,
                       const tnh_buf *R, tnh_buf *theta, int kdim, int maxiter, double tol,
                       double coeff, int prec, int *nmv_out) {
  return eigsolve_ground_op(
      ctx,
      [&](const tnh_buf *x, tnh_buf *y) { heff_apply(ctx, L, W1, W2, R, x, y, coeff, 0.0, prec); },
      theta, kdim, maxiter, tol, nmv_out);
}

double eigsolve_ground_op(tnh_ctx *ctx, const Matvec &matvec, tnh_buf *theta, int kdim,
                          int maxiter, double tol, int *nmv_out) {
  TNH_REQUIRE(theta && theta->ndim == 4, TNH_ERR_SHAPE, "theta must be (chi_l, d, d, chi_r)");
  TNH_REQUIRE(theta->dtype == TNH_F64 || theta->dtype == TNH_C128, TNH_ERR_ARG,
              "theta must be F64 or C128 (Krylov vectors are FP64)");
  TNH_REQUIRE(kdim >= 1 && kdim < kMaxVec, TNH_ERR_ARG, "krylovdim must be in [1, 31]");
  TNH_REQUIRE(maxiter >= 1, TNH_ERR_ARG, "maxiter must be >= 1");
  const bool cplx = is_complex(theta->dtype);
  const int64_t n = theta->numel;
  TNH_REQUIRE(n > 0, TNH_ERR_SHAPE, "empty theta");
  if (int64_t(kdim) > n) kdim = int(n);
  const int64_t vstride = n * (cplx ? 2 : 1);
  double *V = static_cast<double *>(ctx->workspace(700, sizeof(double) * vstride * (kdim + 2)));
  double *Wv = V + vstride * (kdim + 1);
  double *part = static_cast<double *>(ctx->workspace(701, sizeof(double) * 2 * kDotBlocks * kMaxVec));
  double *dres = static_cast<double *>(ctx->workspace(702, sizeof(double) * 2 * kMaxVec));
  // per Krylov step j: the two Gram-Schmidt coefficient passes (2 x 2 kMaxVec values) and
  // ||w||^2 (2 values), kept on the device until the cycle ends
  constexpr int kStep = 4 * kMaxVec + 2;
  double *hdev = static_cast<double *>(ctx->workspace(703, sizeof(double) * kStep * kMaxVec));
  double *hres = ctx->host_scalars;
  const int g = grid1(ctx, n);
  // out[2 i], out[2 i + 1] = <V[i], w> for the nv vectors V[0..nv) (stride vstride)
  auto dots = [&](const double *Vp, int nv, const double *w, double *out) {
    if (n <= kOneBlockDot) {
      hipLaunchKernelGGL(mdot1_kernel, dim3(nv), dim3(1024), 0, ctx->stream, Vp, vstride, w, n, cplx, out);
    } else {
      hipLaunchKernelGGL(mdot_kernel, dim3(kDotBlocks, nv), dim3(256), 0, ctx->stream, Vp, vstride, w, n,
                         cplx, part);
      hipLaunchKernelGGL(mdot_final_kernel, dim3(nv), dim3(64), 0, ctx->stream, part, kDotBlocks, nv, out);
    }
  };

  auto view = [&](double *p) {
    tnh_buf b = *theta;
    b.ptr = p;
    b.owned = false;
    return b;
  };
  auto norm_of = [&](const double *w) {
    // ||w||^2 via the same kernel with V replaced by w itself
    dots(w, 1, w, dres);
    TNH_HIP(hipGetLastError());
    to_host(ctx, hres, dres, sizeof(double) * 2);
    return std::sqrt(std::max(0.0, hres[0]));
  };

  // v0 = theta / |theta|
  copy_device(ctx, V, theta->ptr, sizeof(double) * vstride);
  {
    double nr = norm_of(V);
    TNH_REQUIRE(std::isfinite(nr) && nr > 0, TNH_ERR_NUMERIC, "eigsolve: initial vector has zero or non-finite norm");
    scale(ctx, V, TNH_F64, vstride, 1.0 / nr);
  }

  using cd = std::complex<double>;
  std::vector<cd> H(size_t(kdim + 1) * kdim, cd(0));  // (kdim+1) x kdim, column-major
  auto h = [&](int i, int j) -> cd & { return H[i + size_t(j) * (kdim + 1)]; };
  int start = 0, nmv = 0, m = kdim;
  double theta_e = 0;
  std::vector<cd> y;
  for (int it = 0; it < maxiter; ++it) {
    m = kdim;
    bool breakdown = false;
    // The cycle runs on the device without host round trips: both Gram-Schmidt passes take
    // their coefficients from device memory and V[j+1] = w / ||w|| is formed there. The host
    // then replays the projected matrix and the breakdown test of each step in order, and
    // steps after a breakdown (whose vectors are never used) are not counted.
    for (int j = start; j < kdim; ++j) {
      tnh_buf vj = view(V + j * vstride), wb = view(Wv);
      matvec(&vj, &wb);
      double *hj = hdev + int64_t(j) * kStep;
      for (int pass = 0; pass < 2; ++pass) {
        double *cj = hj + pass * 2 * kMaxVec;
        dots(V, j + 1, Wv, cj);
        hipLaunchKernelGGL(vsub_dev_kernel, dim3(g), dim3(256), 0, ctx->stream, Wv, V, vstride, j + 1, cj,
                           n, cplx);
      }
      double *nj = hj + 4 * kMaxVec;
      dots(Wv, 1, Wv, nj);
      hipLaunchKernelGGL(normalize_copy_kernel, dim3(g), dim3(256), 0, ctx->stream, V + (j + 1) * vstride,
                         Wv, nj, vstride);
      TNH_HIP(hipGetLastError());
    }
    std::vector<double> hh(size_t(kStep) * (kdim - start));
    to_host(ctx, hh.data(), hdev + int64_t(start) * kStep, sizeof(double) * hh.size());
    for (int j = start; j < kdim; ++j) {
      const double *hj = hh.data() + size_t(j - start) * kStep;
      // A non-finite coefficient (NaN or Inf in theta or the environments) is an error: it
      // must not reach the host eigensolver, whose ordering of the eigenvalues needs a
      // consistent comparison.
      bool finite = std::isfinite(hj[4 * kMaxVec]);
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < 2 * (j + 1); ++i) finite = finite && std::isfinite(hj[pass * 2 * kMaxVec + i]);
      TNH_REQUIRE(finite, TNH_ERR_NUMERIC, "eigsolve: non-finite Krylov coefficient");
      ++nmv;
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i <= j; ++i) {
          const double *cj = hj + pass * 2 * kMaxVec;
          h(i, j) += cd(cj[2 * i], cplx ? cj[2 * i + 1] : 0.0);
        }
      const double beta = std::sqrt(std::max(0.0, hj[4 * kMaxVec]));
      TNH_REQUIRE(std::isfinite(beta), TNH_ERR_NUMERIC, "eigsolve: non-finite Krylov vector");
      h(j + 1, j) = beta;
      double scale_h = 1.0;
      for (int jj = 0; jj <= j; ++jj)
        for (int ii = 0; ii <= j + 1; ++ii) scale_h = std::max(scale_h, std::abs(h(ii, jj)));
      if (beta <= 1e-14 * scale_h) {
        m = j + 1;
        breakdown = true;
        break;
      }
    }
    std::vector<cd> Hm(size_t(m) * m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) Hm[i + size_t(j) * m] = h(i, j);
    herm_lowest(m, Hm, &theta_e, y);
    cd b(0);
    for (int jj = 0; jj < m; ++jj) b += h(m, jj) * y[jj];
    if (breakdown || std::abs(b) <= tol || it == maxiter - 1) break;
    // restart: V0 <- Ritz vector, V1 <- V[m]; projected matrix [[theta, conj(b)], [b, .]]
    Coef yc{};
    for (int i = 0; i < m; ++i) {
      yc.re[i] = y[i].real();
      yc.im[i] = y[i].imag();
    }
    hipLaunchKernelGGL(lincomb_kernel, dim3(g), dim3(256), 0, ctx->stream, Wv, V, vstride, m, yc, n,
                       cplx);
    dots(Wv, 1, Wv, dres);
    TNH_HIP(hipGetLastError());
    copy_device(ctx, V + vstride, V + int64_t(m) * vstride, sizeof(double) * vstride);
    hipLaunchKernelGGL(normalize_copy_kernel, dim3(g), dim3(256), 0, ctx->stream, V, Wv, dres, vstride);
    TNH_HIP(hipGetLastError());
    std::fill(H.begin(), H.end(), cd(0));
    h(0, 0) = theta_e;
    h(1, 0) = b;
    start = 1;
  }
  // Ritz vector into theta, normalised
  Coef yc{};
  for (int i = 0; i < m; ++i) {
    yc.re[i] = y[i].real();
    yc.im[i] = y[i].imag();
  }
  hipLaunchKernelGGL(lincomb_kernel, dim3(g), dim3(256), 0, ctx->stream,
                     static_cast<double *>(theta->ptr), V, vstride, m, yc, n, cplx);
  TNH_HIP(hipGetLastError());
  const double rn = norm_of(static_cast<double *>(theta->ptr));
  TNH_REQUIRE(std::isfinite(rn) && rn > 0, TNH_ERR_NUMERIC, "eigsolve: zero Ritz vector");
  scale(ctx, theta->ptr, TNH_F64, vstride, 1.0 / rn);
  if (nmv_out) *nmv_out = nmv;
  return theta_e;
}

}  // namespace tnh

using namespace tnh;

extern "C" int tnh_eigsolve_ground(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1,
                                   const tnh_buf *W2, const tnh_buf *R, tnh_buf *theta,
                                   int krylovdim, int maxiter, double tol, double coeff,
                                   int precision, double *energy, int *nmatvec) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    double e = eigsolve_ground(ctx, L, W1, W2, R, theta, krylovdim, maxiter, tol, coeff,
                               precision, nmatvec);
    if (energy) *energy = e;
  });
}
