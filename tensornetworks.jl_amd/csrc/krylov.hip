// Restarted Lanczos for the lowest eigenpair of H_eff (the KrylovKit eigsolve call at
// src/algorithms/mps/dmrg.jl:50-53): Krylov-Schur restart keeping one Ritz vector,
// full re-orthogonalisation by classical Gram-Schmidt applied twice.
//
// The Krylov basis lives in HBM in FP64 (planar complex when theta is complex); the
// matvecs run in the requested precision. The projected matrix (at most 32 x 32) is
// assembled and diagonalised on the device (ritz_kernel); the host reads back only the Ritz
// value, the residual and the breakdown and finiteness flags of each cycle.
#include <algorithm>
#include <cmath>
#include <vector>

#include "hess_eig.hpp"
#include "tnh_internal.hpp"

namespace tnh {

void heff_apply(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2,
                const tnh_buf *R, const tnh_buf *x, tnh_buf *out, double cr, double ci, int prec);

namespace {

constexpr int kMaxVec = 32;
constexpr int kDotBlocks = 256;

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

// out = sum_i y_i V[i] for i < m, m and the coefficients read from the state ritz_kernel
// writes (y re at st[kSY..], im at st[kSY + kMaxVec..])
__global__ __launch_bounds__(256) void lincomb_dev_kernel(double *__restrict__ out, const double *__restrict__ V,
                                                          int64_t vstride, const double *__restrict__ st,
                                                          int64_t n, bool cplx);

// Per Krylov step j: the two Gram-Schmidt coefficient passes (2 x 2 kMaxVec values) and
// ||w||^2 (2 values), kept on the device until the cycle ends.
constexpr int kStep = 4 * kMaxVec + 2;
// ritz_kernel's state and results: the Ritz value, |b|, b (re, im), m, the breakdown flag,
// the finiteness flag (1 finite, 0 non-finite coefficient, -1 non-finite vector norm), then
// the Ritz coefficients y (re at kSY, im at kSY + kMaxVec). Cycles after the first read the
// previous Ritz value and b from it as their projected matrix's first column.
enum { kSTheta = 0, kSAbsB = 1, kSBr = 2, kSBi = 3, kSM = 4, kSBrk = 5, kSFin = 6, kSY = 8 };
constexpr int kStateLen = kSY + 2 * kMaxVec;
constexpr int kEmb = 2 * kMaxVec;  // largest side of the symmetric problem

// The projected matrix of one restart cycle and its lowest eigenpair, on one wave, with the
// host's former bookkeeping: H (m x m, Hermitian) assembled from the coefficient passes of
// steps start..kdim-1 (and, after a restart, the kept Ritz value and b in column 0); a step
// whose new norm beta is at most 1e-14 times the largest |h| so far is a breakdown and ends
// the cycle there (m = j + 1); non-finite coefficients stop it with a flag. The lowest
// eigenpair comes from a cyclic Jacobi on H (real) or on its real symmetric embedding
// [[Re, -Im], [Im, Re]] (complex: every eigenvalue twice, any vector of the lowest pair is
// the Ritz vector up to a phase), run until the off-diagonal mass is below 1e-32 of the
// total; the Ritz coefficients are normalised and b = sum_j h(m, j) y_j.
__global__ __launch_bounds__(64) void ritz_kernel(const double *__restrict__ hdev, int start, int kdim, int cplx,
                                                  double *__restrict__ st) {
  constexpr int LH = kMaxVec + 1;  // leading dimension of H ((kdim + 1) x kdim)
  __shared__ double Hr[LH * kMaxVec], Hi[LH * kMaxVec];
  __shared__ double A[kEmb * kEmb], Z[kEmb * kEmb];
  __shared__ int sm, sbrk, sfin;
  const int t = threadIdx.x;
  for (int e = t; e < LH * kMaxVec; e += 64) {
    Hr[e] = 0.0;
    Hi[e] = 0.0;
  }
  __syncthreads();
  if (t == 0) {
    if (start == 1) {
      Hr[0] = st[kSTheta];
      Hr[1] = st[kSBr];
      Hi[1] = st[kSBi];
    }
    int m = kdim, brk = 0, fin = 1;
    for (int j = start; j < kdim; ++j) {
      const double *hj = hdev + int64_t(j) * kStep;
      bool ok = isfinite(hj[4 * kMaxVec]);
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < 2 * (j + 1); ++i) ok = ok && isfinite(hj[pass * 2 * kMaxVec + i]);
      if (!ok) {
        fin = 0;
        break;
      }
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i <= j; ++i) {
          const double *cj = hj + pass * 2 * kMaxVec;
          Hr[i + j * LH] += cj[2 * i];
          if (cplx) Hi[i + j * LH] += cj[2 * i + 1];
        }
      const double beta = sqrt(fmax(0.0, hj[4 * kMaxVec]));
      if (!isfinite(beta)) {
        fin = -1;
        break;
      }
      Hr[(j + 1) + j * LH] = beta;
      double scale_h = 1.0;
      for (int jj = 0; jj <= j; ++jj)
        for (int ii = 0; ii <= j + 1; ++ii)
          scale_h = fmax(scale_h, sqrt(Hr[ii + jj * LH] * Hr[ii + jj * LH] + Hi[ii + jj * LH] * Hi[ii + jj * LH]));
      if (beta <= 1e-14 * scale_h) {
        m = j + 1;
        brk = 1;
        break;
      }
    }
    sm = m;
    sbrk = brk;
    sfin = fin;
  }
  __syncthreads();
  const int m = sm;
  if (sfin != 1) {
    if (t == 0) {
      st[kSFin] = sfin;
      st[kSM] = m;
    }
    return;
  }
  // the symmetric problem: n = m (real) or 2 m (embedding), Hermitian part of H_m
  const int n = cplx ? 2 * m : m;
  for (int e = t; e < n * n; e += 64) {
    const int i = e % n, j = e / n;
    double v;
    if (!cplx) {
      v = 0.5 * (Hr[i + j * LH] + Hr[j + i * LH]);
    } else {
      const int bi = i / m, bj = j / m, ii = i % m, jj = j % m;
      const double hre = 0.5 * (Hr[ii + jj * LH] + Hr[jj + ii * LH]);
      const double him = 0.5 * (Hi[ii + jj * LH] - Hi[jj + ii * LH]);
      v = bi == bj ? hre : (bi == 1 ? him : -him);  // [[Re, -Im], [Im, Re]]
    }
    A[i + j * n] = v;
    Z[i + j * n] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int e = t; e < n * n; e += 64) {
      const double a = A[e];
      tot += a * a;
      if (e % n != e / n) off += a * a;
    }
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o, 64);
      tot += __shfl_xor(tot, o, 64);
    }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p + q * n];
        if (fabs(apq) < 1e-300) continue;
        const double tau = (A[q + q * n] - A[p + p * n]) / (2 * apq);
        const double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
        const double c = 1 / sqrt(1 + tt * tt), s = tt * c;
        __syncthreads();
        for (int k = t; k < n; k += 64) {  // A <- A J, Z <- Z J
          const double akp = A[k + p * n], akq = A[k + q * n];
          A[k + p * n] = c * akp - s * akq;
          A[k + q * n] = s * akp + c * akq;
          const double zkp = Z[k + p * n], zkq = Z[k + q * n];
          Z[k + p * n] = c * zkp - s * zkq;
          Z[k + q * n] = s * zkp + c * zkq;
        }
        __syncthreads();
        for (int k = t; k < n; k += 64) {  // A <- J^T A
          const double apk = A[p + k * n], aqk = A[q + k * n];
          A[p + k * n] = c * apk - s * aqk;
          A[q + k * n] = s * apk + c * aqk;
        }
        __syncthreads();
      }
  }
  if (t == 0) {
    int lo = 0;
    for (int i = 1; i < n; ++i)
      if (A[i + i * n] < A[lo + lo * n]) lo = i;
    double nrm = 0;
    for (int i = 0; i < m; ++i) {
      const double yr = Z[i + lo * n], yi = cplx ? Z[(i + m) + lo * n] : 0.0;
      nrm += yr * yr + yi * yi;
    }
    nrm = sqrt(nrm);
    double br = 0, bi = 0;
    for (int i = 0; i < m; ++i) {
      const double yr = Z[i + lo * n] / nrm, yi = cplx ? Z[(i + m) + lo * n] / nrm : 0.0;
      st[kSY + i] = yr;
      st[kSY + kMaxVec + i] = yi;
      const double hr = Hr[m + i * LH], hi = Hi[m + i * LH];  // h(m, i)
      br += hr * yr - hi * yi;
      bi += hr * yi + hi * yr;
    }
    st[kSTheta] = A[lo + lo * n];
    st[kSAbsB] = sqrt(br * br + bi * bi);
    st[kSBr] = br;
    st[kSBi] = bi;
    st[kSM] = m;
    st[kSBrk] = sbrk;
    st[kSFin] = 1;
  }
}

__global__ __launch_bounds__(256) void lincomb_dev_kernel(double *__restrict__ out, const double *__restrict__ V,
                                                          int64_t vstride, const double *__restrict__ st,
                                                          int64_t n, bool cplx) {
  const int nv = int(st[kSM]);
  const double *yr = st + kSY, *yi = st + kSY + kMaxVec;
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    double ar = 0, ai = 0;
    for (int i = 0; i < nv; ++i) {
      const double vr = V[i * vstride + k];
      if (cplx) {
        const double vi = V[i * vstride + n + k];
        ar += yr[i] * vr - yi[i] * vi;
        ai += yr[i] * vi + yi[i] * vr;
      } else {
        ar += yr[i] * vr;
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
  double *hdev = static_cast<double *>(ctx->workspace(703, sizeof(double) * kStep * kMaxVec));
  double *sdev = static_cast<double *>(ctx->workspace(704, sizeof(double) * kStateLen));
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

  int start = 0, nmv = 0;
  double theta_e = 0;
  for (int it = 0; it < maxiter; ++it) {
    // The cycle runs on the device without host round trips: both Gram-Schmidt passes take
    // their coefficients from device memory and V[j+1] = w / ||w|| is formed there; then
    // ritz_kernel replays the projected matrix and the breakdown test of each step in order
    // and solves it. Steps after a breakdown (whose vectors are never used) are not counted.
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
    hipLaunchKernelGGL(ritz_kernel, dim3(1), dim3(64), 0, ctx->stream, hdev, start, kdim, cplx ? 1 : 0, sdev);
    TNH_HIP(hipGetLastError());
    double res[kSY];
    to_host(ctx, res, sdev, sizeof(res));
    // A non-finite coefficient (NaN or Inf in theta or the environments) is an error.
    TNH_REQUIRE(res[kSFin] != 0.0, TNH_ERR_NUMERIC, "eigsolve: non-finite Krylov coefficient");
    TNH_REQUIRE(res[kSFin] == 1.0, TNH_ERR_NUMERIC, "eigsolve: non-finite Krylov vector");
    const int m = int(res[kSM]);
    nmv += m - start;
    theta_e = res[kSTheta];
    TNH_REQUIRE(std::isfinite(theta_e) && std::isfinite(res[kSAbsB]), TNH_ERR_NUMERIC,
                "eigsolve: non-finite Ritz value");
    if (res[kSBrk] != 0.0 || res[kSAbsB] <= tol || it == maxiter - 1) break;
    // restart: V0 <- Ritz vector, V1 <- V[m]; projected matrix [[theta, conj(b)], [b, .]]
    hipLaunchKernelGGL(lincomb_dev_kernel, dim3(g), dim3(256), 0, ctx->stream, Wv, V, vstride, sdev, n, cplx);
    dots(Wv, 1, Wv, dres);
    TNH_HIP(hipGetLastError());
    copy_device(ctx, V + vstride, V + int64_t(m) * vstride, sizeof(double) * vstride);
    hipLaunchKernelGGL(normalize_copy_kernel, dim3(g), dim3(256), 0, ctx->stream, V, Wv, dres, vstride);
    TNH_HIP(hipGetLastError());
    start = 1;
  }
  // Ritz vector into theta, normalised
  hipLaunchKernelGGL(lincomb_dev_kernel, dim3(g), dim3(256), 0, ctx->stream, static_cast<double *>(theta->ptr), V,
                     vstride, sdev, n, cplx);
  TNH_HIP(hipGetLastError());
  const double rn = norm_of(static_cast<double *>(theta->ptr));
  TNH_REQUIRE(std::isfinite(rn) && rn > 0, TNH_ERR_NUMERIC, "eigsolve: zero Ritz vector");
  scale(ctx, theta->ptr, TNH_F64, vstride, 1.0 / rn);
  if (nmv_out) *nmv_out = nmv;
  return theta_e;
}

// Arnoldi for the eigenvalue of smallest real part of a non-Hermitian operator, as KrylovKit's
// eigsolve(f, x0, 1, :SR; ishermitian = false) that dmrg.jl:51-53 calls with ishermitian =
// false: the same device cycle as the Lanczos above (matvec, two Gram-Schmidt passes whose
// coefficients are the Hessenberg columns, normalisation), a restart that keeps the one Ritz
// vector u (A u = lambda u + b v_m, so the next projection starts with [lambda; b] in column
// 0), and per cycle one read-back of the coefficients: the m x m Hessenberg problem is solved
// on the host (hess_eig.hpp). A step whose new norm is at most 1e-14 times the largest |h| is
// a breakdown (invariant subspace). theta must be complex. Returns the Ritz value.
std::complex<double> eigsolve_nonherm_op(tnh_ctx *ctx, const Matvec &matvec, tnh_buf *theta,
                                         int kdim, int maxiter, double tol, int *nmv_out) {
  using cd = std::complex<double>;
  TNH_REQUIRE(theta && theta->ndim == 4, TNH_ERR_SHAPE, "theta must be (chi_l, d, d, chi_r)");
  TNH_REQUIRE(theta->dtype == TNH_C128, TNH_ERR_ARG,
              "the non-Hermitian eigensolve needs a complex (C128) theta");
  TNH_REQUIRE(kdim >= 1 && kdim < kMaxVec, TNH_ERR_ARG, "krylovdim must be in [1, 31]");
  TNH_REQUIRE(maxiter >= 1, TNH_ERR_ARG, "maxiter must be >= 1");
  const int64_t n = theta->numel;
  TNH_REQUIRE(n > 0, TNH_ERR_SHAPE, "empty theta");
  if (int64_t(kdim) > n) kdim = int(n);
  const int64_t vstride = 2 * n;
  double *V = static_cast<double *>(ctx->workspace(700, sizeof(double) * vstride * (kdim + 2)));
  double *Wv = V + vstride * (kdim + 1);
  double *part = static_cast<double *>(ctx->workspace(701, sizeof(double) * 2 * kDotBlocks * kMaxVec));
  double *dres = static_cast<double *>(ctx->workspace(702, sizeof(double) * 2 * kMaxVec));
  double *hdev = static_cast<double *>(ctx->workspace(703, sizeof(double) * kStep * kMaxVec));
  double *sdev = static_cast<double *>(ctx->workspace(704, sizeof(double) * kStateLen));
  double *hres = ctx->host_scalars;
  const int g = grid1(ctx, n);
  auto dots = [&](const double *Vp, int nv, const double *w, double *out) {
    if (n <= kOneBlockDot) {
      hipLaunchKernelGGL(mdot1_kernel, dim3(nv), dim3(1024), 0, ctx->stream, Vp, vstride, w, n, true, out);
    } else {
      hipLaunchKernelGGL(mdot_kernel, dim3(kDotBlocks, nv), dim3(256), 0, ctx->stream, Vp, vstride, w, n,
                         true, part);
      hipLaunchKernelGGL(mdot_final_kernel, dim3(nv), dim3(64), 0, ctx->stream, part, kDotBlocks, nv, out);
    }
  };
  auto view = [&](double *ptr) {
    tnh_buf b = *theta;
    b.ptr = ptr;
    b.owned = false;
    return b;
  };
  auto norm_of = [&](const double *w) {
    dots(w, 1, w, dres);
    TNH_HIP(hipGetLastError());
    to_host(ctx, hres, dres, sizeof(double) * 2);
    return std::sqrt(std::max(0.0, hres[0]));
  };
  copy_device(ctx, V, theta->ptr, sizeof(double) * vstride);
  {
    const double nr = norm_of(V);
    TNH_REQUIRE(std::isfinite(nr) && nr > 0, TNH_ERR_NUMERIC,
                "eigsolve: initial vector has zero or non-finite norm");
    scale(ctx, V, TNH_F64, vstride, 1.0 / nr);
  }
  std::vector<double> hh(size_t(kStep) * kMaxVec), st(kStateLen, 0.0);
  std::vector<cd> H(size_t(kMaxVec) * kMaxVec), beta(kMaxVec), y;
  auto Hat = [&](int i, int j) -> cd & { return H[size_t(i) + size_t(j) * kMaxVec]; };
  cd lam(0, 0), keep_lam(0, 0), keep_b(0, 0);
  int start = 0, nmv = 0;
  for (int it = 0; it < maxiter; ++it) {
    for (int j = start; j < kdim; ++j) {
      tnh_buf vj = view(V + j * vstride), wb = view(Wv);
      matvec(&vj, &wb);
      double *hj = hdev + int64_t(j) * kStep;
      for (int pass = 0; pass < 2; ++pass) {
        double *cj = hj + pass * 2 * kMaxVec;
        dots(V, j + 1, Wv, cj);
        hipLaunchKernelGGL(vsub_dev_kernel, dim3(g), dim3(256), 0, ctx->stream, Wv, V, vstride, j + 1, cj,
                           n, true);
      }
      double *nj = hj + 4 * kMaxVec;
      dots(Wv, 1, Wv, nj);
      hipLaunchKernelGGL(normalize_copy_kernel, dim3(g), dim3(256), 0, ctx->stream, V + (j + 1) * vstride,
                         Wv, nj, vstride);
      TNH_HIP(hipGetLastError());
    }
    to_host(ctx, hh.data() + size_t(start) * kStep, hdev + int64_t(start) * kStep,
            sizeof(double) * kStep * (kdim - start));
    // the Hessenberg projection of this cycle
    double scale_h = 0;
    if (start == 1) {
      Hat(0, 0) = keep_lam;
      beta[0] = keep_b;
      scale_h = std::max(std::abs(keep_lam), std::abs(keep_b));
    }
    int m = kdim;
    bool brk = false;
    for (int j = start; j < kdim; ++j) {
      const double *hj = hh.data() + size_t(j) * kStep;
      for (int i = 0; i <= j; ++i) {
        const cd h(hj[2 * i] + hj[2 * kMaxVec + 2 * i], hj[2 * i + 1] + hj[2 * kMaxVec + 2 * i + 1]);
        TNH_REQUIRE(std::isfinite(h.real()) && std::isfinite(h.imag()), TNH_ERR_NUMERIC,
                    "eigsolve: non-finite Krylov coefficient");
        Hat(i, j) = h;
        scale_h = std::max(scale_h, std::abs(h));
      }
      const double q = hj[4 * kMaxVec];
      TNH_REQUIRE(std::isfinite(q), TNH_ERR_NUMERIC, "eigsolve: non-finite Krylov vector");
      beta[j] = cd(std::sqrt(std::max(q, 0.0)), 0.0);
      if (beta[j].real() <= 1e-14 * scale_h) {  // invariant subspace: stop the cycle here
        m = j + 1;
        brk = true;
        break;
      }
    }
    nmv += m - start;
    std::vector<cd> Hm(size_t(m) * m), ev;
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i)
        Hm[i + size_t(j) * m] = i <= j ? Hat(i, j) : (i == j + 1 ? beta[j] : cd(0, 0));
    TNH_REQUIRE(hess::eigenvalues(m, Hm, ev), TNH_ERR_NOCONV,
                "eigsolve: the Hessenberg QR did not converge");
    int w = 0;
    for (int i = 1; i < m; ++i)
      if (ev[i].real() < ev[w].real()) w = i;  // :SR
    lam = ev[w];
    hess::eigenvector(m, Hm, lam, y);
    const cd b = brk ? cd(0, 0) : beta[m - 1] * y[m - 1];
    std::fill(st.begin(), st.end(), 0.0);
    st[kSM] = m;
    for (int i = 0; i < m; ++i) {
      st[kSY + i] = y[i].real();
      st[kSY + kMaxVec + i] = y[i].imag();
    }
    to_device(ctx, sdev, st.data(), sizeof(double) * kStateLen);
    if (brk || std::abs(b) <= tol || it == maxiter - 1) break;
    // restart: V0 <- u = V y (unit norm), V1 <- v_m; column 0 of the next projection [lam; b]
    hipLaunchKernelGGL(lincomb_dev_kernel, dim3(g), dim3(256), 0, ctx->stream, Wv, V, vstride, sdev, n, true);
    dots(Wv, 1, Wv, dres);
    TNH_HIP(hipGetLastError());
    copy_device(ctx, V + vstride, V + int64_t(m) * vstride, sizeof(double) * vstride);
    hipLaunchKernelGGL(normalize_copy_kernel, dim3(g), dim3(256), 0, ctx->stream, V, Wv, dres, vstride);
    TNH_HIP(hipGetLastError());
    keep_lam = lam;
    keep_b = b;
    start = 1;
  }
  hipLaunchKernelGGL(lincomb_dev_kernel, dim3(g), dim3(256), 0, ctx->stream, static_cast<double *>(theta->ptr), V,
                     vstride, sdev, n, true);
  TNH_HIP(hipGetLastError());
  const double rn = norm_of(static_cast<double *>(theta->ptr));
  TNH_REQUIRE(std::isfinite(rn) && rn > 0, TNH_ERR_NUMERIC, "eigsolve: zero Ritz vector");
  scale(ctx, theta->ptr, TNH_F64, vstride, 1.0 / rn);
  if (nmv_out) *nmv_out = nmv;
  return lam;
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
