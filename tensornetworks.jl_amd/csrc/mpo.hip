// MPOs on the device: construction from an operator list (the finite-state automaton of
// MPO(st, H) followed by its SVD compression sweeps, src/structures/mps/mpo.jl:323-459) and
// the products MPO * MPS and MPO * MPO (MPOMPSProduct / MPOMPOProduct, mpo.jl:125-169):
// one elementwise kernel forms every product site with its bonds combined, then the
// library's centre moves canonicalise it (without truncation left to right, then with
// the caller's truncation back to site 1), exactly the reference's sequence.
#include <algorithm>
#include <complex>
#include <vector>

#include "tnh_internal.hpp"

namespace tnh {
namespace {

// out[(w + Dw a), so, e, (w' + Dw2 b)] = sum_k M[w, so, k, w'] A[a, k, e, b]
// (planar complex; *_im = element offset of the imaginary plane, 0 for a real operand).
// e is the extra physical index of an MPO right factor (de = 1 for an MPS).
__global__ __launch_bounds__(256) void site_product_kernel(const double *__restrict__ M, int64_t Mim,
                                                           const double *__restrict__ A, int64_t Aim,
                                                           double *__restrict__ out, int64_t Oim, int Dw,
                                                           int dout, int dk, int Dw2, int Da, int de, int Db,
                                                           int64_t total) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  // output index order: w fastest, then a, so, e, w', b
  int64_t r = idx;
  const int w = int(r % Dw);
  r /= Dw;
  const int a = int(r % Da);
  r /= Da;
  const int so = int(r % dout);
  r /= dout;
  const int e = int(r % de);
  r /= de;
  const int w2 = int(r % Dw2);
  const int64_t b = r / Dw2;
  double sr = 0.0, si = 0.0;
  for (int k = 0; k < dk; ++k) {
    const int64_t mi = w + int64_t(Dw) * (so + int64_t(dout) * (k + int64_t(dk) * w2));
    const int64_t ai = a + int64_t(Da) * (k + int64_t(dk) * (e + int64_t(de) * b));
    const double mr = M[mi], mim = Mim ? M[Mim + mi] : 0.0;
    const double ar = A[ai], aim = Aim ? A[Aim + ai] : 0.0;
    sr += mr * ar - mim * aim;
    si += mr * aim + mim * ar;
  }
  out[idx] = sr;
  if (Oim) out[Oim + idx] = si;
}

// One product site: M (Dw, dout, dk, Dw2) times A viewed as (Da, dk, de, Db); the result is
// (Dw Da, dout, Dw2 Db) for de = 1 or (Dw Da, dout, de, Dw2 Db) otherwise.
tnh_buf *site_product(tnh_ctx *ctx, const tnh_buf *M, const tnh_buf *A, bool mpo_right) {
  TNH_REQUIRE(M && M->ndim == 4, TNH_ERR_SHAPE, "MPO sites must be (D_l, d, d, D_r)");
  TNH_REQUIRE(is_double(M->dtype) && is_double(A->dtype), TNH_ERR_ARG, "MPO products are FP64");
  const int Dw = int(M->dims[0]), dout = int(M->dims[1]), dk = int(M->dims[2]), Dw2 = int(M->dims[3]);
  int Da, de, Db;
  if (mpo_right) {
    TNH_REQUIRE(A->ndim == 4, TNH_ERR_SHAPE, "MPO sites must be (D_l, d, d, D_r)");
    Da = int(A->dims[0]);
    de = int(A->dims[2]);
    Db = int(A->dims[3]);
  } else {
    TNH_REQUIRE(A->ndim == 3, TNH_ERR_SHAPE, "MPS sites must be (chi_l, d, chi_r)");
    Da = int(A->dims[0]);
    de = 1;
    Db = int(A->dims[2]);
  }
  TNH_REQUIRE(A->dims[1] == dk, TNH_ERR_SHAPE, "physical dimensions of the MPO and its operand differ");
  const bool cplx = is_complex(M->dtype) || is_complex(A->dtype);
  tnh_buf *out = nullptr;
  if (mpo_right) {
    int64_t dims[4] = {int64_t(Dw) * Da, dout, de, int64_t(Dw2) * Db};
    buf_ensure(ctx, &out, cplx ? TNH_C128 : TNH_F64, 4, dims);
  } else {
    int64_t dims[3] = {int64_t(Dw) * Da, dout, int64_t(Dw2) * Db};
    buf_ensure(ctx, &out, cplx ? TNH_C128 : TNH_F64, 3, dims);
  }
  const int64_t total = out->numel;
  if (total > 0) {
    hipLaunchKernelGGL(site_product_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, ctx->stream,
                       static_cast<const double *>(M->ptr), is_complex(M->dtype) ? M->numel : 0,
                       static_cast<const double *>(A->ptr), is_complex(A->dtype) ? A->numel : 0,
                       static_cast<double *>(out->ptr), cplx ? out->numel : 0, Dw, dout, dk, Dw2, Da, de, Db,
                       total);
    TNH_HIP(hipGetLastError());
  }
  return out;
}

// View rank-2 sites (D_l, d, d, D_r) as rank-1 sites (D_l, d*d, D_r) and back: the centre
// moves group every index but the bond, so both views give the same decompositions.
void as_rank1(tnh_buf *b) {
  const int64_t d2 = b->dims[1] * b->dims[2], r = b->dims[3];
  b->ndim = 3;
  b->dims[1] = d2;
  b->dims[2] = r;
  b->dims[3] = 1;
}
void as_rank2(tnh_buf *b, int64_t d_out) {
  const int64_t dd = b->dims[1], r = b->dims[2];
  b->ndim = 4;
  b->dims[1] = d_out;
  b->dims[2] = dd / d_out;
  b->dims[3] = r;
}

// A x B over A's last index and B's first, result dims given (A: rows x k, B: k x cols).
tnh_buf *bond_product(tnh_ctx *ctx, const tnh_buf *A, int64_t arows, const tnh_buf *B, int nd,
                      const int64_t *dims) {
  const int64_t kk = A->numel / arows;
  const int64_t bcols = B->numel / kk;
  TNH_REQUIRE(kk * bcols == B->numel, TNH_ERR_SHAPE, "MPO bond mismatch");
  const bool c = is_complex(A->dtype) || is_complex(B->dtype);
  tnh_buf *res = nullptr;
  buf_ensure(ctx, &res, c ? TNH_C128 : TNH_F64, nd, dims);
  Gemm g;
  g.M = int(arows);
  g.N = int(bcols);
  g.K = int(kk);
  g.lda = arows;
  g.ldb = kk;
  g.ldc = arows;
  COperand a, b;
  a.re = A->re();
  a.im = is_complex(A->dtype) ? A->im() : nullptr;
  b.re = B->re();
  b.im = is_complex(B->dtype) ? B->im() : nullptr;
  try {
    cgemm(ctx, g, a, b, res->re(), c ? res->im() : nullptr, TNH_F64, TNH_PREC_F64);
  } catch (...) {
    buf_delete(ctx, res);
    throw;
  }
  return res;
}

}  // namespace

// The SVD sweeps that close MPO(st, H), mpo.jl:442-457, on the device: left to right,
// W_s <- U S and W_{s+1} <- V W_{s+1} (split over W_s's right bond); right to left,
// W_s <- S V and W_{s-1} <- W_{s-1} U (split over W_s's left bond). Each split applies the
// truncation rule; the one-sided Jacobi gives the small singular values the accuracy that
// a cutoff of 1e-15 on the discarded weight needs.
void mpo_compress(tnh_ctx *ctx, int N, tnh_buf **W, double cutoff, int maxdim, int mindim) {
  for (int s = 0; s + 1 < N; ++s) {
    tnh_buf *L = W[s];
    const int64_t Dl = L->dims[0], d1 = L->dims[1], d2 = L->dims[2], Dr = L->dims[3];
    tnh_buf *US = nullptr, *Vh = nullptr, *next = nullptr;
    SvdOut o;
    int64_t ld[4] = {Dl, d1, d2, 0}, rd[2] = {0, Dr};
    try {
      split_matrix(ctx, L, int(Dl * d1 * d2), int(Dr), 1, cutoff, maxdim, mindim, true, 4, ld, 2, rd, &US, &Vh,
                   &o);
      const tnh_buf *R = W[s + 1];
      int64_t nd[4] = {int64_t(o.k), R->dims[1], R->dims[2], R->dims[3]};
      next = bond_product(ctx, Vh, o.k, R, 4, nd);
    } catch (...) {
      buf_delete(ctx, US);
      buf_delete(ctx, Vh);
      throw;
    }
    buf_delete(ctx, Vh);
    buf_delete(ctx, W[s]);
    buf_delete(ctx, W[s + 1]);
    W[s] = US;
    W[s + 1] = next;
  }
  for (int s = N - 1; s >= 1; --s) {
    tnh_buf *R = W[s];
    const int64_t Dl = R->dims[0], d1 = R->dims[1], d2 = R->dims[2], Dr = R->dims[3];
    tnh_buf *U = nullptr, *SV = nullptr, *prev = nullptr;
    SvdOut o;
    int64_t ld[2] = {Dl, 0}, rd[4] = {0, d1, d2, Dr};
    try {
      split_matrix(ctx, R, int(Dl), int(d1 * d2 * Dr), 0, cutoff, maxdim, mindim, true, 2, ld, 4, rd, &U, &SV,
                   &o);
      const tnh_buf *L = W[s - 1];
      int64_t nd[4] = {L->dims[0], L->dims[1], L->dims[2], int64_t(o.k)};
      prev = bond_product(ctx, L, L->dims[0] * L->dims[1] * L->dims[2], U, 4, nd);
    } catch (...) {
      buf_delete(ctx, U);
      buf_delete(ctx, SV);
      throw;
    }
    buf_delete(ctx, U);
    buf_delete(ctx, W[s]);
    buf_delete(ctx, W[s - 1]);
    W[s] = SV;
    W[s - 1] = prev;
  }
}

}  // namespace tnh

using namespace tnh;

namespace {

// The product sites of O * X (X an MPS, or an MPO when mpo_right), then the centre moved
// from an unset centre to site N without truncation and back to site 1 with truncation.
void mpo_product(tnh_ctx *ctx, int N, const tnh_buf *const *O, const tnh_buf *const *X, bool mpo_right,
                 double cutoff, int maxdim, int mindim, tnh_buf **out) {
  TNH_REQUIRE(N >= 1 && O && X && out, TNH_ERR_ARG, "MPO product: bad arguments");
  std::vector<tnh_buf *> s(N, nullptr);
  try {
    for (int i = 0; i < N; ++i) {
      TNH_REQUIRE(O[i] && X[i], TNH_ERR_ARG, "MPO product: NULL site tensor");
      s[i] = site_product(ctx, O[i], X[i], mpo_right);
    }
    for (int i = 0; i + 1 < N; ++i)
      TNH_REQUIRE(s[i]->dims[s[i]->ndim - 1] == s[i + 1]->dims[0], TNH_ERR_SHAPE,
                  "MPO product: bond dimensions of neighbouring sites differ");
    const int64_t d_out = O[0]->dims[1];
    if (mpo_right)
      for (auto *b : s) as_rank1(b);
    // MPOMPSProduct / MPOMPOProduct (mpo.jl:134-146, :155-168): moveright! through every new
    // site as it is formed, i.e. an untruncated sweep to site N, then movecenter!(phi, 1; kw)
    mps_movecenter(ctx, N, s.data(), 0, N, 0.0, 0, 1);
    mps_movecenter(ctx, N, s.data(), N, 1, cutoff, maxdim, mindim);
    if (mpo_right)
      for (auto *b : s) as_rank2(b, d_out);
  } catch (...) {
    for (auto *b : s) buf_delete(ctx, b);
    throw;
  }
  for (int i = 0; i < N; ++i) out[i] = s[i];
}

}  // namespace

extern "C" {

int tnh_mpo_apply(tnh_ctx *ctx, int N, const tnh_buf *const *O, const tnh_buf *const *psi, double cutoff,
                  int maxdim, int mindim, tnh_buf **phi) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(maxdim >= 0 && mindim >= 0, TNH_ERR_ARG, "maxdim and mindim must be >= 0");
    mpo_product(ctx, N, O, psi, false, cutoff, maxdim, mindim, phi);
  });
}

int tnh_mpo_mpo_product(tnh_ctx *ctx, int N, const tnh_buf *const *O1, const tnh_buf *const *O2,
                        double cutoff, int maxdim, int mindim, tnh_buf **O) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(maxdim >= 0 && mindim >= 0, TNH_ERR_ARG, "maxdim and mindim must be >= 0");
    for (int i = 0; i < N; ++i)
      TNH_REQUIRE(O1 && O2 && O1[i] && O2[i] && O1[i]->dims[2] == O2[i]->dims[1], TNH_ERR_SHAPE,
                  "MPO product: inner physical dimensions differ");
    mpo_product(ctx, N, O1, O2, true, cutoff, maxdim, mindim, O);
  });
}

int tnh_mpo_from_oplist(tnh_ctx *ctx, int N, int d, int nterms, const int *first_site, const int *range,
                        const double *ops, const double *coeffs, double cutoff, int maxdim, int mindim,
                        tnh_buf **W) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(N >= 1 && d >= 1 && nterms >= 0 && W, TNH_ERR_ARG, "mpo_from_oplist: bad arguments");
    TNH_REQUIRE(nterms == 0 || (first_site && range && ops && coeffs), TNH_ERR_ARG,
                "mpo_from_oplist: NULL term arrays");
    TNH_REQUIRE(maxdim >= 0 && mindim >= 0, TNH_ERR_ARG, "maxdim and mindim must be >= 0");
    using cd = std::complex<double>;
    // channels crossing bond b (between sites b and b + 1, 1-based b): one per multi-site
    // term that crosses it, numbered in term order; channel 0 = nothing placed yet, the
    // last = every operator of a term placed (mpo.jl:334-342)
    std::vector<int> nchan(N + 1, 0);
    std::vector<std::vector<int>> chan(nterms);
    std::vector<int64_t> opoff(nterms + 1, 0);
    for (int t = 0; t < nterms; ++t) {
      TNH_REQUIRE(range[t] >= 1 && first_site[t] >= 1 && first_site[t] + range[t] - 1 <= N, TNH_ERR_ARG,
                  "mpo_from_oplist: term " + std::to_string(t) + " lies outside the chain");
      opoff[t + 1] = opoff[t] + int64_t(range[t]) * d * d;
      if (range[t] > 1)
        for (int b = first_site[t]; b < first_site[t] + range[t] - 1; ++b) chan[t].push_back(++nchan[b]);
    }
    std::vector<int64_t> D(N + 1, 1);
    for (int b = 1; b < N; ++b) D[b] = 2 + nchan[b];
    std::vector<std::vector<cd>> T(N);
    auto at = [&](int i, int64_t w, int s, int t, int64_t w2) -> cd & {
      return T[i][w + D[i] * (s + int64_t(d) * (t + int64_t(d) * w2))];
    };
    for (int i = 0; i < N; ++i) {
      T[i].assign(size_t(D[i] * d * d * D[i + 1]), cd(0));
      for (int s = 0; s < d; ++s) {
        if (i > 0) at(i, D[i] - 1, s, s, D[i + 1] - 1) = 1.0;  // finished terms pass through
        if (i + 1 < N) at(i, 0, s, s, 0) = 1.0;                 // nothing placed yet
      }
    }
    auto opm = [&](int t, int k, int s, int tt) {  // entry (s, tt) of term t's k-th site operator
      const double *p = ops + 2 * (opoff[t] + int64_t(k) * d * d + s + int64_t(d) * tt);
      return cd(p[0], p[1]);
    };
    for (int t = 0; t < nterms; ++t) {
      const cd c(coeffs[2 * t], coeffs[2 * t + 1]);
      const int f = first_site[t] - 1;  // 0-based first site
      if (range[t] == 1) {
        for (int s = 0; s < d; ++s)
          for (int tt = 0; tt < d; ++tt) at(f, 0, s, tt, D[f + 1] - 1) += c * opm(t, 0, s, tt);
        continue;
      }
      for (int k = 0; k < range[t]; ++k) {
        const int i = f + k;
        const int64_t row = k == 0 ? 0 : chan[t][k - 1];
        const int64_t col = k == range[t] - 1 ? D[i + 1] - 1 : chan[t][k];
        const cd scale = k == 0 ? c : cd(1.0);
        for (int s = 0; s < d; ++s)
          for (int tt = 0; tt < d; ++tt) at(i, row, s, tt, col) += scale * opm(t, k, s, tt);
      }
    }
    bool real = true;
    for (auto &v : T)
      for (auto &x : v) real = real && x.imag() == 0.0;
    std::vector<tnh_buf *> out(N, nullptr);
    try {
      for (int i = 0; i < N; ++i) {
        int64_t dims[4] = {D[i], d, d, D[i + 1]};
        out[i] = buf_new(ctx, real ? TNH_F64 : TNH_C128, 4, dims);
        const int64_t n = out[i]->numel;
        std::vector<double> h(size_t(n) * (real ? 1 : 2));
        for (int64_t e = 0; e < n; ++e) {
          h[e] = T[i][e].real();
          if (!real) h[n + e] = T[i][e].imag();
        }
        to_device(ctx, out[i]->ptr, h.data(), h.size() * sizeof(double));
      }
      if (cutoff >= 0.0) mpo_compress(ctx, N, out.data(), cutoff, maxdim, mindim);
    } catch (...) {
      for (auto *b : out) buf_delete(ctx, b);
      throw;
    }
    for (int i = 0; i < N; ++i) W[i] = out[i];
  });
}

int tnh_mpo_compress(tnh_ctx *ctx, int N, tnh_buf **W, double cutoff, int maxdim, int mindim) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(N >= 1 && W && maxdim >= 0 && mindim >= 0, TNH_ERR_ARG, "mpo_compress: bad arguments");
    for (int i = 0; i < N; ++i)
      TNH_REQUIRE(W[i] && W[i]->ndim == 4, TNH_ERR_SHAPE, "MPO sites must be (D_l, d, d, D_r)");
    mpo_compress(ctx, N, W, cutoff, maxdim, mindim);
  });
}

}  // extern "C"
