// Local expectation values <psi| O_term |phi> of an operator list, the computation of
// inner(st, psi, oplist, phi), src/structures/mps/mps.jl:87-134: environments from both
// ends (ProjMPS(psi, phi; rank=1) blocks) and, per term, the left block carried across the
// term's sites with the operators inserted. With phi = psi every contraction is the
// environment-update GEMM chain with a 1 x d x d x 1 "MPO" holding the local operator
// (identity between the operators of a term); with two states the blocks are 2-index
// overlaps B[a_psi, a_phi] and the operator is applied to phi's site tensor first.
#include <algorithm>
#include <complex>
#include <vector>

#include "tnh_internal.hpp"

using namespace tnh;

namespace {

// a (1, d, d, 1) device tensor holding the d x d operator O[s_out, s_in]
tnh_buf *op_tensor(tnh_ctx *ctx, int d, const double *interleaved) {
  bool cplx = false;
  for (int i = 0; i < d * d; ++i) cplx |= interleaved[2 * i + 1] != 0.0;
  int64_t dims[4] = {1, d, d, 1};
  tnh_buf *b = buf_new(ctx, cplx ? TNH_C128 : TNH_F64, 4, dims);
  std::vector<double> planar(size_t(d) * d * (cplx ? 2 : 1));
  for (int i = 0; i < d * d; ++i) {
    planar[i] = interleaved[2 * i];
    if (cplx) planar[size_t(d) * d + i] = interleaved[2 * i + 1];
  }
  to_device(ctx, b->ptr, planar.data(), planar.size() * sizeof(double));
  return b;
}

// a d x d device matrix O[s_out, s_in] (F64, or planar C128 when any imaginary part is set)
tnh_buf *op_matrix(tnh_ctx *ctx, int d, const double *interleaved) {
  tnh_buf *b = op_tensor(ctx, d, interleaved);
  b->ndim = 2;
  b->dims[0] = b->dims[1] = d;
  return b;
}

// out[b, t, b'] = sum_s O[t, s] A[b, s, b']: one small GEMM per right-bond index (strided batch)
void apply_op_site(tnh_ctx *ctx, const tnh_buf *O, const tnh_buf *A, tnh_buf **out) {
  const int64_t cb = A->dims[0], d = A->dims[1], cb2 = A->dims[2];
  TNH_REQUIRE(O->dims[0] == d && O->dims[1] == d, TNH_ERR_SHAPE, "operator does not match d");
  TNH_REQUIRE(cb2 <= 65535, TNH_ERR_UNSUPPORTED, "inner: right bond above 65535");
  const bool oc = is_complex(O->dtype) || is_complex(A->dtype);
  buf_ensure(ctx, out, oc ? TNH_C128 : TNH_F64, 3, A->dims);
  Gemm g;
  g.M = int(cb);
  g.N = int(d);
  g.K = int(d);
  g.transB = true;
  g.lda = cb;
  g.sA = cb * d;
  g.ldb = d;
  g.sB = 0;
  g.ldc = cb;
  g.sC = cb * d;
  g.batch = int(cb2);
  COperand a, o;
  a.re = A->re();
  a.im = is_complex(A->dtype) ? A->im() : nullptr;
  o.re = O->re();
  o.im = is_complex(O->dtype) ? O->im() : nullptr;
  cgemm(ctx, g, a, o, (*out)->re(), oc ? (*out)->im() : nullptr, TNH_F64, TNH_PREC_F64);
}

}  // namespace

// inner(phi, W, psi) / inner(phi, psi) of src/structures/mps/mpo.jl:181-217: the left block
// carried from site 1 to N (three-index with an MPO, the two-index overlap without), then the
// single entry of the last block.
extern "C" int tnh_inner_mpo(tnh_ctx *ctx, int N, const tnh_buf *const *phi,
                             const tnh_buf *const *W, const tnh_buf *const *psi, double *re,
                             double *im) {
  if (!ctx) return TNH_ERR_ARG;
  tnh_buf *L0 = nullptr, *L1 = nullptr, *one = nullptr;
  int rc = tnh_guard(ctx, [&] {
    TNH_REQUIRE(N >= 1 && phi && psi, TNH_ERR_ARG, "inner: bad arguments");
    for (int i = 0; i < N; ++i) {
      TNH_REQUIRE(phi[i] && phi[i]->ndim == 3 && psi[i] && psi[i]->ndim == 3, TNH_ERR_SHAPE,
                  "inner: site tensors must be (chi_l, d, chi_r)");
      TNH_REQUIRE(is_double(phi[i]->dtype) && is_double(psi[i]->dtype), TNH_ERR_ARG,
                  "inner: FP64 site tensors");
      TNH_REQUIRE(phi[i]->dims[1] == psi[i]->dims[1], TNH_ERR_SHAPE,
                  "GMPS must share the same physical dim.");
      if (W) TNH_REQUIRE(W[i] && W[i]->ndim == 4, TNH_ERR_SHAPE, "inner: MPO tensors must be (D_l, d, d, D_r)");
    }
    TNH_REQUIRE(phi[0]->dims[0] == 1 && psi[0]->dims[0] == 1 && phi[N - 1]->dims[2] == 1 &&
                    psi[N - 1]->dims[2] == 1,
                TNH_ERR_SHAPE, "inner: edge bonds must be 1");
    if (W)
      TNH_REQUIRE(W[0]->dims[0] == 1 && W[N - 1]->dims[3] == 1, TNH_ERR_SHAPE,
                  "inner: MPO edge bonds must be 1");
    const tnh_buf *prev = nullptr;  // nullptr: the ones block of the chain's left end
    if (!W) {
      const int64_t od[2] = {1, 1};
      one = buf_new(ctx, TNH_F64, 2, od);
      fill(ctx, one->ptr, TNH_F64, 1, 1.0);
      prev = one;
    }
    for (int i = 0; i < N; ++i) {
      if (W)
        env_left_pair(ctx, prev, phi[i], psi[i], W[i], &L1, TNH_PREC_F64);
      else
        proj_env_left(ctx, prev, phi[i], psi[i], &L1);
      std::swap(L0, L1);
      prev = L0;
    }
    TNH_REQUIRE(L0->numel == 1, TNH_ERR_SHAPE, "inner: open right bond");
    double h[2] = {0, 0};
    to_host(ctx, h, L0->ptr, is_complex(L0->dtype) ? 16 : 8);  // planar: re then im
    if (re) *re = h[0];
    if (im) *im = is_complex(L0->dtype) ? h[1] : 0.0;
  });
  for (auto *b : {L0, L1, one}) buf_delete(ctx, b);
  return rc;
}

extern "C" int tnh_inner_oplist_pair(tnh_ctx *ctx, int N, const tnh_buf *const *psi,
                                     const tnh_buf *const *phi, int nterms, const int *first_site,
                                     const int *range, const double *ops, double *out) {
  if (!ctx) return TNH_ERR_ARG;
  std::vector<tnh_buf *> owned;
  int rc = tnh_guard(ctx, [&] {
    TNH_REQUIRE(N >= 1 && psi && phi && (nterms == 0 || (first_site && range && ops && out)),
                TNH_ERR_ARG, "inner: bad arguments");
    for (int i = 0; i < N; ++i) {
      TNH_REQUIRE(psi[i] && psi[i]->ndim == 3 && phi[i] && phi[i]->ndim == 3, TNH_ERR_SHAPE,
                  "site tensors must be (chi_l, d, chi_r)");
      TNH_REQUIRE(is_double(psi[i]->dtype) && is_double(phi[i]->dtype), TNH_ERR_ARG,
                  "inner: FP64 site tensors");
      TNH_REQUIRE(phi[i]->dims[1] == psi[0]->dims[1] && psi[i]->dims[1] == psi[0]->dims[1],
                  TNH_ERR_SHAPE, "GMPS must share the same physical dim.");
    }
    TNH_REQUIRE(psi[0]->dims[0] == 1 && phi[0]->dims[0] == 1 && psi[N - 1]->dims[2] == 1 &&
                    phi[N - 1]->dims[2] == 1,
                TNH_ERR_SHAPE, "inner: edge bonds must be 1");
    const int d = int(psi[0]->dims[1]);
    const int64_t od[2] = {1, 1};
    tnh_buf *one = buf_new(ctx, TNH_F64, 2, od);
    owned.push_back(one);
    fill(ctx, one->ptr, TNH_F64, 1, 1.0);
    // overlap environments: E[i] = left block through sites 1..i, F[i] = right block from i..N
    std::vector<tnh_buf *> E(N + 2, nullptr), F(N + 2, nullptr);
    E[0] = F[N + 1] = one;
    for (int i = 1; i <= N; ++i) {
      proj_env_left(ctx, E[i - 1], psi[i - 1], phi[i - 1], &E[i]);
      owned.push_back(E[i]);
    }
    for (int i = N; i >= 1; --i) {
      proj_env_right(ctx, F[i + 1], psi[i - 1], phi[i - 1], &F[i]);
      owned.push_back(F[i]);
    }
    tnh_buf *X = nullptr, *Y = nullptr, *B = nullptr, *O = nullptr;  // per-term temporaries
    const double *op = ops;
    try {
      for (int t = 0; t < nterms; ++t) {
        const int s = first_site[t], r = range[t];
        TNH_REQUIRE(s >= 1 && r >= 1 && s + r - 1 <= N, TNH_ERR_ARG,
                    "inner: term " + std::to_string(t) + " out of range");
        const tnh_buf *prev = E[s - 1];
        for (int k = 0; k < r; ++k) {  // mps.jl:103-125
          buf_delete(ctx, O);
          O = op_matrix(ctx, d, op + 2 * size_t(d) * d * k);
          apply_op_site(ctx, O, phi[s - 1 + k], &B);
          proj_env_left(ctx, prev, psi[s - 1 + k], B, &Y);
          std::swap(X, Y);
          prev = X;
        }
        op += 2 * size_t(d) * d * r;
        double re = 0, im = 0;
        cdot_host(ctx, X, F[s + r], &re, &im);  // mps.jl:128-130
        out[2 * t] = re;
        out[2 * t + 1] = im;
      }
    } catch (...) {
      for (auto *b : {X, Y, B, O}) buf_delete(ctx, b);
      throw;
    }
    for (auto *b : {X, Y, B, O}) buf_delete(ctx, b);
  });
  for (auto *b : owned) buf_delete(ctx, b);
  return rc;
}

extern "C" int tnh_inner_oplist(tnh_ctx *ctx, int N, const tnh_buf *const *psi, int nterms,
                                const int *first_site, const int *range, const double *ops,
                                double *out) {
  if (!ctx) return TNH_ERR_ARG;
  std::vector<tnh_buf *> owned;
  int rc = tnh_guard(ctx, [&] {
    TNH_REQUIRE(N >= 1 && psi && (nterms == 0 || (first_site && range && ops && out)),
                TNH_ERR_ARG, "inner: bad arguments");
    for (int i = 0; i < N; ++i)
      TNH_REQUIRE(psi[i] && psi[i]->ndim == 3, TNH_ERR_SHAPE, "site tensors must be (chi_l, d, chi_r)");
    const int d = int(psi[0]->dims[1]);
    std::vector<double> eye(2 * size_t(d) * d, 0.0);
    for (int i = 0; i < d; ++i) eye[2 * (i + size_t(i) * d)] = 1.0;
    tnh_buf *Wid = op_tensor(ctx, d, eye.data());
    owned.push_back(Wid);
    // norm environments: E[i] = left block through sites 1..i, F[i] = right block from i..N
    std::vector<tnh_buf *> E(N + 2, nullptr), F(N + 2, nullptr);
    for (int i = 1; i <= N; ++i) {
      env_left(ctx, E[i - 1], psi[i - 1], Wid, &E[i], TNH_PREC_F64);
      owned.push_back(E[i]);
    }
    for (int i = N; i >= 1; --i) {
      env_right(ctx, F[i + 1], psi[i - 1], Wid, &F[i], TNH_PREC_F64);
      owned.push_back(F[i]);
    }
    const double *op = ops;
    for (int t = 0; t < nterms; ++t) {
      const int s = first_site[t], r = range[t];
      TNH_REQUIRE(s >= 1 && r >= 1 && s + r - 1 <= N, TNH_ERR_ARG,
                  "inner: term " + std::to_string(t) + " out of range");
      tnh_buf *X = nullptr;
      const tnh_buf *prev = E[s - 1];
      for (int k = 0; k < r; ++k) {
        tnh_buf *W = op_tensor(ctx, d, op + 2 * size_t(d) * d * k);
        tnh_buf *Y = nullptr;
        try {
          env_left(ctx, prev, psi[s - 1 + k], W, &Y, TNH_PREC_F64);
        } catch (...) {
          buf_delete(ctx, W);
          buf_delete(ctx, X);
          throw;
        }
        buf_delete(ctx, W);
        buf_delete(ctx, X);
        X = Y;
        prev = X;
      }
      op += 2 * size_t(d) * d * r;
      const tnh_buf *R = F[s + r];
      double re = 0, im = 0;
      if (R) {
        cdot_host(ctx, X, R, &re, &im);
      } else {
        // beyond the right end: the block is the 1x1x1 ones block
        TNH_REQUIRE(X->numel == 1, TNH_ERR_SHAPE, "inner: open right bond");
        double h[2] = {0, 0};
        to_host(ctx, h, X->ptr, is_complex(X->dtype) ? 16 : 8);
        re = h[0];
        im = is_complex(X->dtype) ? h[1] : 0.0;
      }
      buf_delete(ctx, X);
      out[2 * t] = re;
      out[2 * t + 1] = im;
    }
  });
  for (auto *b : owned) buf_delete(ctx, b);
  return rc;
}
