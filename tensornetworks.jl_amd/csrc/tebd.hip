// Two-site gate application of TEBD: applygate(psi, site, gate, direction; ...),
// src/structures/mps/gatelist.jl:137-171 (rng = 2, rank 1).
//   theta = A1 A2; theta'[a,o1,o2,b] = sum G[o1,i1,o2,i2] theta[a,i1,i2,b]; split, no
//   normalisation (:157).
// The gate stage is an HBM-bound mixing pass over theta (the same kernel as the MPO
// mixing stage of the matvec, with the d^2 x d^2 gate as its matrix).
#include <algorithm>
#include <cmath>
#include <vector>

#include "tnh_internal.hpp"

namespace tnh {

namespace {

// Wc[(i1,i2),(o1,o2)] = G[o1,i1,o2,i2] (planar complex)
template <typename T>
__global__ void gate_wc_kernel(const T *G, int64_t g_im, int d, double *Wc, bool cplx) {
  const int I = d * d;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= I * I) return;
  const int i = e % I, j = e / I;
  const int i1 = i % d, i2 = i / d, o1 = j % d, o2 = j / d;
  const int64_t q = o1 + int64_t(d) * (i1 + int64_t(d) * (o2 + int64_t(d) * i2));
  Wc[e] = double(G[q]);
  if (cplx) Wc[int64_t(I) * I + e] = g_im ? double(G[g_im + q]) : 0.0;
}

// Wc[I, O] = G[o1, i1, ..., or, ir], I = i1 + d i2 + ..., O = o1 + d o2 + ... (planar
// complex); D = d^r
template <typename T>
__global__ void gate_wc_n_kernel(const T *G, int64_t g_im, int d, int r, int D, double *Wc,
                                 bool cplx) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= D * D) return;
  int i = e % D, j = e / D;
  int64_t q = 0, mult = 1;
  for (int k = 0; k < r; ++k) {
    q += (j % d) * mult;
    mult *= d;
    q += (i % d) * mult;
    mult *= d;
    i /= d;
    j /= d;
  }
  Wc[e] = double(G[q]);
  if (cplx) Wc[int64_t(D) * D + e] = g_im ? double(G[g_im + q]) : 0.0;
}

}  // namespace

// theta_out = G applied to theta (both (chi_l, d, d, chi_r)); theta_out (re)allocated.
void apply_gate_theta(tnh_ctx *ctx, const tnh_buf *theta, const tnh_buf *G, tnh_buf **out) {
  TNH_REQUIRE(G && G->ndim == 4, TNH_ERR_SHAPE, "gate must be (d, d, d, d)");
  const int64_t cl = theta->dims[0], d = theta->dims[1], cr = theta->dims[3];
  TNH_REQUIRE(theta->dims[2] == d && G->dims[0] == d && G->dims[1] == d && G->dims[2] == d &&
                  G->dims[3] == d,
              TNH_ERR_SHAPE, "gate dims must equal the physical dimension");
  TNH_REQUIRE(d * d <= 64, TNH_ERR_UNSUPPORTED, "d^2 > 64");
  const bool gc = is_complex(G->dtype);
  const bool tc = is_complex(theta->dtype);
  const bool oc = gc || tc;
  const int I = int(d * d);
  double *Wc = static_cast<double *>(ctx->workspace(840, sizeof(double) * I * I * 2));
  const int64_t gim = gc ? G->numel : 0;
  if (is_double(G->dtype))
    hipLaunchKernelGGL(gate_wc_kernel<double>, dim3((I * I + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const double *>(G->ptr), gim, int(d), Wc, gc);
  else
    hipLaunchKernelGGL(gate_wc_kernel<float>, dim3((I * I + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const float *>(G->ptr), gim, int(d), Wc, gc);
  TNH_HIP(hipGetLastError());
  const int odt = is_double(theta->dtype) ? (oc ? TNH_C128 : TNH_F64) : (oc ? TNH_C64 : TNH_F32);
  buf_ensure(ctx, out, odt, 4, theta->dims);
  Small s;
  s.na = int(cl);
  s.nb = int(cr);
  s.I = I;
  s.J = I;
  s.in = theta->ptr;
  s.in_im = tc ? theta->numel : 0;
  s.tin = real_of(theta->dtype);
  s.out = (*out)->ptr;
  s.out_im = oc ? (*out)->numel : 0;
  s.tout = real_of(odt);
  s.Wc = Wc;
  s.Wc_cplx = gc;
  s.sbI = cl * I;
  s.sbJ = cl * I;
  s.offI.resize(I);
  s.offJ.resize(I);
  for (int i = 0; i < I; ++i) s.offI[i] = s.offJ[i] = cl * i;
  small_contract(ctx, s);
}

void applygate2(tnh_ctx *ctx, tnh_buf **A1, tnh_buf **A2, const tnh_buf *G, int direction,
                double cutoff, int maxdim, int mindim, int *k, double *disc) {
  TNH_REQUIRE(A1 && A2 && *A1 && *A2, TNH_ERR_ARG, "NULL site tensors");
  tnh_buf *theta = nullptr, *theta2 = nullptr;
  try {
    two_site_theta(ctx, *A1, *A2, &theta);  // :142-145
    apply_gate_theta(ctx, theta, G, &theta2);  // :149-154
    replacesites2(ctx, theta2, direction, 0, cutoff, maxdim, mindim, A1, A2, k, disc, nullptr, 0,
                  false);  // :157
  } catch (...) {
    buf_delete(ctx, theta);
    buf_delete(ctx, theta2);
    throw;
  }
  buf_delete(ctx, theta);
  buf_delete(ctx, theta2);
}

// The r - 1 truncated splits of replacesites!'s r-site branch (gmps.jl:215-257) of *cur, a
// (chi_l, d^r, chi_r) tensor that is consumed (on success *cur ends as a site of A and is set
// to NULL; on failure the caller still owns it). Sweeping right (direction 0) site i becomes
// U_k of [(left bond, d), rest] and the rest S V^H; sweeping left site j becomes V_k^H of
// [rest, (d, right bond)] and the rest U S. k[i] (may be NULL) receives the bond between
// sites i and i + 1, *disc the largest discarded weight.
void split_sites(tnh_ctx *ctx, int r, int64_t d, tnh_buf **cur, int direction, double cutoff,
                 int maxdim, int mindim, tnh_buf **A, int *k, double *disc) {
  const int64_t cl = (*cur)->dims[0], cr = (*cur)->dims[2];
  int64_t D = 1;
  for (int i = 0; i < r; ++i) D *= d;
  TNH_REQUIRE((*cur)->ndim == 3 && (*cur)->dims[1] == D, TNH_ERR_SHAPE, "split: theta shape");
  double dmax = 0;
  SvdOut o;
  tnh_buf *nxt = nullptr;
  try {
    if (!direction) {
      int64_t c = cl, rem = D;
      for (int i = 0; i + 1 < r; ++i) {
        rem /= d;
        const int64_t ld[3] = {c, d, 0}, rd[3] = {0, rem, cr};
        const int kk = split_matrix(ctx, *cur, int(c * d), int(rem * cr), 0, cutoff, maxdim,
                                    mindim, false, 3, ld, 3, rd, &A[i], &nxt, &o, nullptr);
        if (k) k[i] = kk;
        dmax = std::max(dmax, o.discarded);
        std::swap(*cur, nxt);
        buf_delete(ctx, nxt);
        nxt = nullptr;
        c = kk;
      }
      buf_delete(ctx, A[r - 1]);
      A[r - 1] = *cur;
    } else {
      int64_t c = cr, rem = D;
      for (int j = r - 1; j >= 1; --j) {
        rem /= d;
        const int64_t ld[3] = {cl, rem, 0}, rd[3] = {0, d, c};
        const int kk = split_matrix(ctx, *cur, int(cl * rem), int(d * c), 1, cutoff, maxdim,
                                    mindim, false, 3, ld, 3, rd, &nxt, &A[j], &o, nullptr);
        if (k) k[j - 1] = kk;
        dmax = std::max(dmax, o.discarded);
        std::swap(*cur, nxt);
        buf_delete(ctx, nxt);
        nxt = nullptr;
        c = kk;
      }
      buf_delete(ctx, A[0]);
      A[0] = *cur;
    }
  } catch (...) {
    buf_delete(ctx, nxt);
    throw;
  }
  *cur = nullptr;
  if (disc) *disc = dmax;
}

// applygate(psi, site, gate, direction; ...) for a gate on r >= 2 sites, gatelist.jl:137-171:
// theta = A[0] ... A[r-1] (:142-145), theta'[a, O, b] = sum_I G[O, I] theta[a, I, b]
// (:149-154), then replacesites! without normalisation (:157), whose r-site branch
// (gmps.jl:215-257) splits off one site at a time with the truncated SVD: sweeping right
// every site but the last becomes left-orthonormal, sweeping left every site but the first
// right-orthonormal. A[] is replaced in place; k[i] receives the bond between sites i and
// i + 1 of the window, *disc the largest discarded weight.
void applygate_sites(tnh_ctx *ctx, int r, tnh_buf **A, const tnh_buf *G, int direction,
                     double cutoff, int maxdim, int mindim, int *k, double *disc) {
  TNH_REQUIRE(r >= 2 && A, TNH_ERR_ARG, "applygate: at least two sites");
  for (int i = 0; i < r; ++i)
    TNH_REQUIRE(A[i] && A[i]->ndim == 3 && is_double(A[i]->dtype), TNH_ERR_SHAPE,
                "applygate: site tensors must be FP64 (chi_l, d, chi_r)");
  const int64_t d = A[0]->dims[1], cl = A[0]->dims[0], cr = A[r - 1]->dims[2];
  int64_t D = 1;
  for (int i = 0; i < r; ++i) {
    TNH_REQUIRE(A[i]->dims[1] == d, TNH_ERR_SHAPE, "applygate: physical dimensions differ");
    if (i) TNH_REQUIRE(A[i - 1]->dims[2] == A[i]->dims[0], TNH_ERR_SHAPE, "applygate: bonds do not chain");
    D *= d;
    TNH_REQUIRE(D <= 64, TNH_ERR_UNSUPPORTED, "applygate: d^r > 64 (gate too wide)");
  }
  TNH_REQUIRE(G && G->ndim == 2 * r, TNH_ERR_SHAPE,
              "gate must have 2 r indices (o1, i1, ..., or, ir)");
  for (int i = 0; i < 2 * r; ++i)
    TNH_REQUIRE(G->dims[i] == d, TNH_ERR_SHAPE, "gate dims must equal the physical dimension");
  bool tc = false;
  for (int i = 0; i < r; ++i) tc |= is_complex(A[i]->dtype);
  const bool gc = is_complex(G->dtype), oc = tc || gc;
  tnh_buf *t0 = nullptr, *t1 = nullptr, *cur = nullptr, *nxt = nullptr;
  try {
    // theta: (chi_l, d^i, chi) x (chi, d chi') for i = 1 .. r-1
    const tnh_buf *th = A[0];
    for (int i = 1; i < r; ++i) {
      const int64_t K = A[i]->dims[0], m = th->numel / K, n = d * A[i]->dims[2];
      const bool c = is_complex(th->dtype) || is_complex(A[i]->dtype);
      int64_t td[3] = {cl, th->dims[1] * d, A[i]->dims[2]};
      buf_ensure(ctx, &t1, c ? TNH_C128 : TNH_F64, 3, td);
      Gemm g;
      g.M = int(m);
      g.N = int(n);
      g.K = int(K);
      g.lda = m;
      g.ldb = K;
      g.ldc = m;
      COperand a, b;
      a.re = th->re();
      a.im = is_complex(th->dtype) ? th->im() : nullptr;
      b.re = A[i]->re();
      b.im = is_complex(A[i]->dtype) ? A[i]->im() : nullptr;
      cgemm(ctx, g, a, b, t1->re(), c ? t1->im() : nullptr, TNH_F64, TNH_PREC_F64);
      std::swap(t0, t1);
      th = t0;
    }
    // the gate as a D x D mixing matrix, applied to every (a, b) column of theta
    const int Di = int(D);
    double *Wc = static_cast<double *>(ctx->workspace(841, sizeof(double) * Di * Di * 2));
    const int64_t gim = gc ? G->numel : 0;
    if (is_double(G->dtype))
      hipLaunchKernelGGL(gate_wc_n_kernel<double>, dim3((Di * Di + 255) / 256), dim3(256), 0,
                         ctx->stream, static_cast<const double *>(G->ptr), gim, int(d), r, Di, Wc, gc);
    else
      hipLaunchKernelGGL(gate_wc_n_kernel<float>, dim3((Di * Di + 255) / 256), dim3(256), 0,
                         ctx->stream, static_cast<const float *>(G->ptr), gim, int(d), r, Di, Wc, gc);
    TNH_HIP(hipGetLastError());
    int64_t cd[3] = {cl, D, cr};
    buf_ensure(ctx, &cur, oc ? TNH_C128 : TNH_F64, 3, cd);
    Small s;
    s.na = int(cl);
    s.nb = int(cr);
    s.I = Di;
    s.J = Di;
    s.in = th->ptr;
    s.in_im = is_complex(th->dtype) ? th->numel : 0;
    s.tin = TNH_F64;
    s.out = cur->ptr;
    s.out_im = oc ? cur->numel : 0;
    s.tout = TNH_F64;
    s.Wc = Wc;
    s.Wc_cplx = gc;
    s.sbI = cl * D;
    s.sbJ = cl * D;
    s.offI.resize(Di);
    s.offJ.resize(Di);
    for (int i = 0; i < Di; ++i) s.offI[i] = s.offJ[i] = cl * i;
    small_contract(ctx, s);
    buf_delete(ctx, t0);
    buf_delete(ctx, t1);
    t0 = t1 = nullptr;
    split_sites(ctx, r, d, &cur, direction, cutoff, maxdim, mindim, A, k, disc);  // :157
  } catch (...) {
    for (auto *b : {t0, t1, cur, nxt}) buf_delete(ctx, b);
    throw;
  }
}

}  // namespace tnh

using namespace tnh;

extern "C" int tnh_applygate2(tnh_ctx *ctx, tnh_buf **A1, tnh_buf **A2, const tnh_buf *G,
                              int direction, double cutoff, int maxdim, int mindim, int *k,
                              double *discarded) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(maxdim >= 0 && mindim >= 0, TNH_ERR_ARG, "maxdim and mindim must be >= 0");
    applygate2(ctx, A1, A2, G, direction, cutoff, maxdim, mindim, k, discarded);
  });
}

extern "C" int tnh_applygate_sites(tnh_ctx *ctx, int nsites, tnh_buf **A, const tnh_buf *G,
                                   int direction, double cutoff, int maxdim, int mindim, int *k,
                                   double *discarded) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(maxdim >= 0 && mindim >= 0, TNH_ERR_ARG, "maxdim and mindim must be >= 0");
    applygate_sites(ctx, nsites, A, G, direction, cutoff, maxdim, mindim, k, discarded);
  });
}

extern "C" int tnh_replacesites_sites(tnh_ctx *ctx, int nsites, const tnh_buf *theta,
                                      int direction, int normalize, double cutoff, int maxdim,
                                      int mindim, tnh_buf **A, int *k, double *discarded) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(maxdim >= 0 && mindim >= 0, TNH_ERR_ARG, "maxdim and mindim must be >= 0");
    TNH_REQUIRE(nsites >= 2 && nsites <= 6 && A && theta, TNH_ERR_ARG,
                "replacesites: nsites must be in [2, 6]");
    TNH_REQUIRE(theta->ndim == nsites + 2 && is_double(theta->dtype), TNH_ERR_SHAPE,
                "replacesites: theta must be FP64 (chi_l, d, ..., d, chi_r) with nsites d's");
    const int64_t d = theta->dims[1];
    for (int i = 1; i <= nsites; ++i)
      TNH_REQUIRE(theta->dims[i] == d, TNH_ERR_SHAPE, "replacesites: physical dimensions differ");
    int64_t D = 1;
    for (int i = 0; i < nsites; ++i) D *= d;
    const int64_t cd[3] = {theta->dims[0], D, theta->dims[nsites + 1]};
    tnh_buf *cur = buf_new(ctx, theta->dtype, 3, cd);
    try {
      copy_device(ctx, cur->ptr, theta->ptr, theta->bytes);
      split_sites(ctx, nsites, d, &cur, direction, cutoff, maxdim, mindim, A, k, discarded);
    } catch (...) {
      buf_delete(ctx, cur);
      throw;
    }
    if (normalize) {  // normalize!(psi) on the new centre, gmps.jl:262-264
      tnh_buf *c = direction ? A[0] : A[nsites - 1];
      tnh_buf v = *c;
      v.dtype = TNH_F64;
      v.numel = c->numel * (is_complex(c->dtype) ? 2 : 1);
      double nr = 0, ni = 0;
      cdot_host(ctx, &v, &v, &nr, &ni);
      nr = std::sqrt(std::max(nr, 0.0));
      TNH_REQUIRE(std::isfinite(nr) && nr > 0, TNH_ERR_NUMERIC, "replacesites: zero centre tensor");
      scale(ctx, c->ptr, c->dtype, c->numel, 1.0 / nr);
    }
  });
}
