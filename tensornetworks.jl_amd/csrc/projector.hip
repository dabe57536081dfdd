// Rank-1 ("squared") Hamiltonian terms of the two-site DMRG: ProjMPS(phi, psi; rank=2,
// squared=true, coeff), src/structures/mps/projmps.jl:16-42, which dmrg(psi, Hs...) builds for
// every rank-1 argument (src/algorithms/mps/dmrg.jl:144-145). Their blocks are overlaps
// L[a_phi, a_psi] (FP64, planar complex). Per bond step the projection P of phi's two sites
// between the blocks (project, projmps.jl:152-179) is formed once by three GEMMs; each matvec
// then adds coeff * conj(P) * sum(P .* x) (projmps.jl:135-144) in two launches with no host
// round trip: block partials of the inner product, then an update in which every block sums
// the partials itself (same order in every block, so the result is deterministic).
#include "tnh_internal.hpp"

namespace tnh {
namespace {

constexpr int kPT = 256;  // threads per block
constexpr int kPG = 256;  // blocks of the inner-product partials (<= kPT)

__device__ inline double proj_block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < kPT / 64; ++i) s += red[i];
  __syncthreads();
  return s;
}

// part[2 b], part[2 b + 1] = block b's share of sum_k P[k] x[k] (bilinear; planar complex
// with the imaginary plane at offset pim / xim, 0 = real)
__global__ void __launch_bounds__(kPT) proj_dot_kernel(const double *P, int64_t pim,
                                                       const double *x, int64_t xim, int64_t n,
                                                       double *part) {
  __shared__ double red[kPT / 64];
  double sr = 0, si = 0;
  for (int64_t k = blockIdx.x * int64_t(kPT) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kPT) {
    const double pr = P[k], xr = x[k];
    const double pi = pim ? P[pim + k] : 0.0, xi = xim ? x[xim + k] : 0.0;
    sr += pr * xr - pi * xi;
    si += pr * xi + pi * xr;
  }
  sr = proj_block_sum(sr, red);
  si = proj_block_sum(si, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sr;
    part[2 * blockIdx.x + 1] = si;
  }
}

// y[k] += coeff * s * conj(P[k]), s = the sum of the np partials
__global__ void __launch_bounds__(kPT) proj_axpy_kernel(const double *P, int64_t pim, double *y,
                                                        int64_t yim, int64_t n,
                                                        const double *part, int np,
                                                        double coeff) {
  __shared__ double red[kPT / 64];
  const int t = threadIdx.x;
  double sr = proj_block_sum(t < np ? part[2 * t] : 0.0, red);
  double si = proj_block_sum(t < np ? part[2 * t + 1] : 0.0, red);
  sr *= coeff;
  si *= coeff;
  for (int64_t k = blockIdx.x * int64_t(kPT) + t; k < n; k += int64_t(gridDim.x) * kPT) {
    const double pr = P[k], pi = pim ? P[pim + k] : 0.0;
    y[k] += sr * pr + si * pi;
    if (yim) y[yim + k] += si * pr - sr * pi;
  }
}

COperand cop(const tnh_buf *b, bool conj) {
  COperand o;
  o.re = b->re();
  o.im = is_complex(b->dtype) ? b->im() : nullptr;
  o.dtype = real_of(b->dtype);
  o.conj = conj;
  return o;
}

bool any_complex(std::initializer_list<const tnh_buf *> bs) {
  for (const tnh_buf *b : bs)
    if (is_complex(b->dtype)) return true;
  return false;
}

void require_f64(const tnh_buf *b, int nd, const char *name) {
  TNH_REQUIRE(b != nullptr, TNH_ERR_ARG, std::string(name) + " is NULL");
  TNH_REQUIRE(b->ndim == nd, TNH_ERR_SHAPE,
              std::string(name) + " must have " + std::to_string(nd) + " indices, got " +
                  shape_str(b));
  TNH_REQUIRE(is_double(b->dtype), TNH_ERR_ARG, std::string(name) + " must be FP64");
}

char *plane_im(void *p, bool cplx, int64_t n) {
  return cplx ? static_cast<char *>(p) + n * sizeof(double) : nullptr;
}

}  // namespace

// buildleft!(projV, idx) for objects [phi, psi], projmps.jl:50-66:
// L'[f', p'] = sum_{f, p, s} conj(phi[f, s, f']) L[f, p] psi[p, s, p']
void proj_env_left(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *phi, const tnh_buf *psi,
                   tnh_buf **Lout) {
  require_f64(L, 2, "overlap block");
  require_f64(phi, 3, "phi");
  require_f64(psi, 3, "psi");
  const int64_t cf = phi->dims[0], d = phi->dims[1], cf2 = phi->dims[2];
  const int64_t cp = psi->dims[0], cp2 = psi->dims[2];
  TNH_REQUIRE(psi->dims[1] == d, TNH_ERR_SHAPE, "phi and psi physical dimensions differ");
  TNH_REQUIRE(L->dims[0] == cf && L->dims[1] == cp, TNH_ERR_SHAPE,
              "overlap block " + shape_str(L) + " does not match phi " + shape_str(phi) +
                  " / psi " + shape_str(psi));
  const bool xc = any_complex({L, psi}), oc = xc || is_complex(phi->dtype);
  const int64_t nx = cf * d * cp2;
  void *X = ctx->workspace(962, nx * sizeof(double) * (xc ? 2 : 1));
  // (1) X[f, (s, p')] = L[f, p] psi[p, (s, p')]
  Gemm g1;
  g1.M = int(cf);
  g1.N = int(d * cp2);
  g1.K = int(cp);
  g1.lda = cf;
  g1.ldb = cp;
  g1.ldc = cf;
  cgemm(ctx, g1, cop(L, false), cop(psi, false), X, plane_im(X, xc, nx), TNH_F64, TNH_PREC_F64);
  // (2) L'[f', p'] = phi[(f, s), f']^H X[(f, s), p']
  int64_t od[2] = {cf2, cp2};
  buf_ensure(ctx, Lout, oc ? TNH_C128 : TNH_F64, 2, od);
  Gemm g2;
  g2.M = int(cf2);
  g2.N = int(cp2);
  g2.K = int(cf * d);
  g2.transA = true;
  g2.lda = cf * d;
  g2.ldb = cf * d;
  g2.ldc = cf2;
  COperand xop;
  xop.re = X;
  xop.im = plane_im(X, xc, nx);
  xop.dtype = TNH_F64;
  cgemm(ctx, g2, cop(phi, true), xop, (*Lout)->re(), oc ? (*Lout)->im() : nullptr, TNH_F64,
        TNH_PREC_F64);
}

// buildright!(projV, idx) for objects [phi, psi], projmps.jl:74-95:
// R'[f, p] = sum_{s, f', p'} conj(phi[f, s, f']) R[f', p'] psi[p, s, p']
void proj_env_right(tnh_ctx *ctx, const tnh_buf *R, const tnh_buf *phi, const tnh_buf *psi,
                    tnh_buf **Rout) {
  require_f64(R, 2, "overlap block");
  require_f64(phi, 3, "phi");
  require_f64(psi, 3, "psi");
  const int64_t cf = phi->dims[0], d = phi->dims[1], cf2 = phi->dims[2];
  const int64_t cp = psi->dims[0], cp2 = psi->dims[2];
  TNH_REQUIRE(psi->dims[1] == d, TNH_ERR_SHAPE, "phi and psi physical dimensions differ");
  TNH_REQUIRE(R->dims[0] == cf2 && R->dims[1] == cp2, TNH_ERR_SHAPE,
              "overlap block " + shape_str(R) + " does not match phi " + shape_str(phi) +
                  " / psi " + shape_str(psi));
  const bool yc = any_complex({R, psi}), oc = yc || is_complex(phi->dtype);
  const int64_t ny = cp * d * cf2;
  void *Y = ctx->workspace(962, ny * sizeof(double) * (yc ? 2 : 1));
  // (1) Y[(p, s), f'] = psi[(p, s), p'] R[f', p']^T
  Gemm g1;
  g1.M = int(cp * d);
  g1.N = int(cf2);
  g1.K = int(cp2);
  g1.transB = true;
  g1.lda = cp * d;
  g1.ldb = cf2;
  g1.ldc = cp * d;
  cgemm(ctx, g1, cop(psi, false), cop(R, false), Y, plane_im(Y, yc, ny), TNH_F64, TNH_PREC_F64);
  // (2) R'[f, p] = conj(phi)[f, (s, f')] Y[p, (s, f')]^T
  int64_t od[2] = {cf, cp};
  buf_ensure(ctx, Rout, oc ? TNH_C128 : TNH_F64, 2, od);
  Gemm g2;
  g2.M = int(cf);
  g2.N = int(cp);
  g2.K = int(d * cf2);
  g2.transB = true;
  g2.lda = cf;
  g2.ldb = cp;
  g2.ldc = cf;
  COperand yop;
  yop.re = Y;
  yop.im = plane_im(Y, yc, ny);
  yop.dtype = TNH_F64;
  cgemm(ctx, g2, cop(phi, true), yop, (*Rout)->re(), oc ? (*Rout)->im() : nullptr, TNH_F64,
        TNH_PREC_F64);
}

// project(projV, A, direction, 2), projmps.jl:152-179:
// P[a, s1, s2, b] = sum L[f, a] conj(phi1[f, s1, g]) conj(phi2[g, s2, h]) R[h, b]
void proj_project(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *phi1, const tnh_buf *phi2,
                  const tnh_buf *R, tnh_buf **P) {
  require_f64(L, 2, "left overlap block");
  require_f64(R, 2, "right overlap block");
  require_f64(phi1, 3, "phi");
  require_f64(phi2, 3, "phi");
  const int64_t cf = phi1->dims[0], d = phi1->dims[1], cm = phi1->dims[2], ch = phi2->dims[2];
  const int64_t ca = L->dims[1], cb = R->dims[1];
  TNH_REQUIRE(phi2->dims[0] == cm && phi2->dims[1] == d, TNH_ERR_SHAPE,
              "phi sites " + shape_str(phi1) + " and " + shape_str(phi2) + " do not chain");
  TNH_REQUIRE(L->dims[0] == cf && R->dims[0] == ch, TNH_ERR_SHAPE,
              "overlap blocks " + shape_str(L) + " / " + shape_str(R) + " do not match phi");
  const bool c1 = any_complex({L, phi1}), c2 = c1 || is_complex(phi2->dtype);
  const bool c3 = c2 || is_complex(R->dtype);
  const int64_t n1 = ca * d * cm, n2 = ca * d * d * ch;
  void *Q1 = ctx->workspace(960, n1 * sizeof(double) * (c1 ? 2 : 1));
  void *Q2 = ctx->workspace(961, n2 * sizeof(double) * (c2 ? 2 : 1));
  auto op = [](void *p, bool c, int64_t n) {
    COperand o;
    o.re = p;
    o.im = plane_im(p, c, n);
    o.dtype = TNH_F64;
    return o;
  };
  // (1) Q1[a, (s1, g)] = L[f, a]^T conj(phi1)[f, (s1, g)]
  Gemm g1;
  g1.M = int(ca);
  g1.N = int(d * cm);
  g1.K = int(cf);
  g1.transA = true;
  g1.lda = cf;
  g1.ldb = cf;
  g1.ldc = ca;
  cgemm(ctx, g1, cop(L, false), cop(phi1, true), Q1, plane_im(Q1, c1, n1), TNH_F64, TNH_PREC_F64);
  // (2) Q2[(a, s1), (s2, h)] = Q1[(a, s1), g] conj(phi2)[g, (s2, h)]
  Gemm g2;
  g2.M = int(ca * d);
  g2.N = int(d * ch);
  g2.K = int(cm);
  g2.lda = ca * d;
  g2.ldb = cm;
  g2.ldc = ca * d;
  cgemm(ctx, g2, op(Q1, c1, n1), cop(phi2, true), Q2, plane_im(Q2, c2, n2), TNH_F64, TNH_PREC_F64);
  // (3) P[(a, s1, s2), b] = Q2[(a, s1, s2), h] R[h, b]
  int64_t pd[4] = {ca, d, d, cb};
  buf_ensure(ctx, P, c3 ? TNH_C128 : TNH_F64, 4, pd);
  Gemm g3;
  g3.M = int(ca * d * d);
  g3.N = int(cb);
  g3.K = int(ch);
  g3.lda = ca * d * d;
  g3.ldb = ch;
  g3.ldc = ca * d * d;
  cgemm(ctx, g3, op(Q2, c2, n2), cop(R, false), (*P)->re(), c3 ? (*P)->im() : nullptr, TNH_F64,
        TNH_PREC_F64);
}

// project(projV, A, direction, 1), projmps.jl:152-179 with nsites = 1:
// P[a, s, b] = sum L[f, a] conj(phi[f, s, h]) R[h, b]
void proj_project1(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *phi, const tnh_buf *R,
                   tnh_buf **P) {
  require_f64(L, 2, "left overlap block");
  require_f64(R, 2, "right overlap block");
  require_f64(phi, 3, "phi");
  const int64_t cf = phi->dims[0], d = phi->dims[1], ch = phi->dims[2];
  const int64_t ca = L->dims[1], cb = R->dims[1];
  TNH_REQUIRE(L->dims[0] == cf && R->dims[0] == ch, TNH_ERR_SHAPE,
              "overlap blocks " + shape_str(L) + " / " + shape_str(R) + " do not match phi");
  const bool c1 = any_complex({L, phi}), c2 = c1 || is_complex(R->dtype);
  const int64_t n1 = ca * d * ch;
  void *Q1 = ctx->workspace(960, n1 * sizeof(double) * (c1 ? 2 : 1));
  // (1) Q1[a, (s, h)] = L[f, a]^T conj(phi)[f, (s, h)]
  Gemm g1;
  g1.M = int(ca);
  g1.N = int(d * ch);
  g1.K = int(cf);
  g1.transA = true;
  g1.lda = cf;
  g1.ldb = cf;
  g1.ldc = ca;
  cgemm(ctx, g1, cop(L, false), cop(phi, true), Q1, plane_im(Q1, c1, n1), TNH_F64, TNH_PREC_F64);
  // (2) P[(a, s), b] = Q1[(a, s), h] R[h, b]
  int64_t pd[3] = {ca, d, cb};
  buf_ensure(ctx, P, c2 ? TNH_C128 : TNH_F64, 3, pd);
  Gemm g2;
  g2.M = int(ca * d);
  g2.N = int(cb);
  g2.K = int(ch);
  g2.lda = ca * d;
  g2.ldb = ch;
  g2.ldc = ca * d;
  COperand q;
  q.re = Q1;
  q.im = plane_im(Q1, c1, n1);
  q.dtype = TNH_F64;
  cgemm(ctx, g2, q, cop(R, false), (*P)->re(), c2 ? (*P)->im() : nullptr, TNH_F64, TNH_PREC_F64);
}

// y += coeff * conj(P) * sum(P .* x): the squared branch of product, projmps.jl:135-144
void proj_apply(tnh_ctx *ctx, const tnh_buf *P, const tnh_buf *x, tnh_buf *y, double coeff) {
  TNH_REQUIRE(P->numel == x->numel && y->numel == x->numel, TNH_ERR_SHAPE,
              "projection " + shape_str(P) + " does not match theta " + shape_str(x));
  TNH_REQUIRE(is_double(P->dtype) && is_double(x->dtype) && is_double(y->dtype), TNH_ERR_ARG,
              "projector terms are FP64");
  TNH_REQUIRE(is_complex(y->dtype) || !(is_complex(P->dtype) || is_complex(x->dtype)),
              TNH_ERR_ARG, "complex projector term on a real state");
  const int64_t n = x->numel;
  if (n == 0) return;
  double *part = static_cast<double *>(ctx->workspace(963, sizeof(double) * 2 * kPG));
  const int64_t blocks = (n + kPT - 1) / kPT;
  const int gd = int(std::min<int64_t>(blocks, kPG));
  const int ga = int(std::min<int64_t>(blocks, 2048));
  const int64_t pim = is_complex(P->dtype) ? n : 0;
  hipLaunchKernelGGL(proj_dot_kernel, dim3(gd), dim3(kPT), 0, ctx->stream,
                     static_cast<const double *>(P->ptr), pim, static_cast<const double *>(x->ptr),
                     is_complex(x->dtype) ? n : 0, n, part);
  TNH_HIP(hipGetLastError());
  hipLaunchKernelGGL(proj_axpy_kernel, dim3(ga), dim3(kPT), 0, ctx->stream,
                     static_cast<const double *>(P->ptr), pim, static_cast<double *>(y->ptr),
                     is_complex(y->dtype) ? n : 0, n, part, gd, coeff);
  TNH_HIP(hipGetLastError());
}

// calculate(projV) for objects [phi, psi], projmps.jl:192-216 (without coeff): <phi|psi>
// from the blocks either side of site c
void proj_calculate(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *phi, const tnh_buf *psi,
                    const tnh_buf *R, double *re, double *im) {
  tnh_buf *Lp = nullptr;
  try {
    proj_env_left(ctx, L, phi, psi, &Lp);
    TNH_REQUIRE(Lp->numel == R->numel, TNH_ERR_SHAPE,
                "calculate: right overlap block " + shape_str(R) + " != " + shape_str(Lp));
    cdot_host(ctx, Lp, R, re, im);
  } catch (...) {
    buf_delete(ctx, Lp);
    throw;
  }
  buf_delete(ctx, Lp);
}

}  // namespace tnh
