// <psi|H|psi> for an MPO H (inner(psi, H, psi), src/structures/mps/mpo.jl:181-217) and its
// gradient with respect to every MPO tensor, dE/dW_i[w, s, t, w'] =
//   sum L_{i-1}[a, w, a'] conj(A_i[a, s, b]) A_i[a', t, b'] R_{i+1}[b, w', b'],
// so that E = sum W_i . G_i at every site. Couplings that enter W linearly get their
// derivative from G by the chain rule on the host. The reference has no counterpart (it
// has no differentiation); the contraction conventions are those of buildleft!/buildright!,
// src/structures/mps/projmps.jl:50-95.
//
// Memory: the forward sweep keeps the left environment only at block boundaries (every
// `block` sites); the backward sweep recomputes the environments inside one block from its
// checkpoint, then walks the block right to left with a running right environment.
#include <algorithm>
#include <vector>

#include "tnh_internal.hpp"

namespace tnh {
namespace {

// part[bp][o] (planar: re at 2*o, im at 2*o + 1) for o = w + Dl (s + d (t + d w')):
//   sum_b P[(s, b), (w, t, bp)] R[b, w', bp]
// P is the (d cb) x (Dl d cb) matrix A^H (L A); R the right block (cb, Dr, cb).
__global__ void __launch_bounds__(256) site_grad_partial_kernel(
    const double *__restrict__ P, int64_t P_im, const double *__restrict__ R, int64_t R_im,
    int d, int cb, int Dl, int Dr, double *__restrict__ part) {
  __shared__ double red_re[256];
  __shared__ double red_im[256];
  const int bp = blockIdx.x;
  const int nout = Dl * d * d * Dr;
  const int tid = threadIdx.x;
  const int group = nout < 256 ? nout : 256;
  const int nsplit = 256 / group;
  const int oo = tid % group, sp = tid / group;
  const int64_t ldP = int64_t(d) * cb;
  for (int obase = 0; obase < nout; obase += group) {
    const int o = obase + oo;
    double ar = 0.0, ai = 0.0;
    if (o < nout && sp < nsplit) {
      const int w = o % Dl;
      const int s = (o / Dl) % d;
      const int t = (o / (Dl * d)) % d;
      const int wp = o / (Dl * d * d);
      const double *pc = P + s + ldP * (w + int64_t(Dl) * (t + int64_t(d) * bp));
      const double *rc = R + int64_t(cb) * wp + int64_t(cb) * Dr * bp;
      for (int b = sp; b < cb; b += nsplit) {
        const double pr = pc[int64_t(d) * b], rr = rc[b];
        const double pi = P_im ? pc[P_im + int64_t(d) * b] : 0.0;
        const double ri = R_im ? rc[R_im + b] : 0.0;
        ar += pr * rr - pi * ri;
        ai += pr * ri + pi * rr;
      }
    }
    red_re[tid] = ar;
    red_im[tid] = ai;
    __syncthreads();
    if (sp == 0 && o < nout) {
      double sr = 0.0, si = 0.0;
      for (int p = 0; p < nsplit; ++p) {
        sr += red_re[oo + p * group];
        si += red_im[oo + p * group];
      }
      part[(int64_t(bp) * nout + o) * 2] = sr;
      part[(int64_t(bp) * nout + o) * 2 + 1] = si;
    }
    __syncthreads();
  }
}

// G[o] = sum_bp part[bp][o], in a fixed order (deterministic); planar output
__global__ void site_grad_reduce_kernel(const double *__restrict__ part, int nbp, int nout,
                                        double *__restrict__ G, bool cplx) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nout) return;
  double sr = 0.0, si = 0.0;
  for (int b = 0; b < nbp; ++b) {
    sr += part[(int64_t(b) * nout + o) * 2];
    si += part[(int64_t(b) * nout + o) * 2 + 1];
  }
  G[o] = sr;
  if (cplx) G[nout + o] = si;
}

COperand opnd(const tnh_buf *b, bool conj = false) {
  COperand o;
  o.re = b->re();
  o.im = is_complex(b->dtype) ? b->im() : nullptr;
  o.dtype = real_of(b->dtype);
  o.conj = conj;
  return o;
}

tnh_buf *ones_block(tnh_ctx *ctx) {
  int64_t dims[3] = {1, 1, 1};
  tnh_buf *b = buf_new(ctx, TNH_F64, 3, dims);
  const double one = 1.0;
  to_device(ctx, b->ptr, &one, sizeof(double));
  return b;
}

// G = dE/dW for one site from its left block L (ca, Dl, ca), the site A (ca, d, cb) and the
// right block R (cb, Dr, cb); G has shape (Dl, d, d, Dr).
void site_grad(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *A, const tnh_buf *R, int Dl, int Dr,
               tnh_buf **G) {
  const int64_t ca = A->dims[0], d = A->dims[1], cb = A->dims[2];
  TNH_REQUIRE(L->ndim == 3 && L->dims[0] == ca && L->dims[1] == Dl && L->dims[2] == ca, TNH_ERR_SHAPE,
              "grad: left block does not match the site");
  TNH_REQUIRE(R->ndim == 3 && R->dims[0] == cb && R->dims[1] == Dr && R->dims[2] == cb, TNH_ERR_SHAPE,
              "grad: right block does not match the site");
  TNH_REQUIRE(is_double(L->dtype) && is_double(A->dtype) && is_double(R->dtype), TNH_ERR_ARG,
              "grad: FP64 operands required");
  const bool xc = is_complex(L->dtype) || is_complex(A->dtype);
  const bool pc = xc;  // A^H X is complex iff A or X is
  const bool gc = pc || is_complex(R->dtype);
  const int64_t nx = ca * Dl * d * cb, np = d * cb * Dl * d * cb;
  double *X = static_cast<double *>(ctx->workspace(905, sizeof(double) * nx * (xc ? 2 : 1)));
  double *P = static_cast<double *>(ctx->workspace(906, sizeof(double) * np * (pc ? 2 : 1)));
  // X[(a,w),(t,b')] = L[(a,w),a'] A[a',(t,b')]
  Gemm g1;
  g1.M = int(ca * Dl);
  g1.N = int(d * cb);
  g1.K = int(ca);
  g1.lda = ca * Dl;
  g1.ldb = ca;
  g1.ldc = ca * Dl;
  cgemm(ctx, g1, opnd(L), opnd(A), X, xc ? X + nx : nullptr, TNH_F64, TNH_PREC_F64);
  // P[(s,b),(w,t,b')] = A[a,(s,b)]^H X[a,(w,t,b')]
  Gemm g2;
  g2.M = int(d * cb);
  g2.N = int(Dl * d * cb);
  g2.K = int(ca);
  g2.transA = true;
  g2.lda = ca;
  g2.ldb = ca;
  g2.ldc = d * cb;
  COperand xo;
  xo.re = X;
  xo.im = xc ? X + nx : nullptr;
  cgemm(ctx, g2, opnd(A, true), xo, P, pc ? P + np : nullptr, TNH_F64, TNH_PREC_F64);
  const int nout = int(Dl * d * d * Dr);
  double *part = static_cast<double *>(ctx->workspace(907, sizeof(double) * 2 * nout * cb));
  hipLaunchKernelGGL(site_grad_partial_kernel, dim3(unsigned(cb)), dim3(256), 0, ctx->stream, P,
                     pc ? np : 0, static_cast<const double *>(R->ptr), is_complex(R->dtype) ? R->numel : 0,
                     int(d), int(cb), Dl, Dr, part);
  TNH_HIP(hipGetLastError());
  int64_t gd[4] = {Dl, d, d, Dr};
  buf_ensure(ctx, G, gc ? TNH_C128 : TNH_F64, 4, gd);
  hipLaunchKernelGGL(site_grad_reduce_kernel, dim3((nout + 255) / 256), dim3(256), 0, ctx->stream, part,
                     int(cb), nout, static_cast<double *>((*G)->ptr), gc);
  TNH_HIP(hipGetLastError());
}

void mpo_expect_grad(tnh_ctx *ctx, int N, tnh_buf *const *psi, tnh_buf *const *W, int block,
                     double *er, double *ei, tnh_buf **dW) {
  TNH_REQUIRE(N >= 1 && psi && W && block >= 1, TNH_ERR_ARG, "grad: bad arguments");
  for (int i = 0; i < N; ++i) {
    TNH_REQUIRE(psi[i] && psi[i]->ndim == 3 && W[i] && W[i]->ndim == 4, TNH_ERR_SHAPE,
                "grad: sites must be (chi_l, d, chi_r) and MPO tensors (D_l, d, d, D_r)");
    TNH_REQUIRE(W[i]->dims[1] == psi[i]->dims[1] && W[i]->dims[2] == psi[i]->dims[1], TNH_ERR_SHAPE,
                "grad: physical dimensions of psi and the MPO differ at site " + std::to_string(i + 1));
    if (i + 1 < N) {
      TNH_REQUIRE(psi[i]->dims[2] == psi[i + 1]->dims[0], TNH_ERR_SHAPE, "grad: psi bond mismatch");
      TNH_REQUIRE(W[i]->dims[3] == W[i + 1]->dims[0], TNH_ERR_SHAPE, "grad: MPO bond mismatch");
    }
  }
  TNH_REQUIRE(psi[0]->dims[0] == 1 && psi[N - 1]->dims[2] == 1 && W[0]->dims[0] == 1 &&
                  W[N - 1]->dims[3] == 1,
              TNH_ERR_SHAPE, "grad: edge bonds must be 1");
  std::vector<tnh_buf *> owned;
  auto cleanup = [&] {
    for (tnh_buf *b : owned) buf_delete(ctx, b);
    owned.clear();
  };
  try {
    tnh_buf *edge = ones_block(ctx);
    owned.push_back(edge);
    const int nblk = (N + block - 1) / block;
    // forward: ckpt[k] = left block before site k*block (0-based sites)
    std::vector<tnh_buf *> ckpt(nblk, nullptr);
    ckpt[0] = edge;
    tnh_buf *L = edge;
    for (int i = 0; i < N; ++i) {
      tnh_buf *Ln = nullptr;
      env_left(ctx, L, psi[i], W[i], &Ln, TNH_PREC_F64);
      owned.push_back(Ln);
      const bool boundary = (i + 1) % block == 0 && (i + 1) < N;
      if (boundary) ckpt[(i + 1) / block] = Ln;
      // drop the previous block unless it is a checkpoint
      if (L != edge && std::find(ckpt.begin(), ckpt.end(), L) == ckpt.end()) {
        owned.erase(std::find(owned.begin(), owned.end(), L));
        buf_delete(ctx, L);
      }
      L = Ln;
    }
    double val[2] = {0.0, 0.0};
    to_host(ctx, val, L->ptr, sizeof(double));
    if (is_complex(L->dtype)) to_host(ctx, val + 1, L->im(), sizeof(double));
    *er = val[0];
    *ei = val[1];
    if (dW) {
      tnh_buf *R = edge;
      for (int k = nblk - 1; k >= 0; --k) {
        const int start = k * block, end = std::min(N, start + block);
        std::vector<tnh_buf *> Lb(end - start, nullptr);
        Lb[0] = ckpt[k];
        for (int j = start; j + 1 < end; ++j) {
          env_left(ctx, Lb[j - start], psi[j], W[j], &Lb[j - start + 1], TNH_PREC_F64);
          owned.push_back(Lb[j - start + 1]);
        }
        for (int i = end - 1; i >= start; --i) {
          site_grad(ctx, Lb[i - start], psi[i], R, int(W[i]->dims[0]), int(W[i]->dims[3]), &dW[i]);
          if (i > 0) {
            tnh_buf *Rn = nullptr;
            env_right(ctx, R, psi[i], W[i], &Rn, TNH_PREC_F64);
            owned.push_back(Rn);
            if (R != edge) {
              owned.erase(std::find(owned.begin(), owned.end(), R));
              buf_delete(ctx, R);
            }
            R = Rn;
          }
        }
        for (int j = 1; j < end - start; ++j) {
          owned.erase(std::find(owned.begin(), owned.end(), Lb[j]));
          buf_delete(ctx, Lb[j]);
        }
      }
    }
    TNH_HIP(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace
}  // namespace tnh

using namespace tnh;

extern "C" int tnh_mpo_expect_grad(tnh_ctx *ctx, int N, tnh_buf *const *psi, tnh_buf *const *W,
                                   int block, double *energy_re, double *energy_im, tnh_buf **dW) {
  if (!ctx || !energy_re || !energy_im) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] { mpo_expect_grad(ctx, N, psi, W, block, energy_re, energy_im, dW); });
}
