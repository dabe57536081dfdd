// GEMM dispatch: picks the arithmetic mode and tile shape, and builds planar-complex
// products from real MFMA GEMMs.
#include "gemm_kernel.hpp"
#include "tnh_internal.hpp"

namespace tnh {
namespace gk {
void launch_f64(hipStream_t, const Params &, int, bool, bool, int, int, int, bool);
void launch_f32(hipStream_t, const Params &, int, bool, bool, int, int, int, bool);
void launch_bf16p(hipStream_t, const Params &, int, int, bool);
void split_bf16(hipStream_t, const void *, int, bool, int64_t, int64_t, int, int, int, int,
                unsigned short *);
}  // namespace gk

void gemm(tnh_ctx *ctx, const Gemm &g, int prec) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return;
  TNH_REQUIRE(g.K >= 0, TNH_ERR_ARG, "gemm: negative K");
  TNH_REQUIRE(g.ldc >= g.M, TNH_ERR_ARG, "gemm: ldc < M");
  TNH_REQUIRE(g.K == 0 || (g.transA ? g.lda >= g.K : g.lda >= g.M), TNH_ERR_ARG, "gemm: bad lda");
  TNH_REQUIRE(g.K == 0 || (g.transB ? g.ldb >= g.N : g.ldb >= g.K), TNH_ERR_ARG, "gemm: bad ldb");
  TNH_REQUIRE(g.batch <= 65535, TNH_ERR_ARG, "gemm: batch too large");
  // FP64 arithmetic needs FP64 operands; lower-precision modes take either.
  if (prec == TNH_PREC_F64)
    TNH_REQUIRE(g.tA == TNH_F64 && g.tB == TNH_F64 && g.tC == TNH_F64, TNH_ERR_ARG,
                "gemm: FP64 arithmetic with FP32 operands");
  // Large tiles once the product fills the chip with 128x128 blocks.
  const bool big = (int64_t((g.M + 127) / 128) * ((g.N + 127) / 128) * g.batch >= ctx->cu_count) &&
                   g.M >= 96 && g.N >= 96;
  // FP64 keeps its 64-wide accumulators at two waves per SIMD with 128 x 64 blocks
  const int bm = big ? 128 : 64, bn = big ? (prec == TNH_PREC_F64 ? 64 : 128) : 64;
  gk::Params p;
  p.M = g.M;
  p.N = g.N;
  p.K = g.K;
  p.A = g.A;
  p.lda = g.lda;
  p.sA = g.sA;
  p.B = g.B;
  p.ldb = g.ldb;
  p.sB = g.sB;
  p.C = g.C;
  p.ldc = g.ldc;
  p.sC = g.sC;
  p.alpha = g.alpha;
  p.beta = g.beta;
  p.tiles_m = (g.M + bm - 1) / bm;
  p.tiles_n = (g.N + bn - 1) / bn;
  // 16-byte vector paths: every vector must lie wholly inside or outside the matrix and
  // be 16-byte aligned, i.e. the contiguous extent, the leading dimension, the batch
  // stride and the base address are multiples of the vector width.
  auto aligned16 = [](const void *ptr) { return (reinterpret_cast<uintptr_t>(ptr) & 15u) == 0; };
  auto vec_ok = [&](const void *ptr, int dtype, int64_t contiguous_extent, int64_t ld, int64_t stride) {
    const int64_t vw = 16 / int64_t(real_size(dtype));
    return aligned16(ptr) && contiguous_extent % vw == 0 && ld % vw == 0 && stride % vw == 0;
  };
  // A: k-contiguous when transposed (extent K), else row-contiguous (extent M)
  p.vecA = vec_ok(g.A, g.tA, g.transA ? g.K : g.M, g.lda, g.sA);
  // B: k-contiguous when not transposed (extent K), else row-contiguous (extent N)
  p.vecB = vec_ok(g.B, g.tB, g.transB ? g.N : g.K, g.ldb, g.sB);
  {
    const int64_t cw = g.tC == TNH_F64 ? 2 : 4;
    p.vecC = aligned16(g.C) && g.ldc % cw == 0 && g.sC % cw == 0;
  }
  static const char *kFamily[3] = {"gemm_f64", "gemm_f32", "gemm_bf16x3"};
  hipEvent_t ev = ctx->prof_begin();
  switch (prec) {
    case TNH_PREC_F64:
      gk::launch_f64(ctx->stream, p, g.batch, g.transA, g.transB, g.tA, g.tB, g.tC, big);
      break;
    case TNH_PREC_F32:
      gk::launch_f32(ctx->stream, p, g.batch, g.transA, g.transB, g.tA, g.tB, g.tC, big);
      break;
    case TNH_PREC_BF16X3: {
      // split both operands once into k-contiguous hi/lo bf16 planes (k padded to 32)
      const int Kp = (g.K + 31) / 32 * 32;
      if (Kp == 0) {  // empty sum: C = beta C through the FP32 kernel
        gk::launch_f32(ctx->stream, p, g.batch, g.transA, g.transB, g.tA, g.tB, g.tC, big);
        break;
      }
      const int64_t pa = int64_t(g.M) * Kp, pb = int64_t(g.N) * Kp;
      auto *Ap = static_cast<unsigned short *>(
          ctx->workspace(950, sizeof(unsigned short) * 2 * pa * g.batch));
      auto *Bp = static_cast<unsigned short *>(
          ctx->workspace(951, sizeof(unsigned short) * 2 * pb * g.batch));
      // op(A)[m, k]: rows (m) contiguous unless A is transposed; op(B)[k, n]: k contiguous
      // unless B is transposed
      gk::split_bf16(ctx->stream, g.A, g.tA, !g.transA, g.lda, g.sA, g.M, g.K, Kp, g.batch, Ap);
      gk::split_bf16(ctx->stream, g.B, g.tB, g.transB, g.ldb, g.sB, g.N, g.K, Kp, g.batch, Bp);
      gk::Params q = p;
      q.A = Ap;
      q.B = Bp;
      q.K = Kp;
      q.lda = Kp;
      q.ldb = Kp;
      q.sA = 2 * pa;
      q.sB = 2 * pb;
      q.planeA = pa;
      q.planeB = pb;
      gk::launch_bf16p(ctx->stream, q, g.batch, g.tC, big);
      break;
    }
    default:
      throw Error(TNH_ERR_ARG, "unknown precision " + std::to_string(prec));
  }
  TNH_HIP(hipGetLastError());
  if (ev) {
    const double mnk = double(g.M) * g.N * g.K * g.batch;
    const double bytes = (double(g.M) * g.K * real_size(g.tA) + double(g.K) * g.N * real_size(g.tB) +
                          double(g.M) * g.N * real_size(g.tC) * (g.beta != 0.0 ? 2 : 1)) *
                         g.batch;
    ctx->prof_end(kFamily[prec], ev, 2.0 * mnk, bytes);
  }
}

// Planar complex product from real products:
//   Re C = a (Ar Br - sa sb Ai Bi) + b Re C,  Im C = a (sb Ar Bi + sa Ai Br) + b Im C
// with sa = -1 when A is conjugated (likewise sb). Purely real operands drop terms.
void cgemm(tnh_ctx *ctx, Gemm g, COperand A, COperand B, void *Cre, void *Cim, int tC, int prec) {
  const double alpha = g.alpha, beta = g.beta;
  const double sa = A.conj ? -1.0 : 1.0, sb = B.conj ? -1.0 : 1.0;
  g.tA = A.dtype;
  g.tB = B.dtype;
  g.tC = tC;
  auto run = [&](const void *a, const void *b, void *c, double al, double be) {
    Gemm h = g;
    h.A = a;
    h.B = b;
    h.C = c;
    h.alpha = al;
    h.beta = be;
    gemm(ctx, h, prec);
  };
  // real part
  run(A.re, B.re, Cre, alpha, beta);
  if (A.im && B.im) run(A.im, B.im, Cre, -sa * sb * alpha, 1.0);
  if (!Cim) {
    TNH_REQUIRE(!A.im && !B.im, TNH_ERR_ARG, "cgemm: complex operands need a complex output");
    return;
  }
  // imaginary part
  bool started = false;
  if (B.im) {
    run(A.re, B.im, Cim, sb * alpha, beta);
    started = true;
  }
  if (A.im) {
    run(A.im, B.re, Cim, sa * alpha, started ? 1.0 : beta);
    started = true;
  }
  if (!started) {
    // both operands real: the imaginary part of C becomes beta * Im C
    TNH_REQUIRE(g.batch == 1 && g.ldc == g.M, TNH_ERR_ARG,
                "cgemm: real operands with a strided complex output");
    int64_t n = int64_t(g.M) * g.N;
    if (beta == 0.0)
      fill(ctx, Cim, real_of(tC), n, 0.0);
    else
      scale(ctx, Cim, real_of(tC), n, beta);
  }
}

}  // namespace tnh
