// GEMM dispatch: picks the arithmetic mode and tile shape, and builds planar-complex
// products from real MFMA GEMMs.
#include <algorithm>
#include <cstdio>
#include <string>

#include "gemm_kernel.hpp"
#include "tnh_internal.hpp"

namespace tnh {
namespace gk {
void launch_f64(hipStream_t, const Params &, int, bool, bool, int, int, int, int);
void launch_f32(hipStream_t, const Params &, int, bool, bool, int, int, int, int);
void launch_bf16p(hipStream_t, const Params &, int, int, int);
void split_bf16(hipStream_t, const void *, int, bool, int64_t, int64_t, int, int, int, int,
                unsigned short *);
}  // namespace gk

namespace {

// Candidate tiles (BM x BN) per arithmetic mode, indexed as the launchers index them.
struct TileShape {
  int bm, bn;
};
constexpr int kMaxTiles = 3;
constexpr TileShape kTileShapes[3][kMaxTiles] = {
    {{64, 64}, {128, 64}, {128, 128}},  // FP64
    {{64, 64}, {128, 128}, {128, 64}},  // FP32
    {{64, 64}, {128, 128}, {0, 0}},     // split-BF16x3
};
constexpr int kNumTiles[3] = {3, 3, 2};
// Products below this many flops (or with a side under 96, or batched) keep the size rule:
// their launch latency outweighs the tile choice and timing them costs more than it saves.
constexpr double kTuneMinFlops = 268435456.0;  // 2^28
constexpr float kTuneMargin = 0.03f;  // a tile must beat the rule's by 3% to replace it

// Four classes per octave: the exponent and the two bits below the leading one.
int64_t size_class(int64_t x) {
  if (x <= 0) return 0;
  const int e = 63 - __builtin_clzll(uint64_t(x));
  const int64_t m = e >= 2 ? (x >> (e - 2)) & 3 : 0;
  return int64_t(e) * 4 + m;
}

void set_tiles(gk::Params &p, int prec, int tile) {
  const TileShape t = kTileShapes[prec][tile];
  p.tiles_m = (p.M + t.bm - 1) / t.bm;
  p.tiles_n = (p.N + t.bn - 1) / t.bn;
}

}  // namespace

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
  TNH_REQUIRE(prec >= 0 && prec <= 2, TNH_ERR_ARG, "unknown precision " + std::to_string(prec));
  // Size rule (the autotuner's fallback and its first candidate): large tiles once the
  // product fills the chip with 128 x 128 blocks; FP64 then keeps its 64-wide accumulators
  // at two waves per SIMD with 128 x 64 blocks (tile 1 of every mode).
  const bool big = (int64_t((g.M + 127) / 128) * ((g.N + 127) / 128) * g.batch >= ctx->cu_count) &&
                   g.M >= 96 && g.N >= 96;
  const int rule = big ? 1 : 0;
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
  // one launch of tile `tile` of mode `mode` on parameters q
  auto launch = [&](int mode, const gk::Params &q, int tile) {
    gk::Params r = q;
    set_tiles(r, mode, tile);
    switch (mode) {
      case TNH_PREC_F64:
        gk::launch_f64(ctx->stream, r, g.batch, g.transA, g.transB, g.tA, g.tB, g.tC, tile);
        break;
      case TNH_PREC_F32:
        gk::launch_f32(ctx->stream, r, g.batch, g.transA, g.transB, g.tA, g.tB, g.tC, tile);
        break;
      default:
        gk::launch_bf16p(ctx->stream, r, g.batch, g.tC, tile);
    }
  };
  // tile of mode `mode` for parameters q: the size rule, or the autotuner's choice for the
  // size class, timed now if the class is new
  auto pick = [&](int mode, const gk::Params &q) -> int {
    if (!ctx->gemm_autotune || g.batch != 1 || g.M < 96 || g.N < 96 ||
        2.0 * double(q.M) * q.N * q.K < kTuneMinFlops)
      return rule;
    const std::array<int64_t, 12> key = {mode, g.transA, g.transB, g.tA, g.tB, g.tC, q.vecA, q.vecB,
                                         q.vecC, size_class(q.M), size_class(q.N), size_class(q.K)};
    auto it = ctx->gemm_tiles.find(key);
    if (it != ctx->gemm_tiles.end()) return it->second.tile;
    if (ctx->arena) return rule;  // timing needs a scratch output: not inside a no-allocation call
    hipStreamCaptureStatus cap_status = hipStreamCaptureStatusNone;
    TNH_HIP(hipStreamIsCapturing(ctx->stream, &cap_status));
    if (cap_status != hipStreamCaptureStatusNone) return rule;  // no synchronisation while captured
    // time every candidate writing a scratch C (beta = 0, same leading dimension and
    // alignment class as C): one untimed launch, then two timed ones, best kept
    const size_t cbytes = (size_t(g.ldc) * (g.N - 1) + g.M) * real_size(g.tC);
    size_t cap = 0;
    void *scratch = nullptr;
    try {
      scratch = ctx->alloc(cbytes, &cap);
    } catch (const Error &) {
      return rule;  // no memory for a scratch output: the size rule, re-tried on a later call
    }
    tnh_ctx::TileChoice tc;
    tc.rule = rule;
    tc.M = q.M;
    tc.N = q.N;
    tc.K = q.K;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    auto give_back = [&]() {
      for (hipEvent_t e : {ev0, ev1, ev2})
        if (e) ctx->event_pool.push_back(e);
      ctx->release(scratch, cap);
    };
    try {
      ev0 = ctx->take_event();
      ev1 = ctx->take_event();
      ev2 = ctx->take_event();
      gk::Params s = q;
      s.C = scratch;
      s.beta = 0.0;
      float best = 0.0f;
      // the rule's tile first; another tile replaces it only when at least kTuneMargin
      // faster, so that timing noise between near-equal tiles cannot flip the choice
      float rule_ms = 0.0f;
      for (int i = 0; i < kNumTiles[mode]; ++i) {
        const int tile = i == 0 ? rule : (i == rule ? 0 : i);
        launch(mode, s, tile);
        TNH_HIP(hipEventRecord(ev0, ctx->stream));
        launch(mode, s, tile);
        TNH_HIP(hipEventRecord(ev1, ctx->stream));
        launch(mode, s, tile);
        TNH_HIP(hipEventRecord(ev2, ctx->stream));
        TNH_HIP(hipGetLastError());
        TNH_HIP(hipEventSynchronize(ev2));
        float t1 = 0.0f, t2 = 0.0f;
        TNH_HIP(hipEventElapsedTime(&t1, ev0, ev1));
        TNH_HIP(hipEventElapsedTime(&t2, ev1, ev2));
        const float t = std::min(t1, t2);
        tc.ms[tile] = t;
        if (i == 0) {
          rule_ms = best = t;
          tc.tile = tile;
        } else if (t < best && t < (1.0f - kTuneMargin) * rule_ms) {
          best = t;
          tc.tile = tile;
        }
      }
    } catch (...) {
      give_back();
      throw;
    }
    give_back();
    ctx->gemm_tiles[key] = tc;
    return tc.tile;
  };
  hipEvent_t ev = nullptr;
  switch (prec) {
    case TNH_PREC_F64:
    case TNH_PREC_F32: {
      const int tile = pick(prec, p);
      ev = ctx->prof_begin();
      launch(prec, p, tile);
      break;
    }
    case TNH_PREC_BF16X3: {
      // the profiled interval includes the operand split
      ev = ctx->prof_begin();
      // split both operands once into k-contiguous hi/lo bf16 planes (k padded to 32)
      const int Kp = (g.K + 31) / 32 * 32;
      if (Kp == 0) {  // empty sum: C = beta C through the FP32 kernel
        launch(TNH_PREC_F32, p, rule);
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
      const size_t tuned = ctx->gemm_tiles.size();
      const int tile = pick(TNH_PREC_BF16X3, q);
      if (ev && ctx->gemm_tiles.size() != tuned) {  // keep the tuning runs out of the profile
        ctx->event_pool.push_back(ev);
        ev = ctx->prof_begin();
      }
      launch(TNH_PREC_BF16X3, q, tile);
      break;
    }
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

extern "C" {

int tnh_gemm_autotune(tnh_ctx *ctx, int mode) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(mode == 0 || mode == 1 || mode == 2, TNH_ERR_ARG, "autotune mode must be 0, 1 or 2");
    if (mode == 2) {
      ctx->gemm_tiles.clear();
    } else {
      ctx->gemm_autotune = mode == 1;
    }
  });
}

int tnh_gemm_tuning_report(tnh_ctx *ctx, char *buf, int64_t len, int64_t *needed) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    static const char *kMode[3] = {"f64", "f32", "bf16x3"};
    std::string out;
    char line[256];
    for (const auto &kv : ctx->gemm_tiles) {
      const auto &k = kv.first;
      const auto &c = kv.second;
      const int mode = int(k[0]);
      const tnh::TileShape t = tnh::kTileShapes[mode][c.tile], r = tnh::kTileShapes[mode][c.rule];
      int n = std::snprintf(line, sizeof line, "%s %c%c M=%lld N=%lld K=%lld tile=%dx%d rule=%dx%d ms=",
                            kMode[mode], k[1] ? 'T' : 'N', k[2] ? 'T' : 'N', (long long)c.M,
                            (long long)c.N, (long long)c.K, t.bm, t.bn, r.bm, r.bn);
      for (int i = 0; i < tnh::kNumTiles[mode] && n > 0 && n < int(sizeof line); ++i)
        n += std::snprintf(line + n, sizeof line - n, "%s%.4f", i ? "," : "", c.ms[i]);
      out += line;
      out += '\n';
    }
    if (needed) *needed = int64_t(out.size()) + 1;
    if (buf && len > 0) {
      const size_t m = std::min(size_t(len - 1), out.size());
      std::memcpy(buf, out.data(), m);
      buf[m] = '\0';
    }
  });
}

}  // extern "C"
