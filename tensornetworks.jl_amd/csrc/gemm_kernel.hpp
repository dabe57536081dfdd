// Column-major strided-batched GEMM on CDNA4 matrix cores (gfx950).
//
//   C[M,N] = alpha * op(A)[M,K] * op(B)[K,N] + beta * C        (per batch entry)
//
// Three arithmetic modes share one LDS-tiled skeleton:
//   MODE_F64    : v_mfma_f64_16x16x4f64, FP64 in LDS, FP64 accumulate.
//   MODE_F32    : v_mfma_f32_16x16x4f32, FP32 in LDS, FP32 accumulate.
//   MODE_BF16X3 : each FP32 operand is split into hi + lo bf16 when staged to LDS;
//                 v_mfma_f32_16x16x32_bf16 accumulates hi*hi + hi*lo + lo*hi in FP32.
//
// Tiling: 256 threads = 4 waves (2 x 2); a block computes BM x BN, a wave
// (BM/2) x (BN/2) as a grid of 16x16 MFMA tiles. Operand tiles are staged in LDS
// with k contiguous per row (row = m for A, n for B), padded so that the 16 lanes
// of a 16x16 MFMA operand read 16 distinct bank groups. The k order inside an MFMA
// is permuted (lane group g owns k = 4g..4g+3, or 8g..8g+7 for bf16) so every
// operand fragment is one contiguous LDS vector read; the sum over k is unchanged.
// Global tiles are prefetched into registers one k-step ahead and LDS is double
// buffered, so each k-step needs one barrier.
//
// Blocks are remapped so that each XCD (blocks are dealt round-robin to the 8 XCDs)
// owns a contiguous range of tiles, and within it tiles walk M in groups of 8 rows
// of tiles so consecutive blocks share B columns in that XCD's L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tnh {
namespace gk {

enum { MODE_F64 = 0, MODE_F32 = 1, MODE_BF16X3 = 2 };

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned short v8us __attribute__((ext_vector_type(8)));

struct Params {
  int M, N, K;
  const void *A;
  int64_t lda, sA;
  const void *B;
  int64_t ldb, sB;
  void *C;
  int64_t ldc, sC;
  double alpha, beta;
  int tiles_m, tiles_n;
};

template <int MODE>
struct ModeTraits;
template <>
struct ModeTraits<MODE_F64> {
  using S = double;  // LDS staging type
  static constexpr int BK = 16;
  static constexpr int PAD = 2;
};
template <>
struct ModeTraits<MODE_F32> {
  using S = float;
  static constexpr int BK = 16;
  static constexpr int PAD = 4;
};
template <>
struct ModeTraits<MODE_BF16X3> {
  using S = unsigned short;  // bf16 bits, two planes (hi, lo)
  static constexpr int BK = 32;
  static constexpr int PAD = 8;
};

// Round-to-nearest-even float -> bf16 bits (finite inputs).
__device__ __forceinline__ unsigned short f2bf(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) {
  return __uint_as_float((unsigned int)h << 16);
}

// Element (r, k) of an operand tile: KC = k contiguous in global memory.
template <typename T, bool KC>
__device__ __forceinline__ T ld_elem(const T *__restrict__ base, int64_t ld, int r, int k) {
  return KC ? base[int64_t(k) + int64_t(r) * ld] : base[int64_t(r) + int64_t(k) * ld];
}

// Load the ROWS x BK tile starting at (row0, k0) into registers (zero outside).
template <typename T, bool KC, int ROWS, int BK>
__device__ __forceinline__ void load_tile(const T *__restrict__ base, int64_t ld, int row0,
                                          int nrows, int k0, int K, T (&v)[ROWS * BK / 256]) {
  constexpr int E = ROWS * BK / 256;
  const int t = threadIdx.x;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    int idx = t + 256 * e;
    int r = KC ? idx / BK : idx % ROWS;
    int k = KC ? idx % BK : idx / ROWS;
    int gr = row0 + r, gk = k0 + k;
    v[e] = (gr < nrows && gk < K) ? ld_elem<T, KC>(base, ld, gr, gk) : T(0);
  }
}

template <int MODE, typename T, bool KC, int ROWS>
__device__ __forceinline__ void store_tile(typename ModeTraits<MODE>::S *__restrict__ lds,
                                           const T (&v)[ROWS * ModeTraits<MODE>::BK / 256]) {
  constexpr int BK = ModeTraits<MODE>::BK;
  constexpr int LD = BK + ModeTraits<MODE>::PAD;
  constexpr int E = ROWS * BK / 256;
  const int t = threadIdx.x;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    int idx = t + 256 * e;
    int r = KC ? idx / BK : idx % ROWS;
    int k = KC ? idx % BK : idx / ROWS;
    if constexpr (MODE == MODE_BF16X3) {
      float x = float(v[e]);
      unsigned short hi = f2bf(x);
      unsigned short lo = f2bf(x - bf2f(hi));
      lds[r * LD + k] = hi;
      lds[ROWS * LD + r * LD + k] = lo;  // lo plane follows the hi plane
    } else {
      lds[r * LD + k] = typename ModeTraits<MODE>::S(v[e]);
    }
  }
}

template <int MODE, typename TA, typename TB, typename TC, bool TRA, bool TRB, int BM, int BN>
__global__ __launch_bounds__(256) void gemm_kernel(Params p) {
  using S = typename ModeTraits<MODE>::S;
  constexpr int BK = ModeTraits<MODE>::BK;
  constexpr int LD = BK + ModeTraits<MODE>::PAD;
  constexpr int PLANES = (MODE == MODE_BF16X3) ? 2 : 1;
  constexpr int A_ELEMS = BM * LD * PLANES;
  constexpr int B_ELEMS = BN * LD * PLANES;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int EA = BM * BK / 256, EB = BN * BK / 256;
  // A is k-contiguous in memory when transposed; B when not transposed.
  constexpr bool KCA = TRA;
  constexpr bool KCB = !TRB;

  __shared__ S lds[2 * (A_ELEMS + B_ELEMS)];

  // ---- tile scheduling: XCD-contiguous ranges, grouped along M
  const int ntiles = p.tiles_m * p.tiles_n;
  int bid = blockIdx.x;
  if ((ntiles & 7) == 0) bid = (bid & 7) * (ntiles >> 3) + (bid >> 3);
  constexpr int GROUP = 8;
  const int group_span = GROUP * p.tiles_n;
  const int gid = bid / group_span;
  const int first_m = gid * GROUP;
  const int gsz = min(p.tiles_m - first_m, GROUP);
  const int tm = first_m + (bid % group_span) % gsz;
  const int tn = (bid % group_span) / gsz;
  const int m0 = tm * BM, n0 = tn * BN;

  const int batch = blockIdx.y;
  const TA *__restrict__ A = static_cast<const TA *>(p.A) + int64_t(batch) * p.sA;
  const TB *__restrict__ B = static_cast<const TB *>(p.B) + int64_t(batch) * p.sB;
  TC *__restrict__ C = static_cast<TC *>(p.C) + int64_t(batch) * p.sC;

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;
  const int g = lane >> 4, rr = lane & 15;

  using Acc = typename std::conditional<MODE == MODE_F64, v4d, v4f>::type;
  Acc acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = Acc{0, 0, 0, 0};

  TA ra[EA];
  TB rb[EB];
  const int nk = (p.K + BK - 1) / BK;
  if (nk > 0) {
    load_tile<TA, KCA, BM, BK>(A, p.lda, m0, p.M, 0, p.K, ra);
    load_tile<TB, KCB, BN, BK>(B, p.ldb, n0, p.N, 0, p.K, rb);
  }

  for (int kt = 0; kt < nk; ++kt) {
    S *As = lds + (kt & 1) * (A_ELEMS + B_ELEMS);
    S *Bs = As + A_ELEMS;
    store_tile<MODE, TA, KCA, BM>(As, ra);
    store_tile<MODE, TB, KCB, BN>(Bs, rb);
    __syncthreads();
    if (kt + 1 < nk) {
      load_tile<TA, KCA, BM, BK>(A, p.lda, m0, p.M, (kt + 1) * BK, p.K, ra);
      load_tile<TB, KCB, BN, BK>(B, p.ldb, n0, p.N, (kt + 1) * BK, p.K, rb);
    }

    if constexpr (MODE == MODE_F32) {
      v4f af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i)
        af[i] = *reinterpret_cast<const v4f *>(&As[(wm0 + 16 * i + rr) * LD + 4 * g]);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        bf[j] = *reinterpret_cast<const v4f *>(&Bs[(wn0 + 16 * j + rr) * LD + 4 * g]);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][kk], bf[j][kk], acc[i][j],
                                                              0, 0, 0);
    } else if constexpr (MODE == MODE_F64) {
      v2d af[FM][2], bf[FN][2];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const S *q = &As[(wm0 + 16 * i + rr) * LD + 4 * g];
        af[i][0] = *reinterpret_cast<const v2d *>(q);
        af[i][1] = *reinterpret_cast<const v2d *>(q + 2);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const S *q = &Bs[(wn0 + 16 * j + rr) * LD + 4 * g];
        bf[j][0] = *reinterpret_cast<const v2d *>(q);
        bf[j][1] = *reinterpret_cast<const v2d *>(q + 2);
      }
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                af[i][kk >> 1][kk & 1], bf[j][kk >> 1][kk & 1], acc[i][j], 0, 0, 0);
    } else {  // MODE_BF16X3
      v8us ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        int o = (wm0 + 16 * i + rr) * LD + 8 * g;
        ah[i] = *reinterpret_cast<const v8us *>(&As[o]);
        al[i] = *reinterpret_cast<const v8us *>(&As[BM * LD + o]);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        int o = (wn0 + 16 * j + rr) * LD + 8 * g;
        bh[j] = *reinterpret_cast<const v8us *>(&Bs[o]);
        bl[j] = *reinterpret_cast<const v8us *>(&Bs[BN * LD + o]);
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(v8bf, al[i]), __builtin_bit_cast(v8bf, bh[j]), acc[i][j], 0,
              0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(v8bf, ah[i]), __builtin_bit_cast(v8bf, bl[j]), acc[i][j], 0,
              0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(v8bf, ah[i]), __builtin_bit_cast(v8bf, bh[j]), acc[i][j], 0,
              0, 0);
        }
    }
  }

  // ---- epilogue: accumulator register r of lane (g, rr) holds C[row][rr] of each
  // 16x16 tile with row = 4g + r for the FP32-output MFMAs and row = g + 4r for
  // v_mfma_f64_16x16x4f64 (measured on gfx950).
  const bool use_beta = p.beta != 0.0;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn0 + 16 * j + rr;
      if (col >= p.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + 16 * i + (MODE == MODE_F64 ? (g + 4 * r) : (4 * g + r));
        if (row < p.M) {
          TC *c = C + int64_t(row) + int64_t(col) * p.ldc;
          double v = p.alpha * double(acc[i][j][r]);
          if (use_beta) v += p.beta * double(*c);
          *c = TC(v);
        }
      }
    }
}

}  // namespace gk
}  // namespace tnh
