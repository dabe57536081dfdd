// Column-major strided-batched GEMM on CDNA4 matrix cores (gfx950).
//
//   C[M,N] = alpha * op(A)[M,K] * op(B)[K,N] + beta * C        (per batch entry)
//
// Three arithmetic modes share one LDS-tiled skeleton:
//   MODE_F64    : v_mfma_f64_16x16x4f64, FP64 in LDS, FP64 accumulate (BK = 16).
//   MODE_F32    : v_mfma_f32_16x16x4f32, FP32 in LDS, FP32 accumulate (BK = 32).
//   MODE_BF16P  : split-BF16x3 on operands split beforehand into hi + lo bf16 planes
//                 stored k-contiguous ([row][k], k padded to a multiple of 32): tiles are
//                 streamed with 16-byte loads and LDS stores, and
//                 v_mfma_f32_16x16x32_bf16 accumulates hi*hi + hi*lo + lo*hi in FP32.
//
// Tiling: 256 threads = 4 waves (2 x 2); a block computes BM x BN, a wave
// (BM/2) x (BN/2) as a grid of 16x16 MFMA tiles. The k order inside an MFMA is permuted
// (lane group g = lane/16 owns k = 16c + 4g + kk, kk = 0..3, or k = 8g..8g+7 for bf16);
// A and B use the same permutation, so the sum over k is unchanged.
//
// LDS layouts follow the operand's global layout, so nothing is transposed in flight:
//   k-contiguous operand  -> [row][k], row stride BK + KPAD: a fragment is one 16-byte
//                            LDS read (ds_read_b128);
//   row-contiguous operand -> [k][row], row stride ROWS + 4 (FP32) or ROWS + 16 (FP64)
//                            elements: the 16 lanes of a lane group read 16 consecutive
//                            rows, and the FP32 padding puts the 4 lane groups on disjoint
//                            bank ranges.
// (bf16 planes always use [row][k].)
// Global loads are 16-byte vectors when the host has checked alignment and divisibility
// (Params::vecA / vecB), else bounds-checked scalars. Tiles are prefetched into registers
// one k-step ahead and LDS is double buffered: one barrier per k-step.
//
// Blocks are remapped so that each XCD (blocks are dealt round-robin to the 8 XCDs)
// owns a contiguous range of tiles, and within it tiles walk M in groups of 8 rows
// of tiles so concurrently running blocks share operand panels in that XCD's L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace tnh {
namespace gk {

enum { MODE_F64 = 0, MODE_F32 = 1, MODE_BF16P = 2 };

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned short v8us __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

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
  int vecA, vecB, vecC;
  int64_t planeA, planeB;  // MODE_BF16P: element offset of the lo plane from the hi plane
};

template <int MODE>
struct ModeTraits;
template <>
struct ModeTraits<MODE_F64> {
  using S = double;
  static constexpr int BK = 16;
  static constexpr int KPAD = 2;
};
template <>
struct ModeTraits<MODE_F32> {
  using S = float;
  static constexpr int BK = 32;
  static constexpr int KPAD = 0;  // [row][k] tiles: unpadded rows, chunks permuted (kc_off)
};
template <>
struct ModeTraits<MODE_BF16P> {
  using S = unsigned short;  // bf16 bits, two planes (hi, lo)
  static constexpr int BK = 32;
  static constexpr int KPAD = 8;
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

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = v4f;
};
template <>
struct Vec16<double> {
  using type = v2d;
};

// Offset of element (r, k) in a [row][k] LDS tile of row stride LDK. FP32 rows are 32 floats
// (eight 16-byte k-chunks) with chunk c of row r at position c ^ ((r >> 1) & 7): a staging
// store (8 lanes per row, two rows per 16 lanes) and a fragment read (one chunk of 16
// consecutive rows) then both touch 64 distinct banks. The padded 36-float rows this
// replaces kept the reads conflict-free but put the two rows of a store group on
// overlapping banks (one SQ_LDS_BANK_CONFLICT cycle per LDS instruction in the stage-1 GEMM
// of the matvec). FP64 rows keep their padding.
template <int MODE, int LDK>
__device__ __forceinline__ int kc_off(int r, int k) {
  if constexpr (MODE == MODE_F32)
    return r * LDK + ((((k >> 2) ^ ((r >> 1) & 7)) << 2) | (k & 3));
  else
    return r * LDK + k;
}

// One operand tile (ROWS x BK) of type T in global memory; KCG = k-contiguous globally.
template <int MODE, typename T, bool KCG, int ROWS>
struct Tile {
  using S = typename ModeTraits<MODE>::S;
  static constexpr int BK = ModeTraits<MODE>::BK;
  static constexpr bool LDS_KC = KCG;
  static constexpr int LDK = BK + ModeTraits<MODE>::KPAD;  // [row][k] stride
  // [k][row] stride. FP32: ROWS + 4, so the lane groups g = 0..3 of a fragment read (rows
  // k = 4g + kk) start 16 banks apart and each half-wave's 32 four-byte reads hit 32 distinct
  // banks (ROWS + 16 put all four groups on the same 16 banks). FP64: 16 lanes of 8 bytes
  // already fill the banks, so the stride only keeps 16-byte alignment.
  static constexpr int LDR = ROWS + (sizeof(S) == 4 ? 4 : 16);
  static constexpr int ELEMS = LDS_KC ? ROWS * LDK : BK * LDR;
  static constexpr int VW = 16 / sizeof(T);  // elements per 16-byte vector
  static constexpr int NE = ROWS * BK / 256;  // elements per thread
  static constexpr int NV = NE / VW;          // vectors per thread
  static_assert(NE % VW == 0, "tile too small for 16-byte vectors");

  // registers hold the operand already converted to the compute type
  using RT = typename std::conditional<MODE == MODE_F64, double, float>::type;
  RT v[NE];

  // position (r, k) of element `i` of this thread in the tile
  __device__ __forceinline__ static void pos(bool vec, int i, int &r, int &k) {
    const int t = threadIdx.x;
    if (vec) {
      const int vi = t + 256 * (i / VW), w = i % VW;
      if (KCG) {
        k = (vi % (BK / VW)) * VW + w;
        r = vi / (BK / VW);
      } else {
        r = (vi % (ROWS / VW)) * VW + w;
        k = vi / (ROWS / VW);
      }
    } else {
      const int idx = t + 256 * i;
      r = KCG ? idx / BK : idx % ROWS;
      k = KCG ? idx % BK : idx / ROWS;
    }
  }

  __device__ __forceinline__ void load(const T *__restrict__ base, int64_t ld, int row0, int nrows,
                                       int k0, int K, bool vec) {
    if (vec) {
#pragma unroll
      for (int e = 0; e < NV; ++e) {
        int r, k;
        pos(true, e * VW, r, k);
        const int gr = row0 + r, gk = k0 + k;
        const bool in = gr < nrows && gk < K;  // whole vector in or out (host-checked)
        // branch-free: out-of-range vectors read the (always valid) base and are zeroed
        // by a select, so the load's wait is deferred to the LDS store that consumes it
        const int64_t off = KCG ? (int64_t(gk) + int64_t(gr) * ld) : (int64_t(gr) + int64_t(gk) * ld);
        const auto q = *reinterpret_cast<const typename Vec16<T>::type *>(base + (in ? off : 0));
#pragma unroll
        for (int w = 0; w < VW; ++w) v[e * VW + w] = in ? RT(q[w]) : RT(0);
      }
    } else {
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        int r, k;
        pos(false, e, r, k);
        const int gr = row0 + r, gk = k0 + k;
        const bool in = gr < nrows && gk < K;
        const int64_t off = KCG ? (int64_t(gk) + int64_t(gr) * ld) : (int64_t(gr) + int64_t(gk) * ld);
        const T x = base[in ? off : 0];
        v[e] = in ? RT(x) : RT(0);
      }
    }
  }

  __device__ __forceinline__ void store(S *__restrict__ lds, bool vec) const {
    if (vec) {
      // a vector's elements are contiguous in the LDS layout too
#pragma unroll
      for (int e = 0; e < NV; ++e) {
        int r, k;
        pos(true, e * VW, r, k);
        S *d = LDS_KC ? &lds[kc_off<MODE, LDK>(r, k)] : &lds[k * LDR + r];
        if constexpr (std::is_same<S, float>::value && VW == 4) {
          *reinterpret_cast<v4f *>(d) = v4f{v[e * VW], v[e * VW + 1], v[e * VW + 2], v[e * VW + 3]};
        } else if constexpr (std::is_same<S, float>::value) {  // two FP64 inputs
          *reinterpret_cast<v2f *>(d) = v2f{v[e * VW], v[e * VW + 1]};
        } else {  // FP64 staging
          *reinterpret_cast<v2d *>(d) = v2d{v[e * VW], v[e * VW + 1]};
        }
      }
    } else {
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        int r, k;
        pos(false, e, r, k);
        (LDS_KC ? lds[kc_off<MODE, LDK>(r, k)] : lds[k * LDR + r]) = S(v[e]);
      }
    }
  }
};

// One pre-split operand tile (ROWS x 32 bf16, hi and lo planes) from [row][k] planes with
// row stride ld (k padded, so every 16-byte vector lies inside a row); rows beyond nrows
// load the tile base and are zeroed (branch-free).
// LDS layout: unpadded 64-byte rows whose four 16-byte k-chunks are permuted per row,
// chunk c of row r at position c ^ swz(r), swz(r) = ((r >> 1) ^ (r >> 3)) & 3. A fragment
// read (16 rows, one chunk) and a staging store (4 chunks of consecutive rows) then touch
// distinct banks whether the LDS serves 8 or 16 lanes of a 16-byte access per cycle; the
// padded 80-byte rows this replaces stalled ~28% of the kernel on bank conflicts
// (SQ_LDS_BANK_CONFLICT).
__device__ __forceinline__ int bf16_swz(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }

template <int ROWS>
struct TileP {
  static constexpr int BK = 32;
  static constexpr int LDK = BK;  // [row][k] stride (64 bytes), chunks permuted by bf16_swz
  static constexpr int ELEMS = 2 * ROWS * LDK;
  static constexpr int NV = ROWS * BK / 8 / 256;  // 16-byte vectors per plane per thread
  static_assert(NV >= 1, "tile too small");
  v8us hi[NV], lo[NV];

  __device__ __forceinline__ static void pos(int e, int &r, int &k) {
    const int vi = threadIdx.x + 256 * e;
    k = (vi % (BK / 8)) * 8;
    r = vi / (BK / 8);
  }
  __device__ __forceinline__ void load(const unsigned short *__restrict__ base, int64_t ld,
                                       int64_t plane, int row0, int nrows, int k0) {
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      int r, k;
      pos(e, r, k);
      const int gr = row0 + r;
      const bool in = gr < nrows;
      const int64_t off = in ? (int64_t(k0 + k) + int64_t(gr) * ld) : 0;
      const v8us h = *reinterpret_cast<const v8us *>(base + off);
      const v8us l = *reinterpret_cast<const v8us *>(base + plane + off);
      const v8us z = {0, 0, 0, 0, 0, 0, 0, 0};
      hi[e] = in ? h : z;
      lo[e] = in ? l : z;
    }
  }
  __device__ __forceinline__ void store(unsigned short *__restrict__ lds) const {
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      int r, k;
      pos(e, r, k);
      const int o = r * LDK + ((((k >> 3) ^ bf16_swz(r)) & 3) << 3);
      *reinterpret_cast<v8us *>(&lds[o]) = hi[e];
      *reinterpret_cast<v8us *>(&lds[ROWS * LDK + o]) = lo[e];
    }
  }
};

// FP32/FP64 fragment: 4 k-values (k = 16c + 4g + kk) of one row for lane (g, rr).
template <int MODE, bool LDS_KC, int LDK, int LDR>
__device__ __forceinline__ void frag4(const typename ModeTraits<MODE>::S *__restrict__ lds, int row,
                                      int c, int g, typename ModeTraits<MODE>::S (&f)[4]) {
  using S = typename ModeTraits<MODE>::S;
  if constexpr (LDS_KC) {
    const S *q = &lds[kc_off<MODE, LDK>(row, 16 * c + 4 * g)];
    if constexpr (std::is_same<S, float>::value) {
      const v4f x = *reinterpret_cast<const v4f *>(q);
      f[0] = x[0];
      f[1] = x[1];
      f[2] = x[2];
      f[3] = x[3];
    } else {
      const v2d x0 = *reinterpret_cast<const v2d *>(q);
      const v2d x1 = *reinterpret_cast<const v2d *>(q + 2);
      f[0] = x0[0];
      f[1] = x0[1];
      f[2] = x1[0];
      f[3] = x1[1];
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) f[kk] = lds[(16 * c + 4 * g + kk) * LDR + row];
  }
}

template <int MODE, typename TA, typename TB, typename TC, bool TRA, bool TRB, int BM, int BN, bool VEC>
__global__ __launch_bounds__(256) void gemm_kernel(Params p) {
  using S = typename ModeTraits<MODE>::S;
  constexpr int BK = ModeTraits<MODE>::BK;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  // A is k-contiguous in memory when transposed; B when not transposed.
  constexpr bool PRE = MODE == MODE_BF16P;
  using TileA = typename std::conditional<PRE, TileP<BM>, Tile<MODE, TA, TRA, BM>>::type;
  using TileB = typename std::conditional<PRE, TileP<BN>, Tile<MODE, TB, !TRB, BN>>::type;
  constexpr int A_ELEMS = TileA::ELEMS, B_ELEMS = TileB::ELEMS;

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
  constexpr bool va = VEC, vb = VEC;  // 16-byte loads for both operands (host-checked)

  using Acc = typename std::conditional<MODE == MODE_F64, v4d, v4f>::type;
  Acc acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = Acc{0, 0, 0, 0};

  TileA ta;
  TileB tb;
  const int nk = (p.K + BK - 1) / BK;
  if (nk > 0) {
    if constexpr (PRE) {
      ta.load(A, p.lda, p.planeA, m0, p.M, 0);
      tb.load(B, p.ldb, p.planeB, n0, p.N, 0);
    } else {
      ta.load(A, p.lda, m0, p.M, 0, p.K, va);
      tb.load(B, p.ldb, n0, p.N, 0, p.K, vb);
    }
  }

  for (int kt = 0; kt < nk; ++kt) {
    S *As = lds + (kt & 1) * (A_ELEMS + B_ELEMS);
    S *Bs = As + A_ELEMS;
    if constexpr (PRE) {
      ta.store(As);
      tb.store(Bs);
    } else {
      ta.store(As, va);
      tb.store(Bs, vb);
    }
    __syncthreads();
    if (kt + 1 < nk) {
      if constexpr (PRE) {
        ta.load(A, p.lda, p.planeA, m0, p.M, (kt + 1) * BK);
        tb.load(B, p.ldb, p.planeB, n0, p.N, (kt + 1) * BK);
      } else {
        ta.load(A, p.lda, m0, p.M, (kt + 1) * BK, p.K, va);
        tb.load(B, p.ldb, n0, p.N, (kt + 1) * BK, p.K, vb);
      }
    }

    // FP32-output modes run the matrix-core section at raised wave priority: the CU's
    // other workgroup then issues its loads and LDS stores around it instead of delaying
    // the MFMA stream (+1.9% for FP32; FP64 64 x 64 tiles lost 11% with it).
    if constexpr (MODE != MODE_F64) __builtin_amdgcn_s_setprio(1);
    if constexpr (MODE == MODE_BF16P) {
      constexpr int LD = TileA::LDK;
      v8us ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int row = wm0 + 16 * i + rr;
        const int o = row * LD + (((g ^ bf16_swz(row)) & 3) << 3);
        ah[i] = *reinterpret_cast<const v8us *>(&As[o]);
        al[i] = *reinterpret_cast<const v8us *>(&As[BM * LD + o]);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int row = wn0 + 16 * j + rr;
        const int o = row * LD + (((g ^ bf16_swz(row)) & 3) << 3);
        bh[j] = *reinterpret_cast<const v8us *>(&Bs[o]);
        bl[j] = *reinterpret_cast<const v8us *>(&Bs[BN * LD + o]);
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(v8bf, al[i]), __builtin_bit_cast(v8bf, bh[j]), acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(v8bf, ah[i]), __builtin_bit_cast(v8bf, bl[j]), acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              __builtin_bit_cast(v8bf, ah[i]), __builtin_bit_cast(v8bf, bh[j]), acc[i][j], 0, 0, 0);
        }
    } else {
#pragma unroll
      for (int c = 0; c < BK / 16; ++c) {
        S af[FM][4], bf[FN][4];
#pragma unroll
        for (int i = 0; i < FM; ++i)
          frag4<MODE, TileA::LDS_KC, TileA::LDK, TileA::LDR>(As, wm0 + 16 * i + rr, c, g, af[i]);
#pragma unroll
        for (int j = 0; j < FN; ++j)
          frag4<MODE, TileB::LDS_KC, TileB::LDK, TileB::LDR>(Bs, wn0 + 16 * j + rr, c, g, bf[j]);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
              if constexpr (MODE == MODE_F32)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][kk], bf[j][kk], acc[i][j], 0, 0, 0);
              else
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][kk], bf[j][kk], acc[i][j], 0, 0, 0);
            }
      }
    }
    if constexpr (MODE != MODE_F64) __builtin_amdgcn_s_setprio(0);
  }

  // ---- epilogue: accumulator register r of lane (g, rr) holds C[row][rr] of each
  // 16x16 tile with row = 4g + r for the FP32-output MFMAs and row = g + 4r for
  // v_mfma_f64_16x16x4f64 (measured on gfx950).
  const bool use_beta = p.beta != 0.0;
  const bool vc = p.vecC != 0;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn0 + 16 * j + rr;
      if (col >= p.N) continue;
      TC *ccol = C + int64_t(col) * p.ldc;
      if constexpr (MODE != MODE_F64) {
        const int row0 = m0 + wm0 + 16 * i + 4 * g;
        if (vc && row0 + 3 < p.M) {
          // four consecutive rows of one column: one 16- or 32-byte store
          TC out[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            double vv = p.alpha * double(acc[i][j][r]);
            if (use_beta) vv += p.beta * double(ccol[row0 + r]);
            out[r] = TC(vv);
          }
          if constexpr (std::is_same<TC, float>::value) {
            *reinterpret_cast<v4f *>(ccol + row0) = v4f{out[0], out[1], out[2], out[3]};
          } else {
            *reinterpret_cast<v2d *>(ccol + row0) = v2d{out[0], out[1]};
            *reinterpret_cast<v2d *>(ccol + row0 + 2) = v2d{out[2], out[3]};
          }
          continue;
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + 16 * i + (MODE == MODE_F64 ? (g + 4 * r) : (4 * g + r));
        if (row < p.M) {
          double vv = p.alpha * double(acc[i][j][r]);
          if (use_beta) vv += p.beta * double(ccol[row]);
          ccol[row] = TC(vv);
        }
      }
    }
}

}  // namespace gk
}  // namespace tnh
