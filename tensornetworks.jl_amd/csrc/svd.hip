// Truncated SVD of theta on the GPU (replaces svd(x, idx; ...) of src/tensors.jl:168-227 and
// the two-site split of replacesites!, src/structures/mps/gmps.jl:199-267).
//
// Primary path: FP64 Gram matrix G = M M^H (or M^H M, the smaller side) on the matrix
// cores, then a two-sided block-Jacobi eigensolver on G. Fallback (non-convergence,
// non-finite values, or on request) -- mirroring gesdd -> gesvd at src/tensors.jl:190-195:
// one-sided block-Jacobi (Hestenes) directly on M, which keeps relative accuracy of
// small singular values.
//
// Block Jacobi: columns are cut into blocks of JB = 16; a round-robin (circle)
// schedule pairs all blocks per round, and every pair (32 columns) is handled by one
// workgroup that diagonalises its 32 x 32 sub-problem in LDS with a parallel cyclic
// Jacobi: 16 disjoint rotations per step, applied in one in-place pass over the 2 x 2
// blocks they couple (exact zeroing of the rotated entry), at most `max_inner` inner
// sweeps unless the pair is the whole matrix. The accumulated 32 x 32 rotation of each
// pair is then applied to the full columns (and rows, two-sided) of G / M and to the
// accumulated unitary V. Pairs already orthogonal to tolerance are skipped.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "dc_eig.hpp"
#include "jacobi_lds.hpp"
#include "tnh_internal.hpp"

namespace tnh {

namespace {

using namespace jac;
constexpr int kGramRows = 64;
constexpr int kGramPer = JP * JP / 256;  // Gram entries per thread

// One-sided mode: partial Grams of the column pairs, split over row chunks.
// partial[((pair * nsplit + split) * 2 + plane) * JP * JP + (i + JP * j)] = sum_r conj(x_ri) x_rj
__global__ __launch_bounds__(256) void jac_gram_kernel(const double *__restrict__ X, int rows,
                                                       int64_t imX, int cols, bool cplx,
                                                       const int *__restrict__ pairs, int nsplit,
                                                       double *__restrict__ partial) {
  const int pair = blockIdx.x, split = blockIdx.y;
  __shared__ double tr[kGramRows][JP + 1], ti[kGramRows][JP + 1];
  __shared__ int colmap[JP];
  const int t = threadIdx.x;
  if (t < JP) colmap[t] = pair_col(pairs + 2 * pair, t, cols);
  __syncthreads();
  int chunk = (rows + nsplit - 1) / nsplit;
  chunk = (chunk + kGramRows - 1) / kGramRows * kGramRows;
  const int r0 = split * chunk, r1 = min(rows, r0 + chunk);
  double ar[kGramPer], ai[kGramPer];
#pragma unroll
  for (int q = 0; q < kGramPer; ++q) ar[q] = ai[q] = 0.0;
  for (int rb = r0; rb < r1; rb += kGramRows) {
    for (int e = t; e < kGramRows * JP; e += 256) {
      const int rr = e % kGramRows, k = e / kGramRows;
      const int r = rb + rr, c = colmap[k];
      double vr = 0, vi = 0;
      if (r < r1 && c >= 0) {
        vr = X[r + int64_t(c) * rows];
        if (cplx) vi = X[imX + r + int64_t(c) * rows];
      }
      tr[rr][k] = vr;
      ti[rr][k] = vi;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kGramPer; ++q) {
      const int e = t + 256 * q, i = e % JP, j = e / JP;
      double sr = 0, si = 0;
      for (int rr = 0; rr < kGramRows; ++rr) {
        const double xr = tr[rr][i], xi = ti[rr][i], yr = tr[rr][j], yi = ti[rr][j];
        sr += xr * yr + xi * yi;
        si += xr * yi - xi * yr;
      }
      ar[q] += sr;
      ai[q] += si;
    }
    __syncthreads();
  }
  double *out = partial + int64_t(pair * nsplit + split) * 2 * JP * JP;
#pragma unroll
  for (int q = 0; q < kGramPer; ++q) {
    const int e = t + 256 * q;
    out[e] = ar[q];
    out[JP * JP + e] = ai[q];
  }
}


// Diagonalise the 2B x 2B Hermitian sub-problem of each pair. Two-sided mode reads
// G[P,P] from the global matrix; one-sided mode sums the partial Grams. Writes the
// accumulated unitary Q (row-major, planar) and, in two-sided mode, the diagonalised
// block D (column-major, planar), an
// `active` flag (0 = pair already converged, nothing to apply) and the pair's
// off-diagonal measure into conv (atomic max over the sweep).
template <bool CPLX, int B, int NT>
__global__ __launch_bounds__(NT) void jac_eig_kernel(const double *__restrict__ G, int64_t imG,
                                                      int n, const double *__restrict__ partial,
                                                      int nsplit, bool two_sided,
                                                      const int *__restrict__ pairs, int cols,
                                                      double tol, double abstol, int max_inner,
                                                      bool cross, double *__restrict__ Qout,
                                                      double *__restrict__ Dout,
                                                      int *__restrict__ active,
                                                      unsigned long long *__restrict__ conv) {
  constexpr int JP = 2 * B;
  const int pair = blockIdx.x, t = threadIdx.x;
  __shared__ double gr[JP][JP + 1], gi[CPLX ? JP : 1][JP + 1], qr[JP][JP + 1], qi[CPLX ? JP : 1][JP + 1];
  __shared__ int colmap[JP];
  if (t < JP) colmap[t] = pair_col<B>(pairs + 2 * pair, t, cols);
  __syncthreads();
  for (int e = t; e < JP * JP; e += NT) {
    const int i = e % JP, j = e / JP;
    double vr = 0, vi = 0;
    if (two_sided) {
      const int ci = colmap[i], cj = colmap[j];
      if (ci >= 0 && cj >= 0) {
        vr = G[ci + int64_t(cj) * n];
        if (CPLX) vi = G[imG + ci + int64_t(cj) * n];
      }
    } else {
      for (int s = 0; s < nsplit; ++s) {
        const double *pp = partial + int64_t(pair * nsplit + s) * 2 * JP * JP;
        vr += pp[e];
        if (CPLX) vi += pp[JP * JP + e];
      }
    }
    gr[i][j] = vr;
    qr[i][j] = (i == j) ? 1.0 : 0.0;
    if (CPLX) {
      gi[i][j] = (i == j) ? 0.0 : vi;
      qi[i][j] = 0.0;
    }
  }
  __syncthreads();
  // off-diagonal measure: max |g_ij| / sqrt(|g_ii g_jj|) over i != j above the noise floor
  double m = 0;
  for (int e = t; e < JP * JP; e += NT) {
    const int i = e % JP, j = e / JP;
    if (i == j) continue;
    const double b = CPLX ? hypot(gr[i][j], gi[i][j]) : fabs(gr[i][j]);
    if (b <= abstol) continue;
    const double s = sqrt(fabs(gr[i][i] * gr[j][j]));
    m = fmax(m, s > 0 ? b / s : 1.0);
  }
  m = block_max<NT>(m);
  if (t == 0) atomicMax(conv, (unsigned long long)__double_as_longlong(m));
  if (m <= tol) {
    if (t == 0) active[pair] = 0;
    return;
  }
  (void)jacobi_lds<CPLX, B, NT>(gr, gi, qr, qi, tol, abstol, max_inner, cross);
  double *Q = Qout + int64_t(pair) * 2 * JP * JP;  // row-major: Q[i][j] at i * JP + j
  for (int e = t; e < JP * JP; e += NT) {
    const int i = e / JP, j = e % JP;
    Q[e] = qr[i][j];
    Q[JP * JP + e] = CPLX ? qi[i][j] : 0.0;
  }
  if (two_sided) {
    double *D = Dout + int64_t(pair) * 2 * JP * JP;
    for (int e = t; e < JP * JP; e += NT) {
      const int i = e % JP, j = e / JP;
      D[e] = gr[i][j];
      D[JP * JP + e] = CPLX ? gi[i][j] : 0.0;
    }
  }
  if (t == 0) active[pair] = 1;
}

// Y[:, P] <- Y[:, P] Q_P for Y in {X (rows_x rows), V (rows_v rows)}.
template <bool CPLX, int B>
__global__ __launch_bounds__(256) void jac_colupd_kernel(double *__restrict__ X, int rows_x,
                                                         int64_t imX, double *__restrict__ V,
                                                         int rows_v, int64_t imV,
                                                         const int *__restrict__ pairs, int cols,
                                                         const double *__restrict__ Qin,
                                                         const int *__restrict__ active) {
  constexpr int JP = 2 * B;
  const int pair = blockIdx.y;
  if (!active[pair]) return;
  __shared__ double qr[JP * JP], qi[CPLX ? JP * JP : 1];
  __shared__ int colmap[JP];
  const int t = threadIdx.x;
  const double *Q = Qin + int64_t(pair) * 2 * JP * JP;
  for (int e = t; e < JP * JP; e += 256) {
    qr[e] = Q[e];
    if (CPLX) qi[e] = Q[JP * JP + e];
  }
  if (t < JP) colmap[t] = pair_col<B>(pairs + 2 * pair, t, cols);
  __syncthreads();
  const int r = blockIdx.x * 256 + t;
  for (int which = 0; which < 2; ++which) {
    double *Y = which ? V : X;
    const int rows = which ? rows_v : rows_x;
    const int64_t im = which ? imV : imX;
    if (!Y || r >= rows) continue;
    double yr[JP], yi[CPLX ? JP : 1];
#pragma unroll
    for (int k = 0; k < JP; ++k) {
      const int c = colmap[k];
      yr[k] = c >= 0 ? Y[r + int64_t(c) * rows] : 0.0;
      if (CPLX) yi[k] = c >= 0 ? Y[im + r + int64_t(c) * rows] : 0.0;
    }
    for (int j = 0; j < JP; ++j) {
      const int c = colmap[j];
      if (c < 0) continue;
      double sr = 0, si = 0;
#pragma unroll
      for (int k = 0; k < JP; ++k) {
        const double a = qr[k * JP + j];
        sr += yr[k] * a;
        if (CPLX) {
          const double b = qi[k * JP + j];
          sr -= yi[k] * b;
          si += yr[k] * b + yi[k] * a;
        }
      }
      Y[r + int64_t(c) * rows] = sr;
      if (CPLX) Y[im + r + int64_t(c) * rows] = si;
    }
  }
}

typedef double jv4d __attribute__((ext_vector_type(4)));

// Real two-sided mode on the matrix cores, exploiting the symmetry of G: one workgroup per
// tile of the upper triangle of pair tiles (P <= K, blockIdx.x = K (K + 1) / 2 + P) forms
// Q_P^T G[P, K] Q_K with v_mfma_f64_16x16x4f64 (each wave one quadrant of the 2B x 2B tile,
// the intermediate product staged through LDS) and writes it and its transpose G[K, P], so G
// stays bitwise symmetric and half the tiles are read. The Q of an inactive pair is the
// identity (exact in FP64). v_mfma_f64_16x16x4f64 lane (g = lane / 16, rr = lane % 16)
// supplies A[rr][k_g] and B[k_g][rr]; accumulator register r holds C[g + 4 r][rr].
template <int B>
__global__ __launch_bounds__(256) void jac_twoside_sym_kernel(double *__restrict__ G, int n,
                                                              const int *__restrict__ pairs,
                                                              const double *__restrict__ Qin,
                                                              const double *__restrict__ Din,
                                                              const int *__restrict__ active) {
  constexpr int JP = 2 * B, H = JP / 2, FT = H / 16, KS = JP / 4;  // quadrant, MFMA tiles, k steps
  const int idx = blockIdx.x, t = threadIdx.x;
  int K = int((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while ((K + 1) * (K + 2) / 2 <= idx) ++K;
  while (K * (K + 1) / 2 > idx) --K;
  const int P = idx - K * (K + 1) / 2;
  const bool aP = active[P] != 0, aK = active[K] != 0;
  if (!aP && !aK) return;
  __shared__ double T[JP][JP + 1];
  __shared__ int rowmap[JP], colmap[JP];
  if (t < JP) {
    rowmap[t] = pair_col<B>(pairs + 2 * P, t, n);
    colmap[t] = pair_col<B>(pairs + 2 * K, t, n);
  }
  __syncthreads();
  if (P == K) {  // aP holds here
    const double *D = Din + int64_t(P) * 2 * JP * JP;
    for (int e = t; e < JP * JP; e += 256) {
      const int i = e % JP, j = e / JP;
      const int r = rowmap[i], c = colmap[j];
      if (r >= 0 && c >= 0) G[r + int64_t(c) * n] = D[e];
    }
    return;
  }
  // row-major rotations: Q[k][j] at k * JP + j
  const double *qp = Qin + int64_t(P) * 2 * JP * JP;
  const double *qk = Qin + int64_t(K) * 2 * JP * JP;
  for (int e = t; e < JP * JP; e += 256) {
    const int i = e % JP, j = e / JP;
    const int r = rowmap[i], c = colmap[j];
    T[i][j] = (r >= 0 && c >= 0) ? G[r + int64_t(c) * n] : 0.0;
  }
  const int lane = t & 63, w = t >> 6, g = lane >> 4, rr = lane & 15;
  const int i0 = H * (w & 1), j0 = H * (w >> 1);
  // B operands of T Q_K: Q_K[4 s + g][j0 + 16 y + rr] (the identity when K is inactive)
  double fb[KS][FT];
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int y = 0; y < FT; ++y) {
      const int k = 4 * s + g, j = j0 + 16 * y + rr;
      fb[s][y] = aK ? qk[k * JP + j] : (k == j ? 1.0 : 0.0);
    }
  __syncthreads();
  jv4d acc[FT][FT];
#pragma unroll
  for (int x = 0; x < FT; ++x)
#pragma unroll
    for (int y = 0; y < FT; ++y) acc[x][y] = jv4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < KS; ++s) {  // U = T Q_K
    double fa[FT];
#pragma unroll
    for (int x = 0; x < FT; ++x) fa[x] = T[i0 + 16 * x + rr][4 * s + g];
#pragma unroll
    for (int x = 0; x < FT; ++x)
#pragma unroll
      for (int y = 0; y < FT; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[x], fb[s][y], acc[x][y], 0, 0, 0);
  }
  // A operands of Q_P^T U: Q_P[4 s + g][i0 + 16 x + rr]
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int x = 0; x < FT; ++x) {
      const int k = 4 * s + g, i = i0 + 16 * x + rr;
      fb[s][x] = aP ? qp[k * JP + i] : (k == i ? 1.0 : 0.0);
    }
  __syncthreads();
#pragma unroll
  for (int x = 0; x < FT; ++x)
#pragma unroll
    for (int y = 0; y < FT; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        T[i0 + 16 * x + g + 4 * r][j0 + 16 * y + rr] = acc[x][y][r];
        acc[x][y][r] = 0.0;
      }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < KS; ++s) {  // Q_P^T U
    double fu[FT];
#pragma unroll
    for (int y = 0; y < FT; ++y) fu[y] = T[4 * s + g][j0 + 16 * y + rr];
#pragma unroll
    for (int x = 0; x < FT; ++x)
#pragma unroll
      for (int y = 0; y < FT; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[s][x], fu[y], acc[x][y], 0, 0, 0);
  }
  __syncthreads();
#pragma unroll
  for (int x = 0; x < FT; ++x)
#pragma unroll
    for (int y = 0; y < FT; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) T[i0 + 16 * x + g + 4 * r][j0 + 16 * y + rr] = acc[x][y][r];
  __syncthreads();
  for (int e = t; e < JP * JP; e += 256) {
    {  // G[P, K], rows fastest
      const int i = e % JP, j = e / JP;
      const int r = rowmap[i], c = colmap[j];
      if (r >= 0 && c >= 0) G[r + int64_t(c) * n] = T[i][j];
    }
    {  // G[K, P] = G[P, K]^T, its rows (the columns of the tile) fastest
      const int j = e % JP, i = e / JP;
      const int r = rowmap[i], c = colmap[j];
      if (r >= 0 && c >= 0) G[c + int64_t(r) * n] = T[i][j];
    }
  }
}

// Real V[:, P] <- V[:, P] Q_P on the matrix cores, as V[:, P]^T <- Q_P^T V[:, P]^T: the
// v_mfma_f64_16x16x4f64 B operand is then 16 consecutive rows of one column of V, read
// straight from memory, and each accumulator register of a lane group holds 16 consecutive
// rows of one output column. One workgroup per 128 rows and pair, wave w on rows 32 w..;
// Q_P staged in LDS. Workgroups touch disjoint (rows, columns) blocks, so V updates in place.
template <int B>
__global__ __launch_bounds__(256) void jac_colupd_mfma_kernel(double *__restrict__ Y, int rows,
                                                              const int *__restrict__ pairs, int cols,
                                                              const double *__restrict__ Qin,
                                                              const int *__restrict__ active) {
  constexpr int JP = 2 * B, FJ = JP / 16, KS = JP / 4;
  const int pair = blockIdx.y;
  if (!active[pair]) return;
  __shared__ double Qs[JP][JP + 1];
  __shared__ int colmap[JP];
  const int t = threadIdx.x;
  const double *Q = Qin + int64_t(pair) * 2 * JP * JP;  // row-major Q[k][j]
  for (int e = t; e < JP * JP; e += 256) Qs[e / JP][e % JP] = Q[e];
  if (t < JP) colmap[t] = pair_col<B>(pairs + 2 * pair, t, cols);
  __syncthreads();
  const int lane = t & 63, w = t >> 6, g = lane >> 4, rr = lane & 15;
  const int r0 = blockIdx.x * 128 + 32 * w;
  double vb[KS][2];
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int r = r0 + 16 * y + rr, c = colmap[4 * s + g];
      vb[s][y] = (r < rows && c >= 0) ? Y[r + int64_t(c) * rows] : 0.0;
    }
  jv4d acc[FJ][2];
#pragma unroll
  for (int x = 0; x < FJ; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = jv4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
#pragma unroll
    for (int x = 0; x < FJ; ++x) {
      const double fa = Qs[4 * s + g][16 * x + rr];  // Q^T[j][k] = Q[k][j]
#pragma unroll
      for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa, vb[s][y], acc[x][y], 0, 0, 0);
    }
  }
  // acc[x][y][r] = new V[r0 + 16 y + rr][column 16 x + g + 4 r of the pair]
#pragma unroll
  for (int x = 0; x < FJ; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = r0 + 16 * y + rr, c = colmap[16 * x + g + 4 * r];
        if (i < rows && c >= 0) Y[i + int64_t(c) * rows] = acc[x][y][r];
      }
}

// Complex (planar) counterparts of the two MFMA kernels above, for Hermitian G: each complex
// product is four real MFMA products (a subtraction passes the negated operand). The Q
// operands are read from L2 inside the k loop; G tiles of complex problems are few (their
// p rarely exceeds a few thousand), so these launches are latency- rather than
// throughput-bound. The mirrored tile is the conjugate transpose.
template <int B>
__global__ __launch_bounds__(256) void jac_twoside_herm_kernel(double *__restrict__ G, int n, int64_t imG,
                                                               const int *__restrict__ pairs,
                                                               const double *__restrict__ Qin,
                                                               const double *__restrict__ Din,
                                                               const int *__restrict__ active) {
  constexpr int JP = 2 * B, H = JP / 2, FT = H / 16, KS = JP / 4, JJ = JP * JP;
  const int idx = blockIdx.x, t = threadIdx.x;
  int K = int((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while ((K + 1) * (K + 2) / 2 <= idx) ++K;
  while (K * (K + 1) / 2 > idx) --K;
  const int P = idx - K * (K + 1) / 2;
  const bool aP = active[P] != 0, aK = active[K] != 0;
  if (!aP && !aK) return;
  __shared__ double Tr[JP][JP + 1], Ti[JP][JP + 1];
  __shared__ int rowmap[JP], colmap[JP];
  if (t < JP) {
    rowmap[t] = pair_col<B>(pairs + 2 * P, t, n);
    colmap[t] = pair_col<B>(pairs + 2 * K, t, n);
  }
  __syncthreads();
  if (P == K) {  // aP holds here
    const double *D = Din + int64_t(P) * 2 * JJ;
    for (int e = t; e < JJ; e += 256) {
      const int i = e % JP, j = e / JP;
      const int r = rowmap[i], c = colmap[j];
      if (r >= 0 && c >= 0) {
        G[r + int64_t(c) * n] = D[e];
        G[imG + r + int64_t(c) * n] = D[JJ + e];
      }
    }
    return;
  }
  const double *qp = Qin + int64_t(P) * 2 * JJ;  // row-major re plane, then im plane
  const double *qk = Qin + int64_t(K) * 2 * JJ;
  for (int e = t; e < JJ; e += 256) {
    const int i = e % JP, j = e / JP;
    const int r = rowmap[i], c = colmap[j];
    const bool ok = r >= 0 && c >= 0;
    Tr[i][j] = ok ? G[r + int64_t(c) * n] : 0.0;
    Ti[i][j] = ok ? G[imG + r + int64_t(c) * n] : 0.0;
  }
  __syncthreads();
  const int lane = t & 63, w = t >> 6, g = lane >> 4, rr = lane & 15;
  const int i0 = H * (w & 1), j0 = H * (w >> 1);
  auto qre = [&](const double *q, bool act, int k, int j) { return act ? q[k * JP + j] : (k == j ? 1.0 : 0.0); };
  auto qim = [&](const double *q, bool act, int k, int j) { return act ? q[JJ + k * JP + j] : 0.0; };
  jv4d ar[FT][FT], ai[FT][FT];
#pragma unroll
  for (int x = 0; x < FT; ++x)
#pragma unroll
    for (int y = 0; y < FT; ++y) ar[x][y] = ai[x][y] = jv4d{0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < KS; ++s) {  // U = T Q_K
    const int k = 4 * s + g;
    double fr[FT], fi[FT], br[FT], bi[FT];
#pragma unroll
    for (int x = 0; x < FT; ++x) {
      fr[x] = Tr[i0 + 16 * x + rr][k];
      fi[x] = Ti[i0 + 16 * x + rr][k];
      br[x] = qre(qk, aK, k, j0 + 16 * x + rr);
      bi[x] = qim(qk, aK, k, j0 + 16 * x + rr);
    }
#pragma unroll
    for (int x = 0; x < FT; ++x)
#pragma unroll
      for (int y = 0; y < FT; ++y) {
        ar[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fr[x], br[y], ar[x][y], 0, 0, 0);
        ar[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-fi[x], bi[y], ar[x][y], 0, 0, 0);
        ai[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fr[x], bi[y], ai[x][y], 0, 0, 0);
        ai[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fi[x], br[y], ai[x][y], 0, 0, 0);
      }
  }
  __syncthreads();
#pragma unroll
  for (int x = 0; x < FT; ++x)
#pragma unroll
    for (int y = 0; y < FT; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Tr[i0 + 16 * x + g + 4 * r][j0 + 16 * y + rr] = ar[x][y][r];
        Ti[i0 + 16 * x + g + 4 * r][j0 + 16 * y + rr] = ai[x][y][r];
        ar[x][y][r] = 0.0;
        ai[x][y][r] = 0.0;
      }
  __syncthreads();
  for (int s = 0; s < KS; ++s) {  // Q_P^H U
    const int k = 4 * s + g;
    double pr[FT], pi[FT], ur[FT], ui[FT];
#pragma unroll
    for (int x = 0; x < FT; ++x) {
      pr[x] = qre(qp, aP, k, i0 + 16 * x + rr);
      pi[x] = qim(qp, aP, k, i0 + 16 * x + rr);
      ur[x] = Tr[k][j0 + 16 * x + rr];
      ui[x] = Ti[k][j0 + 16 * x + rr];
    }
#pragma unroll
    for (int x = 0; x < FT; ++x)
#pragma unroll
      for (int y = 0; y < FT; ++y) {  // conj(Q_P[k][i]) U[k][j]
        ar[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[x], ur[y], ar[x][y], 0, 0, 0);
        ar[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(pi[x], ui[y], ar[x][y], 0, 0, 0);
        ai[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[x], ui[y], ai[x][y], 0, 0, 0);
        ai[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-pi[x], ur[y], ai[x][y], 0, 0, 0);
      }
  }
  __syncthreads();
#pragma unroll
  for (int x = 0; x < FT; ++x)
#pragma unroll
    for (int y = 0; y < FT; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Tr[i0 + 16 * x + g + 4 * r][j0 + 16 * y + rr] = ar[x][y][r];
        Ti[i0 + 16 * x + g + 4 * r][j0 + 16 * y + rr] = ai[x][y][r];
      }
  __syncthreads();
  for (int e = t; e < JJ; e += 256) {
    {  // G[P, K], rows fastest
      const int i = e % JP, j = e / JP;
      const int r = rowmap[i], c = colmap[j];
      if (r >= 0 && c >= 0) {
        G[r + int64_t(c) * n] = Tr[i][j];
        G[imG + r + int64_t(c) * n] = Ti[i][j];
      }
    }
    {  // G[K, P] = G[P, K]^H
      const int j = e % JP, i = e / JP;
      const int r = rowmap[i], c = colmap[j];
      if (r >= 0 && c >= 0) {
        G[c + int64_t(r) * n] = Tr[i][j];
        G[imG + c + int64_t(r) * n] = -Ti[i][j];
      }
    }
  }
}

// Complex V[:, P] <- V[:, P] Q_P, as V^T <- Q^T V^T on the matrix cores (see the real
// kernel): workgroups of 64 rows (16 per wave) and one pair, Q_P staged in LDS.
template <int B>
__global__ __launch_bounds__(256) void jac_colupd_herm_kernel(double *__restrict__ Y, int rows, int64_t imY,
                                                              const int *__restrict__ pairs, int cols,
                                                              const double *__restrict__ Qin,
                                                              const int *__restrict__ active) {
  constexpr int JP = 2 * B, FJ = JP / 16, KS = JP / 4, JJ = JP * JP;
  const int pair = blockIdx.y;
  if (!active[pair]) return;
  __shared__ double Qr[JP][JP + 1], Qi[JP][JP + 1];
  __shared__ int colmap[JP];
  const int t = threadIdx.x;
  const double *Q = Qin + int64_t(pair) * 2 * JJ;
  for (int e = t; e < JJ; e += 256) {
    Qr[e / JP][e % JP] = Q[e];
    Qi[e / JP][e % JP] = Q[JJ + e];
  }
  if (t < JP) colmap[t] = pair_col<B>(pairs + 2 * pair, t, cols);
  __syncthreads();
  const int lane = t & 63, w = t >> 6, g = lane >> 4, rr = lane & 15;
  const int r0 = blockIdx.x * 64 + 16 * w;
  const int row = r0 + rr;
  double vr[KS], vi[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int c = colmap[4 * s + g];
    const bool ok = row < rows && c >= 0;
    vr[s] = ok ? Y[row + int64_t(c) * rows] : 0.0;
    vi[s] = ok ? Y[imY + row + int64_t(c) * rows] : 0.0;
  }
  jv4d ar[FJ], ai[FJ];
#pragma unroll
  for (int x = 0; x < FJ; ++x) ar[x] = ai[x] = jv4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
#pragma unroll
    for (int x = 0; x < FJ; ++x) {
      const double qr_ = Qr[4 * s + g][16 * x + rr], qi_ = Qi[4 * s + g][16 * x + rr];
      ar[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(qr_, vr[s], ar[x], 0, 0, 0);
      ar[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(-qi_, vi[s], ar[x], 0, 0, 0);
      ai[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(qr_, vi[s], ai[x], 0, 0, 0);
      ai[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(qi_, vr[s], ai[x], 0, 0, 0);
    }
  }
  // ar[x][r] = new V[row][column 16 x + g + 4 r of the pair]
#pragma unroll
  for (int x = 0; x < FJ; ++x)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int c = colmap[16 * x + g + 4 * r];
      if (row < rows && c >= 0) {
        Y[row + int64_t(c) * rows] = ar[x][r];
        Y[imY + row + int64_t(c) * rows] = ai[x][r];
      }
    }
}

// G <- (G + G^H) / 2: the MFMA product M M^H is not bitwise Hermitian (G[i,j] and G[j,i]
// come from different tiles), and the Jacobi rotations assume it is.
__global__ void hermitize_kernel(double *G, int n, int64_t imG, bool cplx) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= int64_t(n) * n) return;
  const int i = int(e % n), j = int(e / n);
  if (i > j) return;
  const int64_t a = i + int64_t(j) * n, b = j + int64_t(i) * n;
  const double re = 0.5 * (G[a] + G[b]);
  G[a] = re;
  G[b] = re;
  if (cplx) {
    const double im = 0.5 * (G[imG + a] - G[imG + b]);
    G[imG + a] = (i == j) ? 0.0 : im;
    G[imG + b] = (i == j) ? 0.0 : -im;
  }
}

__global__ void eye_kernel(double *V, int n, int64_t imV, bool cplx) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t nn = int64_t(n) * n;
  if (e >= nn) return;
  V[e] = (e % n == e / n) ? 1.0 : 0.0;
  if (cplx) V[imV + e] = 0.0;
}

// diag[j] = Re G[j, j]  or  colnorm2[j] = sum_r |X[r, j]|^2 (one block per column)
__global__ void diag_kernel(const double *G, int n, double *out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = G[j + int64_t(j) * n];
}

__global__ void colnorm2_kernel(const double *X, int rows, int64_t imX, bool cplx, double *out) {
  const int j = blockIdx.x;
  double s = 0;
  for (int r = threadIdx.x; r < rows; r += blockDim.x) {
    const double a = X[r + int64_t(j) * rows];
    s += a * a;
    if (cplx) {
      const double b = X[imX + r + int64_t(j) * rows];
      s += b * b;
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[j] = red[0] + red[1] + red[2] + red[3];
}

// out[:, j] = scale_j * src[:, perm[j]]  (rows x k, planar complex when cplx)
__global__ void gather_cols_kernel(const double *src, int rows, int64_t imS, const int *perm,
                                   const double *scale, int k, double *out, int64_t imO,
                                   bool cplx) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= int64_t(rows) * k) return;
  const int r = int(e % rows), j = int(e / rows);
  const int64_t s = r + int64_t(perm[j]) * rows;
  const double f = scale ? scale[j] : 1.0;
  out[e] = f * src[s];
  if (cplx) out[imO + e] = f * src[imS + s];
}

// Fixes the gauge of each column of a (planar) basis: the column is multiplied by the phase
// that makes its entry of largest modulus (the first such entry) real and positive. The
// eigensolver's own sign or phase choice then no longer reaches the split factors, so a bond
// basis comes out the same each time the same subspace is split. One workgroup per column.
__global__ __launch_bounds__(256) void canon_phase_kernel(double *U, int rows, int64_t imU, bool cplx) {
  const int j = blockIdx.x;
  double *ur = U + int64_t(j) * rows;
  double *ui = cplx ? U + imU + int64_t(j) * rows : nullptr;
  __shared__ double bm[256];
  __shared__ int bi[256];
  double best = -1.0;
  int at = 0;
  for (int r = threadIdx.x; r < rows; r += 256) {
    const double m = cplx ? ur[r] * ur[r] + ui[r] * ui[r] : ur[r] * ur[r];
    if (m > best) {
      best = m;
      at = r;
    }
  }
  bm[threadIdx.x] = best;
  bi[threadIdx.x] = at;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double o = bm[threadIdx.x + w];
      const int oi = bi[threadIdx.x + w];
      if (o > bm[threadIdx.x] || (o == bm[threadIdx.x] && oi < bi[threadIdx.x])) {
        bm[threadIdx.x] = o;
        bi[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  const int r0 = bi[0];
  const double mr = ur[r0], mi = cplx ? ui[r0] : 0.0;
  const double mag = sqrt(mr * mr + mi * mi);
  __syncthreads();  // every thread has read the pivot before any thread rescales it
  if (!(mag > 0.0)) return;
  const double fr = mr / mag, fi = -mi / mag;  // conj(u) / |u|
  for (int r = threadIdx.x; r < rows; r += 256) {
    if (cplx) {
      const double a = ur[r], b = ui[r];
      ur[r] = a * fr - b * fi;
      ui[r] = a * fi + b * fr;
    } else {
      ur[r] = ur[r] * fr;
    }
  }
}

// X[i, c] *= f[i] for a column-major matrix with `rows` rows (tot elements over all planes)
__global__ void scale_rows_kernel(double *X, int rows, int64_t tot, const double *f) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e < tot) X[e] *= f[e % rows];
}

// out[j, r] = scale_j * conj(src[r, perm[j]])  (k x rows), via a 32 x 32 LDS tile
__global__ void gather_cols_adjoint_kernel(const double *src, int rows, int64_t imS,
                                           const int *perm, const double *scale, int k,
                                           double *out, int64_t imO, bool cplx) {
  __shared__ double tr[32][33], ti[32][33];
  const int r0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 256 threads: ty in [0, 8)
  for (int jj = ty; jj < 32; jj += 8) {
    const int j = j0 + jj, r = r0 + tx;
    double a = 0, b = 0;
    if (j < k && r < rows) {
      const int64_t s = r + int64_t(perm[j]) * rows;
      const double f = scale ? scale[j] : 1.0;
      a = f * src[s];
      b = cplx ? -f * src[imS + s] : 0.0;
    }
    tr[jj][tx] = a;
    ti[jj][tx] = b;
  }
  __syncthreads();
  for (int rr = ty; rr < 32; rr += 8) {
    const int r = r0 + rr, j = j0 + tx;
    if (j < k && r < rows) {
      out[j + int64_t(r) * k] = tr[tx][rr];
      if (cplx) out[imO + j + int64_t(r) * k] = ti[tx][rr];
    }
  }
}

// The whole eigenproblem of a Hermitian p x p Gram matrix (p <= 2B) in one workgroup of NT
// threads: G is made exactly Hermitian in LDS and padded with zeros to 2B x 2B, the noise
// floor 4 eps sqrt(p) max_i G_ii is taken from its diagonal, the cyclic Jacobi runs until a
// sweep rotates nothing, and V (p x p, planar), the eigenvalues out[0..p) and the converged
// flag out[p] (1 or 0) are written. The host picks the smallest width that holds p (a sweep
// is 2B - 1 steps).
template <bool CPLX, int B, int NT>
__global__ __launch_bounds__(NT) void tiny_eig_kernel(const double *__restrict__ G, int p, int64_t imG,
                                                      double tol, double *__restrict__ V, int64_t imV,
                                                      double *__restrict__ out) {
  constexpr int PB = 2 * B;
  __shared__ double gr[PB][PB + 1], gi[CPLX ? PB : 1][PB + 1], qr[PB][PB + 1], qi[CPLX ? PB : 1][PB + 1];
  const int t = threadIdx.x;
  for (int e = t; e < PB * PB; e += NT) {
    const int i = e % PB, j = e / PB;
    double vr = 0, vi = 0;
    if (i < p && j < p) {
      vr = 0.5 * (G[i + int64_t(j) * p] + G[j + int64_t(i) * p]);
      if (CPLX) vi = (i == j) ? 0.0 : 0.5 * (G[imG + i + int64_t(j) * p] - G[imG + j + int64_t(i) * p]);
    }
    gr[i][j] = vr;
    qr[i][j] = (i == j) ? 1.0 : 0.0;
    if (CPLX) {
      gi[i][j] = vi;
      qi[i][j] = 0.0;
    }
  }
  __syncthreads();
  double dm = t < PB ? fabs(gr[t][t]) : 0.0;
  dm = block_max<NT>(dm);
  // exact power-of-two scaling to max_i G_ii in [1, 2): the rotation test forms G_ii G_jj,
  // which would overflow (or underflow) for Gram matrices of very large (or small) norm
  const int ex = (dm > 0 && isfinite(dm)) ? ilogb(dm) : 0;
  if (ex != 0) {
    const double sc = ldexp(1.0, -ex);
    for (int e = t; e < PB * PB; e += NT) {
      gr[e % PB][e / PB] *= sc;
      if (CPLX) gi[e % PB][e / PB] *= sc;
    }
    dm *= sc;
    __syncthreads();
  }
  const double abstol = 4.0 * 2.220446049250313e-16 * sqrt(double(p)) * dm;
  const bool conv = jacobi_lds<CPLX, B, NT>(gr, gi, qr, qi, tol, abstol, kMaxInnerCap);
  for (int e = t; e < p * p; e += NT) {
    const int i = e % p, j = e / p;
    V[i + int64_t(j) * p] = qr[i][j];
    if (CPLX) V[imV + i + int64_t(j) * p] = qi[i][j];
  }
  if (t < p) out[t] = ldexp(gr[t][t], ex);
  if (t == 0) out[p] = conv ? 1.0 : 0.0;
}

template <bool CPLX>
void launch_tiny_eig(hipStream_t st, const double *G, int p, int64_t imG, double tol, double *V,
                     int64_t imV, double *out) {
  if (p <= 4)
    hipLaunchKernelGGL((tiny_eig_kernel<CPLX, 2, 64>), dim3(1), dim3(64), 0, st, G, p, imG, tol, V, imV, out);
  else if (p <= 8)
    hipLaunchKernelGGL((tiny_eig_kernel<CPLX, 4, 64>), dim3(1), dim3(64), 0, st, G, p, imG, tol, V, imV, out);
  else if (p <= 16)
    hipLaunchKernelGGL((tiny_eig_kernel<CPLX, 8, 64>), dim3(1), dim3(64), 0, st, G, p, imG, tol, V, imV, out);
  else if (p <= 32)
    hipLaunchKernelGGL((tiny_eig_kernel<CPLX, 16, 256>), dim3(1), dim3(256), 0, st, G, p, imG, tol, V, imV, out);
  else  // p <= 64: 2 x 64 x 65 doubles of LDS per plane pair (133 KB complex)
    hipLaunchKernelGGL((tiny_eig_kernel<CPLX, 32, 1024>), dim3(1), dim3(1024), 0, st, G, p, imG, tol, V, imV, out);
}

// ---- small Hermitian eigenproblems by reduction and divide and conquer (dc_eig.hpp)

// Reduction and divide and conquer of the p x p Hermitian G (planar; p <= kDcPMax): the
// eigenvalues (ascending, out[0..p)), the success flag out[p], the eigenvectors of the
// tridiagonal matrix in Q, the Householder vectors in G and their factors in taus[0..2p).
template <bool CPLX, int NT>
__global__ __launch_bounds__(NT) void dc_eig_kernel(double *G, int p, int64_t imG, double *taus, double *Q,
                                                     double *Q2, double *U, double *out) {
  __shared__ dc::DcShared sh;
  dc::dc_eig<CPLX>(dc::DcDeviceEx<NT>{}, sh, G, CPLX ? G + imG : nullptr, p, taus, taus + p, Q, Q2, U, out);
}

// One 1024-thread workgroup for every size: measured, 64 threads for p <= 64 and 256 for
// p <= 128 were 1.3-2x slower (the serial chains inside the phases, not the barriers, set
// the time).
template <bool CPLX>
void launch_dc_eig(hipStream_t st, double *G, int p, int64_t imG, double *taus, double *Q, double *Q2,
                   double *U, double *out) {
  hipLaunchKernelGGL((dc_eig_kernel<CPLX, dc::kDcThreads>), dim3(1), dim3(dc::kDcThreads), 0, st, G, p, imG, taus,
                     Q, Q2, U, out);
}

// V = H_0 H_1 ... H_{p-2} Z: the tridiagonal eigenvectors Z taken back through the
// Householder reflectors H_k = I - tau_k v_k v_k^H (v_k in column k of G below the diagonal).
// One wave per 4 columns, 16 lanes per column; lane l keeps rows l, l + 16, ... (R of them,
// 16 R >= p) of its column in registers for the whole product, so the reflectors are the only
// memory traffic and the column's dot product is a 16-lane shuffle reduction: no barriers.
// The next reflector is loaded while the current one is applied.
template <bool CPLX, int R>
__global__ __launch_bounds__(256) void dc_backtransform_kernel(const double *__restrict__ G, int p, int64_t imG,
                                                               const double *__restrict__ taus,
                                                               const double *__restrict__ Z, double *__restrict__ V,
                                                               int64_t imV) {
  const int lane = threadIdx.x & 63, l = lane & 15;
  const int c = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 4 + (lane >> 4);
  const bool act = c < p;
  const int64_t P = p;
  double xr[R], xi[R], nr[R], ni[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = l + 16 * r;
    xr[r] = (act && i < p) ? Z[i + c * P] : 0.0;
    xi[r] = 0.0;
  }
  auto load = [&](int k, double *ar_, double *ai_) {  // v_k on this lane's rows (0 off its support)
    const double *vr = G + k * P, *vi = CPLX ? G + imG + k * P : nullptr;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = l + 16 * r;
      const bool on = i > k && i < p;
      ar_[r] = on ? vr[i] : 0.0;
      ai_[r] = (CPLX && on) ? vi[i] : 0.0;
    }
  };
  int k = p - 2;
  double tnr = 0, tni = 0;
  if (k >= 0) {
    load(k, nr, ni);
    tnr = taus[k];
    tni = CPLX ? taus[p + k] : 0.0;
  }
  for (; k >= 0; --k) {
    double cr[R], ci[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cr[r] = nr[r];
      ci[r] = ni[r];
    }
    const double tr = tnr, ti = tni;
    if (k > 0) {
      load(k - 1, nr, ni);
      tnr = taus[k - 1];
      tni = CPLX ? taus[p + k - 1] : 0.0;
    }
    if (tr == 0.0 && ti == 0.0) continue;
    double sr = 0, si = 0;  // s = v^H x
#pragma unroll
    for (int r = 0; r < R; ++r) {
      sr += cr[r] * xr[r] + ci[r] * xi[r];
      if (CPLX) si += cr[r] * xi[r] - ci[r] * xr[r];
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
      sr += __shfl_xor(sr, o, 64);
      if (CPLX) si += __shfl_xor(si, o, 64);
    }
    const double fr = tr * sr - ti * si, fi = tr * si + ti * sr;  // tau s
#pragma unroll
    for (int r = 0; r < R; ++r) {
      xr[r] -= cr[r] * fr - ci[r] * fi;
      if (CPLX) xi[r] -= cr[r] * fi + ci[r] * fr;
    }
  }
  if (!act) return;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = l + 16 * r;
    if (i < p) {
      V[i + c * P] = xr[r];
      if (CPLX) V[imV + i + c * P] = xi[r];
    }
  }
}

template <bool CPLX>
void launch_dc_backtransform(hipStream_t st, const double *G, int p, int64_t imG, const double *taus,
                             const double *Z, double *V, int64_t imV) {
  const dim3 grid(unsigned(((p + 3) / 4 + 3) / 4)), block(256);
  if (p <= 64)
    hipLaunchKernelGGL((dc_backtransform_kernel<CPLX, 4>), grid, block, 0, st, G, p, imG, taus, Z, V, imV);
  else if (p <= 128)
    hipLaunchKernelGGL((dc_backtransform_kernel<CPLX, 8>), grid, block, 0, st, G, p, imG, taus, Z, V, imV);
  else if (p <= 256)
    hipLaunchKernelGGL((dc_backtransform_kernel<CPLX, 16>), grid, block, 0, st, G, p, imG, taus, Z, V, imV);
  else
    hipLaunchKernelGGL((dc_backtransform_kernel<CPLX, 32>), grid, block, 0, st, G, p, imG, taus, Z, V, imV);
}

struct Schedule {
  int nb = 0, nbp = 0, rounds = 0, npairs = 0;
  int *dev = nullptr;  // [round][pair][2]
};

}  // namespace

// ------------------------------------------------------------------ engine

struct JacobiResult {
  int sweeps = 0;
  bool converged = false;
  double measure = 0;
};

// The circle schedule depends on the block count only, whatever the block width.
static Schedule schedule_for(tnh_ctx *ctx, int cols, int width) {
  Schedule s;
  s.nb = (cols + width - 1) / width;
  s.nbp = s.nb + (s.nb & 1);
  if (s.nbp < 2) s.nbp = 2;
  s.rounds = s.nbp - 1;
  s.npairs = s.nbp / 2;
  auto it = ctx->schedules.find(s.nb);
  if (it == ctx->schedules.end()) {  // built once per block count, kept on the device
    std::vector<int> h(size_t(s.rounds) * s.npairs * 2);
    for (int r = 0; r < s.rounds; ++r)
      for (int i = 0; i < s.npairs; ++i) {
        int p, q;
        circle_pair(s.nbp, r, i, &p, &q);
        h[(size_t(r) * s.npairs + i) * 2] = p < s.nb ? p : -1;
        h[(size_t(r) * s.npairs + i) * 2 + 1] = q < s.nb ? q : -1;
      }
    void *d = nullptr;
    TNH_HIP(hipMalloc(&d, sizeof(int) * h.size()));
    to_device(ctx, d, h.data(), sizeof(int) * h.size());
    it = ctx->schedules.emplace(s.nb, std::make_pair(d, sizeof(int) * h.size())).first;
  }
  s.dev = static_cast<int *>(it->second.first);
  return s;
}

// two_sided: diagonalise Hermitian G (n x n) in place, V <- V * rotations.
// one-sided: orthogonalise the columns of X (rows x n) in place, V (n x n) accumulates.
// Columns are cut into blocks of width B (pairs of 2B columns); the two-sided problem (real
// or Hermitian) uses B = 32, which halves the rounds per sweep and so the passes over G and
// V, the one-sided fallback B = JB = 16.
// Two-sided problems up to this size use 16-wide blocks: twice the rounds of 32-wide ones,
// but a round is then latency- rather than throughput-bound and its 32 x 32 pair solver
// takes a fraction of the 64 x 64 one's time.
constexpr int kNarrowMax = 512;

template <int B>
static JacobiResult jacobi_blocks(tnh_ctx *ctx, bool two_sided, double *X, int rows, int n, bool cplx,
                                  double *V, double tol, double abstol, int max_sweeps) {
  constexpr int JPB = 2 * B;
  constexpr int NT = B == JB ? 256 : 1024;  // threads of the pair eigensolver (1024 measured faster than 512 at B = 32)
  static_assert(B == JB || B == 2 * JB, "block widths 16 and 32");
  if constexpr (B != JB)
    TNH_REQUIRE(two_sided, TNH_ERR_ARG, "wide Jacobi blocks serve the two-sided problem");
  JacobiResult res;
  const Schedule s = schedule_for(ctx, n, B);
  const int64_t imX = int64_t(rows) * n, imV = int64_t(n) * n;
  int nsplit = 1;
  if (!two_sided) {
    nsplit = std::max(1, std::min((2 * ctx->cu_count + s.npairs - 1) / s.npairs,
                                  (rows + kGramRows - 1) / kGramRows));
  }
  double *partial = nullptr;
  if (!two_sided)
    partial = static_cast<double *>(
        ctx->workspace(601, sizeof(double) * size_t(s.npairs) * nsplit * 2 * JPB * JPB));
  // Two-sided problems overlap the column update of V (it needs only the round's rotations)
  // with the next round's pair solver, a latency-bound launch of one workgroup per pair:
  // V is updated on the helper stream once G's tile update has run, the rotations Q and the activity flags are double-buffered by round parity, and
  // the pair solver of round g + 2 waits for the V update of round g before it overwrites
  // that round's buffer. The helper stream starts after, and is joined back into, the
  // context's stream.
  const bool overlap = two_sided;
  const int nbuf = overlap ? 2 : 1;
  const size_t qn = size_t(s.npairs) * 2 * JPB * JPB;
  double *Q = static_cast<double *>(ctx->workspace(602, sizeof(double) * qn * nbuf));
  double *D = two_sided ? static_cast<double *>(ctx->workspace(603, sizeof(double) * size_t(s.npairs) * 2 * JPB * JPB))
                        : nullptr;
  int *active = static_cast<int *>(ctx->workspace(604, sizeof(int) * s.npairs * nbuf));
  unsigned long long *conv = static_cast<unsigned long long *>(
      ctx->workspace(605, sizeof(unsigned long long) * (max_sweeps + 1)));
  TNH_HIP(hipMemsetAsync(conv, 0, sizeof(unsigned long long) * (max_sweeps + 1), ctx->stream));
  const int rmax = std::max(rows, n);
  // a single pair covering every column is the whole problem: solve it to convergence.
  // Otherwise one inner sweep per pair and round: measured fastest overall (the outer
  // sweeps revisit every pair, so fully converged sub-problems buy little).
  const int max_inner = (s.nb <= 2) ? kMaxInnerCap : 1;
  hipStream_t vs = ctx->stream;
  hipEvent_t evq[2] = {nullptr, nullptr}, evv[2] = {nullptr, nullptr}, evj = nullptr;
  if (overlap) {
    vs = ctx->aux_stream();
    for (int i = 0; i < 2; ++i) {
      evq[i] = ctx->take_event();
      evv[i] = ctx->take_event();
    }
    evj = ctx->take_event();
    TNH_HIP(hipEventRecord(evj, ctx->stream));
    TNH_HIP(hipStreamWaitEvent(vs, evj, 0));
  }
  auto sweeps = [&]() {
    for (int sw = 0; sw < max_sweeps; ++sw) {
      for (int r = 0; r < s.rounds; ++r) {
        const int *pr = s.dev + size_t(r) * s.npairs * 2;
        const int64_t gr = int64_t(sw) * s.rounds + r;
        const int b = overlap ? int(gr & 1) : 0;
        double *Qb = Q + qn * b;
        int *act = active + s.npairs * b;
        if (overlap && gr >= 2) TNH_HIP(hipStreamWaitEvent(ctx->stream, evv[b], 0));
        // the first round of a sweep visits every block once: it also rotates the index pairs
        // inside the blocks; the other rounds only the pairs across the two blocks of a pair
        const bool cross = s.nb > 2 && r > 0;
        if constexpr (B == JB) {
          if (!two_sided)
            hipLaunchKernelGGL(jac_gram_kernel, dim3(s.npairs, nsplit), dim3(256), 0, ctx->stream, X,
                               rows, imX, n, cplx, pr, nsplit, partial);
        }
        if (cplx)
          hipLaunchKernelGGL((jac_eig_kernel<true, B, NT>), dim3(s.npairs), dim3(NT), 0, ctx->stream, X,
                             imX, n, partial, nsplit, two_sided, pr, n, tol, abstol, max_inner, cross, Qb, D,
                             act, conv + sw);
        else
          hipLaunchKernelGGL((jac_eig_kernel<false, B, NT>), dim3(s.npairs), dim3(NT), 0, ctx->stream, X,
                             imX, n, partial, nsplit, two_sided, pr, n, tol, abstol, max_inner, cross, Qb, D,
                             act, conv + sw);
        if (two_sided && cplx)
          hipLaunchKernelGGL(jac_twoside_herm_kernel<B>, dim3(s.npairs * (s.npairs + 1) / 2), dim3(256), 0,
                             ctx->stream, X, n, imX, pr, Qb, D, act);
        if (two_sided && !cplx)
          hipLaunchKernelGGL(jac_twoside_sym_kernel<B>, dim3(s.npairs * (s.npairs + 1) / 2), dim3(256), 0,
                             ctx->stream, X, n, pr, Qb, D, act);
        if (overlap) {
          TNH_HIP(hipEventRecord(evq[b], ctx->stream));
          TNH_HIP(hipStreamWaitEvent(vs, evq[b], 0));
        }
        // two-sided: V takes the column rotations on the helper stream after the tile update,
        // alongside the next round's pair solver (which fills a quarter of the CUs); one-sided:
        // X and V take the column rotations
        double *Xc = two_sided ? nullptr : X;
        if constexpr (B == JB) {
          if (cplx)
            hipLaunchKernelGGL((jac_colupd_kernel<true, B>), dim3((rmax + 255) / 256, s.npairs), dim3(256), 0,
                               vs, Xc, rows, imX, V, n, imV, pr, n, Qb, act);
        } else {
          if (cplx)
            hipLaunchKernelGGL(jac_colupd_herm_kernel<B>, dim3((n + 63) / 64, s.npairs), dim3(256), 0, vs, V,
                               n, imV, pr, n, Qb, act);
        }
        if (!cplx && two_sided)
          hipLaunchKernelGGL(jac_colupd_mfma_kernel<B>, dim3((n + 127) / 128, s.npairs), dim3(256), 0, vs, V, n,
                             pr, n, Qb, act);
        else if (!cplx)
          hipLaunchKernelGGL((jac_colupd_kernel<false, B>), dim3((rmax + 255) / 256, s.npairs), dim3(256), 0,
                             vs, Xc, rows, imX, V, n, imV, pr, n, Qb, act);
        if (overlap) TNH_HIP(hipEventRecord(evv[b], vs));
      }
      TNH_HIP(hipGetLastError());
      unsigned long long bits = 0;
      to_host(ctx, &bits, conv + sw, sizeof(bits));
      double m;
      std::memcpy(&m, &bits, sizeof(m));
      res.sweeps = sw + 1;
      res.measure = m;
      if (!std::isfinite(m)) return;
      if (m <= tol) {
        res.converged = true;
        return;
      }
    }
  };
  // the context's stream waits for the last V update; events go back to the pool
  auto join = [&]() {
    if (!overlap) return;
    hipError_t e1 = hipEventRecord(evj, vs);
    hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(ctx->stream, evj, 0) : e1;
    for (int i = 0; i < 2; ++i) {
      ctx->event_pool.push_back(evq[i]);
      ctx->event_pool.push_back(evv[i]);
    }
    ctx->event_pool.push_back(evj);
    TNH_HIP(e2);
  };
  try {
    sweeps();
  } catch (...) {
    if (overlap) (void)hipStreamSynchronize(vs);
    try {
      join();
    } catch (...) {
    }
    throw;
  }
  join();
  return res;
}

static JacobiResult jacobi(tnh_ctx *ctx, bool two_sided, double *X, int rows, int n, bool cplx,
                           double *V, double tol, double abstol, int max_sweeps) {
  if (two_sided && n > kNarrowMax)
    return jacobi_blocks<2 * JB>(ctx, two_sided, X, rows, n, cplx, V, tol, abstol, max_sweeps);
  return jacobi_blocks<JB>(ctx, two_sided, X, rows, n, cplx, V, tol, abstol, max_sweeps);
}

// ------------------------------------------------------------- decomposition


// Host copy of the truncation rule, src/tensors.jl:202-215.
static int truncation_rank(const std::vector<double> &S, double cutoff, int maxdim, int mindim,
                           double *disc) {
  const int sz = int(S.size());
  mindim = std::min(mindim, sz);
  maxdim = (maxdim == 0 || maxdim > sz) ? sz : maxdim;
  maxdim = maxdim == 0 ? 1 : maxdim;
  std::vector<double> S2(sz);
  double tot = 0;
  for (int i = 0; i < sz; ++i) {
    S2[i] = S[i] * S[i];
    tot += S2[i];
  }
  if (cutoff != 0) {
    // reverse cumulative sum divided by the total, as the reference computes it
    std::vector<double> cum(sz);
    double acc = 0;
    for (int i = sz - 1; i >= 0; --i) {
      acc += S2[i];
      cum[i] = acc;
    }
    int idxs = 0;
    for (int i = 0; i < sz; ++i)
      if (cum[i] / tot > cutoff) idxs = i + 1;
    if (idxs == 0) idxs = 1;
    maxdim = std::min(maxdim, idxs);
  }
  const int k = std::max(maxdim, mindim);
  double d = 0;
  for (int i = k; i < sz; ++i) d += S2[i];
  *disc = tot > 0 ? d / tot : 0.0;
  return k;
}

// Decompose the m x n matrix Mp (device, column-major, planar if cplx).
// side 0 ("left"):  returns the left factor basis Uk (m x k, orthonormal columns, in
//                   order of descending s) in *basis;
// side 1 ("right"): returns the right factor basis Vk (n x k) with M ~ (M Vk) Vk^H.
// The basis buffer is workspace slot 620 (valid until the next SVD).
static double *decompose(tnh_ctx *ctx, const double *Mp, int m, int n, bool cplx, int side,
                         double cutoff, int maxdim, int mindim, bool force_fallback, SvdOut *o,
                         tnh_buf **warm = nullptr) {
  o->m = m;
  o->n = n;
  const int p = side == 0 ? m : n;  // basis dimension
  const int64_t pp = int64_t(p) * p;
  const int planes = cplx ? 2 : 1;
  double *G = static_cast<double *>(ctx->workspace(610, sizeof(double) * pp * planes));
  double *V = static_cast<double *>(ctx->workspace(611, sizeof(double) * pp * planes));
  std::vector<double> val(p);
  bool ok = false;
  // relative off-diagonal tolerance 4 sqrt(p) eps (at least 1e-14), the scale of the
  // rounding that p-term inner products leave
  const double tol = std::max(1e-14, 4.0 * 2.220446049250313e-16 * std::sqrt(double(p)));
  // G = M M^H (side 0) or M^H M (side 1)
  auto form_gram = [&]() {
    Gemm g;
    g.M = p;
    g.N = p;
    g.K = side == 0 ? n : m;
    g.lda = m;
    g.ldb = m;
    g.ldc = p;
    COperand A, B;
    A.re = Mp;
    A.im = cplx ? Mp + int64_t(m) * n : nullptr;
    B = A;
    if (side == 0) {
      g.transB = true;  // M * M^H
      B.conj = true;
    } else {
      g.transA = true;  // M^H * M
      A.conj = true;
    }
    cgemm(ctx, g, A, B, G, cplx ? G + pp : nullptr, TNH_F64, TNH_PREC_F64);
  };
  auto take_eigenvalues = [&](const double *dout) {
    std::vector<double> h(p + 1);
    to_host(ctx, h.data(), dout, sizeof(double) * (p + 1));
    if (h[p] != 1.0) return false;
    for (int i = 0; i < p; ++i) {
      if (!std::isfinite(h[i])) return false;
      val[i] = std::sqrt(std::max(h[i], 0.0));
    }
    return true;
  };
  const int dtp = cplx ? TNH_C128 : TNH_F64;
  // a usable warm-start basis of a block-Jacobi split: Q^H G Q is then formed from M
  // directly, and G itself is never needed
  const bool use_warm = !force_fallback && p > dc::kDcPMax && warm && *warm && (*warm)->ndim == 2 &&
                        (*warm)->dims[0] == p && (*warm)->dims[1] == p && (*warm)->dtype == dtp;
  if (!force_fallback && !use_warm) {
    form_gram();
    if (p > 8 && p <= dc::kDcPMax) {
      // reduction to tridiagonal form and divide and conquer in one workgroup, then the
      // back-transformation of the eigenvectors through the Householder reflectors
      double *taus = static_cast<double *>(ctx->workspace(624, sizeof(double) * 2 * p));
      double *Qd = static_cast<double *>(ctx->workspace(625, sizeof(double) * pp));
      double *Q2d = static_cast<double *>(ctx->workspace(626, sizeof(double) * pp));
      double *Ud = static_cast<double *>(ctx->workspace(627, sizeof(double) * pp));
      double *dout = static_cast<double *>(ctx->workspace(616, sizeof(double) * (p + 1)));
      hipEvent_t pd = ctx->prof_begin();
      if (cplx) {
        launch_dc_eig<true>(ctx->stream, G, p, pp, taus, Qd, Q2d, Ud, dout);
        launch_dc_backtransform<true>(ctx->stream, G, p, pp, taus, Qd, V, pp);
      } else {
        launch_dc_eig<false>(ctx->stream, G, p, pp, taus, Qd, Q2d, Ud, dout);
        launch_dc_backtransform<false>(ctx->stream, G, p, pp, taus, Qd, V, pp);
      }
      TNH_HIP(hipGetLastError());
      ctx->prof_end("dc_eig", pd, 0.0, double(p));
      ok = take_eigenvalues(dout);
      if (!ok) form_gram();  // the reduction overwrote G: the Jacobi paths below start afresh
    }
    if (!ok && p <= 2 * JP) {
      // the whole problem fits one workgroup: a single launch and a single download
      double *dout = static_cast<double *>(ctx->workspace(616, sizeof(double) * (p + 1)));
      if (cplx)
        launch_tiny_eig<true>(ctx->stream, G, p, pp, tol, V, pp, dout);
      else
        launch_tiny_eig<false>(ctx->stream, G, p, pp, tol, V, pp, dout);
      TNH_HIP(hipGetLastError());
      ok = take_eigenvalues(dout);
      if (!ok) form_gram();
    }
  }
  if (!force_fallback && !ok && p > JP) {
    if (use_warm) {
      // Warm start: with an orthonormal basis Q of this Gram's space from an earlier split of
      // the same bond (its eigenvectors then), the Jacobi runs on Q^H G Q, nearly diagonal
      // when the matrix has changed little, and accumulates its rotations onto V = Q, so that
      // V ends as the eigenvectors of G itself. Same convergence test, so the same accuracy.
      // Q^H G Q = X X^H with X = Q^H M (side 0), or X^H X with X = M Q (side 1): two products
      // instead of the Gram matrix and two more.
      const double *Q = static_cast<const double *>((*warm)->ptr);
      const int q = side == 0 ? n : m;  // the other side of M
      const int64_t xq = int64_t(p) * q;
      double *X = static_cast<double *>(ctx->workspace(618, sizeof(double) * xq * planes));
      COperand Mo, Qo, Xo;
      Mo.re = Mp;
      Mo.im = cplx ? Mp + int64_t(m) * n : nullptr;
      Qo.re = Q;
      Qo.im = cplx ? Q + pp : nullptr;
      Xo.re = X;
      Xo.im = cplx ? X + xq : nullptr;
      Gemm g1, g2;
      COperand A1, B1, A2, B2;
      if (side == 0) {
        g1.M = p;  // X (p x n) = Q^H M
        g1.N = n;
        g1.K = p;
        g1.transA = true;
        g1.lda = p;
        g1.ldb = m;
        g1.ldc = p;
        A1 = Qo;
        A1.conj = true;
        B1 = Mo;
        g2.M = p;  // G = X X^H
        g2.N = p;
        g2.K = n;
        g2.transB = true;
        g2.lda = p;
        g2.ldb = p;
        g2.ldc = p;
        A2 = Xo;
        B2 = Xo;
        B2.conj = true;
      } else {
        g1.M = m;  // X (m x p) = M Q
        g1.N = p;
        g1.K = p;
        g1.lda = m;
        g1.ldb = p;
        g1.ldc = m;
        A1 = Mo;
        B1 = Qo;
        g2.M = p;  // G = X^H X
        g2.N = p;
        g2.K = m;
        g2.transA = true;
        g2.lda = m;
        g2.ldb = m;
        g2.ldc = p;
        A2 = Xo;
        A2.conj = true;
        B2 = Xo;
      }
      cgemm(ctx, g1, A1, B1, X, cplx ? X + xq : nullptr, TNH_F64, TNH_PREC_F64);
      cgemm(ctx, g2, A2, B2, G, cplx ? G + pp : nullptr, TNH_F64, TNH_PREC_F64);
      copy_device(ctx, V, Q, sizeof(double) * pp * planes);
    } else {
      hipLaunchKernelGGL(eye_kernel, dim3(unsigned((pp + 255) / 256)), dim3(256), 0, ctx->stream, V, p,
                         pp, cplx);
    }
    hipLaunchKernelGGL(hermitize_kernel, dim3(unsigned((pp + 255) / 256)), dim3(256), 0, ctx->stream,
                       G, p, pp, cplx);
    // Entries below 4 eps sqrt(p) max_i G_ii are rounding noise: each rotation of a
    // large pair leaves ~eps * lambda_max in the rows it touches, so a purely relative
    // test on noise-level eigenvalues never settles.
    double dmax = 0;
    {
      double *dv = static_cast<double *>(ctx->workspace(612, sizeof(double) * p));
      hipLaunchKernelGGL(diag_kernel, dim3((p + 255) / 256), dim3(256), 0, ctx->stream, G, p, dv);
      TNH_HIP(hipGetLastError());
      to_host(ctx, val.data(), dv, sizeof(double) * p);
      for (double v : val) dmax = std::max(dmax, std::fabs(v));
    }
    // exact power-of-two scaling to max_i G_ii in [1, 2): the rotation test forms G_ii G_jj,
    // which overflows for Gram matrices above ~1e154 (products of unnormalised tensors)
    const int ex = (dmax > 0 && std::isfinite(dmax)) ? std::ilogb(dmax) : 0;
    if (ex != 0) {
      scale(ctx, G, TNH_F64, pp * planes, std::ldexp(1.0, -ex));
      dmax = std::ldexp(dmax, -ex);
    }
    const double abstol = std::max(4.0 * 2.220446049250313e-16 * std::sqrt(double(p)), ctx->gram_floor) * dmax;
    hipEvent_t pj = ctx->prof_begin();
    JacobiResult jr = jacobi(ctx, true, G, p, p, cplx, V, tol, abstol, 40);
    ctx->prof_end(use_warm ? "jacobi_warm" : "jacobi", pj, double(jr.sweeps), double(p));
    if (jr.converged) {
      double *dv = static_cast<double *>(ctx->workspace(612, sizeof(double) * p));
      hipLaunchKernelGGL(diag_kernel, dim3((p + 255) / 256), dim3(256), 0, ctx->stream, G, p, dv);
      TNH_HIP(hipGetLastError());
      to_host(ctx, val.data(), dv, sizeof(double) * p);
      ok = true;
      for (auto &v : val) {
        if (!std::isfinite(v)) ok = false;
        v = std::sqrt(std::max(std::ldexp(v, ex), 0.0));
      }
    }
    if (ok && warm && p > dc::kDcPMax) {
      // The eigenvectors become the next split's warm start, ordered by descending
      // eigenvalue: the next Q^H G Q then holds its large eigenvalues in the leading column
      // blocks, so pairs of two trailing (noise-level) blocks are below the floor from the
      // first sweep on and their workgroups skip all work.
      std::vector<int> order(p);
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return val[a] > val[b]; });
      int *dord = static_cast<int *>(ctx->workspace(623, sizeof(int) * p));
      to_device(ctx, dord, order.data(), sizeof(int) * p);
      const int64_t wd[2] = {p, p};
      buf_ensure(ctx, warm, dtp, 2, wd);
      hipLaunchKernelGGL(gather_cols_kernel, dim3(unsigned((pp + 255) / 256)), dim3(256), 0, ctx->stream, V, p,
                         pp, dord, nullptr, p, static_cast<double *>((*warm)->ptr), pp, cplx);
      TNH_HIP(hipGetLastError());
    }
  }
  double *W = nullptr;  // the matrix whose columns are (unnormalised) singular directions
  if (!ok) {
    // one-sided Jacobi on M (side 1) or on M^H (side 0): its accumulated V is the basis
    o->fallback = true;
    const int rows = side == 0 ? n : m;
    double *X = static_cast<double *>(ctx->workspace(613, sizeof(double) * int64_t(rows) * p * planes));
    if (side == 1) {
      copy_device(ctx, X, Mp, sizeof(double) * int64_t(m) * n * planes);
    } else {
      // X = M^H (n x m): X[c, r] = conj(M[r, c]), i.e. the adjoint gather of all n
      // columns of the m-row matrix M
      std::vector<int> id(n);
      std::iota(id.begin(), id.end(), 0);
      int *dperm = static_cast<int *>(ctx->workspace(614, sizeof(int) * n));
      to_device(ctx, dperm, id.data(), sizeof(int) * n);
      dim3 grid((m + 31) / 32, (n + 31) / 32);
      hipLaunchKernelGGL(gather_cols_adjoint_kernel, grid, dim3(256), 0, ctx->stream, Mp, m,
                         int64_t(m) * n, dperm, nullptr, n, X, int64_t(n) * m, cplx);
      TNH_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(eye_kernel, dim3(unsigned((pp + 255) / 256)), dim3(256), 0, ctx->stream, V, p,
                       pp, cplx);
    // noise floor in Gram units: (4 eps sqrt(rows))^2 ||M||_F^2 (columns below
    // ~4 eps sqrt(rows) ||M||_F are re-seeded by every rotation with a large column)
    double fro2 = 0;
    {
      tnh_buf v;
      v.dtype = TNH_F64;
      v.ndim = 1;
      v.numel = int64_t(m) * n * planes;
      v.dims[0] = v.numel;
      v.ptr = const_cast<double *>(Mp);
      v.owned = false;
      double im = 0;
      cdot_host(ctx, &v, &v, &fro2, &im);
    }
    const double f1 = 4.0 * 2.220446049250313e-16 * std::sqrt(double(rows));
    JacobiResult jr = jacobi(ctx, false, X, rows, p, cplx, V, 1e-13, f1 * f1 * fro2, 60);
    TNH_REQUIRE(jr.converged, TNH_ERR_NOCONV,
                "svd: one-sided Jacobi did not converge (measure " + std::to_string(jr.measure) + ")");
    double *dv = static_cast<double *>(ctx->workspace(612, sizeof(double) * p));
    hipLaunchKernelGGL(colnorm2_kernel, dim3(p), dim3(256), 0, ctx->stream, X, rows,
                       int64_t(rows) * p, cplx, dv);
    TNH_HIP(hipGetLastError());
    to_host(ctx, val.data(), dv, sizeof(double) * p);
    for (auto &v : val) {
      TNH_REQUIRE(std::isfinite(v), TNH_ERR_NUMERIC, "svd: non-finite singular value");
      v = std::sqrt(std::max(v, 0.0));
    }
    W = X;
    (void)W;
  }
  // order by descending singular value (stable)
  std::vector<int> perm(p);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return val[a] > val[b]; });
  // the reference's S has min(m, n) entries
  const int r = std::min(m, n);
  o->s.resize(r);
  for (int i = 0; i < r; ++i) o->s[i] = val[perm[i]];
  o->k = truncation_rank(o->s, cutoff, maxdim, mindim, &o->discarded);
  const int k = o->k;
  int *dperm = static_cast<int *>(ctx->workspace(615, sizeof(int) * k));
  to_device(ctx, dperm, perm.data(), sizeof(int) * k);
  double *basis = static_cast<double *>(ctx->workspace(620, sizeof(double) * int64_t(p) * k * planes));
  hipLaunchKernelGGL(gather_cols_kernel, dim3(unsigned((int64_t(p) * k + 255) / 256)), dim3(256), 0,
                     ctx->stream, V, p, pp, dperm, nullptr, k, basis, int64_t(p) * k, cplx);
  hipLaunchKernelGGL(canon_phase_kernel, dim3(unsigned(k)), dim3(256), 0, ctx->stream, basis, p,
                     int64_t(p) * k, cplx);
  TNH_HIP(hipGetLastError());
  // One Newton-Schulz step toward the nearest orthonormal basis, U <- U (3I - U^H U) / 2:
  // the accumulated Jacobi rotations leave U^H U - I ~ 1e-13, this squares it away.
  {
    const int64_t pk = int64_t(p) * k, kk = int64_t(k) * k;
    double *X = static_cast<double *>(ctx->workspace(621, sizeof(double) * kk * planes));
    double *W = static_cast<double *>(ctx->workspace(622, sizeof(double) * pk * planes));
    COperand Ub, Uh;
    Ub.re = basis;
    Ub.im = cplx ? basis + pk : nullptr;
    Uh = Ub;
    Uh.conj = true;
    Gemm g1;  // X = U^H U
    g1.M = k;
    g1.N = k;
    g1.K = p;
    g1.transA = true;
    g1.lda = p;
    g1.ldb = p;
    g1.ldc = k;
    cgemm(ctx, g1, Uh, Ub, X, cplx ? X + kk : nullptr, TNH_F64, TNH_PREC_F64);
    Gemm g2;  // W = U X
    g2.M = p;
    g2.N = k;
    g2.K = k;
    g2.lda = p;
    g2.ldb = k;
    g2.ldc = p;
    COperand Xo;
    Xo.re = X;
    Xo.im = cplx ? X + kk : nullptr;
    cgemm(ctx, g2, Ub, Xo, W, cplx ? W + pk : nullptr, TNH_F64, TNH_PREC_F64);
    axpby(ctx, pk * planes, -0.5, W, 1.5, basis);  // U = 1.5 U - 0.5 W
  }
  return basis;
}

// Real-or-complex FP64 view of a 2-D device tensor; converts other dtypes into workspace.
static const double *as_f64(tnh_ctx *ctx, const tnh_buf *b, int slot) {
  if (is_double(b->dtype)) return static_cast<const double *>(b->ptr);
  void *w = ctx->workspace(slot, sizeof(double) * b->numel * (is_complex(b->dtype) ? 2 : 1));
  convert(ctx, b->ptr, b->dtype, w, is_complex(b->dtype) ? TNH_C128 : TNH_F64, b->numel);
  return static_cast<const double *>(w);
}

void cscale(tnh_ctx *ctx, void *x, int t, int64_t n, double cr, double ci);
void cdot_host(tnh_ctx *ctx, const tnh_buf *x, const tnh_buf *y, double *re, double *im);

static double frob(tnh_ctx *ctx, const tnh_buf *b) {
  double re = 0, im = 0;
  // sum conj(x) x = sum |x|^2 : use the real planes, which hold re then im
  tnh_buf v = *b;
  v.dtype = TNH_F64;
  v.numel = b->numel * (is_complex(b->dtype) ? 2 : 1);
  cdot_host(ctx, &v, &v, &re, &im);
  return std::sqrt(std::max(re, 0.0));
}

// Factor the m x n matrix of `src` (any dtype; its first `split_at` dims are rows) as
// left * right with truncation: side 0 -> left has orthonormal columns (U_k) and
// right = U_k^H M (= S V^H); side 1 -> right has orthonormal rows (V_k^H) and
// left = M V_k (= U S). left / right are (re)allocated with the given dims.
// Returns the kept rank; all singular values go to o->s.
int split_matrix(tnh_ctx *ctx, const tnh_buf *src, int m, int n, int side, double cutoff,
                 int maxdim, int mindim, bool force_fallback, int nl, const int64_t *ldims,
                 int nr, const int64_t *rdims, tnh_buf **A1, tnh_buf **A2, SvdOut *o,
                 tnh_buf **warm);

void replacesites2(tnh_ctx *ctx, const tnh_buf *theta, int direction, int normalize,
                   double cutoff, int maxdim, int mindim, tnh_buf **A1, tnh_buf **A2, int *k_out,
                   double *disc_out, double *schmidt, int schmidt_cap, bool force_fallback,
                   tnh_buf **warm) {
  TNH_REQUIRE(theta && theta->ndim == 4, TNH_ERR_SHAPE, "theta must be (chi_l, d, d, chi_r)");
  TNH_REQUIRE(A1 && A2, TNH_ERR_ARG, "A1/A2 out pointers are NULL");
  const int64_t cl = theta->dims[0], d1 = theta->dims[1], d2 = theta->dims[2], cr = theta->dims[3];
  const int m = int(cl * d1), n = int(d2 * cr);
  TNH_REQUIRE(m > 0 && n > 0, TNH_ERR_SHAPE, "empty theta");
  SvdOut o;
  const int side = direction ? 1 : 0;
  int64_t d1v[3] = {cl, d1, 0}, d2v[3] = {0, d2, cr};
  const int k = split_matrix(ctx, theta, m, n, side, cutoff, maxdim, mindim, force_fallback, 3,
                             d1v, 3, d2v, A1, A2, &o, warm);
  if (normalize) {
    // normalize!(psi): divide the centre tensor by its norm, gmps.jl:46-51
    tnh_buf *c = side == 0 ? *A2 : *A1;
    const double nr = frob(ctx, c);
    TNH_REQUIRE(std::isfinite(nr) && nr > 0, TNH_ERR_NUMERIC, "replacesites: zero centre tensor");
    scale(ctx, c->ptr, c->dtype, c->numel, 1.0 / nr);
  }
  if (k_out) *k_out = k;
  if (disc_out) *disc_out = o.discarded;
  if (schmidt)
    for (int i = 0; i < std::min(k, schmidt_cap); ++i) schmidt[i] = o.s[i];
}

int split_matrix(tnh_ctx *ctx, const tnh_buf *src, int m, int n, int side, double cutoff,
                 int maxdim, int mindim, bool force_fallback, int nl, const int64_t *ldims_in,
                 int nr, const int64_t *rdims_in, tnh_buf **A1, tnh_buf **A2, SvdOut *o,
                 tnh_buf **warm) {
  TNH_REQUIRE(int64_t(m) * n == src->numel, TNH_ERR_SHAPE, "split: matrix size != tensor size");
  const bool cplx = is_complex(src->dtype);
  const double *M = as_f64(ctx, src, 630);
  double *basis = decompose(ctx, M, m, n, cplx, side, cutoff, maxdim, mindim, force_fallback, o, warm);
  const int k = o->k;
  const int dt = cplx ? TNH_C128 : TNH_F64;
  int64_t d1v[8], d2v[8];
  for (int i = 0; i < nl; ++i) d1v[i] = ldims_in[i];
  for (int i = 0; i < nr; ++i) d2v[i] = rdims_in[i];
  d1v[nl - 1] = k;  // left factor: rows..., k
  d2v[0] = k;       // right factor: k, cols...
  buf_ensure(ctx, A1, dt, nl, d1v);
  buf_ensure(ctx, A2, dt, nr, d2v);
  COperand Mo;
  Mo.re = M;
  Mo.im = cplx ? M + int64_t(m) * n : nullptr;
  if (side == 0) {
    // A1 = U_k (left-orthonormal), A2 = U_k^H M (= S V^H)
    copy_device(ctx, (*A1)->ptr, basis, (*A1)->bytes);
    Gemm g;
    g.M = k;
    g.N = n;
    g.K = m;
    g.transA = true;
    g.lda = m;
    g.ldb = m;
    g.ldc = k;
    COperand U;
    U.re = basis;
    U.im = cplx ? basis + int64_t(m) * k : nullptr;
    U.conj = true;
    cgemm(ctx, g, U, Mo, (*A2)->re(), cplx ? (*A2)->im() : nullptr, TNH_F64, TNH_PREC_F64);
  } else {
    // A2 = V_k^H (right-orthonormal), A1 = M V_k (= U S)
    std::vector<int> id(k);
    std::iota(id.begin(), id.end(), 0);
    int *dperm = static_cast<int *>(ctx->workspace(631, sizeof(int) * k));
    to_device(ctx, dperm, id.data(), sizeof(int) * k);
    dim3 grid((n + 31) / 32, (k + 31) / 32);
    hipLaunchKernelGGL(gather_cols_adjoint_kernel, grid, dim3(256), 0, ctx->stream, basis, n,
                       int64_t(n) * k, dperm, nullptr, k, static_cast<double *>((*A2)->ptr),
                       int64_t(k) * n, cplx);
    TNH_HIP(hipGetLastError());
    Gemm g;
    g.M = m;
    g.N = k;
    g.K = n;
    g.lda = m;
    g.ldb = n;
    g.ldc = m;
    COperand Vk;
    Vk.re = basis;
    Vk.im = cplx ? basis + int64_t(n) * k : nullptr;
    cgemm(ctx, g, Mo, Vk, (*A1)->re(), cplx ? (*A1)->im() : nullptr, TNH_F64, TNH_PREC_F64);
  }
  return k;
}

void svd_trunc(tnh_ctx *ctx, const tnh_buf *Mb, double cutoff, int maxdim, int mindim,
               bool force_fallback, tnh_buf **U, tnh_buf **S, tnh_buf **Vt, int *k_out,
               double *disc_out, int *fb_out) {
  TNH_REQUIRE(Mb && Mb->ndim == 2, TNH_ERR_SHAPE, "svd_trunc: M must be a 2-D tensor");
  const int m = int(Mb->dims[0]), n = int(Mb->dims[1]);
  TNH_REQUIRE(m > 0 && n > 0, TNH_ERR_SHAPE, "svd_trunc: empty matrix");
  const bool cplx = is_complex(Mb->dtype);
  const double *M = as_f64(ctx, Mb, 630);
  SvdOut o;
  const int side = m <= n ? 0 : 1;
  double *basis = decompose(ctx, M, m, n, cplx, side, cutoff, maxdim, mindim, force_fallback, &o);
  const int k = o.k;
  const int dt = cplx ? TNH_C128 : TNH_F64;
  // singular values and their reciprocals (zero rows for s = 0)
  std::vector<double> sv(o.s.begin(), o.s.begin() + k), inv(k);
  for (int i = 0; i < k; ++i) inv[i] = sv[i] > 0 ? 1.0 / sv[i] : 0.0;
  double *dinv = static_cast<double *>(ctx->workspace(632, sizeof(double) * k));
  to_device(ctx, dinv, inv.data(), sizeof(double) * k);
  std::vector<int> id(k);
  std::iota(id.begin(), id.end(), 0);
  int *dperm = static_cast<int *>(ctx->workspace(631, sizeof(int) * k));
  to_device(ctx, dperm, id.data(), sizeof(int) * k);
  COperand Mo;
  Mo.re = M;
  Mo.im = cplx ? M + int64_t(m) * n : nullptr;
  double *other = static_cast<double *>(ctx->workspace(633, sizeof(double) * int64_t(std::max(m, n)) * k * (cplx ? 2 : 1)));
  if (side == 0) {
    // U = basis (m x k); Vt = diag(1/s) U^H M
    if (U) {
      int64_t dU[2] = {m, k};
      buf_ensure(ctx, U, dt, 2, dU);
      copy_device(ctx, (*U)->ptr, basis, (*U)->bytes);
    }
    if (Vt) {
      int64_t dV[2] = {k, n};
      buf_ensure(ctx, Vt, dt, 2, dV);
      Gemm g;
      g.M = k;
      g.N = n;
      g.K = m;
      g.transA = true;
      g.lda = m;
      g.ldb = m;
      g.ldc = k;
      COperand Ub;
      Ub.re = basis;
      Ub.im = cplx ? basis + int64_t(m) * k : nullptr;
      Ub.conj = true;
      cgemm(ctx, g, Ub, Mo, (*Vt)->re(), cplx ? (*Vt)->im() : nullptr, TNH_F64, TNH_PREC_F64);
      // row scaling Vt[i, :] *= 1 / s_i (both planes)
      const int64_t tot = int64_t(k) * n * (cplx ? 2 : 1);
      hipLaunchKernelGGL(scale_rows_kernel, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, ctx->stream,
                         static_cast<double *>((*Vt)->ptr), k, tot, dinv);
      TNH_HIP(hipGetLastError());
    }
  } else {
    // V = basis (n x k): Vt = basis^H; U = M V diag(1/s)
    if (Vt) {
      int64_t dV[2] = {k, n};
      buf_ensure(ctx, Vt, dt, 2, dV);
      dim3 grid((n + 31) / 32, (k + 31) / 32);
      hipLaunchKernelGGL(gather_cols_adjoint_kernel, grid, dim3(256), 0, ctx->stream, basis, n,
                         int64_t(n) * k, dperm, nullptr, k, static_cast<double *>((*Vt)->ptr),
                         int64_t(k) * n, cplx);
    }
    if (U) {
      int64_t dU[2] = {m, k};
      buf_ensure(ctx, U, dt, 2, dU);
      Gemm g;
      g.M = m;
      g.N = k;
      g.K = n;
      g.lda = m;
      g.ldb = n;
      g.ldc = m;
      COperand Vb;
      Vb.re = basis;
      Vb.im = cplx ? basis + int64_t(n) * k : nullptr;
      cgemm(ctx, g, Mo, Vb, other, cplx ? other + int64_t(m) * k : nullptr, TNH_F64, TNH_PREC_F64);
      hipLaunchKernelGGL(gather_cols_kernel, dim3(unsigned((int64_t(m) * k + 255) / 256)), dim3(256), 0,
                         ctx->stream, other, m, int64_t(m) * k, dperm, dinv, k,
                         static_cast<double *>((*U)->ptr), int64_t(m) * k, cplx);
    }
  }
  TNH_HIP(hipGetLastError());
  if (S) {
    int64_t dS[1] = {k};
    buf_ensure(ctx, S, TNH_F64, 1, dS);
    to_device(ctx, (*S)->ptr, sv.data(), sizeof(double) * k);
  }
  if (k_out) *k_out = k;
  if (disc_out) *disc_out = o.discarded;
  if (fb_out) *fb_out = o.fallback ? 1 : 0;
}

// theta = A * B over the shared bond for a rank-1 pair of site tensors, output (k, d, chi)
// as a plain matrix product (rows of A fused, columns of B fused).
static void matmul_sites(tnh_ctx *ctx, const tnh_buf *A, int64_t arows, const tnh_buf *B,
                         int64_t bcols, int nd, const int64_t *dims, tnh_buf **out) {
  const int64_t kk = A->numel / arows;
  TNH_REQUIRE(kk * bcols == B->numel, TNH_ERR_SHAPE, "matmul_sites: bond mismatch");
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
  a.dtype = real_of(A->dtype);
  b.re = B->re();
  b.im = is_complex(B->dtype) ? B->im() : nullptr;
  b.dtype = real_of(B->dtype);
  try {
    cgemm(ctx, g, a, b, res->re(), c ? res->im() : nullptr, TNH_F64, TNH_PREC_F64);
  } catch (...) {
    buf_delete(ctx, res);
    throw;
  }
  buf_delete(ctx, *out);
  *out = res;
}

// moveright!(psi, idx), src/structures/mps/gmps.jl:75-82: A -> U, B -> (S V) B.
void mps_moveright(tnh_ctx *ctx, tnh_buf **A, tnh_buf **B, double cutoff, int maxdim, int mindim) {
  const int64_t ca = (*A)->dims[0], d = (*A)->dims[1], cb = (*A)->dims[2];
  TNH_REQUIRE((*B)->ndim == 3 && (*B)->dims[0] == cb, TNH_ERR_SHAPE, "moveright: bond mismatch");
  SvdOut o;
  tnh_buf *U = nullptr, *SV = nullptr;
  int64_t ld[3] = {ca, d, 0}, rd[2] = {0, cb};
  try {
    split_matrix(ctx, *A, int(ca * d), int(cb), 0, cutoff, maxdim, mindim, false, 3, ld, 2, rd,
                 &U, &SV, &o);
    int64_t nb[3] = {int64_t(o.k), (*B)->dims[1], (*B)->dims[2]};
    matmul_sites(ctx, SV, o.k, *B, (*B)->dims[1] * (*B)->dims[2], 3, nb, B);
  } catch (...) {
    buf_delete(ctx, U);
    buf_delete(ctx, SV);
    throw;
  }
  buf_delete(ctx, SV);
  buf_delete(ctx, *A);
  *A = U;
}

// moveleft!(psi, idx), src/structures/mps/gmps.jl:60-67: B -> V^H, A -> A (U S).
void mps_moveleft(tnh_ctx *ctx, tnh_buf **A, tnh_buf **B, double cutoff, int maxdim, int mindim) {
  const int64_t ca = (*B)->dims[0], d = (*B)->dims[1], cb = (*B)->dims[2];
  TNH_REQUIRE((*A)->ndim == 3 && (*A)->dims[2] == ca, TNH_ERR_SHAPE, "moveleft: bond mismatch");
  SvdOut o;
  tnh_buf *US = nullptr, *Vh = nullptr;
  int64_t ld[2] = {ca, 0}, rd[3] = {0, d, cb};
  try {
    split_matrix(ctx, *B, int(ca), int(d * cb), 1, cutoff, maxdim, mindim, false, 2, ld, 3, rd,
                 &US, &Vh, &o);
    int64_t na[3] = {(*A)->dims[0], (*A)->dims[1], int64_t(o.k)};
    matmul_sites(ctx, *A, (*A)->dims[0] * (*A)->dims[1], US, o.k, 3, na, A);
  } catch (...) {
    buf_delete(ctx, US);
    buf_delete(ctx, Vh);
    throw;
  }
  buf_delete(ctx, US);
  buf_delete(ctx, *B);
  *B = Vh;
}

// movecenter!(psi, idx; kwargs), src/structures/mps/gmps.jl:90-112 (1-based sites).
void mps_movecenter(tnh_ctx *ctx, int N, tnh_buf **s, int center, int idx, double cutoff,
                    int maxdim, int mindim) {
  TNH_REQUIRE(idx >= 1 && idx <= N, TNH_ERR_ARG, "The index is out of range.");
  for (int i = 0; i < N; ++i)
    TNH_REQUIRE(s[i] && s[i]->ndim == 3, TNH_ERR_SHAPE, "site tensors must be (chi_l, d, chi_r)");
  auto right = [&](int i) { mps_moveright(ctx, &s[i - 1], &s[i], cutoff, maxdim, mindim); };
  auto left = [&](int i) { mps_moveleft(ctx, &s[i - 2], &s[i - 1], cutoff, maxdim, mindim); };
  if (center == 0) {
    for (int i = 1; i < idx; ++i) right(i);
    for (int i = 1; i <= N - idx; ++i) left(N + 1 - i);
  } else if (idx > center) {
    for (int i = center; i < idx; ++i) right(i);
  } else if (idx < center) {
    for (int i = 1; i <= center - idx; ++i) left(center + 1 - i);
  }
}

}  // namespace tnh

using namespace tnh;

extern "C" {

int tnh_mps_movecenter(tnh_ctx *ctx, int N, tnh_buf **sites, int center, int new_center,
                       double cutoff, int maxdim, int mindim) {
  if (!ctx || !sites) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(N >= 1 && center >= 0 && center <= N, TNH_ERR_ARG, "bad N or centre");
    mps_movecenter(ctx, N, sites, center, new_center, cutoff, maxdim, mindim);
  });
}

int tnh_svd_trunc(tnh_ctx *ctx, const tnh_buf *M, double cutoff, int maxdim, int mindim,
                  int force_fallback, tnh_buf **U, tnh_buf **S, tnh_buf **Vt, int *k,
                  double *discarded, int *fallback) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(maxdim >= 0 && mindim >= 0, TNH_ERR_ARG, "maxdim and mindim must be >= 0");
    svd_trunc(ctx, M, cutoff, maxdim, mindim, force_fallback != 0, U, S, Vt, k, discarded, fallback);
  });
}

int tnh_replacesites2(tnh_ctx *ctx, const tnh_buf *theta, int direction, int normalize,
                      double cutoff, int maxdim, int mindim, tnh_buf **A1, tnh_buf **A2, int *k,
                      double *discarded, double *schmidt, int schmidt_cap) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(maxdim >= 0 && mindim >= 0, TNH_ERR_ARG, "maxdim and mindim must be >= 0");
    replacesites2(ctx, theta, direction, normalize, cutoff, maxdim, mindim, A1, A2, k, discarded,
                  schmidt, schmidt_cap, false);
  });
}

}  // extern "C"
