// Batched quantum-jump trajectories: every trajectory of qjmc_simulation
// (src/algorithms/mps/qjmc.jl:28-167, classical branch, single-site jump operators) advances
// together, one workgroup per trajectory in each launch. The site tensors of all trajectories
// live in fixed-capacity slots in HBM (bond dimension at most CH = 32 / d, so every Gram
// matrix of a split is at most 32 x 32 and is diagonalised in LDS by the cyclic Jacobi of
// jacobi_lds.hpp). The host keeps each trajectory's orthogonality centre and turns the
// reference's step (applygates!, gatelist.jl:191-223; normalize!; the emission rates of
// :170-220; the jump of :103-107) into a queue of primitives per trajectory: centre moves
// (gmps.jl:90-112), two-site gates (gatelist.jl:137-171) and one-site operators. A round
// launches one kernel per primitive kind over the trajectories whose queue front is of that
// kind, so trajectories with different centres still advance in the same launches. Uniforms
// are consumed per trajectory in the reference's order (one discarded draw, the jump test,
// the channel). A split whose kept rank would exceed CH, or whose Jacobi does not converge,
// stops the run with an error instead of truncating differently from the reference.
#include <algorithm>
#include <cmath>
#include <complex>
#include <deque>
#include <vector>

#include "jacobi_lds.hpp"
#include "tnh_internal.hpp"

namespace tnh {

namespace {

using jac::JB;
using jac::JP;
using jac::kMaxInnerCap;

constexpr int QT = 256;    // threads per trajectory workgroup
constexpr int QM = JP;     // largest matrix side of a split (32)
constexpr int QMM = QM * QM;

enum Prim { P_RIGHT = 0, P_LEFT = 1, P_GATE = 2, P_SITEOP = 3, P_NORM = 4 };
enum Flag { F_OVERFLOW = 1, F_NOCONV = 2, F_NUMERIC = 4 };

struct Batch {
  double *S;        // site slots: [traj][site][re CAP | im CAP]
  int *bond;        // [traj][N + 1], bond[0] = bond[N] = 1
  int *flags;       // [traj]
  const double *op; // operator store, interleaved complex
  int N, d, CH;
  int64_t cap;      // CH * d * CH
  double cutoff;
  int maxdim, mindim;
};

__device__ inline double *slot(const Batch &b, int t, int site) {
  return b.S + (int64_t(t) * b.N + (site - 1)) * 2 * b.cap;
}

// The reference's truncation rule (src/tensors.jl:202-215) on the descending values s[0..r).
__device__ int trunc_rank(const double *s, int r, double cutoff, int maxdim, int mindim, double *disc) {
  mindim = min(mindim, r);
  maxdim = (maxdim == 0 || maxdim > r) ? r : maxdim;
  maxdim = maxdim == 0 ? 1 : maxdim;
  double tot = 0;
  for (int i = 0; i < r; ++i) tot += s[i] * s[i];
  if (cutoff != 0) {
    double acc = 0;
    int idxs = 0;
    for (int i = r - 1; i >= 0; --i) {
      acc += s[i] * s[i];
      if (idxs == 0 && acc / tot > cutoff) idxs = i + 1;
    }
    if (idxs == 0) idxs = 1;
    maxdim = min(maxdim, idxs);
  }
  const int k = max(maxdim, mindim);
  double dd = 0;
  for (int i = k; i < r; ++i) dd += s[i] * s[i];
  *disc = tot > 0 ? dd / tot : 0.0;
  return k;
}

constexpr int QG = QM * (QM + 1);  // one padded Jacobi matrix plane
// Factors and site tensors have one side <= CH <= 16 and the other <= 32 (d CH <= 32), so
// 512 entries hold them; with the 32 x 32 split matrix and the Jacobi planes the workgroup
// needs ~70 KB of LDS, two workgroups per CU.
constexpr int QAB = QMM / 2;

struct Lds {
  double Mr[QMM], Mi[QMM];  // the matrix being split, column-major m x n
  double Ar[QAB], Ai[QAB];  // left factor (m x k), a site tensor, Newton-Schulz W (p x k)
  double Br[QAB], Bi[QAB];  // right factor (k x n), a site tensor, Newton-Schulz X (k x k)
  double gr[QG], gi[QG], qr[QG], qi[QG];  // Jacobi planes, row pitch 2B + 1 of the chosen width
  double val[QM], sv[QM];  // singular values by index, and in descending order
  double red[QT / 64];
  int perm[QM];
  int k;
  int flag;
};

// Truncated split of M (m x n, in L.Mr/Mi) as in split_matrix of the SVD module, side 0:
// A = U_k (orthonormal columns, m x k), B = U_k^H M; side 1: B = V_k^H (orthonormal rows,
// k x n), A = M V_k. Gram of the basis side (p = m or n <= 2B), made exactly Hermitian,
// power-of-two rescaling, cyclic Jacobi to convergence on the p x p problem padded to
// 2B x 2B, stable descending order, the truncation rule, one Newton-Schulz step on the kept
// basis. Returns k (0 on failure, with L.flag set).
template <int B>
__device__ int split_lds_w(Lds &L, int m, int n, int side, double cutoff, int maxdim, int mindim, int CH) {
  constexpr int PB = 2 * B;
  typedef double Row[PB + 1];
  Row *gr = reinterpret_cast<Row *>(L.gr), *gi = reinterpret_cast<Row *>(L.gi);
  Row *qr = reinterpret_cast<Row *>(L.qr), *qi = reinterpret_cast<Row *>(L.qi);
  const int t = threadIdx.x;
  const int p = side == 0 ? m : n;
  // G = M M^H (side 0) or M^H M (side 1), zero-padded to PB x PB
  for (int e = t; e < PB * PB; e += QT) {
    const int i = e % PB, j = e / PB;
    double sr = 0, si = 0;
    if (i < p && j < p) {
      if (side == 0) {
        for (int c = 0; c < n; ++c) {  // M[i,c] conj(M[j,c])
          const double ar = L.Mr[i + m * c], ai = L.Mi[i + m * c];
          const double br = L.Mr[j + m * c], bi = L.Mi[j + m * c];
          sr += ar * br + ai * bi;
          si += ai * br - ar * bi;
        }
      } else {
        for (int r = 0; r < m; ++r) {  // conj(M[r,i]) M[r,j]
          const double ar = L.Mr[r + m * i], ai = L.Mi[r + m * i];
          const double br = L.Mr[r + m * j], bi = L.Mi[r + m * j];
          sr += ar * br + ai * bi;
          si += ar * bi - ai * br;
        }
      }
    }
    gr[i][j] = sr;
    gi[i][j] = (i == j) ? 0.0 : si;
    qr[i][j] = (i == j) ? 1.0 : 0.0;
    qi[i][j] = 0.0;
  }
  __syncthreads();
  // exactly Hermitian, (G + G^H) / 2, as the single-workgroup eigensolve of the SVD module
  for (int e = t; e < PB * PB; e += QT) {
    const int i = e % PB, j = e / PB;
    if (i >= j) continue;
    const double re = 0.5 * (gr[i][j] + gr[j][i]), im = 0.5 * (gi[i][j] - gi[j][i]);
    gr[i][j] = gr[j][i] = re;
    gi[i][j] = im;
    gi[j][i] = -im;
  }
  __syncthreads();
  double dm = t < PB ? fabs(gr[t][t]) : 0.0;
  dm = jac::block_max<QT>(dm);
  const int ex = (dm > 0 && isfinite(dm)) ? ilogb(dm) : 0;
  if (ex != 0) {
    const double sc = ldexp(1.0, -ex);
    for (int e = t; e < PB * PB; e += QT) {
      gr[e % PB][e / PB] *= sc;
      gi[e % PB][e / PB] *= sc;
    }
    dm *= sc;
    __syncthreads();
  }
  const double eps = 2.220446049250313e-16;
  const double tol = fmax(1e-14, 4.0 * eps * sqrt(double(p)));
  const double abstol = 4.0 * eps * sqrt(double(p)) * dm;
  const bool conv = jac::jacobi_lds<true, B, QT>(gr, gi, qr, qi, tol, abstol, kMaxInnerCap);
  if (t < p) L.val[t] = sqrt(fmax(ldexp(gr[t][t], ex), 0.0));
  __syncthreads();
  if (t < p) {  // stable descending rank
    const double v = L.val[t];
    int rk = 0;
    for (int j = 0; j < p; ++j) rk += (L.val[j] > v || (L.val[j] == v && j < t)) ? 1 : 0;
    L.perm[rk] = t;
  }
  __syncthreads();
  if (t == 0) {
    int ok = conv ? 1 : 0;
    for (int i = 0; i < p; ++i) {
      L.sv[i] = L.val[L.perm[i]];
      if (!isfinite(L.sv[i])) ok = 0;
    }
    double disc = 0;
    int k = ok ? trunc_rank(L.sv, min(m, n), cutoff, maxdim, mindim, &disc) : 0;
    if (!conv) L.flag |= F_NOCONV;
    if (conv && !ok) L.flag |= F_NUMERIC;
    if (ok && k > CH) {
      L.flag |= F_OVERFLOW;
      ok = 0;
    }
    L.k = ok ? k : 0;
  }
  __syncthreads();
  const int k = L.k;
  if (k == 0) return 0;
  // basis U (p x k) = columns perm[0..k) of Q, kept in gr/gi (Q is only read, G no longer)
  for (int e = t; e < p * k; e += QT) {
    const int r = e % p, j = e / p, c = L.perm[j];
    gr[r][j] = qr[r][c];
    gi[r][j] = qi[r][c];
  }
  __syncthreads();
  // Newton-Schulz: X = U^H U (k x k, in B), W = U X (p x k, in A), U <- 1.5 U - 0.5 W
  for (int e = t; e < k * k; e += QT) {
    const int i = e % k, j = e / k;
    double sr = 0, si = 0;
    for (int r = 0; r < p; ++r) {  // conj(U[r,i]) U[r,j]
      const double ar = gr[r][i], ai = gi[r][i], br = gr[r][j], bi = gi[r][j];
      sr += ar * br + ai * bi;
      si += ar * bi - ai * br;
    }
    L.Br[e] = sr;
    L.Bi[e] = si;
  }
  __syncthreads();
  for (int e = t; e < p * k; e += QT) {
    const int r = e % p, j = e / p;
    double sr = 0, si = 0;
    for (int l = 0; l < k; ++l) {  // U[r,l] X[l,j]
      const double ar = gr[r][l], ai = gi[r][l], br = L.Br[l + k * j], bi = L.Bi[l + k * j];
      sr += ar * br - ai * bi;
      si += ar * bi + ai * br;
    }
    L.Ar[e] = sr;
    L.Ai[e] = si;
  }
  __syncthreads();
  for (int e = t; e < p * k; e += QT) {
    const int r = e % p, j = e / p;
    gr[r][j] = 1.5 * gr[r][j] - 0.5 * L.Ar[e];
    gi[r][j] = 1.5 * gi[r][j] - 0.5 * L.Ai[e];
  }
  __syncthreads();
  if (side == 0) {
    // A = U (m x k); B = U^H M (k x n)
    for (int e = t; e < m * k; e += QT) {
      L.Ar[e] = gr[e % m][e / m];
      L.Ai[e] = gi[e % m][e / m];
    }
    for (int e = t; e < k * n; e += QT) {
      const int j = e % k, c = e / k;
      double sr = 0, si = 0;
      for (int r = 0; r < m; ++r) {  // conj(U[r,j]) M[r,c]
        const double ar = gr[r][j], ai = gi[r][j], br = L.Mr[r + m * c], bi = L.Mi[r + m * c];
        sr += ar * br + ai * bi;
        si += ar * bi - ai * br;
      }
      L.Br[e] = sr;
      L.Bi[e] = si;
    }
  } else {
    // B = V_k^H (k x n); A = M V_k (m x k)
    for (int e = t; e < k * n; e += QT) {
      const int j = e % k, c = e / k;
      L.Br[e] = gr[c][j];
      L.Bi[e] = -gi[c][j];
    }
    for (int e = t; e < m * k; e += QT) {
      const int r = e % m, j = e / m;
      double sr = 0, si = 0;
      for (int c = 0; c < n; ++c) {  // M[r,c] V[c,j]
        const double ar = L.Mr[r + m * c], ai = L.Mi[r + m * c], br = gr[c][j], bi = gi[c][j];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      L.Ar[e] = sr;
      L.Ai[e] = si;
    }
  }
  __syncthreads();
  return k;
}

// the Jacobi width follows the basis dimension: padded sides 4, 8, 16 or 32
__device__ int split_lds(Lds &L, int m, int n, int side, double cutoff, int maxdim, int mindim, int CH) {
  const int p = side == 0 ? m : n;
  if (p <= 4) return split_lds_w<2>(L, m, n, side, cutoff, maxdim, mindim, CH);
  if (p <= 8) return split_lds_w<4>(L, m, n, side, cutoff, maxdim, mindim, CH);
  if (p <= 16) return split_lds_w<8>(L, m, n, side, cutoff, maxdim, mindim, CH);
  return split_lds_w<16>(L, m, n, side, cutoff, maxdim, mindim, CH);
}

__device__ inline void load_slot(const double *src, int64_t cap, int numel, double *re, double *im) {
  for (int e = threadIdx.x; e < numel; e += QT) {
    re[e] = src[e];
    im[e] = src[cap + e];
  }
}

__device__ inline void store_slot(double *dst, int64_t cap, int numel, const double *re, const double *im) {
  for (int e = threadIdx.x; e < numel; e += QT) {
    dst[e] = re[e];
    dst[cap + e] = im[e];
  }
}

// Two-site gate at (s, s + 1): theta = A_s A_{s+1}, theta'[a,o1,o2,b] = sum G[o1,i1,o2,i2]
// theta[a,i1,i2,b], split with side = direction (gatelist.jl:137-171, no normalisation).
__device__ void prim_gate(const Batch &b, Lds &L, int tr, int s, const double *G, int dir) {
  const int d = b.d, t = threadIdx.x;
  int *bd = b.bond + int64_t(tr) * (b.N + 1);
  const int Dl = bd[s - 1], Dm = bd[s], Dr = bd[s + 1];
  double *sa = slot(b, tr, s), *sb = slot(b, tr, s + 1);
  load_slot(sa, b.cap, Dl * d * Dm, L.Ar, L.Ai);
  load_slot(sb, b.cap, Dm * d * Dr, L.Br, L.Bi);
  __syncthreads();
  const int m = Dl * d, n = d * Dr;
  // theta[(a,i1),(i2,b)] = sum_c A[(a,i1),c] B[c,(i2,b)], into the Jacobi planes viewed flat
  double *Tr = L.gr, *Ti = L.gi;
  for (int e = t; e < m * n; e += QT) {
    const int r = e % m, c = e / m;
    double sr = 0, si = 0;
    for (int l = 0; l < Dm; ++l) {
      const double ar = L.Ar[r + m * l], ai = L.Ai[r + m * l], br = L.Br[l + Dm * c], bi = L.Bi[l + Dm * c];
      sr += ar * br - ai * bi;
      si += ar * bi + ai * br;
    }
    Tr[e] = sr;
    Ti[e] = si;
  }
  __syncthreads();
  // M[(a,o1),(o2,b)] = sum_{i1,i2} G[o1,i1,o2,i2] theta[(a,i1),(i2,b)]
  for (int e = t; e < m * n; e += QT) {
    const int r = e % m, c = e / m;
    const int a = r % Dl, o1 = r / Dl, o2 = c % d, bb = c / d;
    double sr = 0, si = 0;
    for (int i2 = 0; i2 < d; ++i2)
      for (int i1 = 0; i1 < d; ++i1) {
        const int64_t q = o1 + int64_t(d) * (i1 + int64_t(d) * (o2 + int64_t(d) * i2));
        const double gr_ = G[2 * q], gi_ = G[2 * q + 1];
        const int te = (a + Dl * i1) + m * (i2 + d * bb);
        sr += gr_ * Tr[te] - gi_ * Ti[te];
        si += gr_ * Ti[te] + gi_ * Tr[te];
      }
    L.Mr[e] = sr;
    L.Mi[e] = si;
  }
  __syncthreads();
  const int k = split_lds(L, m, n, dir ? 1 : 0, b.cutoff, b.maxdim, b.mindim, b.CH);
  if (k == 0) return;
  store_slot(sa, b.cap, m * k, L.Ar, L.Ai);
  store_slot(sb, b.cap, k * n, L.Br, L.Bi);
  if (t == 0) bd[s] = k;
  __syncthreads();
}

// Centre move by one site: DIR 0 from s to s + 1 (split A_s with side 0, absorb U^H A_s into
// A_{s+1}); DIR 1 from s to s - 1 (split A_s with side 1, absorb A_s V_k into A_{s-1}).
// trunc = 0 is movecenter! without cutoff or maxdim.
__device__ void prim_move(const Batch &b, Lds &L, int tr, int s, int dir, int trunc) {
  const int d = b.d, t = threadIdx.x;
  int *bd = b.bond + int64_t(tr) * (b.N + 1);
  const int Dl = bd[s - 1], Dr = bd[s];
  double *sa = slot(b, tr, s);
  const int m = dir == 0 ? Dl * d : Dl, n = dir == 0 ? Dr : d * Dr;
  load_slot(sa, b.cap, Dl * d * Dr, L.Mr, L.Mi);
  __syncthreads();
  const int k = split_lds(L, m, n, dir, trunc ? b.cutoff : 0.0, trunc ? b.maxdim : 0, trunc ? b.mindim : 1,
                          b.CH);
  if (k == 0) return;
  double *Pr = L.gr, *Pi = L.gi;
  if (dir == 0) {
    // A_{s+1} <- B (k x Dr) A_{s+1} (Dr x d Dr2)
    const int Dr2 = bd[s + 1];
    double *sn = slot(b, tr, s + 1);
    load_slot(sn, b.cap, Dr * d * Dr2, L.Mr, L.Mi);
    __syncthreads();
    const int nn = d * Dr2;
    for (int e = t; e < k * nn; e += QT) {
      const int j = e % k, c = e / k;
      double sr = 0, si = 0;
      for (int l = 0; l < Dr; ++l) {
        const double ar = L.Br[j + k * l], ai = L.Bi[j + k * l], br = L.Mr[l + Dr * c], bi = L.Mi[l + Dr * c];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      Pr[e] = sr;
      Pi[e] = si;
    }
    __syncthreads();
    store_slot(sa, b.cap, m * k, L.Ar, L.Ai);
    store_slot(sn, b.cap, k * nn, Pr, Pi);
    if (t == 0) bd[s] = k;
  } else {
    // A_{s-1} <- A_{s-1} (Dl2 d x Dl) A (Dl x k)
    const int Dl2 = bd[s - 2];
    double *sp = slot(b, tr, s - 1);
    load_slot(sp, b.cap, Dl2 * d * Dl, L.Mr, L.Mi);
    __syncthreads();
    const int mm = Dl2 * d;
    for (int e = t; e < mm * k; e += QT) {
      const int r = e % mm, j = e / mm;
      double sr = 0, si = 0;
      for (int l = 0; l < Dl; ++l) {
        const double ar = L.Mr[r + mm * l], ai = L.Mi[r + mm * l], br = L.Ar[l + Dl * j], bi = L.Ai[l + Dl * j];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      Pr[e] = sr;
      Pi[e] = si;
    }
    __syncthreads();
    store_slot(sa, b.cap, k * n, L.Br, L.Bi);
    store_slot(sp, b.cap, mm * k, Pr, Pi);
    if (t == 0) bd[s - 1] = k;
  }
  __syncthreads();
}

// A[a,o,b] <- sum_i O[o,i] A[a,i,b] (applyop!, mps.jl:141-151, as tnh_apply_site_op).
__device__ void prim_siteop(const Batch &b, Lds &L, int tr, int s, const double *O) {
  const int d = b.d, t = threadIdx.x;
  const int *bd = b.bond + int64_t(tr) * (b.N + 1);
  const int Dl = bd[s - 1], Dr = bd[s];
  double *sa = slot(b, tr, s);
  const int ne = Dl * d * Dr;
  load_slot(sa, b.cap, ne, L.Mr, L.Mi);
  __syncthreads();
  for (int e = t; e < ne; e += QT) {
    const int a = e % Dl, oo = (e / Dl) % d, bb = e / (Dl * d);
    double sr = 0, si = 0;
    for (int i = 0; i < d; ++i) {
      const double gr_ = O[2 * (oo + d * i)], gi_ = O[2 * (oo + d * i) + 1];
      const int x = a + Dl * (i + d * bb);
      sr += gr_ * L.Mr[x] - gi_ * L.Mi[x];
      si += gr_ * L.Mi[x] + gi_ * L.Mr[x];
    }
    sa[e] = sr;
    sa[b.cap + e] = si;
  }
  __syncthreads();
}

// A_c <- A_c / ||A_c|| (normalize!, gmps.jl:46-51). Returns false on a zero or non-finite norm.
__device__ bool prim_normalize(const Batch &b, Lds &L, int tr, int s) {
  const int d = b.d, t = threadIdx.x;
  const int *bd = b.bond + int64_t(tr) * (b.N + 1);
  const int ne = bd[s - 1] * d * bd[s];
  double *sa = slot(b, tr, s);
  double acc = 0;
  for (int e = t; e < ne; e += QT) acc += sa[e] * sa[e] + sa[b.cap + e] * sa[b.cap + e];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((t & 63) == 0) L.red[t >> 6] = acc;
  __syncthreads();
  double tot = 0;
  for (int w = 0; w < QT / 64; ++w) tot += L.red[w];
  __syncthreads();
  const double nr = sqrt(tot);
  if (!(nr > 0) || !isfinite(nr)) return false;
  const double inv = 1.0 / nr;
  for (int e = t; e < ne; e += QT) {
    sa[e] *= inv;
    sa[b.cap + e] *= inv;
  }
  __syncthreads();
  return true;
}

// One workgroup per listed trajectory runs that trajectory's whole primitive list
// items[start[w] .. start[w + 1]) in order ({kind, site, arg, flag} each): trajectories are
// independent, so a phase of the step is a single launch. A failed split or normalisation
// records its flag and ends that trajectory's list.
__global__ __launch_bounds__(QT) void qb_run_kernel(Batch b, const int *__restrict__ trajs,
                                                     const int *__restrict__ start,
                                                     const int *__restrict__ items) {
  __shared__ Lds L;
  const int w = blockIdx.x, tr = trajs[w], t = threadIdx.x;
  if (t == 0) L.flag = b.flags[tr];
  __syncthreads();
  if (L.flag) return;
  for (int i = start[w]; i < start[w + 1]; ++i) {
    const int *it = items + 4 * i;
    const int kind = it[0], s = it[1], arg = it[2], fl = it[3];
    bool ok = true;
    switch (kind) {
      case P_RIGHT:
        prim_move(b, L, tr, s, 0, fl);
        break;
      case P_LEFT:
        prim_move(b, L, tr, s, 1, fl);
        break;
      case P_GATE:
        prim_gate(b, L, tr, s, b.op + 2 * int64_t(arg), fl);
        break;
      case P_SITEOP:
        prim_siteop(b, L, tr, s, b.op + 2 * int64_t(arg));
        break;
      default:
        ok = prim_normalize(b, L, tr, s);
    }
    if (t == 0 && !ok) L.flag |= F_NUMERIC;
    __syncthreads();
    if (L.flag) {
      if (t == 0) b.flags[tr] = L.flag;
      return;
    }
  }
}

// <psi|O|psi> for every (site, operator) term and every listed trajectory: left norm
// environments E_1..E_{N-1} into env scratch, then a right-to-left sweep carrying F, each
// term contracted from E_{s-1}, A_s, O and F_{s+1} (inner(st, psi, oplist, psi)).
// out[(traj list index) * nterms + term] interleaved complex.
__global__ __launch_bounds__(QT) void qb_expect_kernel(Batch b, const int *__restrict__ trajs, int nterms,
                                                        const int *__restrict__ term_site,
                                                        const int64_t *__restrict__ term_off,
                                                        double *__restrict__ env, double *__restrict__ out) {
  __shared__ double Er[QMM / 4], Ei[QMM / 4], Fr[QMM / 4], Fi[QMM / 4];
  __shared__ double Ar[QMM], Ai[QMM], Tr[QMM], Ti[QMM], Ur[QMM], Ui[QMM];
  __shared__ double red[2][QT / 64];
  const int tr = trajs[blockIdx.x], d = b.d, t = threadIdx.x, N = b.N;
  const int *bd = b.bond + int64_t(tr) * (N + 1);
  const int CC = b.CH * b.CH;
  double *envt = env + int64_t(blockIdx.x) * (N + 1) * 2 * CC;  // E_i at envt + i * 2 CC
  // E_0 = [1]
  if (t == 0) {
    envt[0] = 1.0;
    envt[CC] = 0.0;
  }
  __syncthreads();
  for (int s = 1; s < N; ++s) {
    const int Dl = bd[s - 1], Dr = bd[s];
    load_slot(envt + int64_t(s - 1) * 2 * CC, CC, Dl * Dl, Er, Ei);
    load_slot(slot(b, tr, s), b.cap, Dl * d * Dr, Ar, Ai);
    __syncthreads();
    // T[x,(sg,y')] = sum_y E[x,y] A[y,(sg,y')]
    for (int e = t; e < Dl * d * Dr; e += QT) {
      const int x = e % Dl, c = e / Dl;
      double sr = 0, si = 0;
      for (int y = 0; y < Dl; ++y) {
        const double ar = Er[x + Dl * y], ai = Ei[x + Dl * y], br = Ar[y + Dl * c], bi = Ai[y + Dl * c];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      Tr[e] = sr;
      Ti[e] = si;
    }
    __syncthreads();
    // E_s[x',y'] = sum_{x,sg} conj(A[x,sg,x']) T[x,sg,y']
    double *En = envt + int64_t(s) * 2 * CC;
    for (int e = t; e < Dr * Dr; e += QT) {
      const int xp = e % Dr, yp = e / Dr;
      double sr = 0, si = 0;
      for (int sg = 0; sg < d; ++sg)
        for (int x = 0; x < Dl; ++x) {
          const int ia = x + Dl * (sg + d * xp), it = x + Dl * (sg + d * yp);
          const double ar = Ar[ia], ai = -Ai[ia], br = Tr[it], bi = Ti[it];
          sr += ar * br - ai * bi;
          si += ar * bi + ai * br;
        }
      En[e] = sr;
      En[CC + e] = si;
    }
    __syncthreads();
  }
  // right sweep: F_{N+1} = [1]
  if (t == 0) {
    Fr[0] = 1.0;
    Fi[0] = 0.0;
  }
  __syncthreads();
  for (int s = N; s >= 1; --s) {
    const int Dl = bd[s - 1], Dr = bd[s];
    const int na = Dl * d * Dr;
    load_slot(envt + int64_t(s - 1) * 2 * CC, CC, Dl * Dl, Er, Ei);
    load_slot(slot(b, tr, s), b.cap, na, Ar, Ai);
    __syncthreads();
    // T[x,(tau,y')] = sum_y E[x,y] A[y,(tau,y')]
    for (int e = t; e < na; e += QT) {
      const int x = e % Dl, c = e / Dl;
      double sr = 0, si = 0;
      for (int y = 0; y < Dl; ++y) {
        const double ar = Er[x + Dl * y], ai = Ei[x + Dl * y], br = Ar[y + Dl * c], bi = Ai[y + Dl * c];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      Tr[e] = sr;
      Ti[e] = si;
    }
    // U[x,(sg,x')] = sum_{y'} conj(A[x,sg,x']) ... formed per term below; here W[x,sg,y'] =
    // sum_{x'} conj(A[x,sg,x']) F[x',y']
    for (int e = t; e < na; e += QT) {
      const int x = e % Dl, sg = (e / Dl) % d, yp = e / (Dl * d);
      double sr = 0, si = 0;
      for (int xp = 0; xp < Dr; ++xp) {
        const int ia = x + Dl * (sg + d * xp);
        const double ar = Ar[ia], ai = -Ai[ia], br = Fr[xp + Dr * yp], bi = Fi[xp + Dr * yp];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      Ur[e] = sr;
      Ui[e] = si;
    }
    __syncthreads();
    for (int q = 0; q < nterms; ++q) {
      if (term_site[q] != s) continue;  // uniform across the workgroup
      const double *O = b.op + 2 * term_off[q];
      // val = sum_{x,sg,y'} W[x,sg,y'] sum_tau O[sg,tau] T[x,tau,y']
      double vr = 0, vi = 0;
      for (int e = t; e < na; e += QT) {
        const int x = e % Dl, sg = (e / Dl) % d, yp = e / (Dl * d);
        double pr = 0, pi = 0;
        for (int tau = 0; tau < d; ++tau) {
          const double gr_ = O[2 * (sg + d * tau)], gi_ = O[2 * (sg + d * tau) + 1];
          const int it = x + Dl * (tau + d * yp);
          pr += gr_ * Tr[it] - gi_ * Ti[it];
          pi += gr_ * Ti[it] + gi_ * Tr[it];
        }
        vr += Ur[e] * pr - Ui[e] * pi;
        vi += Ur[e] * pi + Ui[e] * pr;
      }
      for (int off = 32; off > 0; off >>= 1) {
        vr += __shfl_down(vr, off, 64);
        vi += __shfl_down(vi, off, 64);
      }
      if ((t & 63) == 0) {
        red[0][t >> 6] = vr;
        red[1][t >> 6] = vi;
      }
      __syncthreads();
      if (t == 0) {
        double sr = 0, si = 0;
        for (int w = 0; w < QT / 64; ++w) {
          sr += red[0][w];
          si += red[1][w];
        }
        out[2 * (int64_t(blockIdx.x) * nterms + q)] = sr;
        out[2 * (int64_t(blockIdx.x) * nterms + q) + 1] = si;
      }
      __syncthreads();
    }
    // F_s[x,y] = sum_{sg,y'} W[x,sg,y'] A[y,sg,y']
    __syncthreads();
    for (int e = t; e < Dl * Dl; e += QT) {
      const int x = e % Dl, y = e / Dl;
      double sr = 0, si = 0;
      for (int sg = 0; sg < d; ++sg)
        for (int yp = 0; yp < Dr; ++yp) {
          const int iw = x + Dl * (sg + d * yp), ia = y + Dl * (sg + d * yp);
          sr += Ur[iw] * Ar[ia] - Ui[iw] * Ai[ia];
          si += Ur[iw] * Ai[ia] + Ui[iw] * Ar[ia];
        }
      Tr[e] = sr;
      Ti[e] = si;
    }
    __syncthreads();
    for (int e = t; e < Dl * Dl; e += QT) {
      Fr[e] = Tr[e];
      Fi[e] = Ti[e];
    }
    __syncthreads();
  }
}

// product state: every slot (1, d, 1) from init (N x d interleaved complex), bonds 1
__global__ void qb_init_kernel(Batch b, const double *__restrict__ init, int ntraj) {
  const int tr = blockIdx.x;
  if (tr >= ntraj) return;
  for (int e = threadIdx.x; e < b.N * b.d; e += blockDim.x) {
    const int s = e / b.d, i = e % b.d;
    double *sa = slot(b, tr, s + 1);
    sa[i] = init[2 * e];
    sa[b.cap + i] = init[2 * e + 1];
  }
  for (int e = threadIdx.x; e <= b.N; e += blockDim.x) b.bond[int64_t(tr) * (b.N + 1) + e] = 1;
  if (threadIdx.x == 0) b.flags[tr] = 0;
}

struct Item {
  int kind, site, arg, flag;
};

// Host side: one trajectory's centre and queue of primitives.
struct TrajQ {
  int center = 0;
  std::deque<Item> q;
};

struct Engine {
  tnh_ctx *ctx;
  const tnh_qjmc_problem *p;
  int T;
  Batch b{};
  std::vector<TrajQ> tq;
  std::vector<int64_t> gate_off;  // operator-store offsets (complex elements) of the gates
  std::vector<int64_t> jump_off, esc_off, obs_off;
  void *ops_dev = nullptr;
  size_t ops_cap = 0;

  // movecenter! (gmps.jl:90-112) as primitives, from the tracked centre
  void push_move(TrajQ &x, int idx, int trunc) {
    const int N = p->N;
    auto right = [&](int i) { x.q.push_back({P_RIGHT, i, 0, trunc}); };
    auto left = [&](int i) { x.q.push_back({P_LEFT, i, 0, trunc}); };
    if (x.center == 0) {
      for (int i = 1; i < idx; ++i) right(i);
      for (int i = 1; i <= N - idx; ++i) left(N + 1 - i);
    } else if (idx > x.center) {
      for (int i = x.center; i < idx; ++i) right(i);
    } else if (idx < x.center) {
      for (int i = 1; i <= x.center - idx; ++i) left(x.center + 1 - i);
    }
    x.center = idx;
  }

  // applygates! (gatelist.jl:191-223) with applygate (:137-171)
  void push_applygates(TrajQ &x) {
    const int N = p->N;
    int base = 0;
    for (int l = 0; l < p->nlayers; ++l) {
      const int n = p->layer_count[l];
      if (n == 0) continue;
      const int first = p->gate_site[base];
      const int last = p->gate_site[base + n - 1] + p->gate_range[base + n - 1] - 1;
      const bool direction = !(std::abs(x.center - first) < std::abs(x.center - last));
      for (int i = 1; i <= n; ++i) {
        const int gi = direction ? n + 1 - i : i;
        const int g = base + gi - 1;
        const int site = p->gate_site[g];
        const int ctr = direction ? site + p->gate_range[g] - 1 : site;
        push_move(x, ctr, 1);
        if (p->gate_range[g] == 2) {
          x.q.push_back({P_GATE, site, int(gate_off[g]), direction ? 1 : 0});
          x.center = direction ? site : site + 1;
        } else {
          x.q.push_back({P_SITEOP, site, int(gate_off[g]), 0});
          x.center = site;
          const int nxt = site + 1 - 2 * (direction ? 1 : 0);
          if (nxt > 0 && nxt <= N) push_move(x, nxt, 0);
        }
      }
      base += n;
    }
  }

  // run every queue to empty in one launch: each trajectory with work gets a workgroup that
  // walks its own list; the lists of all trajectories go up in one upload
  void run_queues() {
    std::vector<int> host, trajs, start;
    for (int t = 0; t < T; ++t) {
      auto &x = tq[t];
      if (x.q.empty()) continue;
      trajs.push_back(t);
      start.push_back(int(host.size() / 4));
      for (const Item &it : x.q) host.insert(host.end(), {it.kind, it.site, it.arg, it.flag});
      x.q.clear();
    }
    const int nw = int(trajs.size());
    if (nw == 0) return;
    start.push_back(int(host.size() / 4));
    // one buffer: trajs | start | items
    std::vector<int> buf;
    buf.reserve(trajs.size() + start.size() + host.size());
    buf.insert(buf.end(), trajs.begin(), trajs.end());
    buf.insert(buf.end(), start.begin(), start.end());
    buf.insert(buf.end(), host.begin(), host.end());
    const size_t bytes = buf.size() * sizeof(int);
    if (bytes > ops_cap) {  // the caching allocator orders reuse after queued work
      if (ops_dev) ctx->release(ops_dev, ops_cap);
      ops_dev = ctx->alloc(std::max(bytes, ops_cap * 2), &ops_cap);
    }
    to_device(ctx, ops_dev, buf.data(), bytes);
    const int *od = static_cast<const int *>(ops_dev);
    hipLaunchKernelGGL(qb_run_kernel, dim3(nw), dim3(QT), 0, ctx->stream, b, od, od + nw, od + 2 * nw + 1);
    TNH_HIP(hipGetLastError());
  }
};

// Reads the trajectories' flags: a trajectory that outgrew its slots is marked in `dead` (it
// stops; the caller reruns it elsewhere), any other flag is an error of the whole call.
void check_flags(tnh_ctx *ctx, const Engine &E, std::vector<char> &dead) {
  std::vector<int> f(E.T);
  to_host(ctx, f.data(), E.b.flags, sizeof(int) * E.T);
  for (int t = 0; t < E.T; ++t) {
    if (f[t] & F_OVERFLOW) dead[t] = 1;
    TNH_REQUIRE(!(f[t] & F_NOCONV), TNH_ERR_NOCONV,
                "qjmc batched: Gram eigensolve did not converge (trajectory " + std::to_string(t) + ")");
    TNH_REQUIRE(!(f[t] & F_NUMERIC), TNH_ERR_NUMERIC,
                "qjmc batched: non-finite or zero state (trajectory " + std::to_string(t) + ")");
  }
}

}  // namespace

}  // namespace tnh

using namespace tnh;

extern "C" int tnh_qjmc_run_batched(tnh_ctx *ctx, const tnh_qjmc_problem *p, int ntraj,
                                    const double *uniforms, int64_t nuniform, double *meas, int *njumps,
                                    int *jumps, int *jump_steps, int maxjumps, int *maxbond) {
  if (!ctx) return TNH_ERR_ARG;
  std::vector<void *> owned;
  std::vector<size_t> owned_sz;
  Engine eng;
  int rc = tnh_guard(ctx, [&] {
    TNH_REQUIRE(p && ntraj >= 1 && uniforms && njumps && maxbond, TNH_ERR_ARG, "qjmc: bad arguments");
    TNH_REQUIRE(p->N >= 2 && p->d >= 1 && p->nsteps >= 0 && p->savesteps >= 1 && p->dt > 0, TNH_ERR_ARG,
                "qjmc: bad problem sizes");
    TNH_REQUIRE(p->d <= QM / 2, TNH_ERR_UNSUPPORTED, "qjmc batched: d > 16");
    TNH_REQUIRE(p->nobs == 0 || meas, TNH_ERR_ARG, "qjmc: meas is NULL");
    TNH_REQUIRE(maxjumps == 0 || (jumps && jump_steps), TNH_ERR_ARG, "qjmc: jump arrays are NULL");
    for (int t = 0; t < p->njump; ++t)
      TNH_REQUIRE(p->jump_site[t] >= 1 && p->jump_site[t] <= p->N, TNH_ERR_ARG, "qjmc: jump site out of range");
    for (int o = 0; o < p->nobs; ++o)
      TNH_REQUIRE(p->obs_site[o] >= 1 && p->obs_site[o] <= p->N, TNH_ERR_ARG, "qjmc: observable site out of range");
    int ng = 0;
    for (int l = 0; l < p->nlayers; ++l) ng += p->layer_count[l];
    for (int g = 0; g < ng; ++g)
      TNH_REQUIRE((p->gate_range[g] == 1 || p->gate_range[g] == 2) && p->gate_site[g] >= 1 &&
                      p->gate_site[g] + p->gate_range[g] - 1 <= p->N,
                  TNH_ERR_ARG, "qjmc: gate out of range");
    const int N = p->N, d = p->d, dd = d * d;
    Engine &E = eng;
    E.ctx = ctx;
    E.p = p;
    E.T = ntraj;
    E.tq.resize(ntraj);
    auto alloc = [&](size_t bytes) {
      size_t cap = 0;
      void *q = ctx->alloc(bytes, &cap);
      owned.push_back(q);
      owned_sz.push_back(cap);
      return q;
    };
    // operator store: gates, jump operators, escape operators c-free L^H L, observables
    std::vector<double> ops;
    auto add_op = [&](const double *src, int64_t n) {
      const int64_t off = int64_t(ops.size() / 2);
      ops.insert(ops.end(), src, src + 2 * n);
      return off;
    };
    for (int g = 0; g < ng; ++g)
      E.gate_off.push_back(add_op(p->gate_data + 2 * p->gate_offset[g], p->gate_range[g] == 2 ? dd * dd : dd));
    std::vector<double> esc(2 * dd);
    for (int t = 0; t < p->njump; ++t) {
      const double *L = p->jump_op + 2 * size_t(dd) * t;
      E.jump_off.push_back(add_op(L, dd));
      for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i) {  // (L^H L)[i, j] = sum_k conj(L[k, i]) L[k, j]
          std::complex<double> s = 0;
          for (int k = 0; k < d; ++k)
            s += std::conj(std::complex<double>(L[2 * (k + d * i)], L[2 * (k + d * i) + 1])) *
                 std::complex<double>(L[2 * (k + d * j)], L[2 * (k + d * j) + 1]);
          esc[2 * (i + d * j)] = s.real();
          esc[2 * (i + d * j) + 1] = s.imag();
        }
      E.esc_off.push_back(add_op(esc.data(), dd));
    }
    for (int o = 0; o < p->nobs; ++o) E.obs_off.push_back(add_op(p->obs_op + 2 * size_t(dd) * o, dd));
    TNH_REQUIRE(ops.size() / 2 < size_t(1) << 30, TNH_ERR_ARG, "qjmc batched: operator store too large");
    double *op_dev = static_cast<double *>(alloc(sizeof(double) * std::max<size_t>(ops.size(), 2)));
    to_device(ctx, op_dev, ops.data(), sizeof(double) * ops.size());
    // state
    E.b.N = N;
    E.b.d = d;
    E.b.CH = std::min(16, QM / d);  // Gram side d * CH <= 32; environments CH x CH <= 256
    E.b.cap = int64_t(E.b.CH) * d * E.b.CH;
    E.b.cutoff = p->cutoff;
    E.b.maxdim = p->maxdim;
    E.b.mindim = p->mindim;
    E.b.op = op_dev;
    E.b.S = static_cast<double *>(alloc(sizeof(double) * size_t(ntraj) * N * 2 * E.b.cap));
    E.b.bond = static_cast<int *>(alloc(sizeof(int) * size_t(ntraj) * (N + 1)));
    E.b.flags = static_cast<int *>(alloc(sizeof(int) * size_t(ntraj)));
    double *init_dev = static_cast<double *>(alloc(sizeof(double) * 2 * N * d));
    to_device(ctx, init_dev, p->init_state, sizeof(double) * 2 * N * d);
    hipLaunchKernelGGL(qb_init_kernel, dim3(ntraj), dim3(64), 0, ctx->stream, E.b, init_dev, ntraj);
    TNH_HIP(hipGetLastError());
    // expectation terms: escape operators, then observables
    const int nesc = p->njump, nobs = p->nobs;
    std::vector<int> tsite;
    std::vector<int64_t> toff;
    for (int t = 0; t < nesc; ++t) {
      tsite.push_back(p->jump_site[t]);
      toff.push_back(E.esc_off[t]);
    }
    for (int o = 0; o < nobs; ++o) {
      tsite.push_back(p->obs_site[o]);
      toff.push_back(E.obs_off[o]);
    }
    const int nterm = int(tsite.size());
    int *tsite_dev = static_cast<int *>(alloc(sizeof(int) * std::max(1, nterm)));
    int64_t *toff_dev = static_cast<int64_t *>(alloc(sizeof(int64_t) * std::max(1, nterm)));
    if (nterm) {
      to_device(ctx, tsite_dev, tsite.data(), sizeof(int) * nterm);
      to_device(ctx, toff_dev, toff.data(), sizeof(int64_t) * nterm);
    }
    std::vector<int> all(ntraj);
    for (int t = 0; t < ntraj; ++t) all[t] = t;
    int *all_dev = static_cast<int *>(alloc(sizeof(int) * ntraj));
    to_device(ctx, all_dev, all.data(), sizeof(int) * ntraj);
    double *env = static_cast<double *>(alloc(sizeof(double) * size_t(ntraj) * (N + 1) * 2 * E.b.CH * E.b.CH));
    double *outd = static_cast<double *>(alloc(sizeof(double) * 2 * size_t(ntraj) * std::max(1, nterm)));
    std::vector<double> outh(2 * size_t(ntraj) * std::max(1, nterm));
    // all terms of all trajectories in one launch; terms [q0, q0 + nq) are read back
    auto expect = [&](int q0, int nq) {
      hipLaunchKernelGGL(qb_expect_kernel, dim3(ntraj), dim3(QT), 0, ctx->stream, E.b, all_dev, nq,
                         tsite_dev + q0, toff_dev + q0, env, outd);
      TNH_HIP(hipGetLastError());
      to_host(ctx, outh.data(), outd, sizeof(double) * 2 * size_t(ntraj) * nq);
    };
    const int nsave = p->nsteps / p->savesteps + 1;
    std::vector<int64_t> pos(ntraj, 0);
    std::vector<int> save(ntraj, 0), nj(ntraj, 0);
    auto draw = [&](int t) {
      TNH_REQUIRE(pos[t] < nuniform, TNH_ERR_ARG, "qjmc: not enough uniforms for a trajectory");
      return uniforms[size_t(t) * nuniform + pos[t]++];
    };
    auto measure = [&]() {
      if (nobs > 0) {
        expect(nesc, nobs);
        for (int t = 0; t < ntraj; ++t)
          for (int o = 0; o < nobs; ++o) {
            double *m = meas + (size_t(t) * nsave + save[t]) * 2 * nobs;
            m[2 * o] = outh[2 * (size_t(t) * nobs + o)];
            m[2 * o + 1] = outh[2 * (size_t(t) * nobs + o) + 1];
          }
      }
      for (int t = 0; t < ntraj; ++t) ++save[t];
    };
    auto normalize_all = [&](const std::vector<int> &which) {
      for (int t : which) {
        auto &x = E.tq[t];
        if (x.center == 0) E.push_move(x, 1, 0);  // norm(): gmps.jl:29-37
        x.q.push_back({P_NORM, x.center, 0, 0});
      }
      E.run_queues();
    };
    measure();
    std::vector<double> rates(std::max(1, nesc));
    std::vector<char> dead(ntraj, 0);  // trajectories that outgrew their slots
    for (int step = 1; step <= p->nsteps; ++step) {
      for (int t = 0; t < ntraj; ++t)
        if (!dead[t]) E.push_applygates(E.tq[t]);
      E.run_queues();
      std::vector<int> live;
      for (int t = 0; t < ntraj; ++t)
        if (!dead[t]) live.push_back(t);
      for (int t : live) (void)draw(t);  // qjmc.jl:64: drawn and unused
      normalize_all(live);
      check_flags(ctx, E, dead);
      std::vector<int> jumpers;
      if (nesc > 0) {
        expect(0, nesc);
        for (int t = 0; t < ntraj; ++t) {
          if (dead[t]) continue;
          double er = 0;
          for (int q = 0; q < nesc; ++q) {
            const double c = p->jump_coeff[q];
            const double *v = &outh[2 * (size_t(t) * nesc + q)];
            rates[q] = std::abs(std::complex<double>(c * c * v[0], c * c * v[1]));
            er += rates[q];
          }
          const double prob = std::exp(-er * p->dt);
          if (draw(t) > prob) {
            const double r = draw(t);
            int idx = -1;
            double cum = 0;
            for (int q = 0; q < nesc && idx < 0; ++q) {
              cum += rates[q];
              if (r < cum / er) idx = q;
            }
            TNH_REQUIRE(idx >= 0, TNH_ERR_NUMERIC, "qjmc: no jump channel selected");
            auto &x = E.tq[t];
            E.push_move(x, 1, 0);
            x.q.push_back({P_SITEOP, p->jump_site[idx], int(E.jump_off[idx]), 0});
            E.push_move(x, N, 0);
            E.push_move(x, 1, 1);
            jumpers.push_back(t);
            if (nj[t] < maxjumps) {
              jumps[size_t(t) * maxjumps + nj[t]] = idx + 1;
              jump_steps[size_t(t) * maxjumps + nj[t]] = step;
            }
            ++nj[t];
          }
        }
      } else {
        for (int t = 0; t < ntraj; ++t)
          if (!dead[t]) (void)draw(t);  // the jump test draws even with no channels
      }
      if (!jumpers.empty()) {
        E.run_queues();
        normalize_all(jumpers);
      }
      if (step % p->savesteps == 0) measure();
    }
    check_flags(ctx, E, dead);
    std::vector<int> bonds(size_t(ntraj) * (N + 1));
    to_host(ctx, bonds.data(), E.b.bond, sizeof(int) * bonds.size());
    for (int t = 0; t < ntraj; ++t) {
      njumps[t] = dead[t] ? 0 : nj[t];
      int mb = 0;
      for (int i = 1; i < N; ++i) mb = std::max(mb, bonds[size_t(t) * (N + 1) + i]);
      maxbond[t] = dead[t] ? -1 : mb;
    }
  });
  // stream-ordered cache: later work on the stream reuses these only after this run's kernels
  for (size_t i = 0; i < owned.size(); ++i) ctx->release(owned[i], owned_sz[i]);
  if (eng.ops_dev) ctx->release(eng.ops_dev, eng.ops_cap);
  return rc;
}
