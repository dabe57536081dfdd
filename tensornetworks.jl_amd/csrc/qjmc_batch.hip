// Batched quantum-jump trajectories: every trajectory of qjmc_simulation
// (src/algorithms/mps/qjmc.jl:28-167, classical branch, single-site jump operators) advances
// together, one workgroup per trajectory in each launch. The site tensors of all trajectories
// live in fixed-capacity slots in HBM. Narrow slots hold bond dimension CH <= 32 / d, so every
// Gram matrix of a split is at most 32 x 32 and is diagonalised in LDS by the cyclic Jacobi
// of jacobi_lds.hpp. Wide slots (CH <= 512 / d) give each trajectory a 1024-thread
// workgroup, work matrices in HBM and the divide-and-conquer eigensolver of dc_eig.hpp; a
// call runs every trajectory on narrow slots and continues the ones that outgrew them, each
// from the start of the step it outgrew them in (its state and bookkeeping are kept at every
// step start; the step is replayed on its own uniforms), on wide slots of bond dimension 64
// and then of the largest the chain and the truncation allow. The host keeps each trajectory's orthogonality centre
// and turns the reference's step (applygates!, gatelist.jl:191-223; normalize!; the emission
// rates of :170-220; the jump of :103-107) into a list of primitives per trajectory: centre
// moves (gmps.jl:90-112), two-site gates (gatelist.jl:137-171) and one-site operators. A
// launch gives every trajectory with work a workgroup that walks its own list, so
// trajectories with different centres and bond dimensions advance in the same launch.
// Uniforms are consumed per trajectory in the reference's order (one discarded draw, the
// jump test, the channel). A split whose kept rank would exceed CH stops that trajectory
// (rerun on the next slot size, or marked for the caller) instead of truncating differently
// from the reference.
#include <algorithm>
#include <cmath>
#include <complex>
#include <deque>
#include <memory>
#include <numeric>
#include <vector>

#include "dc_eig.hpp"
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
// F_FALLBACK: a wide-slot split whose divide and conquer failed; the trajectory is rerun on
// the per-trajectory path, whose splits fall back to the Jacobi solvers.
enum Flag { F_OVERFLOW = 1, F_NOCONV = 2, F_NUMERIC = 4, F_FALLBACK = 8 };

struct Batch {
  double *S;        // site slots: [traj][site][re CAP | im CAP]
  int *bond;        // [traj][N + 1], bond[0] = bond[N] = 1
  int *flags;       // [traj]
  const double *op; // operator store, interleaved complex
  int N, d, CH;
  int64_t cap;      // CH * d * CH
  double cutoff;
  int maxdim, mindim;
  double *scratch;  // wide slots: per-workgroup work matrices, sstride doubles each
  int64_t sstride;
  int P;            // wide slots: d * CH, the largest side of a split
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

// Trajectory src[w]'s site tensors and bond dimensions, from slots of capacity scap (S, sbond)
// into trajectory dst[w]'s slots of capacity dcap >= its tensors (D, dbond): the state every
// live trajectory has at the start of a step (kept for a continuation), and the move of an
// overflowed trajectory into larger slots. Only each site's Dl d Dr entries are copied.
__global__ void qb_copy_state_kernel(const double *__restrict__ S, int64_t scap, const int *__restrict__ sbond,
                                     const int *__restrict__ src, double *__restrict__ D, int64_t dcap,
                                     int *__restrict__ dbond, const int *__restrict__ dst, int N, int d) {
  const int ts = src[blockIdx.x], td = dst[blockIdx.x];
  const int *bs = sbond + int64_t(ts) * (N + 1);
  for (int site = 1; site <= N; ++site) {
    const int ne = bs[site - 1] * d * bs[site];
    const double *a = S + (int64_t(ts) * N + (site - 1)) * 2 * scap;
    double *b = D + (int64_t(td) * N + (site - 1)) * 2 * dcap;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      b[e] = a[e];
      b[dcap + e] = a[scap + e];
    }
  }
  for (int e = threadIdx.x; e <= N; e += blockDim.x) dbond[int64_t(td) * (N + 1) + e] = bs[e];
}

// ---- wide slots: bond dimension up to kDcPMax / d, one 1024-thread workgroup per trajectory
//
// The same primitives as above with every matrix in HBM (per-workgroup work matrices of side
// P = d CH) instead of LDS: products on the matrix cores (wg_zgemm), and the Gram eigensolve
// of each split by the reduction and divide and conquer of dc_eig.hpp, which needs the whole
// LDS of the workgroup. Above a bond dimension of 16 the eigensolve is most of a step, and it
// is a one-workgroup algorithm either way: running one trajectory per workgroup fills the
// GPU, where the per-trajectory path keeps one workgroup busy per host thread.

constexpr int WT = dc::kDcThreads;  // threads per trajectory workgroup
constexpr int WW = WT / 64;         // waves

typedef double wv4d __attribute__((ext_vector_type(4)));  // an FP64 MFMA accumulator

// work matrices of one workgroup: complex regions of 2 P^2 doubles (planes P^2 apart), real
// regions of P^2
struct WScr {
  double *M, *T, *G;  // the matrix being split; theta, then eigenvectors; Gram, then reflectors
  double *Q, *Q2;     // (real) divide-and-conquer work, then the basis and U^H U
  double *U;          // divide-and-conquer work, then U X, then the right factor (complex)
  double *taus, *dout, *val;
  int *perm;
  int64_t P2;
};

// doubles per workgroup for side P (a multiple of 32, so every region stays 256-byte aligned)
__host__ __device__ inline int64_t wide_stride(int P) {
  const int64_t P2 = int64_t(P) * P, P32 = (int64_t(P) + 32) & ~int64_t(31);
  return ((10 * P2 + 31) & ~int64_t(31)) + 5 * P32;
}

__device__ inline WScr wscr(const Batch &b, int w) {
  WScr s;
  const int64_t P2 = int64_t(b.P) * b.P, P32 = (int64_t(b.P) + 32) & ~int64_t(31);
  double *x = b.scratch + int64_t(w) * b.sstride;
  s.P2 = P2;
  s.M = x;
  s.T = x + 2 * P2;
  s.G = x + 4 * P2;
  s.Q = x + 6 * P2;
  s.Q2 = x + 7 * P2;
  s.U = x + 8 * P2;
  double *t = x + ((10 * P2 + 31) & ~int64_t(31));
  s.taus = t;                // 2 P (real, then imaginary factors)
  s.dout = t + 2 * P32;      // P + 1
  s.val = t + 3 * P32;       // P
  s.perm = reinterpret_cast<int *>(t + 4 * P32);  // P ints
  return s;
}

// Operand of a workgroup product: element (i, l) at re[i * rs + l * cs] and im[...],
// conjugated when conj is set.
struct ZOp {
  const double *re, *im;
  int64_t rs, cs;
  bool conj;
};

// C (M x N, column-major, leading dimension ldc, planes Cr and Ci) = A (M x K) B (K x N) by
// the whole workgroup on the matrix cores: each wave takes 32 x 16 output tiles in turn; per
// 16 x 16 half and k-step four v_mfma_f64_16x16x4f64 (re += ar br - ai bi, im += ar bi + ai br),
// the operands read per lane from memory (L1/L2). Out-of-range rows, columns and k read the
// operand base and are zeroed. Ends with a barrier.
__device__ void wg_zgemm(int M, int N, int K, const ZOp &A, const ZOp &B, double *Cr, double *Ci, int64_t ldc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, rr = lane & 15;
  const int tm = (M + 31) / 32, tn = (N + 15) / 16;
  const double sa = A.conj ? -1.0 : 1.0, sb = B.conj ? -1.0 : 1.0;
  for (int tile = wave; tile < tm * tn; tile += WW) {
    const int row0 = 32 * (tile % tm), col0 = 16 * (tile / tm);
    wv4d accr[2], acci[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      accr[x] = wv4d{0.0, 0.0, 0.0, 0.0};
      acci[x] = wv4d{0.0, 0.0, 0.0, 0.0};
    }
    const int c = col0 + rr;
    for (int k0 = 0; k0 < K; k0 += 4) {
      const int k = k0 + g;
      const bool bon = k < K && c < N;
      const int64_t boff = bon ? int64_t(k) * B.rs + int64_t(c) * B.cs : 0;
      const double bvr = B.re[boff], bvi = B.im[boff];
      const double br = bon ? bvr : 0.0, bi = bon ? sb * bvi : 0.0;
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const int r = row0 + 16 * x + rr;
        const bool on = k < K && r < M;
        const int64_t off = on ? int64_t(r) * A.rs + int64_t(k) * A.cs : 0;
        const double vr = A.re[off], vi = A.im[off];
        const double ar = on ? vr : 0.0, ai = on ? sa * vi : 0.0;
        accr[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, accr[x], 0, 0, 0);
        accr[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, accr[x], 0, 0, 0);
        acci[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, acci[x], 0, 0, 0);
        acci[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, acci[x], 0, 0, 0);
      }
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = row0 + 16 * x + g + 4 * i;
        if (r < M && c < N) {
          Cr[r + c * ldc] = accr[x][i];
          Ci[r + c * ldc] = acci[x][i];
        }
      }
  }
  __syncthreads();
}

// V = H_0 ... H_{p-2} Z (as dc_backtransform_kernel of the SVD module, inside the
// workgroup): one wave per column, lane l holding rows l, l + 64, ... (R of them) of it; each
// reflector is a wave-wide shuffle reduction and an update. Ends with a barrier.
template <int R>
__device__ void wg_backtransform_r(const double *Gr, const double *Gi, int p, const double *taus, const double *Z,
                                   double *Vr, double *Vi) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t P = p;
  for (int c = wave; c < p; c += WW) {
    double xr[R], xi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = lane + 64 * r;
      xr[r] = i < p ? Z[i + c * P] : 0.0;
      xi[r] = 0.0;
    }
    for (int k = p - 2; k >= 0; --k) {
      const double tr = taus[k], ti = taus[p + k];
      if (tr == 0.0 && ti == 0.0) continue;
      double sr = 0, si = 0;  // s = v^H x
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        const bool on = i > k && i < p;
        const int64_t off = on ? i + k * P : 0;
        const double a = Gr[off], b = Gi[off];
        const double vr = on ? a : 0.0, vi = on ? b : 0.0;
        sr += vr * xr[r] + vi * xi[r];
        si += vr * xi[r] - vi * xr[r];
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o, 64);
        si += __shfl_xor(si, o, 64);
      }
      const double fr = tr * sr - ti * si, fi = tr * si + ti * sr;  // tau s
#pragma unroll
      for (int r = 0; r < R; ++r) {  // the reflector again (L1), not kept in registers
        const int i = lane + 64 * r;
        const bool on = i > k && i < p;
        const int64_t off = on ? i + k * P : 0;
        const double a = Gr[off], b = Gi[off];
        const double vr = on ? a : 0.0, vi = on ? b : 0.0;
        xr[r] -= vr * fr - vi * fi;
        xi[r] -= vr * fi + vi * fr;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = lane + 64 * r;
      if (i < p) {
        Vr[i + c * P] = xr[r];
        Vi[i + c * P] = xi[r];
      }
    }
  }
  __syncthreads();
}

__device__ void wg_backtransform(const double *Gr, const double *Gi, int p, const double *taus, const double *Z,
                                 double *Vr, double *Vi) {
  if (p <= 128)
    wg_backtransform_r<2>(Gr, Gi, p, taus, Z, Vr, Vi);
  else
    wg_backtransform_r<8>(Gr, Gi, p, taus, Z, Vr, Vi);
}

// workgroup sum of (a, b), returned to every thread (two barriers)
__device__ void wg_sum2(double &a, double &b, double (*red)[WW]) {
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  a = 0;
  b = 0;
  for (int w = 0; w < WW; ++w) {
    a += red[0][w];
    b += red[1][w];
  }
  __syncthreads();
}

struct WShared {
  dc::DcShared dc;
  double red[2][WW];
  int k, flag;
};

// Truncated split of M (m x n, column-major, leading dimension m, planes Mr / Mi) with the
// same steps and outcome as split_lds_w: Gram matrix of the basis side p, eigenvalues and
// eigenvectors by dc_eig (reduction, divide and conquer, back-transformation), stable
// descending order, the truncation rule, one Newton-Schulz step on the kept basis. On return
// the left factor (m x k, leading dimension m) is in w.G and the right factor (k x n, leading
// dimension k) in w.U, planes P^2 apart; returns k (0 on failure, with sh.flag set).
__device__ int split_w(const Batch &b, WShared &sh, const WScr &w, const double *Mr, const double *Mi, int m, int n,
                       int side, double cutoff, int maxdim, int mindim) {
  const int t = threadIdx.x;
  const int p = side == 0 ? m : n;
  const int64_t P2 = w.P2;
  // G = M M^H (side 0) or M^H M (side 1)
  {
    ZOp A, B;
    if (side == 0) {
      A = ZOp{Mr, Mi, 1, m, false};
      B = ZOp{Mr, Mi, m, 1, true};
    } else {
      A = ZOp{Mr, Mi, m, 1, true};
      B = ZOp{Mr, Mi, 1, m, false};
    }
    wg_zgemm(p, p, side == 0 ? n : m, A, B, w.G, w.G + P2, p);
  }
  dc::dc_eig<true>(dc::DcDeviceEx<WT>{}, sh.dc, w.G, w.G + P2, p, w.taus, w.taus + p, w.Q, w.Q2, w.U, w.dout);
  __syncthreads();
  if (w.dout[p] != 1.0) {
    if (t == 0) sh.flag |= F_FALLBACK;
    __syncthreads();
    return 0;
  }
  // eigenvectors of G into w.T (the reflectors in w.G, the tridiagonal eigenvectors in w.Q)
  wg_backtransform(w.G, w.G + P2, p, w.taus, w.Q, w.T, w.T + P2);
  for (int i = t; i < p; i += WT) w.val[i] = sqrt(fmax(w.dout[i], 0.0));
  __syncthreads();
  for (int i = t; i < p; i += WT) {  // stable descending rank
    const double v = w.val[i];
    int rk = 0;
    for (int j = 0; j < p; ++j) rk += (w.val[j] > v || (w.val[j] == v && j < i)) ? 1 : 0;
    w.perm[rk] = i;
  }
  __syncthreads();
  if (t == 0) {
    // descending values into w.dout (no longer needed), then the rule
    int ok = 1;
    for (int i = 0; i < p; ++i) {
      w.dout[i] = w.val[w.perm[i]];
      if (!isfinite(w.dout[i])) ok = 0;
    }
    double disc = 0;
    int k = ok ? trunc_rank(w.dout, min(m, n), cutoff, maxdim, mindim, &disc) : 0;
    if (!ok) sh.flag |= F_NUMERIC;
    if (ok && k > b.CH) {
      sh.flag |= F_OVERFLOW;
      k = 0;
    }
    sh.k = k;
  }
  __syncthreads();
  const int k = sh.k;
  if (k == 0) return 0;
  // basis U (p x k) = columns perm[0..k) of V, into w.Q
  double *Ur = w.Q, *Ui = w.Q + int64_t(p) * k;
  for (int e = t; e < p * k; e += WT) {
    const int r = e % p, j = e / p, c = w.perm[j];
    Ur[e] = w.T[r + int64_t(p) * c];
    Ui[e] = w.T[P2 + r + int64_t(p) * c];
  }
  __syncthreads();
  // Newton-Schulz: X = U^H U (k x k, w.Q2), W = U X (p x k, w.U), U <- 1.5 U - 0.5 W
  double *Xr = w.Q2, *Xi = w.Q2 + int64_t(k) * k;
  wg_zgemm(k, k, p, ZOp{Ur, Ui, p, 1, true}, ZOp{Ur, Ui, 1, p, false}, Xr, Xi, k);
  wg_zgemm(p, k, k, ZOp{Ur, Ui, 1, p, false}, ZOp{Xr, Xi, 1, k, false}, w.U, w.U + P2, p);
  for (int e = t; e < p * k; e += WT) {
    Ur[e] = 1.5 * Ur[e] - 0.5 * w.U[e];
    Ui[e] = 1.5 * Ui[e] - 0.5 * w.U[P2 + e];
  }
  __syncthreads();
  if (side == 0) {
    // left factor U (m x k); right factor U^H M (k x n)
    for (int e = t; e < m * k; e += WT) {
      w.G[e] = Ur[e];
      w.G[P2 + e] = Ui[e];
    }
    wg_zgemm(k, n, m, ZOp{Ur, Ui, m, 1, true}, ZOp{Mr, Mi, 1, m, false}, w.U, w.U + P2, k);
  } else {
    // right factor V_k^H (k x n); left factor M V_k (m x k)
    for (int e = t; e < k * n; e += WT) {
      const int j = e % k, c = e / k;
      w.U[e] = Ur[c + int64_t(n) * j];
      w.U[P2 + e] = -Ui[c + int64_t(n) * j];
    }
    wg_zgemm(m, k, n, ZOp{Mr, Mi, 1, m, false}, ZOp{Ur, Ui, 1, n, false}, w.G, w.G + P2, m);
  }
  return k;
}

__device__ inline void copy_to_slot(double *dst, int64_t cap, int numel, const double *src, int64_t plane) {
  for (int e = threadIdx.x; e < numel; e += WT) {
    dst[e] = src[e];
    dst[cap + e] = src[plane + e];
  }
}

// The split primitives (two-site gate, centre moves) in three parts around one split_w call
// site, so the kernel holds a single copy of the eigensolver. split_setup prepares the matrix
// to split and returns its description; split_finish writes the factors back.
struct WSplit {
  const double *Mr, *Mi;
  int m, n, side;
  double cutoff;
  int maxdim, mindim;
};

// two-site gate at (s, s + 1), as prim_gate: theta = A_s A_{s+1}, the gate applied into w.M
__device__ WSplit gate_setup(const Batch &b, const WScr &w, int tr, int s, const double *Gop, int dir) {
  const int d = b.d;
  const int *bd = b.bond + int64_t(tr) * (b.N + 1);
  const int Dl = bd[s - 1], Dm = bd[s], Dr = bd[s + 1];
  const double *sa = slot(b, tr, s), *sb = slot(b, tr, s + 1);
  const int m = Dl * d, n = d * Dr;
  const int64_t P2 = w.P2;
  // theta = A_s (m x Dm) A_{s+1} (Dm x n), into w.T
  wg_zgemm(m, n, Dm, ZOp{sa, sa + b.cap, 1, m, false}, ZOp{sb, sb + b.cap, 1, Dm, false}, w.T, w.T + P2, m);
  // M[(a,o1),(o2,b)] = sum_{i1,i2} G[o1,i1,o2,i2] theta[(a,i1),(i2,b)]
  for (int e = threadIdx.x; e < m * n; e += WT) {
    const int r = e % m, c = e / m;
    const int a = r % Dl, o1 = r / Dl, o2 = c % d, bb = c / d;
    double sr = 0, si = 0;
    for (int i2 = 0; i2 < d; ++i2)
      for (int i1 = 0; i1 < d; ++i1) {
        const int64_t q = o1 + int64_t(d) * (i1 + int64_t(d) * (o2 + int64_t(d) * i2));
        const double gr_ = Gop[2 * q], gi_ = Gop[2 * q + 1];
        const int64_t te = (a + Dl * i1) + int64_t(m) * (i2 + d * bb);
        const double xr = w.T[te], xi = w.T[P2 + te];
        sr += gr_ * xr - gi_ * xi;
        si += gr_ * xi + gi_ * xr;
      }
    w.M[e] = sr;
    w.M[P2 + e] = si;
  }
  __syncthreads();
  return WSplit{w.M, w.M + P2, m, n, dir ? 1 : 0, b.cutoff, b.maxdim, b.mindim};
}

// centre move by one site (dir 0: s -> s + 1, dir 1: s -> s - 1), as prim_move: A_s is split
// in place
__device__ WSplit move_setup(const Batch &b, int tr, int s, int dir, int trunc) {
  const int d = b.d;
  const int *bd = b.bond + int64_t(tr) * (b.N + 1);
  const int Dl = bd[s - 1], Dr = bd[s];
  const double *sa = slot(b, tr, s);
  const int m = dir == 0 ? Dl * d : Dl, n = dir == 0 ? Dr : d * Dr;
  return WSplit{sa, sa + b.cap, m, n, dir, trunc ? b.cutoff : 0.0, trunc ? b.maxdim : 0, trunc ? b.mindim : 1};
}

// the kept factors (k > 0) back into the slots: left factor in w.G, right factor in w.U
__device__ void split_finish(const Batch &b, const WScr &w, int tr, int kind, int s, const WSplit &q, int k) {
  const int d = b.d;
  int *bd = b.bond + int64_t(tr) * (b.N + 1);
  double *sa = slot(b, tr, s);
  const int64_t P2 = w.P2;
  const int m = q.m, n = q.n;
  if (kind == P_GATE) {
    double *sb = slot(b, tr, s + 1);
    copy_to_slot(sa, b.cap, m * k, w.G, P2);
    copy_to_slot(sb, b.cap, k * n, w.U, P2);
    if (threadIdx.x == 0) bd[s] = k;
  } else if (kind == P_RIGHT) {
    // A_{s+1} <- (U^H A_s) (k x Dr) A_{s+1} (Dr x d Dr2)
    const int Dr = bd[s], Dr2 = bd[s + 1];
    double *sn = slot(b, tr, s + 1);
    const int nn = d * Dr2;
    wg_zgemm(k, nn, Dr, ZOp{w.U, w.U + P2, 1, k, false}, ZOp{sn, sn + b.cap, 1, Dr, false}, w.T, w.T + P2, k);
    copy_to_slot(sa, b.cap, m * k, w.G, P2);
    copy_to_slot(sn, b.cap, k * nn, w.T, P2);
    if (threadIdx.x == 0) bd[s] = k;
  } else {
    // A_{s-1} <- A_{s-1} (Dl2 d x Dl) (A_s V_k) (Dl x k)
    const int Dl = bd[s - 1], Dl2 = bd[s - 2];
    double *sp = slot(b, tr, s - 1);
    const int mm = Dl2 * d;
    wg_zgemm(mm, k, Dl, ZOp{sp, sp + b.cap, 1, mm, false}, ZOp{w.G, w.G + P2, 1, m, false}, w.T, w.T + P2, mm);
    copy_to_slot(sa, b.cap, k * n, w.U, P2);
    copy_to_slot(sp, b.cap, mm * k, w.T, P2);
    if (threadIdx.x == 0) bd[s - 1] = k;
  }
  __syncthreads();
}

// A[a,o,b] <- sum_i O[o,i] A[a,i,b], in place: each thread owns the d entries of its (a, b)
__device__ void prim_siteop_w(const Batch &b, int tr, int s, const double *O) {
  const int d = b.d;
  const int *bd = b.bond + int64_t(tr) * (b.N + 1);
  const int Dl = bd[s - 1], Dr = bd[s];
  double *sa = slot(b, tr, s);
  for (int e = threadIdx.x; e < Dl * Dr; e += WT) {
    const int a = e % Dl, bb = e / Dl;
    double xr[QM / 2], xi[QM / 2];
    for (int i = 0; i < d; ++i) {
      const int64_t x = a + int64_t(Dl) * (i + int64_t(d) * bb);
      xr[i] = sa[x];
      xi[i] = sa[b.cap + x];
    }
    for (int o = 0; o < d; ++o) {
      double sr = 0, si = 0;
      for (int i = 0; i < d; ++i) {
        const double gr_ = O[2 * (o + d * i)], gi_ = O[2 * (o + d * i) + 1];
        sr += gr_ * xr[i] - gi_ * xi[i];
        si += gr_ * xi[i] + gi_ * xr[i];
      }
      const int64_t x = a + int64_t(Dl) * (o + int64_t(d) * bb);
      sa[x] = sr;
      sa[b.cap + x] = si;
    }
  }
  __syncthreads();
}

// A_c <- A_c / ||A_c||; false on a zero or non-finite norm
__device__ bool prim_normalize_w(const Batch &b, WShared &sh, int tr, int s) {
  const int *bd = b.bond + int64_t(tr) * (b.N + 1);
  const int ne = bd[s - 1] * b.d * bd[s];
  double *sa = slot(b, tr, s);
  double acc = 0, dummy = 0;
  for (int e = threadIdx.x; e < ne; e += WT) acc += sa[e] * sa[e] + sa[b.cap + e] * sa[b.cap + e];
  wg_sum2(acc, dummy, sh.red);
  const double nr = sqrt(acc);
  if (!(nr > 0) || !isfinite(nr)) return false;
  const double inv = 1.0 / nr;
  for (int e = threadIdx.x; e < ne; e += WT) {
    sa[e] *= inv;
    sa[b.cap + e] *= inv;
  }
  __syncthreads();
  return true;
}

// qb_run_kernel on wide slots: one 1024-thread workgroup per listed trajectory, work matrices
// at b.scratch + (list index) * b.sstride
__global__ __launch_bounds__(WT) void qw_run_kernel(Batch b, const int *__restrict__ trajs,
                                                    const int *__restrict__ start, const int *__restrict__ items) {
  __shared__ WShared sh;
  const int w = blockIdx.x, tr = trajs[w], t = threadIdx.x;
  const WScr ws = wscr(b, w);
  if (t == 0) sh.flag = b.flags[tr];
  __syncthreads();
  if (sh.flag) return;
  for (int i = start[w]; i < start[w + 1]; ++i) {
    const int *it = items + 4 * i;
    const int kind = it[0], s = it[1], arg = it[2], fl = it[3];
    if (kind == P_SITEOP) {
      prim_siteop_w(b, tr, s, b.op + 2 * int64_t(arg));
    } else if (kind == P_NORM) {
      if (!prim_normalize_w(b, sh, tr, s) && t == 0) sh.flag |= F_NUMERIC;
    } else {
      const WSplit q = kind == P_GATE ? gate_setup(b, ws, tr, s, b.op + 2 * int64_t(arg), fl)
                                      : move_setup(b, tr, s, kind == P_RIGHT ? 0 : 1, fl);
      const int k = split_w(b, sh, ws, q.Mr, q.Mi, q.m, q.n, q.side, q.cutoff, q.maxdim, q.mindim);
      if (k > 0) split_finish(b, ws, tr, kind, s, q, k);
    }
    __syncthreads();
    if (sh.flag) {
      if (t == 0) b.flags[tr] = sh.flag;
      return;
    }
  }
}

// sum over the centre-form tensor of the term at site j: val = sum_{x,sg,b} conj(A[x,sg,b])
// sum_tau O[sg,tau] W[x,tau,b], A and W of shape Dl x d x Dr
__device__ void term_value(const double *A, int64_t capA, const double *W, int64_t capW, int Dl, int d, int Dr,
                           const double *O, double (*red)[WW], double *out) {
  const int ne = Dl * d * Dr;
  double vr = 0, vi = 0;
  for (int e = threadIdx.x; e < ne; e += WT) {
    const int x = e % Dl, sg = (e / Dl) % d, bb = e / (Dl * d);
    double pr = 0, pi = 0;
    for (int tau = 0; tau < d; ++tau) {
      const double gr_ = O[2 * (sg + d * tau)], gi_ = O[2 * (sg + d * tau) + 1];
      const int64_t it = x + int64_t(Dl) * (tau + int64_t(d) * bb);
      pr += gr_ * W[it] - gi_ * W[capW + it];
      pi += gr_ * W[capW + it] + gi_ * W[it];
    }
    const double ar = A[e], ai = -A[capA + e];
    vr += ar * pr - ai * pi;
    vi += ar * pi + ai * pr;
  }
  wg_sum2(vr, vi, red);
  if (threadIdx.x == 0) {
    out[0] = vr;
    out[1] = vi;
  }
}

// <psi|O|psi> for every (site, operator) term of every listed trajectory, from the mixed
// canonical form that the step loop keeps (centre c = centre[list index], sites left of it
// left-orthonormal, right of it right-orthonormal): the terms at c need only the centre
// tensor, those right of c the chain L_j = A_j^H (L_{j-1} A_j) started from L_c = A_c^H A_c,
// those left of c the chain R_j = (A_j R_{j+1}) A_j^H started from R_c = A_c A_c^H. One
// environment is live at a time. Centre 0 (the initial product state, every bond 1) takes
// the product of the site norms. out as qb_expect_kernel.
__global__ __launch_bounds__(WT) void qw_expect_kernel(Batch b, const int *__restrict__ trajs,
                                                       const int *__restrict__ centre, int nterms,
                                                       const int *__restrict__ term_site,
                                                       const int64_t *__restrict__ term_off, double *__restrict__ out) {
  __shared__ double red[2][WW];
  const int w = blockIdx.x, tr = trajs[w], d = b.d, N = b.N, c = centre[w];
  const int *bd = b.bond + int64_t(tr) * (N + 1);
  const WScr ws = wscr(b, w);
  const int64_t P2 = ws.P2;
  double *o = out + 2 * int64_t(w) * nterms;
  if (c == 0) {
    if (threadIdx.x == 0) {
      bool prod = true;
      for (int j = 0; j <= N; ++j) prod = prod && bd[j] == 1;
      for (int q = 0; q < nterms; ++q) {
        const int s = term_site[q];
        const double *O = b.op + 2 * term_off[q];
        double nr = 1.0;
        for (int j = 1; j <= N; ++j) {
          if (j == s) continue;
          const double *a = slot(b, tr, j);
          double x = 0;
          for (int i = 0; i < d; ++i) x += a[i] * a[i] + a[b.cap + i] * a[b.cap + i];
          nr *= x;
        }
        const double *a = slot(b, tr, s);
        double vr = 0, vi = 0;
        for (int sg = 0; sg < d; ++sg)
          for (int tau = 0; tau < d; ++tau) {
            const double gr_ = O[2 * (sg + d * tau)], gi_ = O[2 * (sg + d * tau) + 1];
            const double xr = a[sg], xi = -a[b.cap + sg], yr = a[tau], yi = a[b.cap + tau];
            const double pr = gr_ * yr - gi_ * yi, pi = gr_ * yi + gi_ * yr;
            vr += xr * pr - xi * pi;
            vi += xr * pi + xi * pr;
          }
        o[2 * q] = prod ? nr * vr : NAN;
        o[2 * q + 1] = prod ? nr * vi : NAN;
      }
    }
    return;
  }
  int smin = c, smax = c;
  for (int q = 0; q < nterms; ++q) {
    smin = min(smin, term_site[q]);
    smax = max(smax, term_site[q]);
  }
  double *Er = ws.M, *Ei = ws.M + P2;  // the live environment
  double *Wr = ws.T, *Wi = ws.T + P2;
  // sites c .. smax
  for (int j = c; j <= smax; ++j) {
    const int Dl = bd[j - 1], Dr = bd[j], m = Dl * d;
    const double *A = slot(b, tr, j);
    const double *W = A;
    int64_t capW = b.cap;
    if (j > c) {  // W = L_{j-1} (Dl x Dl) A_j (Dl x d Dr)
      wg_zgemm(Dl, d * Dr, Dl, ZOp{Er, Ei, 1, Dl, false}, ZOp{A, A + b.cap, 1, Dl, false}, Wr, Wi, Dl);
      W = Wr;
      capW = P2;
    }
    for (int q = 0; q < nterms; ++q)
      if (term_site[q] == j)
        term_value(A, b.cap, W, capW, Dl, d, Dr, b.op + 2 * term_off[q], red, o + 2 * q);
    if (j < smax)  // L_j = A_j^H W (both (Dl d) x Dr)
      wg_zgemm(Dr, Dr, m, ZOp{A, A + b.cap, m, 1, true}, ZOp{W, W + capW, 1, m, false}, Er, Ei, Dr);
  }
  // sites c - 1 .. smin
  if (smin < c) {
    {  // R_c = A_c (Dl x d Dr) A_c^H
      const int Dl = bd[c - 1], Dr = bd[c];
      const double *A = slot(b, tr, c);
      wg_zgemm(Dl, Dl, d * Dr, ZOp{A, A + b.cap, 1, Dl, false}, ZOp{A, A + b.cap, Dl, 1, true}, Er, Ei, Dl);
    }
    for (int j = c - 1; j >= smin; --j) {
      const int Dl = bd[j - 1], Dr = bd[j], m = Dl * d;
      const double *A = slot(b, tr, j);
      // W = A_j ((Dl d) x Dr) R_{j+1} (Dr x Dr)
      wg_zgemm(m, Dr, Dr, ZOp{A, A + b.cap, 1, m, false}, ZOp{Er, Ei, 1, Dr, false}, Wr, Wi, m);
      for (int q = 0; q < nterms; ++q)
        if (term_site[q] == j)
          term_value(A, b.cap, Wr, P2, Dl, d, Dr, b.op + 2 * term_off[q], red, o + 2 * q);
      if (j > smin)  // R_j = W (Dl x d Dr) A_j^H
        wg_zgemm(Dl, Dl, d * Dr, ZOp{Wr, Wi, 1, Dl, false}, ZOp{A, A + b.cap, Dl, 1, true}, Er, Ei, Dl);
    }
  }
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
  bool wide = false;  // wide slots: qw_* kernels
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
    if (wide)
      hipLaunchKernelGGL(qw_run_kernel, dim3(nw), dim3(WT), 0, ctx->stream, b, od, od + nw, od + 2 * nw + 1);
    else
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
    if (f[t] & (F_OVERFLOW | F_FALLBACK)) dead[t] = 1;
    TNH_REQUIRE(!(f[t] & F_NOCONV), TNH_ERR_NOCONV,
                "qjmc batched: Gram eigensolve did not converge (trajectory " + std::to_string(t) + ")");
    TNH_REQUIRE(!(f[t] & F_NUMERIC), TNH_ERR_NUMERIC,
                "qjmc batched: non-finite or zero state (trajectory " + std::to_string(t) + ")");
  }
}


// Device buffers of one engine run, returned to the context's stream-ordered cache at the end
// (also when the run throws): later work on the stream reuses them only after this run's
// kernels.
struct Owned {
  tnh_ctx *ctx;
  std::vector<std::pair<void *, size_t>> v;
  void *take(size_t bytes) {
    size_t cap = 0;
    void *q = ctx->alloc(bytes, &cap);
    v.emplace_back(q, cap);
    return q;
  }
  ~Owned() {
    for (auto &x : v) ctx->release(x.first, x.second);
  }
};

// A trajectory's state at the start of a step: where a run on larger slots continues it.
struct Resume {
  int step0 = 1;       // the first step to run
  int64_t pos = 0;     // uniforms consumed before it
  int save = 0, nj = 0;  // saves written and jumps recorded before it
  int center = 0;      // orthogonality centre
};

// Trajectories that outgrew their slots, each with its Resume point and its site tensors and
// bonds at that point (packed, slot capacity cap), held in context memory until released.
struct Pack {
  tnh_ctx *ctx = nullptr;
  std::vector<int> rows;
  std::vector<Resume> res;
  int64_t cap = 0;
  double *S = nullptr;
  int *bond = nullptr;
  size_t scap = 0, bcap = 0;
  Pack() = default;
  Pack(const Pack &) = delete;
  Pack &operator=(const Pack &) = delete;
  ~Pack() {
    if (S) ctx->release(S, scap);
    if (bond) ctx->release(bond, bcap);
  }
};

// A view of trajectories [i0, i1) of a Pack, as the input of a run.
struct PackView {
  const std::vector<int> *rows;
  const std::vector<Resume> *res;
  int i0, i1;
  int64_t cap;
  const double *S;
  const int *bond;
};

// The step loop of qjmc_simulation over the trajectories `rows` (their indices into the
// caller's uniform and output arrays) on slots of bond dimension CH: narrow slots (the whole
// state of a split in LDS, CH <= 32 / d) or wide ones (work matrices in HBM, d CH <= 512).
// Fresh trajectories start from the product state at step 1; with `in`, each continues from
// its Resume point with its packed state. At the start of every step the state of each live
// trajectory is kept; a trajectory that outgrows its slots in a step (or whose wide split
// needs the general solvers) stops, and goes into `out` with its state and bookkeeping from
// the start of that step (maxbond[row] = -1 until a larger run completes it).
void run_engine(tnh_ctx *ctx, const tnh_qjmc_problem *p, std::vector<int> rows, const double *uniforms,
                int64_t nuniform, double *meas, int *njumps, int *jumps, int *jump_steps, int maxjumps, int *maxbond,
                int CH, bool wide, const PackView *in, Pack *out) {
  if (in) rows.assign(in->rows->begin() + in->i0, in->rows->begin() + in->i1);
  const int N = p->N, d = p->d, dd = d * d, ntraj = int(rows.size());
  Owned own{ctx, {}};
  Engine E;
  E.ctx = ctx;
  E.p = p;
  E.T = ntraj;
  E.wide = wide;
  E.tq.resize(ntraj);
  struct OpsGuard {
    Engine &e;
    ~OpsGuard() {
      if (e.ops_dev) e.ctx->release(e.ops_dev, e.ops_cap);
    }
  } ops_guard{E};
  // operator store: gates, jump operators, escape operators c-free L^H L, observables
  int ng = 0;
  for (int l = 0; l < p->nlayers; ++l) ng += p->layer_count[l];
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
  double *op_dev = static_cast<double *>(own.take(sizeof(double) * std::max<size_t>(ops.size(), 2)));
  to_device(ctx, op_dev, ops.data(), sizeof(double) * ops.size());
  // state
  E.b.N = N;
  E.b.d = d;
  E.b.CH = CH;
  E.b.cap = int64_t(CH) * d * CH;
  E.b.cutoff = p->cutoff;
  E.b.maxdim = p->maxdim;
  E.b.mindim = p->mindim;
  E.b.op = op_dev;
  E.b.S = static_cast<double *>(own.take(sizeof(double) * size_t(ntraj) * N * 2 * E.b.cap));
  E.b.bond = static_cast<int *>(own.take(sizeof(int) * size_t(ntraj) * (N + 1)));
  E.b.flags = static_cast<int *>(own.take(sizeof(int) * size_t(ntraj)));
  if (wide) {
    E.b.P = d * CH;
    E.b.sstride = wide_stride(E.b.P);
    E.b.scratch = static_cast<double *>(own.take(sizeof(double) * size_t(ntraj) * E.b.sstride));
  }
  double *init_dev = static_cast<double *>(own.take(sizeof(double) * 2 * N * d));
  to_device(ctx, init_dev, p->init_state, sizeof(double) * 2 * N * d);
  hipLaunchKernelGGL(qb_init_kernel, dim3(ntraj), dim3(64), 0, ctx->stream, E.b, init_dev, ntraj);
  TNH_HIP(hipGetLastError());
  // index lists of the state copies: [0, ntraj) for both sides, then a live list
  std::vector<int> iota_h(ntraj);
  std::iota(iota_h.begin(), iota_h.end(), 0);
  int *idx_dev = static_cast<int *>(own.take(sizeof(int) * 2 * size_t(ntraj)));
  int *live_dev = idx_dev + ntraj;
  to_device(ctx, idx_dev, iota_h.data(), sizeof(int) * ntraj);
  std::vector<Resume> rs(ntraj);  // each trajectory's latest step start
  if (in) {
    TNH_REQUIRE(in->cap <= E.b.cap, TNH_ERR_ARG, "qjmc batched: continuation into smaller slots");
    for (int t = 0; t < ntraj; ++t) rs[t] = (*in->res)[in->i0 + t];
    hipLaunchKernelGGL(qb_copy_state_kernel, dim3(ntraj), dim3(256), 0, ctx->stream,
                       in->S + int64_t(in->i0) * N * 2 * in->cap, in->cap, in->bond + int64_t(in->i0) * (N + 1),
                       idx_dev, E.b.S, E.b.cap, E.b.bond, idx_dev, N, d);
    TNH_HIP(hipGetLastError());
  }
  // the state at the start of the current step, per trajectory
  double *snap = static_cast<double *>(own.take(sizeof(double) * size_t(ntraj) * N * 2 * E.b.cap));
  int *snap_bond = static_cast<int *>(own.take(sizeof(int) * size_t(ntraj) * (N + 1)));
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
  int *tsite_dev = static_cast<int *>(own.take(sizeof(int) * std::max(1, nterm)));
  int64_t *toff_dev = static_cast<int64_t *>(own.take(sizeof(int64_t) * std::max(1, nterm)));
  if (nterm) {
    to_device(ctx, tsite_dev, tsite.data(), sizeof(int) * nterm);
    to_device(ctx, toff_dev, toff.data(), sizeof(int64_t) * nterm);
  }
  std::vector<int> all(ntraj);
  for (int t = 0; t < ntraj; ++t) all[t] = t;
  int *all_dev = static_cast<int *>(own.take(sizeof(int) * ntraj));
  to_device(ctx, all_dev, all.data(), sizeof(int) * ntraj);
  double *env = nullptr;
  int *centre_dev = nullptr;
  if (wide)
    centre_dev = static_cast<int *>(own.take(sizeof(int) * ntraj));
  else
    env = static_cast<double *>(own.take(sizeof(double) * size_t(ntraj) * (N + 1) * 2 * CH * CH));
  double *outd = static_cast<double *>(own.take(sizeof(double) * 2 * size_t(ntraj) * std::max(1, nterm)));
  std::vector<double> outh(2 * size_t(ntraj) * std::max(1, nterm));
  std::vector<int> centres(ntraj);
  // all terms of all trajectories in one launch; terms [q0, q0 + nq) are read back
  auto expect = [&](int q0, int nq) {
    if (wide) {
      for (int t = 0; t < ntraj; ++t) centres[t] = E.tq[t].center;
      to_device(ctx, centre_dev, centres.data(), sizeof(int) * ntraj);
      hipLaunchKernelGGL(qw_expect_kernel, dim3(ntraj), dim3(WT), 0, ctx->stream, E.b, all_dev, centre_dev, nq,
                         tsite_dev + q0, toff_dev + q0, outd);
    } else {
      hipLaunchKernelGGL(qb_expect_kernel, dim3(ntraj), dim3(QT), 0, ctx->stream, E.b, all_dev, nq, tsite_dev + q0,
                         toff_dev + q0, env, outd);
    }
    TNH_HIP(hipGetLastError());
    to_host(ctx, outh.data(), outd, sizeof(double) * 2 * size_t(ntraj) * nq);
  };
  const int nsave = p->nsteps / p->savesteps + 1;
  std::vector<int64_t> pos(ntraj, 0);
  std::vector<int> save(ntraj, 0), nj(ntraj, 0);
  int first = 1;
  if (in) {
    first = p->nsteps + 1;
    for (int t = 0; t < ntraj; ++t) {
      pos[t] = rs[t].pos;
      save[t] = rs[t].save;
      nj[t] = rs[t].nj;
      E.tq[t].center = rs[t].center;
      first = std::min(first, rs[t].step0);
    }
  }
  auto draw = [&](int t) {
    TNH_REQUIRE(pos[t] < nuniform, TNH_ERR_ARG, "qjmc: not enough uniforms for a trajectory");
    return uniforms[size_t(rows[t]) * nuniform + pos[t]++];
  };
  // the observables of the trajectories that have run `step` (step 0, fresh runs only: the
  // initial state of every trajectory)
  auto measure = [&](int step) {
    auto started = [&](int t) { return step == 0 || rs[t].step0 <= step; };
    if (nobs > 0) {
      expect(nesc, nobs);
      for (int t = 0; t < ntraj; ++t) {
        if (!started(t)) continue;
        for (int o = 0; o < nobs; ++o) {
          double *m = meas + (size_t(rows[t]) * nsave + save[t]) * 2 * nobs;
          m[2 * o] = outh[2 * (size_t(t) * nobs + o)];
          m[2 * o + 1] = outh[2 * (size_t(t) * nobs + o) + 1];
        }
      }
    }
    for (int t = 0; t < ntraj; ++t)
      if (started(t)) ++save[t];
  };
  auto normalize_all = [&](const std::vector<int> &which) {
    for (int t : which) {
      auto &x = E.tq[t];
      if (x.center == 0) E.push_move(x, 1, 0);  // norm(): gmps.jl:29-37
      x.q.push_back({P_NORM, x.center, 0, 0});
    }
    E.run_queues();
  };
  if (!in) measure(0);
  std::vector<double> rates(std::max(1, nesc));
  std::vector<char> dead(ntraj, 0);  // trajectories that outgrew their slots
  std::vector<int> live_up;          // the live list on the device
  for (int step = first; step <= p->nsteps; ++step) {
    std::vector<int> live;
    for (int t = 0; t < ntraj; ++t)
      if (!dead[t] && rs[t].step0 <= step) live.push_back(t);
    // every live trajectory's state at the start of this step: a trajectory that outgrows
    // its slots in the step continues from here on larger ones
    if (!live.empty()) {
      if (live != live_up) {
        to_device(ctx, live_dev, live.data(), sizeof(int) * live.size());
        live_up = live;
      }
      hipLaunchKernelGGL(qb_copy_state_kernel, dim3(unsigned(live.size())), dim3(256), 0, ctx->stream, E.b.S,
                         E.b.cap, E.b.bond, live_dev, snap, E.b.cap, snap_bond, live_dev, N, d);
      TNH_HIP(hipGetLastError());
    }
    for (int t : live) rs[t] = Resume{step, pos[t], save[t], nj[t], E.tq[t].center};
    for (int t : live) E.push_applygates(E.tq[t]);
    E.run_queues();
    for (int t : live) (void)draw(t);  // qjmc.jl:64: drawn and unused
    normalize_all(live);
    check_flags(ctx, E, dead);
    std::vector<int> jumpers;
    if (nesc > 0) {
      expect(0, nesc);
      for (int t : live) {
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
          const size_t row = size_t(rows[t]);
          if (nj[t] < maxjumps) {
            jumps[row * maxjumps + nj[t]] = idx + 1;
            jump_steps[row * maxjumps + nj[t]] = step;
          }
          ++nj[t];
        }
      }
    } else {
      for (int t : live)
        if (!dead[t]) (void)draw(t);  // the jump test draws even with no channels
    }
    if (!jumpers.empty()) {
      E.run_queues();
      normalize_all(jumpers);
      check_flags(ctx, E, dead);  // a split of the jump's sweeps may overflow: same step
    }
    if (step % p->savesteps == 0) measure(step);
  }
  check_flags(ctx, E, dead);
  std::vector<int> bonds(size_t(ntraj) * (N + 1));
  to_host(ctx, bonds.data(), E.b.bond, sizeof(int) * bonds.size());
  std::vector<int> gone;
  for (int t = 0; t < ntraj; ++t) {
    njumps[rows[t]] = dead[t] ? 0 : nj[t];
    int mb = 0;
    for (int i = 1; i < N; ++i) mb = std::max(mb, bonds[size_t(t) * (N + 1) + i]);
    maxbond[rows[t]] = dead[t] ? -1 : mb;
    if (dead[t]) gone.push_back(t);
  }
  if (out && !gone.empty()) {
    // the overflowed trajectories' states at the start of the step they outgrew the slots in
    const int ng = int(gone.size());
    out->ctx = ctx;
    out->cap = E.b.cap;
    out->S = static_cast<double *>(ctx->alloc(sizeof(double) * size_t(ng) * N * 2 * E.b.cap, &out->scap));
    out->bond = static_cast<int *>(ctx->alloc(sizeof(int) * size_t(ng) * (N + 1), &out->bcap));
    to_device(ctx, live_dev, gone.data(), sizeof(int) * ng);
    hipLaunchKernelGGL(qb_copy_state_kernel, dim3(ng), dim3(256), 0, ctx->stream, snap, E.b.cap, snap_bond,
                       live_dev, out->S, E.b.cap, out->bond, idx_dev, N, d);
    TNH_HIP(hipGetLastError());
    for (int t : gone) {
      out->rows.push_back(rows[t]);
      out->res.push_back(rs[t]);
    }
  }
}

// device memory of the wide reruns held at once (slots and work matrices); more overflowed
// trajectories than fit run in consecutive groups
constexpr double kWideBudgetBytes = 24.0 * (1 << 30);

}  // namespace

}  // namespace tnh

using namespace tnh;


extern "C" int tnh_qjmc_run_batched(tnh_ctx *ctx, const tnh_qjmc_problem *p, int ntraj,
                                    const double *uniforms, int64_t nuniform, double *meas, int *njumps,
                                    int *jumps, int *jump_steps, int maxjumps, int *maxbond) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
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
    const int N = p->N, d = p->d;
    std::vector<int> rows(ntraj);
    std::iota(rows.begin(), rows.end(), 0);
    // narrow slots: Gram side d CH <= 32 in LDS; environments CH x CH <= 256
    const int chn = std::min(16, QM / d);
    std::vector<std::unique_ptr<Pack>> packs;
    packs.emplace_back(new Pack());
    run_engine(ctx, p, rows, uniforms, nuniform, meas, njumps, jumps, jump_steps, maxjumps, maxbond, chn, false,
               nullptr, packs.back().get());
    if (d < 2) return;  // every bond of a d = 1 chain is 1
    // Wide slots for the trajectories that outgrew the narrow ones, each continued from the
    // start of the step it outgrew them in: bond dimension 64, then the largest the chain
    // and the truncation allow (at most 512 / d, the divide and conquer's side).
    int cmax = dc::kDcPMax / d;
    int64_t full = 1;  // d^floor(N / 2), the largest Schmidt rank of the chain
    for (int i = 0; i < N / 2 && full < cmax; ++i) full *= d;
    cmax = int(std::min<int64_t>(cmax, full));
    if (p->maxdim > 0) cmax = std::min(cmax, std::max(p->maxdim, p->mindim));
    int prev = chn;
    for (const int ch : {std::min(cmax, 64), cmax}) {
      if (ch <= prev) continue;
      prev = ch;
      std::vector<std::unique_ptr<Pack>> next;
      // slots and their step-start copies, work matrices, per trajectory
      const double per = 8.0 * (2.0 * N * 2 * ch * d * ch + double(wide_stride(d * ch)) + 2.0 * (N + 1) + 64);
      const int group = int(std::max(1.0, std::floor(kWideBudgetBytes / per)));
      for (auto &pk : packs)
        for (int g0 = 0; g0 < int(pk->rows.size()); g0 += group) {
          const PackView v{&pk->rows, &pk->res, g0, std::min(int(pk->rows.size()), g0 + group), pk->cap, pk->S,
                           pk->bond};
          next.emplace_back(new Pack());
          run_engine(ctx, p, {}, uniforms, nuniform, meas, njumps, jumps, jump_steps, maxjumps, maxbond, ch, true,
                     &v, next.back().get());
        }
      packs.swap(next);
    }
  });
}
