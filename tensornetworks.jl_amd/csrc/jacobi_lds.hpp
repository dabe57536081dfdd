// Parallel cyclic Jacobi eigensolver of a small Hermitian matrix held in LDS, shared by the
// block-Jacobi SVD (one pair sub-problem per workgroup) and the batched trajectory kernels
// (one whole Gram matrix per workgroup).
#pragma once

#include <hip/hip_runtime.h>

namespace tnh {
namespace jac {

constexpr int JB = 16;
constexpr int JP = 2 * JB;
constexpr int NS = JP / 2;  // rotations per inner step
constexpr int kMaxInnerCap = 32;

// circle-method pairing of n (even) players in round `step`, pair `i`
__host__ __device__ inline void circle_pair(int n, int step, int i, int *p, int *q) {
  if (i == 0) {
    *p = 0;
    *q = (step % (n - 1)) + 1;
  } else {
    *p = ((step + i) % (n - 1)) + 1;
    *q = ((step - i + (n - 1)) % (n - 1)) + 1;
  }
}

template <int B = JB>
__device__ inline int pair_col(const int *pr, int k, int cols) {
  const int blk = k < B ? pr[0] : pr[1];
  const int c = blk * B + (k % B);
  return (blk >= 0 && c < cols) ? c : -1;
}

template <int NT = 256>
__device__ inline double block_max(double v) {
  __shared__ double red[NT / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = red[0];
#pragma unroll
  for (int w = 1; w < NT / 64; ++w) m = fmax(m, red[w]);
  __syncthreads();
  return m;
}

// Cyclic parallel Jacobi on the 2B x 2B Hermitian matrix (gr, gi) held in LDS, accumulating
// the rotations into (qr, qi): at most `sweeps` sweeps, stopping after a sweep that rotated
// nothing (the return value says whether that happened). Every thread of the NT-thread
// block calls it. One step applies the B rotations of a circle-method round: they act on
// disjoint index pairs, so G <- J^H G J splits into B x B independent 2 x 2 blocks
// {p_I, q_I} x {p_J, q_J}, updated in place (columns by J_J, then rows by J_I^H; the rotated
// diagonal blocks take their exact values). One barrier after the rotation parameters, one
// after the update. CPLX = false skips the imaginary planes.
// With `cross` set, a sweep rotates only the JB x JB index pairs that couple the pair's two
// column blocks (local indices < JB against >= JB): JB steps of the disjoint rotations
// (i, JB + (i + step) mod JB) instead of the JP - 1 steps of the full cyclic order.
template <int B>
__device__ inline void step_pair(bool cross, int step, int i, int *p, int *q) {
  if (cross) {
    *p = i;
    *q = B + ((i + step) % B);
  } else {
    circle_pair(2 * B, step, i, p, q);
  }
}

template <bool CPLX, int B, int NT>
__device__ bool jacobi_lds(double (*gr)[2 * B + 1], double (*gi)[2 * B + 1], double (*qr)[2 * B + 1],
                           double (*qi)[2 * B + 1], double tol, double abstol, int sweeps,
                           bool cross = false) {
  constexpr int JP = 2 * B, NS = B;
  static_assert(NT % NS == 0, "thread layout of the step tasks");  // threads beyond the tasks idle
  const int t = threadIdx.x;
  __shared__ int rrot[NS];
  __shared__ double rcs[NS], rsn[NS], rer[NS], rei[NS], ra[NS], rc[NS], rtb[NS];
  __shared__ int anyrot[kMaxInnerCap];
  if (t < kMaxInnerCap) anyrot[t] = 0;
  __syncthreads();
  if (sweeps > kMaxInnerCap) sweeps = kMaxInnerCap;
  for (int sw = 0; sw < sweeps; ++sw) {
    const int nsteps = cross ? B : JP - 1;
    for (int step = 0; step < nsteps; ++step) {
      if (t < NS) {
        int p, q;
        step_pair<B>(cross, step, t, &p, &q);
        const double a = gr[p][p], c = gr[q][q];
        const double br = gr[p][q], bi = CPLX ? gi[p][q] : 0.0;
        // the thresholds |b| > abstol and |b| > tol sqrt(|a c|), compared squared (callers
        // scale G to max_i G_ii ~ 1, so no square overflows)
        const double b2 = CPLX ? br * br + bi * bi : br * br;
        const bool rot = b2 > abstol * abstol && b2 > tol * tol * fabs(a * c);
        double cs = 1, sn = 0, tt = 0, er = 1, ei = 0, bb = 0;
        if (rot) {
          // t = sign(c - a) / (|tau| + sqrt(1 + tau^2)) with tau = (c - a) / (2 |b|), in the
          // form 2 |b| sign(c - a) / (|c - a| + sqrt((c - a)^2 + 4 |b|^2)): one square root
          // and one division fewer on the step's critical path
          bb = CPLX ? sqrt(b2) : fabs(br);
          const double dd = c - a;
          tt = (dd >= 0 ? 2.0 : -2.0) * bb / (fabs(dd) + sqrt(dd * dd + 4.0 * b2));
          cs = rsqrt(1 + tt * tt);
          sn = tt * cs;
          if (CPLX) {  // e^{-i phi} = conj(b) / |b|
            const double ib = 1.0 / bb;
            er = br * ib;
            ei = -bi * ib;
          } else {
            er = br >= 0 ? 1.0 : -1.0;
          }
          anyrot[sw] = 1;
        }
        rrot[t] = rot;
        rcs[t] = cs;
        rsn[t] = sn;
        rer[t] = er;
        rei[t] = ei;
        ra[t] = a;
        rc[t] = c;
        rtb[t] = tt * bb;
      }
      __syncthreads();
      {
        // Each thread keeps one rotation I = t % NS in registers: it updates the 2 x 2 blocks
        // (I, J) of G for J = t / NS + k NT / NS (rows by J_I^H, columns by J_J) and the rows
        // t / NS + k NT / NS of the column pair of I in Q.
        const int I = t % NS;
        int pI, qI;  // the pairing, recomputed instead of read from LDS
        step_pair<B>(cross, step, I, &pI, &qI);
        const bool rI = rrot[I];
        const double csI = rcs[I], snI = rsn[I], erI = rer[I], eiI = rei[I];
        for (int J = t / NS; J < NS; J += NT / NS) {
          const bool rJ = rrot[J];
          if (!rI && !rJ) continue;
          int pJ, qJ;
          step_pair<B>(cross, step, J, &pJ, &qJ);
          if (I == J) {  // rotated diagonal block: exact values
            gr[pI][qI] = gr[qI][pI] = 0.0;
            gr[pI][pI] = ra[I] - rtb[I];
            gr[qI][qI] = rc[I] + rtb[I];
            if (CPLX) gi[pI][qI] = gi[qI][pI] = gi[pI][pI] = gi[qI][qI] = 0.0;
            continue;
          }
          // x = [[g(pI,pJ), g(pI,qJ)], [g(qI,pJ), g(qI,qJ)]]
          double x00r = gr[pI][pJ], x01r = gr[pI][qJ], x10r = gr[qI][pJ], x11r = gr[qI][qJ];
          double x00i = 0, x01i = 0, x10i = 0, x11i = 0;
          if (CPLX) {
            x00i = gi[pI][pJ];
            x01i = gi[pI][qJ];
            x10i = gi[qI][pJ];
            x11i = gi[qI][qJ];
          }
          if (rJ) {  // columns: [x_p x_q] <- [x_p x_q] [[cs, sn], [-sn e, cs e]]
            const double cs = rcs[J], sn = rsn[J], er = rer[J], ei = rei[J];
            double eqr = er * x01r - ei * x01i, eqi = er * x01i + ei * x01r;
            double npr = cs * x00r - sn * eqr, npi = cs * x00i - sn * eqi;
            double nqr = sn * x00r + cs * eqr, nqi = sn * x00i + cs * eqi;
            x00r = npr; x00i = npi; x01r = nqr; x01i = nqi;
            eqr = er * x11r - ei * x11i;
            eqi = er * x11i + ei * x11r;
            npr = cs * x10r - sn * eqr;
            npi = cs * x10i - sn * eqi;
            nqr = sn * x10r + cs * eqr;
            nqi = sn * x10i + cs * eqi;
            x10r = npr; x10i = npi; x11r = nqr; x11i = nqi;
          }
          if (rI) {  // rows: [x_p; x_q] <- J^H [x_p; x_q]  (conj(e))
            const double cs = csI, sn = snI, er = erI, ei = -eiI;
            double eqr = er * x10r - ei * x10i, eqi = er * x10i + ei * x10r;
            double npr = cs * x00r - sn * eqr, npi = cs * x00i - sn * eqi;
            double nqr = sn * x00r + cs * eqr, nqi = sn * x00i + cs * eqi;
            x00r = npr; x00i = npi; x10r = nqr; x10i = nqi;
            eqr = er * x11r - ei * x11i;
            eqi = er * x11i + ei * x11r;
            npr = cs * x01r - sn * eqr;
            npi = cs * x01i - sn * eqi;
            nqr = sn * x01r + cs * eqr;
            nqi = sn * x01i + cs * eqi;
            x01r = npr; x01i = npi; x11r = nqr; x11i = nqi;
          }
          gr[pI][pJ] = x00r;
          gr[pI][qJ] = x01r;
          gr[qI][pJ] = x10r;
          gr[qI][qJ] = x11r;
          if (CPLX) {
            gi[pI][pJ] = x00i;
            gi[pI][qJ] = x01i;
            gi[qI][pJ] = x10i;
            gi[qI][qJ] = x11i;
          }
        }
        if (rI) {  // Q <- Q J_I on this thread's rows
          for (int r = t / NS; r < JP; r += NT / NS) {
            const double pr_ = qr[r][pI], qr_ = qr[r][qI];
            const double pi_ = CPLX ? qi[r][pI] : 0.0, qi_ = CPLX ? qi[r][qI] : 0.0;
            const double eqr = erI * qr_ - eiI * qi_, eqi = erI * qi_ + eiI * qr_;
            qr[r][pI] = csI * pr_ - snI * eqr;
            qr[r][qI] = snI * pr_ + csI * eqr;
            if (CPLX) {
              qi[r][pI] = csI * pi_ - snI * eqi;
              qi[r][qI] = snI * pi_ + csI * eqi;
            }
          }
        }
      }
      __syncthreads();
    }
    // anyrot[sw] is complete after the step's final barrier and never rewritten
    if (!anyrot[sw]) return true;
  }
  return false;
}

}  // namespace jac
}  // namespace tnh
