// Two-site DMRG contractions as MFMA GEMM chains with an MPO-mixing stage between them:
//   H_eff matvec   (product(::ProjMPS), src/structures/mps/projmps.jl:103-144)
//   env updates    (buildleft!/buildright!, projmps.jl:50-95)
//   calculate      (projmps.jl:192-216)
//   theta = A1*A2  (dmrg.jl:44-47)
//
// No tensor is permuted in memory: every stage reads the previous stage's
// column-major output through a GEMM leading dimension or a per-index offset table.
#include <algorithm>
#include <cmath>

#include "tnh_internal.hpp"

namespace tnh {

namespace {

constexpr int kMaxIJ = 64;

struct SmallParams {
  int64_t total;  // na * nb
  int na, I, J;
  const void *in;
  int64_t in_im;
  void *out;
  int64_t out_im;
  const double *Wc;
  int64_t sbI, sbJ;
  double beta;
  int64_t offI[kMaxIJ];
  int64_t offJ[kMaxIJ];
};

// out[a, j, b] = sum_i in[a, i, b] * Wc[i, j]; one thread per (a, b), a fastest so
// that every load and store of a wave is a contiguous run along a.
template <typename TI, typename TO, bool CIN, bool CW>
__global__ __launch_bounds__(256) void small_kernel(SmallParams p) {
  using R = typename std::conditional<std::is_same<TI, double>::value ||
                                          std::is_same<TO, double>::value,
                                      double, float>::type;
  extern __shared__ double wsm[];
  R *W = reinterpret_cast<R *>(wsm);
  const int IJ = p.I * p.J;
  for (int e = threadIdx.x; e < IJ * (CW ? 2 : 1); e += blockDim.x) W[e] = R(p.Wc[e]);
  __syncthreads();

  const int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (idx >= p.total) return;
  const int a = int(idx % p.na);
  const int64_t b = idx / p.na;
  const TI *in = static_cast<const TI *>(p.in) + a + b * p.sbI;
  TO *out = static_cast<TO *>(p.out) + a + b * p.sbJ;

  R vr[kMaxIJ], vi[CIN ? kMaxIJ : 1];
#pragma unroll
  for (int i = 0; i < kMaxIJ; ++i) {
    if (i < p.I) {
      vr[i] = R(in[p.offI[i]]);
      if constexpr (CIN) vi[i] = R(in[p.in_im + p.offI[i]]);
    }
  }
  for (int j = 0; j < p.J; ++j) {
    const R *wr = W + j * p.I;
    const R *wi = W + IJ + j * p.I;
    R sr = 0, si = 0;
#pragma unroll
    for (int i = 0; i < kMaxIJ; ++i) {
      if (i < p.I) {
        sr += vr[i] * wr[i];
        if constexpr (CIN && CW) {
          sr -= vi[i] * wi[i];
          si += vr[i] * wi[i] + vi[i] * wr[i];
        } else if constexpr (CIN) {
          si += vi[i] * wr[i];
        } else if constexpr (CW) {
          si += vr[i] * wi[i];
        }
      }
    }
    TO *o = out + p.offJ[j];
    if (p.beta != 0.0) {
      sr += R(p.beta) * R(o[0]);
      if constexpr (CIN || CW) si += R(p.beta) * R(o[p.out_im]);
    }
    o[0] = TO(sr);
    if constexpr (CIN || CW) o[p.out_im] = TO(si);
  }
}

// Vector form of small_kernel for a real input and I, J <= 32: each thread takes VA
// consecutive a (one 16-byte load per input column), so a wave moves 1 KB per load or
// store instruction and every mixing coefficient read from LDS serves VA multiply-adds.
// The host checks that na, the offsets, the batch strides and the imaginary-plane offset
// are multiples of VA and that both base pointers are 16-byte aligned.
constexpr int kVecIJ = 32;
template <typename TI, typename TO, bool CW, int VA>
__global__ __launch_bounds__(256) void small_vec_kernel(SmallParams p) {
  using R = typename std::conditional<std::is_same<TI, double>::value ||
                                          std::is_same<TO, double>::value,
                                      double, float>::type;
  typedef TI VI __attribute__((ext_vector_type(VA)));
  typedef TO VO __attribute__((ext_vector_type(VA)));
  extern __shared__ double wsm[];
  R *W = reinterpret_cast<R *>(wsm);
  const int IJ = p.I * p.J;
  for (int e = threadIdx.x; e < IJ * (CW ? 2 : 1); e += blockDim.x) W[e] = R(p.Wc[e]);
  __syncthreads();

  const int nav = p.na / VA;
  const int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (idx >= p.total / VA) return;
  const int a = int(idx % nav) * VA;
  const int64_t b = idx / nav;
  const TI *in = static_cast<const TI *>(p.in) + a + b * p.sbI;
  TO *out = static_cast<TO *>(p.out) + a + b * p.sbJ;

  R v[kVecIJ][VA];
#pragma unroll
  for (int i = 0; i < kVecIJ; ++i) {
    if (i < p.I) {
      const VI x = *reinterpret_cast<const VI *>(in + p.offI[i]);
#pragma unroll
      for (int u = 0; u < VA; ++u) v[i][u] = R(x[u]);
    }
  }
  for (int j = 0; j < p.J; ++j) {
    const R *wr = W + j * p.I;
    const R *wi = W + IJ + j * p.I;
    R sr[VA], si[VA];
#pragma unroll
    for (int u = 0; u < VA; ++u) sr[u] = si[u] = R(0);
#pragma unroll
    for (int i = 0; i < kVecIJ; ++i) {
      if (i < p.I) {
        const R c = wr[i];
#pragma unroll
        for (int u = 0; u < VA; ++u) sr[u] += v[i][u] * c;
        if constexpr (CW) {
          const R ci = wi[i];
#pragma unroll
          for (int u = 0; u < VA; ++u) si[u] += v[i][u] * ci;
        }
      }
    }
    TO *o = out + p.offJ[j];
    if (p.beta != 0.0) {
      const VO x = *reinterpret_cast<const VO *>(o);
#pragma unroll
      for (int u = 0; u < VA; ++u) sr[u] += R(p.beta) * R(x[u]);
      if constexpr (CW) {
        const VO y = *reinterpret_cast<const VO *>(o + p.out_im);
#pragma unroll
        for (int u = 0; u < VA; ++u) si[u] += R(p.beta) * R(y[u]);
      }
    }
    VO rv, iv;
#pragma unroll
    for (int u = 0; u < VA; ++u) {
      rv[u] = TO(sr[u]);
      iv[u] = TO(si[u]);
    }
    *reinterpret_cast<VO *>(o) = rv;
    if constexpr (CW) *reinterpret_cast<VO *>(o + p.out_im) = iv;
  }
}

// Wc[(w,s1,s2),(o1,o2,w2)] = sum_w1 W1[w,o1,s1,w1] W2[w1,o2,s2,w2] (planar complex out).
template <typename T>
__global__ void wc_pair_kernel(const T *W1, int64_t w1_im, const T *W2, int64_t w2_im, int D1,
                               int d, int Dm, int D3, double *Wc, bool cplx) {
  const int I = D1 * d * d, J = d * d * D3;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= I * J) return;
  const int i = e % I, j = e / I;
  const int w = i % D1, s1 = (i / D1) % d, s2 = i / (D1 * d);
  const int o1 = j % d, o2 = (j / d) % d, w2 = j / (d * d);
  double sr = 0, si = 0;
  for (int w1 = 0; w1 < Dm; ++w1) {
    int64_t a = w + int64_t(D1) * (o1 + int64_t(d) * (s1 + int64_t(d) * w1));
    int64_t b = w1 + int64_t(Dm) * (o2 + int64_t(d) * (s2 + int64_t(d) * w2));
    double ar = double(W1[a]), br = double(W2[b]);
    double ai = w1_im ? double(W1[w1_im + a]) : 0.0;
    double bi = w2_im ? double(W2[w2_im + b]) : 0.0;
    sr += ar * br - ai * bi;
    si += ar * bi + ai * br;
  }
  Wc[e] = sr;
  if (cplx) Wc[int64_t(I) * J + e] = si;
}

// Environment mixing matrices from one MPO tensor W[w,s,t,w'] (Dl, d, d, Dr):
//   left : Wl[(w,t),(s,w')]   right: Wr[(t,w'),(w,s)]
template <typename T>
__global__ void wc_env_kernel(const T *W, int64_t w_im, int Dl, int d, int Dr, double *Wc,
                              bool cplx, bool left) {
  const int I = left ? Dl * d : d * Dr, J = left ? d * Dr : Dl * d;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= I * J) return;
  const int i = e % I, j = e / I;
  int w, s, t, wp;
  if (left) {
    w = i % Dl;
    t = i / Dl;
    s = j % d;
    wp = j / d;
  } else {
    t = i % d;
    wp = i / d;
    w = j % Dl;
    s = j / Dl;
  }
  int64_t q = w + int64_t(Dl) * (s + int64_t(d) * (t + int64_t(d) * wp));
  Wc[e] = double(W[q]);
  if (cplx) Wc[int64_t(I) * J + e] = w_im ? double(W[w_im + q]) : 0.0;
}

// planar complex in-place scale by (cr + i ci)
template <typename T>
__global__ void cscale_kernel(T *x, int64_t n, double cr, double ci) {
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    double xr = double(x[k]), xi = double(x[n + k]);
    x[k] = T(cr * xr - ci * xi);
    x[n + k] = T(cr * xi + ci * xr);
  }
}

__device__ inline double block_reduce(double v) {
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) s = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return s;
}

// sum_k x[k] y[k] over planar complex / real data: out[0] = re, out[1] = im (partials)
__global__ void cdot_partial_kernel(const double *x, int64_t xim, const double *y, int64_t yim,
                                    int64_t n, double *part) {
  double sr = 0, si = 0;
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    double xr = x[k], yr = y[k];
    double xi = xim ? x[xim + k] : 0.0, yi = yim ? y[yim + k] : 0.0;
    sr += xr * yr - xi * yi;
    si += xr * yi + xi * yr;
  }
  sr = block_reduce(sr);
  si = block_reduce(si);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sr;
    part[2 * blockIdx.x + 1] = si;
  }
}

}  // namespace

// ------------------------------------------------------------------ launchers

void small_contract(tnh_ctx *ctx, const Small &s) {
  TNH_REQUIRE(s.I <= kMaxIJ && s.J <= kMaxIJ, TNH_ERR_UNSUPPORTED,
              "MPO bond dimension too large for the mixing stage (I, J <= 64)");
  TNH_REQUIRE((int)s.offI.size() >= s.I && (int)s.offJ.size() >= s.J, TNH_ERR_ARG,
              "small_contract: offsets");
  if (s.na <= 0 || s.nb <= 0) return;
  SmallParams p{};
  p.total = int64_t(s.na) * s.nb;
  p.na = s.na;
  p.I = s.I;
  p.J = s.J;
  p.in = s.in;
  p.in_im = s.in_im;
  p.out = s.out;
  p.out_im = s.out_im;
  p.Wc = s.Wc;
  p.sbI = s.sbI;
  p.sbJ = s.sbJ;
  p.beta = s.beta;
  for (int i = 0; i < s.I; ++i) p.offI[i] = s.offI[i];
  for (int j = 0; j < s.J; ++j) p.offJ[j] = s.offJ[j];
  const bool cin = s.in_im != 0, cw = s.Wc_cplx;
  TNH_REQUIRE(!(cin || cw) || s.out_im != 0, TNH_ERR_ARG, "small_contract: complex output needed");
  const bool i64 = s.tin == TNH_F64, o64 = s.tout == TNH_F64;
  const size_t shm = size_t(s.I) * s.J * (cw ? 2 : 1) * ((i64 || o64) ? 8 : 4);
  dim3 grid(unsigned((p.total + 255) / 256));
  hipEvent_t ev = ctx->prof_begin();
  // vector form: real input, I, J <= 32, everything a multiple of the 16-byte vector width
  const int va = (i64 || o64) ? 2 : 4;  // FP64 arithmetic: two values per thread (register budget)
  bool vec = !cin && s.I <= kVecIJ && s.J <= kVecIJ && s.na % va == 0 && s.sbI % va == 0 &&
             s.sbJ % va == 0 && (!cw || s.out_im % va == 0) &&
             (reinterpret_cast<uintptr_t>(s.in) & 15u) == 0 && (reinterpret_cast<uintptr_t>(s.out) & 15u) == 0;
  for (int i = 0; vec && i < s.I; ++i) vec = s.offI[i] % va == 0;
  for (int j = 0; vec && j < s.J; ++j) vec = s.offJ[j] % va == 0;
  if (vec) {
    dim3 gv(unsigned((p.total / va + 255) / 256));
#define TNH_SMALLV(TI, TO, VA)                                                                 \
  do {                                                                                          \
    if (cw)                                                                                     \
      hipLaunchKernelGGL((small_vec_kernel<TI, TO, true, VA>), gv, dim3(256), shm, ctx->stream, p); \
    else                                                                                        \
      hipLaunchKernelGGL((small_vec_kernel<TI, TO, false, VA>), gv, dim3(256), shm, ctx->stream, p); \
  } while (0)
    if (i64 && o64)
      TNH_SMALLV(double, double, 2);
    else if (i64)
      TNH_SMALLV(double, float, 2);
    else if (o64)
      TNH_SMALLV(float, double, 2);
    else
      TNH_SMALLV(float, float, 4);
#undef TNH_SMALLV
  } else {
#define TNH_SMALL(TI, TO)                                                                     \
  do {                                                                                          \
    if (cin && cw)                                                                              \
      hipLaunchKernelGGL((small_kernel<TI, TO, true, true>), grid, dim3(256), shm, ctx->stream, p); \
    else if (cin)                                                                               \
      hipLaunchKernelGGL((small_kernel<TI, TO, true, false>), grid, dim3(256), shm, ctx->stream, p); \
    else if (cw)                                                                                \
      hipLaunchKernelGGL((small_kernel<TI, TO, false, true>), grid, dim3(256), shm, ctx->stream, p); \
    else                                                                                        \
      hipLaunchKernelGGL((small_kernel<TI, TO, false, false>), grid, dim3(256), shm, ctx->stream, p); \
  } while (0)
    if (i64 && o64)
      TNH_SMALL(double, double);
    else if (i64)
      TNH_SMALL(double, float);
    else if (o64)
      TNH_SMALL(float, double);
    else
      TNH_SMALL(float, float);
#undef TNH_SMALL
  }
  TNH_HIP(hipGetLastError());
  if (ev) {
    const double planes_in = cin ? 2 : 1, planes_out = (cin || cw) ? 2 : 1;
    const double cmul = (cin && cw) ? 4 : ((cin || cw) ? 2 : 1);
    ctx->prof_end("mix", ev, 2.0 * cmul * double(p.total) * s.I * s.J,
                  double(p.total) * (s.I * planes_in * real_size(s.tin) +
                                     s.J * planes_out * real_size(s.tout) * (s.beta != 0.0 ? 2 : 1)));
  }
}

namespace {

// The 1x1x1 block of ones used outside the chain (abstractprojmps.jl:33-46).
const tnh_buf *edge_block(tnh_ctx *ctx) {
  if (ctx->arena) {  // inside a call that allocates nothing: a block carved from its scratch
    tnh_buf &e = ctx->arena->edge;
    if (!e.ptr) {
      e.ptr = ctx->workspace(800, 8);
      fill(ctx, e.ptr, TNH_F64, 1, 1.0);
      e.dtype = TNH_F64;
      e.ndim = 3;
      e.dims[0] = e.dims[1] = e.dims[2] = 1;
      e.numel = 1;
      e.bytes = 8;
      e.owned = false;
    }
    return &e;
  }
  if (!ctx->edge.ptr) {
    void *p = ctx->workspace(800, 8);
    double one = 1.0;
    TNH_HIP(hipMemcpyAsync(p, &one, 8, hipMemcpyHostToDevice, ctx->stream));
    TNH_HIP(hipStreamSynchronize(ctx->stream));
    ctx->edge.dtype = TNH_F64;
    ctx->edge.ndim = 3;
    ctx->edge.dims[0] = ctx->edge.dims[1] = ctx->edge.dims[2] = 1;
    ctx->edge.numel = 1;
    ctx->edge.bytes = 8;
    ctx->edge.owned = false;
    ctx->edge.ptr = p;
  }
  return &ctx->edge;
}

COperand cop(const tnh_buf *b, bool conj = false) {
  COperand o;
  o.re = b->re();
  o.im = is_complex(b->dtype) ? b->im() : nullptr;
  o.dtype = real_of(b->dtype);
  o.conj = conj;
  return o;
}

void require_dims(const tnh_buf *b, int nd, const char *name) {
  TNH_REQUIRE(b != nullptr, TNH_ERR_ARG, std::string(name) + " is NULL");
  TNH_REQUIRE(b->ndim == nd, TNH_ERR_SHAPE,
              std::string(name) + " must have " + std::to_string(nd) + " dims, got " +
                  shape_str(b));
}

int inter_dtype(int prec, bool cplx) {
  int r = prec == TNH_PREC_F64 ? TNH_F64 : TNH_F32;
  return cplx ? complex_of(r) : r;
}

void check_prec_operand(int prec, const tnh_buf *b, const char *name) {
  if (prec == TNH_PREC_F64)
    TNH_REQUIRE(is_double(b->dtype), TNH_ERR_ARG,
                std::string(name) + " is FP32 but FP64 arithmetic was requested");
}

// Device buffer of the mixing matrix for an operator pair (slot = workspace slot).
double *make_wc_pair(tnh_ctx *ctx, const tnh_buf *W1, const tnh_buf *W2, int slot, bool *cplx) {
  const int D1 = int(W1->dims[0]), d = int(W1->dims[1]), Dm = int(W1->dims[3]);
  const int D3 = int(W2->dims[3]);
  const int I = D1 * d * d, J = d * d * D3;
  *cplx = is_complex(W1->dtype) || is_complex(W2->dtype);
  double *Wc = static_cast<double *>(ctx->workspace(slot, sizeof(double) * I * J * 2));
  int n = I * J;
  TNH_REQUIRE(W1->dtype == W2->dtype || (real_of(W1->dtype) == real_of(W2->dtype)),
              TNH_ERR_ARG, "W1 and W2 must share a precision");
  int64_t im1 = is_complex(W1->dtype) ? W1->numel : 0;
  int64_t im2 = is_complex(W2->dtype) ? W2->numel : 0;
  if (is_double(W1->dtype))
    hipLaunchKernelGGL(wc_pair_kernel<double>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const double *>(W1->ptr), im1,
                       static_cast<const double *>(W2->ptr), im2, D1, d, Dm, D3, Wc, *cplx);
  else
    hipLaunchKernelGGL(wc_pair_kernel<float>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const float *>(W1->ptr), im1,
                       static_cast<const float *>(W2->ptr), im2, D1, d, Dm, D3, Wc, *cplx);
  TNH_HIP(hipGetLastError());
  return Wc;
}

double *make_wc_env(tnh_ctx *ctx, const tnh_buf *W, bool left, int slot, bool *cplx) {
  const int Dl = int(W->dims[0]), d = int(W->dims[1]), Dr = int(W->dims[3]);
  const int n = Dl * d * d * Dr;
  *cplx = is_complex(W->dtype);
  double *Wc = static_cast<double *>(ctx->workspace(slot, sizeof(double) * n * 2));
  int64_t im = *cplx ? W->numel : 0;
  if (is_double(W->dtype))
    hipLaunchKernelGGL(wc_env_kernel<double>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const double *>(W->ptr), im, Dl, d, Dr, Wc, *cplx, left);
  else
    hipLaunchKernelGGL(wc_env_kernel<float>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const float *>(W->ptr), im, Dl, d, Dr, Wc, *cplx, left);
  TNH_HIP(hipGetLastError());
  return Wc;
}

}  // namespace

void cscale(tnh_ctx *ctx, void *x, int t, int64_t n, double cr, double ci) {
  if (n <= 0) return;
  if (!is_complex(t)) {
    scale(ctx, x, t, n, cr);
    return;
  }
  int g = int(std::min<int64_t>((n + 255) / 256, int64_t(ctx->cu_count) * 8));
  if (is_double(t))
    hipLaunchKernelGGL(cscale_kernel<double>, dim3(g), dim3(256), 0, ctx->stream,
                       static_cast<double *>(x), n, cr, ci);
  else
    hipLaunchKernelGGL(cscale_kernel<float>, dim3(g), dim3(256), 0, ctx->stream,
                       static_cast<float *>(x), n, cr, ci);
  TNH_HIP(hipGetLastError());
}

// sum_k x[k] y[k] (no conjugation) of two FP64 device tensors -> (re, im) on the host
void cdot_host(tnh_ctx *ctx, const tnh_buf *x, const tnh_buf *y, double *re, double *im) {
  TNH_REQUIRE(x->numel == y->numel, TNH_ERR_SHAPE, "dot: size mismatch");
  TNH_REQUIRE(is_double(x->dtype) && is_double(y->dtype), TNH_ERR_ARG, "dot: FP64 operands");
  const int g = 256;
  double *part = static_cast<double *>(ctx->workspace(803, sizeof(double) * 2 * g));
  hipLaunchKernelGGL(cdot_partial_kernel, dim3(g), dim3(256), 0, ctx->stream,
                     static_cast<const double *>(x->ptr), is_complex(x->dtype) ? x->numel : 0,
                     static_cast<const double *>(y->ptr), is_complex(y->dtype) ? y->numel : 0,
                     x->numel, part);
  TNH_HIP(hipGetLastError());
  double h[2 * g];
  to_host(ctx, h, part, sizeof(h));
  double sr = 0, si = 0;
  for (int i = 0; i < g; ++i) {
    sr += h[2 * i];
    si += h[2 * i + 1];
  }
  *re = sr;
  *im = si;
}

// ------------------------------------------------------------- H_eff matvec

void heff_apply(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2,
                const tnh_buf *R, const tnh_buf *x, tnh_buf *out, double cr, double ci,
                int prec) {
  if (!L) L = edge_block(ctx);
  if (!R) R = edge_block(ctx);
  require_dims(L, 3, "L");
  require_dims(W1, 4, "W1");
  require_dims(W2, 4, "W2");
  require_dims(R, 3, "R");
  require_dims(x, 4, "theta");
  require_dims(out, 4, "out");
  const int64_t cl = x->dims[0], d = x->dims[1], cr_ = x->dims[3];
  const int64_t D1 = W1->dims[0], Dm = W1->dims[3], D3 = W2->dims[3];
  TNH_REQUIRE(x->dims[2] == d && W1->dims[1] == d && W1->dims[2] == d && W2->dims[1] == d &&
                  W2->dims[2] == d,
              TNH_ERR_SHAPE, "physical dimensions of theta, W1, W2 differ");
  TNH_REQUIRE(L->dims[0] == cl && L->dims[2] == cl && L->dims[1] == D1, TNH_ERR_SHAPE,
              "L " + shape_str(L) + " does not match theta " + shape_str(x) + " / W1 " +
                  shape_str(W1));
  TNH_REQUIRE(W2->dims[0] == Dm, TNH_ERR_SHAPE, "W1 right bond != W2 left bond");
  TNH_REQUIRE(R->dims[0] == cr_ && R->dims[2] == cr_ && R->dims[1] == D3, TNH_ERR_SHAPE,
              "R " + shape_str(R) + " does not match theta " + shape_str(x) + " / W2 " +
                  shape_str(W2));
  for (int i = 0; i < 4; ++i)
    TNH_REQUIRE(out->dims[i] == x->dims[i], TNH_ERR_SHAPE, "out shape != theta shape");
  check_prec_operand(prec, L, "L");
  check_prec_operand(prec, R, "R");
  check_prec_operand(prec, x, "theta");
  check_prec_operand(prec, out, "out");
  TNH_REQUIRE(out != x, TNH_ERR_ARG, "out must not alias theta");
  TNH_REQUIRE(D1 * d * d <= kMaxIJ && d * d * D3 <= kMaxIJ, TNH_ERR_UNSUPPORTED,
              "D*d^2 > 64 is not supported by the mixing stage");

  bool wc_cplx = false;
  double *Wc = make_wc_pair(ctx, W1, W2, 810, &wc_cplx);
  const bool t1c = is_complex(L->dtype) || is_complex(x->dtype);
  const bool t2c = t1c || wc_cplx;
  const bool oc = t2c || is_complex(R->dtype);
  TNH_REQUIRE(!oc || is_complex(out->dtype), TNH_ERR_ARG,
              "complex operands need a complex output tensor");
  const int tT1 = inter_dtype(prec, t1c), tT2 = inter_dtype(prec, t2c);
  const int64_t n1 = cl * D1 * d * d * cr_, n2 = cl * d * d * D3 * cr_;
  const size_t rs = real_size(tT1);
  char *T1 = static_cast<char *>(ctx->workspace(811, n1 * rs * (t1c ? 2 : 1)));
  char *T2 = static_cast<char *>(ctx->workspace(812, n2 * rs * (t2c ? 2 : 1)));

  // (1) T1[(a,w),(s1,s2,b)] = L[(a,w),a'] x[a',(s1,s2,b)]
  // FP32 arithmetic on an FP64 theta: one rounding pass to an FP32 copy first (a streaming
  // kernel at HBM rate) lets the GEMM stage 16-byte FP32 vectors without converting inside
  // its k loop; skipped for small theta, where the extra launch would cost more.
  COperand xop = cop(x);
  if (prec == TNH_PREC_F32 && is_double(x->dtype) && x->numel >= (int64_t(1) << 20)) {
    const int tx = is_complex(x->dtype) ? complex_of(TNH_F32) : TNH_F32;
    char *xs = static_cast<char *>(ctx->workspace(813, size_t(x->numel) * 4 * (is_complex(x->dtype) ? 2 : 1)));
    reformat(ctx, x->ptr, x->dtype, false, xs, tx, false, x->numel);
    xop.re = xs;
    xop.im = is_complex(x->dtype) ? xs + size_t(x->numel) * 4 : nullptr;
    xop.dtype = TNH_F32;
  }
  Gemm g1;
  g1.M = int(cl * D1);
  g1.N = int(d * d * cr_);
  g1.K = int(cl);
  g1.lda = cl * D1;
  g1.ldb = cl;
  g1.ldc = cl * D1;
  cgemm(ctx, g1, cop(L), xop, T1, t1c ? T1 + n1 * rs : nullptr, real_of(tT1), prec);

  // (2) T2[a,(o1,o2,w2),b] = T1[a,(w,s1,s2),b] Wc[(w,s1,s2),(o1,o2,w2)]
  Small s;
  s.na = int(cl);
  s.nb = int(cr_);
  s.I = int(D1 * d * d);
  s.J = int(d * d * D3);
  s.in = T1;
  s.in_im = t1c ? n1 : 0;
  s.tin = real_of(tT1);
  s.out = T2;
  s.out_im = t2c ? n2 : 0;
  s.tout = real_of(tT2);
  s.Wc = Wc;
  s.Wc_cplx = wc_cplx;
  s.sbI = cl * D1 * d * d;
  s.sbJ = cl * d * d * D3;
  s.offI.resize(s.I);
  s.offJ.resize(s.J);
  for (int i = 0; i < s.I; ++i) s.offI[i] = cl * i;
  for (int j = 0; j < s.J; ++j) s.offJ[j] = cl * j;
  small_contract(ctx, s);

  // (3) out[(a,o1,o2),b'] = coeff * T2[(a,o1,o2),(w2,b)] R[b',(w2,b)]^T
  Gemm g3;
  g3.M = int(cl * d * d);
  g3.N = int(cr_);
  g3.K = int(D3 * cr_);
  g3.lda = cl * d * d;
  g3.ldb = cr_;
  g3.transB = true;
  g3.ldc = cl * d * d;
  g3.alpha = (ci == 0.0) ? cr : 1.0;
  COperand a3;
  a3.re = T2;
  a3.im = t2c ? T2 + n2 * rs : nullptr;
  a3.dtype = real_of(tT2);
  cgemm(ctx, g3, a3, cop(R), out->re(), is_complex(out->dtype) ? out->im() : nullptr,
        real_of(out->dtype), prec);
  if (ci != 0.0) {
    TNH_REQUIRE(is_complex(out->dtype), TNH_ERR_ARG, "complex coefficient needs a complex output");
    cscale(ctx, out->ptr, out->dtype, out->numel, cr, ci);
  }
}

// Upper bound of the device workspace heff_apply takes for these operands: every slot it
// requests counted separately (no reuse assumed), each rounded up to 256 bytes. Callers
// that bring their own scratch size it with this before the call.
size_t heff_scratch_bound(const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2, const tnh_buf *R,
                          const tnh_buf *x, int prec) {
  require_dims(W1, 4, "W1");
  require_dims(W2, 4, "W2");
  require_dims(x, 4, "theta");
  auto up = [](size_t b) { return (std::max<size_t>(b, 1) + 255) & ~size_t(255); };
  const int64_t cl = x->dims[0], d = x->dims[1], cr_ = x->dims[3];
  const int64_t D1 = W1->dims[0], D3 = W2->dims[3];
  const bool wc_cplx = is_complex(W1->dtype) || is_complex(W2->dtype);
  const bool t1c = (L && is_complex(L->dtype)) || is_complex(x->dtype);
  const bool t2c = t1c || wc_cplx;
  const size_t rs = prec == TNH_PREC_F64 ? 8 : 4;
  const size_t I = size_t(D1 * d * d), J = size_t(d * d * D3);
  const size_t n1 = size_t(cl * D1 * d * d * cr_), n2 = size_t(cl * d * d * D3 * cr_);
  size_t b = up(sizeof(double) * I * J * 2) + up(n1 * rs * (t1c ? 2 : 1)) + up(n2 * rs * (t2c ? 2 : 1));
  if (!L || !R) b += up(8);
  if (prec == TNH_PREC_F32 && is_double(x->dtype) && x->numel >= (int64_t(1) << 20))
    b += up(size_t(x->numel) * 4 * (is_complex(x->dtype) ? 2 : 1));
  if (prec == TNH_PREC_BF16X3) {
    // hi/lo bf16 planes of both operands of both GEMM stages, k padded to 32
    auto planes = [&](int64_t rows, int64_t k) { return up(size_t(2) * rows * ((k + 31) / 32 * 32) * 2); };
    b += planes(cl * D1, cl) + planes(d * d * cr_, cl);           // stage 1: L, theta
    b += planes(cl * d * d, D3 * cr_) + planes(cr_, D3 * cr_);    // stage 3: T2, R
  }
  return b;
}

// ------------------------------------------------------------ environments

// L'[b,w',b'] = sum L[a,w,a'] conj(B[a,s,b]) W[w,s,t,w'] K[a',t,b'] with bra B and ket K (their
// bonds may differ); buildleft!'s contraction (projmps.jl:50-66) between two states
void env_left_pair(tnh_ctx *ctx, const tnh_buf *Lp, const tnh_buf *B, const tnh_buf *K,
                   const tnh_buf *W, tnh_buf **Lout, int prec) {
  if (!Lp) Lp = edge_block(ctx);
  require_dims(Lp, 3, "L");
  require_dims(B, 3, "A");
  require_dims(K, 3, "A");
  require_dims(W, 4, "W");
  const int64_t ca = B->dims[0], d = B->dims[1], cb = B->dims[2];
  const int64_t ka = K->dims[0], kb = K->dims[2];
  const int64_t Dl = W->dims[0], Dr = W->dims[3];
  TNH_REQUIRE(W->dims[1] == d && W->dims[2] == d && K->dims[1] == d, TNH_ERR_SHAPE,
              "W physical dims != A's");
  TNH_REQUIRE(Lp->dims[0] == ca && Lp->dims[2] == ka && Lp->dims[1] == Dl, TNH_ERR_SHAPE,
              "L " + shape_str(Lp) + " does not match A " + shape_str(B) + " / W " +
                  shape_str(W));
  check_prec_operand(prec, Lp, "L");
  check_prec_operand(prec, B, "A");
  check_prec_operand(prec, K, "A");
  TNH_REQUIRE(Dl * d <= kMaxIJ && d * Dr <= kMaxIJ, TNH_ERR_UNSUPPORTED, "D*d > 64");
  bool wcc = false;
  double *Wc = make_wc_env(ctx, W, true, 820, &wcc);
  const bool xc = is_complex(Lp->dtype) || is_complex(K->dtype);
  const bool yc = xc || wcc;
  const bool oc = yc || is_complex(B->dtype);
  const int tX = inter_dtype(prec, xc), tY = inter_dtype(prec, yc);
  const int64_t nx = ca * Dl * d * kb, ny = ca * d * Dr * kb;
  const size_t rs = real_size(tX);
  char *X = static_cast<char *>(ctx->workspace(821, nx * rs * (xc ? 2 : 1)));
  char *Y = static_cast<char *>(ctx->workspace(822, ny * rs * (yc ? 2 : 1)));
  const int odt = oc ? TNH_C128 : TNH_F64;  // environment blocks are kept in FP64
  int64_t od[3] = {cb, Dr, kb};
  buf_ensure(ctx, Lout, odt, 3, od);

  // (1) X[(a,w),(t,b')] = L[(a,w),a'] K[a',(t,b')]
  Gemm g1;
  g1.M = int(ca * Dl);
  g1.N = int(d * kb);
  g1.K = int(ka);
  g1.lda = ca * Dl;
  g1.ldb = ka;
  g1.ldc = ca * Dl;
  cgemm(ctx, g1, cop(Lp), cop(K), X, xc ? X + nx * rs : nullptr, real_of(tX), prec);
  // (2) Y[a,(s,w'),b'] = X[a,(w,t),b'] Wl[(w,t),(s,w')]
  Small s;
  s.na = int(ca);
  s.nb = int(kb);
  s.I = int(Dl * d);
  s.J = int(d * Dr);
  s.in = X;
  s.in_im = xc ? nx : 0;
  s.tin = real_of(tX);
  s.out = Y;
  s.out_im = yc ? ny : 0;
  s.tout = real_of(tY);
  s.Wc = Wc;
  s.Wc_cplx = wcc;
  s.sbI = ca * Dl * d;
  s.sbJ = ca * d * Dr;
  s.offI.resize(s.I);
  s.offJ.resize(s.J);
  for (int i = 0; i < s.I; ++i) s.offI[i] = ca * i;
  for (int j = 0; j < s.J; ++j) s.offJ[j] = ca * j;
  small_contract(ctx, s);
  // (3) L'[b,(w',b')] = B[(a,s),b]^H Y[(a,s),(w',b')]
  Gemm g3;
  g3.M = int(cb);
  g3.N = int(Dr * kb);
  g3.K = int(ca * d);
  g3.transA = true;
  g3.lda = ca * d;
  g3.ldb = ca * d;
  g3.ldc = cb;
  COperand y;
  y.re = Y;
  y.im = yc ? Y + ny * rs : nullptr;
  y.dtype = real_of(tY);
  cgemm(ctx, g3, cop(B, true), y, (*Lout)->re(), is_complex((*Lout)->dtype) ? (*Lout)->im() : nullptr,
        real_of((*Lout)->dtype), prec);
}

void env_left(tnh_ctx *ctx, const tnh_buf *Lp, const tnh_buf *A, const tnh_buf *W,
              tnh_buf **Lout, int prec) {
  env_left_pair(ctx, Lp, A, A, W, Lout, prec);
}

void env_right(tnh_ctx *ctx, const tnh_buf *Rp, const tnh_buf *A, const tnh_buf *W,
               tnh_buf **Rout, int prec) {
  if (!Rp) Rp = edge_block(ctx);
  require_dims(Rp, 3, "R");
  require_dims(A, 3, "A");
  require_dims(W, 4, "W");
  const int64_t ca = A->dims[0], d = A->dims[1], cb = A->dims[2];
  const int64_t Dl = W->dims[0], Dr = W->dims[3];
  TNH_REQUIRE(W->dims[1] == d && W->dims[2] == d, TNH_ERR_SHAPE, "W physical dims != A's");
  TNH_REQUIRE(Rp->dims[0] == cb && Rp->dims[2] == cb && Rp->dims[1] == Dr, TNH_ERR_SHAPE,
              "R " + shape_str(Rp) + " does not match A " + shape_str(A) + " / W " +
                  shape_str(W));
  check_prec_operand(prec, Rp, "R");
  check_prec_operand(prec, A, "A");
  TNH_REQUIRE(Dl * d <= kMaxIJ && d * Dr <= kMaxIJ, TNH_ERR_UNSUPPORTED, "D*d > 64");
  bool wcc = false;
  double *Wc = make_wc_env(ctx, W, false, 830, &wcc);
  const bool xc = is_complex(Rp->dtype) || is_complex(A->dtype);
  const bool yc = xc || wcc;
  const bool oc = yc || is_complex(A->dtype);
  const int tX = inter_dtype(prec, xc), tY = inter_dtype(prec, yc);
  const int64_t nx = ca * d * cb * Dr, ny = ca * Dl * d * cb;
  const size_t rs = real_size(tX);
  char *X = static_cast<char *>(ctx->workspace(831, nx * rs * (xc ? 2 : 1)));
  char *Y = static_cast<char *>(ctx->workspace(832, ny * rs * (yc ? 2 : 1)));
  int odt = oc ? TNH_C128 : TNH_F64;
  int64_t od[3] = {ca, Dl, ca};
  buf_ensure(ctx, Rout, odt, 3, od);

  // (1) X[(a',t),(b,w')] = A[(a',t),b'] R[(b,w'),b']^T
  Gemm g1;
  g1.M = int(ca * d);
  g1.N = int(cb * Dr);
  g1.K = int(cb);
  g1.lda = ca * d;
  g1.ldb = cb * Dr;
  g1.transB = true;
  g1.ldc = ca * d;
  cgemm(ctx, g1, cop(A), cop(Rp), X, xc ? X + nx * rs : nullptr, real_of(tX), prec);
  // (2) Y[a',(w,s),b] = X[a',t,b,w'] Wr[(t,w'),(w,s)]
  Small s;
  s.na = int(ca);
  s.nb = int(cb);
  s.I = int(d * Dr);
  s.J = int(Dl * d);
  s.in = X;
  s.in_im = xc ? nx : 0;
  s.tin = real_of(tX);
  s.out = Y;
  s.out_im = yc ? ny : 0;
  s.tout = real_of(tY);
  s.Wc = Wc;
  s.Wc_cplx = wcc;
  s.sbI = ca * d;
  s.sbJ = ca * Dl * d;
  s.offI.resize(s.I);
  s.offJ.resize(s.J);
  for (int i = 0; i < s.I; ++i) {
    int t = i % int(d), wp = i / int(d);
    s.offI[i] = ca * t + ca * d * cb * wp;
  }
  for (int j = 0; j < s.J; ++j) {
    int w = j % int(Dl), sp = j / int(Dl);
    s.offJ[j] = ca * w + ca * Dl * sp;
  }
  small_contract(ctx, s);
  // (3) per w: R'[a,w,a'] = conj(A)[a,(s,b)] Y_w[a',(s,b)]^T
  Gemm g3;
  g3.M = int(ca);
  g3.N = int(ca);
  g3.K = int(d * cb);
  g3.lda = ca;
  g3.ldb = ca * Dl;
  g3.sB = ca;
  g3.transB = true;
  g3.ldc = ca * Dl;
  g3.sC = ca;
  g3.batch = int(Dl);
  COperand y;
  y.re = Y;
  y.im = yc ? Y + ny * rs : nullptr;
  y.dtype = real_of(tY);
  cgemm(ctx, g3, cop(A, true), y, (*Rout)->re(), is_complex((*Rout)->dtype) ? (*Rout)->im() : nullptr,
        real_of((*Rout)->dtype), prec);
}

void two_site_theta(tnh_ctx *ctx, const tnh_buf *A1, const tnh_buf *A2, tnh_buf **theta) {
  require_dims(A1, 3, "A1");
  require_dims(A2, 3, "A2");
  TNH_REQUIRE(A1->dims[2] == A2->dims[0], TNH_ERR_SHAPE, "A1 right bond != A2 left bond");
  const int64_t cl = A1->dims[0], d1 = A1->dims[1], k = A1->dims[2], d2 = A2->dims[1],
                cr = A2->dims[2];
  const bool c = is_complex(A1->dtype) || is_complex(A2->dtype);
  const int prec = (is_double(A1->dtype) && is_double(A2->dtype)) ? TNH_PREC_F64 : TNH_PREC_F32;
  int dt = prec == TNH_PREC_F64 ? (c ? TNH_C128 : TNH_F64) : (c ? TNH_C64 : TNH_F32);
  int64_t td[4] = {cl, d1, d2, cr};
  buf_ensure(ctx, theta, dt, 4, td);
  Gemm g;
  g.M = int(cl * d1);
  g.N = int(d2 * cr);
  g.K = int(k);
  g.lda = cl * d1;
  g.ldb = k;
  g.ldc = cl * d1;
  cgemm(ctx, g, cop(A1), cop(A2), (*theta)->re(), c ? (*theta)->im() : nullptr, real_of(dt), prec);
}

// One-site H_eff matvec, product(projV, A, direction, nsites=1), projmps.jl:103-144:
// out[a', s, b'] = coeff sum L[a', w, a] W[w, s, t, w'] x[a, t, b] R[b', w', b], as the first
// two stages of the environment update (X = L x, Y = X Wl) followed by out = Y R^T.
// x and out hold chi_l d chi_r values (any shape: the DMRG passes a (chi_l, d, 1, chi_r) view).
void heff1_apply(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W, const tnh_buf *R,
                 const tnh_buf *x, tnh_buf *out, double coeff, int prec) {
  if (!L) L = edge_block(ctx);
  if (!R) R = edge_block(ctx);
  require_dims(L, 3, "L");
  require_dims(W, 4, "W");
  require_dims(R, 3, "R");
  TNH_REQUIRE(x && out, TNH_ERR_ARG, "heff1: NULL vector");
  const int64_t ca = L->dims[0], cb = R->dims[0], d = W->dims[1];
  const int64_t Dl = W->dims[0], Dr = W->dims[3];
  TNH_REQUIRE(W->dims[2] == d && L->dims[2] == ca && L->dims[1] == Dl && R->dims[2] == cb &&
                  R->dims[1] == Dr,
              TNH_ERR_SHAPE, "heff1: L " + shape_str(L) + ", W " + shape_str(W) + ", R " +
                                 shape_str(R) + " do not chain");
  TNH_REQUIRE(x->numel == ca * d * cb && out->numel == x->numel, TNH_ERR_SHAPE,
              "heff1: site tensor " + shape_str(x) + " does not match the blocks");
  TNH_REQUIRE(is_double(out->dtype), TNH_ERR_ARG, "heff1: FP64 output");
  check_prec_operand(prec, L, "L");
  check_prec_operand(prec, R, "R");
  TNH_REQUIRE(Dl * d <= kMaxIJ && d * Dr <= kMaxIJ, TNH_ERR_UNSUPPORTED, "D*d > 64");
  bool wcc = false;
  double *Wc = make_wc_env(ctx, W, true, 853, &wcc);
  const bool xc = is_complex(L->dtype) || is_complex(x->dtype);
  const bool yc = xc || wcc;
  TNH_REQUIRE(is_complex(out->dtype) || !(yc || is_complex(R->dtype)), TNH_ERR_ARG,
              "heff1: complex operands need a complex output");
  const int tX = inter_dtype(prec, xc), tY = inter_dtype(prec, yc);
  const int64_t nx = ca * Dl * d * cb, ny = ca * d * Dr * cb;
  const size_t rs = real_size(tX);
  char *X = static_cast<char *>(ctx->workspace(854, nx * rs * (xc ? 2 : 1)));
  char *Y = static_cast<char *>(ctx->workspace(855, ny * rs * (yc ? 2 : 1)));
  // (1) X[(a',w),(t,b)] = L[(a',w),a] x[a,(t,b)]
  Gemm g1;
  g1.M = int(ca * Dl);
  g1.N = int(d * cb);
  g1.K = int(ca);
  g1.lda = ca * Dl;
  g1.ldb = ca;
  g1.ldc = ca * Dl;
  cgemm(ctx, g1, cop(L), cop(x), X, xc ? X + nx * rs : nullptr, real_of(tX), prec);
  // (2) Y[a',(s,w'),b] = X[a',(w,t),b] Wl[(w,t),(s,w')]
  Small sm;
  sm.na = int(ca);
  sm.nb = int(cb);
  sm.I = int(Dl * d);
  sm.J = int(d * Dr);
  sm.in = X;
  sm.in_im = xc ? nx : 0;
  sm.tin = real_of(tX);
  sm.out = Y;
  sm.out_im = yc ? ny : 0;
  sm.tout = real_of(tY);
  sm.Wc = Wc;
  sm.Wc_cplx = wcc;
  sm.sbI = ca * Dl * d;
  sm.sbJ = ca * d * Dr;
  sm.offI.resize(sm.I);
  sm.offJ.resize(sm.J);
  for (int i = 0; i < sm.I; ++i) sm.offI[i] = ca * i;
  for (int j = 0; j < sm.J; ++j) sm.offJ[j] = ca * j;
  small_contract(ctx, sm);
  // (3) out[(a',s),b'] = coeff Y[(a',s),(w',b)] R[b',(w',b)]^T
  Gemm g3;
  g3.M = int(ca * d);
  g3.N = int(cb);
  g3.K = int(Dr * cb);
  g3.transB = true;
  g3.lda = ca * d;
  g3.ldb = cb;
  g3.ldc = ca * d;
  g3.alpha = coeff;
  COperand y;
  y.re = Y;
  y.im = yc ? Y + ny * rs : nullptr;
  y.dtype = real_of(tY);
  cgemm(ctx, g3, y, cop(R), out->re(), is_complex(out->dtype) ? out->im() : nullptr,
        real_of(out->dtype), prec);
}

double env_calculate(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *A, const tnh_buf *W,
                     const tnh_buf *R, double *im_out) {
  tnh_buf *Lp = nullptr;
  env_left(ctx, L, A, W, &Lp, TNH_PREC_F64);
  const tnh_buf *Rr = R ? R : edge_block(ctx);
  double re = 0, im = 0;
  try {
    TNH_REQUIRE(Lp->numel == Rr->numel, TNH_ERR_SHAPE,
                "calculate: right block " + shape_str(Rr) + " != " + shape_str(Lp));
    cdot_host(ctx, Lp, Rr, &re, &im);
  } catch (...) {
    buf_delete(ctx, Lp);
    throw;
  }
  buf_delete(ctx, Lp);
  if (im_out) *im_out = im;
  return re;
}

}  // namespace tnh

using namespace tnh;

extern "C" {

int tnh_heff_apply(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2,
                   const tnh_buf *R, const tnh_buf *x, tnh_buf *out, double coeff_re,
                   double coeff_im, int precision) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    heff_apply(ctx, L, W1, W2, R, x, out, coeff_re, coeff_im, precision);
  });
}

int tnh_env_left(tnh_ctx *ctx, const tnh_buf *Lprev, const tnh_buf *A, const tnh_buf *W,
                 tnh_buf **Lout, int precision) {
  if (!ctx || !Lout) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] { env_left(ctx, Lprev, A, W, Lout, precision); });
}

int tnh_env_right(tnh_ctx *ctx, const tnh_buf *Rprev, const tnh_buf *A, const tnh_buf *W,
                  tnh_buf **Rout, int precision) {
  if (!ctx || !Rout) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] { env_right(ctx, Rprev, A, W, Rout, precision); });
}

int tnh_env_calculate(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *A, const tnh_buf *W,
                      const tnh_buf *R, double coeff, double *re, double *im) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    double i = 0;
    double r = env_calculate(ctx, L, A, W, R, &i);
    if (re) *re = coeff * r;
    if (im) *im = coeff * i;
  });
}

int tnh_two_site_theta(tnh_ctx *ctx, const tnh_buf *A1, const tnh_buf *A2, tnh_buf **theta) {
  if (!ctx || !theta) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] { two_site_theta(ctx, A1, A2, theta); });
}

int tnh_buf_norm(tnh_ctx *ctx, const tnh_buf *b, double *norm) {
  if (!ctx || !b || !norm) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    const tnh_buf *x = b;
    tnh_buf *tmp = nullptr;
    if (!is_double(b->dtype)) {
      tmp = buf_new(ctx, is_complex(b->dtype) ? TNH_C128 : TNH_F64, b->ndim, b->dims);
      convert(ctx, b->ptr, b->dtype, tmp->ptr, tmp->dtype, b->numel);
      x = tmp;
    }
    tnh_buf v = *x;
    v.dtype = TNH_F64;
    v.numel = x->numel * (is_complex(x->dtype) ? 2 : 1);
    double re = 0, im = 0;
    try {
      cdot_host(ctx, &v, &v, &re, &im);
    } catch (...) {
      buf_delete(ctx, tmp);
      throw;
    }
    buf_delete(ctx, tmp);
    *norm = std::sqrt(std::max(re, 0.0));
  });
}

int tnh_buf_scale(tnh_ctx *ctx, tnh_buf *b, double re, double im) {
  if (!ctx || !b) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    if (im != 0.0 && !is_complex(b->dtype)) {
      // promote to complex in place (new storage), then scale
      tnh_buf *c = buf_new(ctx, complex_of(b->dtype), b->ndim, b->dims);
      convert(ctx, b->ptr, b->dtype, c->ptr, c->dtype, b->numel);
      if (b->owned) ctx->release(b->ptr, b->cap);
      b->dtype = c->dtype;
      b->ptr = c->ptr;
      b->bytes = c->bytes;
      b->cap = c->cap;
      b->owned = true;
      c->owned = false;
      buf_delete(ctx, c);
    }
    cscale(ctx, b->ptr, b->dtype, b->numel, re, im);
  });
}

int tnh_apply_site_op(tnh_ctx *ctx, tnh_buf **A, const tnh_buf *G) {
  if (!ctx || !A || !*A || !G) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    tnh_buf *a = *A;
    TNH_REQUIRE(a->ndim == 3, TNH_ERR_SHAPE, "site tensor must be (chi_l, d, chi_r)");
    const int64_t cl = a->dims[0], d = a->dims[1], cr = a->dims[2];
    TNH_REQUIRE(G->ndim == 2 && G->dims[0] == d && G->dims[1] == d, TNH_ERR_SHAPE,
                "operator must be (d, d)");
    TNH_REQUIRE(d <= 64, TNH_ERR_UNSUPPORTED, "d > 64");
    const bool gc = is_complex(G->dtype), ac = is_complex(a->dtype), oc = gc || ac;
    // Wc[i, o] = G[o, i]
    std::vector<double> hg(size_t(d) * d * (gc ? 2 : 1));
    {
      tnh_buf *g64 = nullptr;
      const tnh_buf *gsrc = G;
      if (!is_double(G->dtype)) {
        g64 = buf_new(ctx, gc ? TNH_C128 : TNH_F64, 2, G->dims);
        convert(ctx, G->ptr, G->dtype, g64->ptr, g64->dtype, G->numel);
        gsrc = g64;
      }
      to_host(ctx, hg.data(), gsrc->ptr, hg.size() * sizeof(double));
      buf_delete(ctx, g64);
    }
    std::vector<double> wc(size_t(d) * d * (gc ? 2 : 1));
    for (int64_t o = 0; o < d; ++o)
      for (int64_t i = 0; i < d; ++i) {
        wc[i + d * o] = hg[o + d * i];
        if (gc) wc[d * d + i + d * o] = hg[d * d + o + d * i];
      }
    double *Wc = static_cast<double *>(ctx->workspace(850, wc.size() * sizeof(double)));
    to_device(ctx, Wc, wc.data(), wc.size() * sizeof(double));
    tnh_buf *out = nullptr;
    const int odt = is_double(a->dtype) ? (oc ? TNH_C128 : TNH_F64) : (oc ? TNH_C64 : TNH_F32);
    buf_ensure(ctx, &out, odt, 3, a->dims);
    Small s;
    s.na = int(cl);
    s.nb = int(cr);
    s.I = int(d);
    s.J = int(d);
    s.in = a->ptr;
    s.in_im = ac ? a->numel : 0;
    s.tin = real_of(a->dtype);
    s.out = out->ptr;
    s.out_im = oc ? out->numel : 0;
    s.tout = real_of(odt);
    s.Wc = Wc;
    s.Wc_cplx = gc;
    s.sbI = cl * d;
    s.sbJ = cl * d;
    s.offI.resize(d);
    s.offJ.resize(d);
    for (int i = 0; i < d; ++i) s.offI[i] = s.offJ[i] = cl * i;
    try {
      small_contract(ctx, s);
    } catch (...) {
      buf_delete(ctx, out);
      throw;
    }
    buf_delete(ctx, a);
    *A = out;
  });
}

int tnh_gemm(tnh_ctx *ctx, int opA, int opB, int64_t M, int64_t N, int64_t K, double alpha,
             const tnh_buf *A, int64_t lda, const tnh_buf *B, int64_t ldb, double beta,
             tnh_buf *Cb, int64_t ldc, int precision) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(A && B && Cb, TNH_ERR_ARG, "gemm: NULL buffer");
    TNH_REQUIRE(opA >= 0 && opA <= 2 && opB >= 0 && opB <= 2, TNH_ERR_ARG, "gemm: op must be 0, 1 or 2");
    TNH_REQUIRE(M >= 0 && N >= 0 && K >= 0 && M < (1LL << 31) && N < (1LL << 31) && K < (1LL << 31),
                TNH_ERR_ARG, "gemm: bad sizes");
    // every element the product touches must lie inside its buffer
    auto need = [](bool tr, int64_t rows, int64_t cols, int64_t ld) -> int64_t {
      int64_t r = tr ? cols : rows, c = tr ? rows : cols;  // stored shape
      return (r == 0 || c == 0) ? 0 : (c - 1) * ld + r;
    };
    TNH_REQUIRE(lda >= std::max<int64_t>(1, opA ? K : M) && ldb >= std::max<int64_t>(1, opB ? N : K) &&
                    ldc >= std::max<int64_t>(1, M),
                TNH_ERR_ARG, "gemm: leading dimension too small");
    TNH_REQUIRE(need(opA != 0, M, K, lda) <= A->numel && need(opB != 0, K, N, ldb) <= B->numel &&
                    need(false, M, N, ldc) <= Cb->numel,
                TNH_ERR_SHAPE, "gemm: matrix view exceeds its buffer");
    const bool cc = is_complex(A->dtype) || is_complex(B->dtype);
    TNH_REQUIRE(!cc || is_complex(Cb->dtype), TNH_ERR_ARG, "gemm: complex product needs complex C");
    if (precision == TNH_PREC_F64)
      TNH_REQUIRE(is_double(A->dtype) && is_double(B->dtype) && is_double(Cb->dtype), TNH_ERR_ARG,
                  "gemm: FP64 arithmetic needs FP64 buffers");
    Gemm g;
    g.M = int(M);
    g.N = int(N);
    g.K = int(K);
    g.transA = opA != 0;
    g.transB = opB != 0;
    TNH_REQUIRE(!(g.transA && g.transB), TNH_ERR_UNSUPPORTED, "gemm: op(A) and op(B) both transposed");
    g.lda = lda;
    g.ldb = ldb;
    g.ldc = ldc;
    g.alpha = alpha;
    g.beta = beta;
    if (!is_complex(Cb->dtype)) {
      g.A = A->ptr;
      g.B = B->ptr;
      g.C = Cb->ptr;
      g.tA = A->dtype;
      g.tB = B->dtype;
      g.tC = Cb->dtype;
      gemm(ctx, g, precision);
      return;
    }
    if (!cc) {
      // real product into a complex C: the imaginary part is scaled by beta
      TNH_REQUIRE(ldc == M, TNH_ERR_ARG, "gemm: real operands into strided complex C");
    }
    cgemm(ctx, g, cop(A, opA == 2), cop(B, opB == 2), Cb->re(), Cb->im(), real_of(Cb->dtype),
          precision);
  });
}

}  // extern "C"
