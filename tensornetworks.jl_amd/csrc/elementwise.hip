// Elementwise and reduction kernels: dtype/layout conversion, scaling, FP64 dots.
// All of these are HBM-bound streams; they use grid-stride loops over 256-thread
// blocks sized to a few waves per CU.
#include "tnh_internal.hpp"

namespace tnh {

namespace {

constexpr int kThreads = 256;

inline int grid_for(tnh_ctx *ctx, int64_t n, int per_thread = 1) {
  int64_t blocks = (n + int64_t(kThreads) * per_thread - 1) / (int64_t(kThreads) * per_thread);
  int64_t cap = int64_t(ctx->cu_count) * 8;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  return int(blocks);
}

template <typename S, typename D>
__global__ void reformat_kernel(const S *__restrict__ src, bool s_cplx, bool s_inter,
                                D *__restrict__ dst, bool d_cplx, bool d_inter, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    S re, im = S(0);
    if (s_cplx) {
      if (s_inter) {
        re = src[2 * i];
        im = src[2 * i + 1];
      } else {
        re = src[i];
        im = src[n + i];
      }
    } else {
      re = src[i];
    }
    if (d_cplx) {
      if (d_inter) {
        dst[2 * i] = D(re);
        dst[2 * i + 1] = D(im);
      } else {
        dst[i] = D(re);
        dst[n + i] = D(im);
      }
    } else {
      dst[i] = D(re);
    }
  }
}

template <typename T>
__global__ void scale_kernel(T *__restrict__ x, int64_t n, T a) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[i] *= a;
}

template <typename T>
__global__ void fill_kernel(T *__restrict__ x, int64_t n, T v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[i] = v;
}

template <typename T>
__global__ void copy_kernel(T *__restrict__ dst, const T *__restrict__ src, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

__global__ void axpby_kernel(int64_t n, double a, const double *__restrict__ x, double b,
                             double *__restrict__ y) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

__device__ inline double block_sum(double v) {
  __shared__ double red[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kThreads / 64; ++i) s += red[i];
  __syncthreads();
  return s;
}

__global__ void dot_partial_kernel(int64_t n, const double *__restrict__ x,
                                   const double *__restrict__ y, double *__restrict__ part) {
  double acc = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    acc += x[i] * y[i];
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void sum_final_kernel(int np, const double *__restrict__ part, double *__restrict__ dst) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
  double s = block_sum(acc);
  if (threadIdx.x == 0) *dst = s;
}

template <typename S>
void reformat_from(tnh_ctx *ctx, const S *src, bool sc, bool si, void *dst, int tdst, bool di,
                   int64_t n) {
  int g = grid_for(ctx, n);
  bool dc = is_complex(tdst);
  if (is_double(tdst))
    hipLaunchKernelGGL((reformat_kernel<S, double>), dim3(g), dim3(kThreads), 0, ctx->stream,
                       src, sc, si, static_cast<double *>(dst), dc, di, n);
  else
    hipLaunchKernelGGL((reformat_kernel<S, float>), dim3(g), dim3(kThreads), 0, ctx->stream,
                       src, sc, si, static_cast<float *>(dst), dc, di, n);
}

}  // namespace

void reformat(tnh_ctx *ctx, const void *src, int tsrc, bool src_interleaved, void *dst, int tdst,
              bool dst_interleaved, int64_t n) {
  if (n <= 0) return;
  bool sc = is_complex(tsrc);
  if (is_double(tsrc))
    reformat_from(ctx, static_cast<const double *>(src), sc, src_interleaved, dst, tdst,
                  dst_interleaved, n);
  else
    reformat_from(ctx, static_cast<const float *>(src), sc, src_interleaved, dst, tdst,
                  dst_interleaved, n);
  TNH_HIP(hipGetLastError());
}

void convert(tnh_ctx *ctx, const void *src, int tsrc, void *dst, int tdst, int64_t n) {
  reformat(ctx, src, tsrc, false, dst, tdst, false, n);
}

void scale(tnh_ctx *ctx, void *x, int t, int64_t n, double a) {
  int64_t m = n * (is_complex(t) ? 2 : 1);
  if (m <= 0) return;
  int g = grid_for(ctx, m);
  if (is_double(t))
    hipLaunchKernelGGL(scale_kernel<double>, dim3(g), dim3(kThreads), 0, ctx->stream,
                       static_cast<double *>(x), m, a);
  else
    hipLaunchKernelGGL(scale_kernel<float>, dim3(g), dim3(kThreads), 0, ctx->stream,
                       static_cast<float *>(x), m, float(a));
  TNH_HIP(hipGetLastError());
}

void fill(tnh_ctx *ctx, void *x, int t, int64_t n, double v) {
  int64_t m = n * (is_complex(t) ? 2 : 1);
  if (m <= 0) return;
  int g = grid_for(ctx, m);
  if (is_double(t))
    hipLaunchKernelGGL(fill_kernel<double>, dim3(g), dim3(kThreads), 0, ctx->stream,
                       static_cast<double *>(x), m, v);
  else
    hipLaunchKernelGGL(fill_kernel<float>, dim3(g), dim3(kThreads), 0, ctx->stream,
                       static_cast<float *>(x), m, float(v));
  TNH_HIP(hipGetLastError());
}

void copy_device(tnh_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (bytes == 0 || dst == src) return;
  const uintptr_t a = reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src);
  if ((a & 15u) == 0 && bytes % 16 == 0) {  // 16-byte elements
    const int64_t n = int64_t(bytes / 16);
    hipLaunchKernelGGL(copy_kernel<double2>, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream,
                       static_cast<double2 *>(dst), static_cast<const double2 *>(src), n);
  } else if ((a & 7u) == 0 && bytes % 8 == 0) {
    const int64_t n = int64_t(bytes / 8);
    hipLaunchKernelGGL(copy_kernel<double>, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream,
                       static_cast<double *>(dst), static_cast<const double *>(src), n);
  } else {
    const int64_t n = int64_t(bytes);
    hipLaunchKernelGGL(copy_kernel<unsigned char>, dim3(grid_for(ctx, n)), dim3(kThreads), 0,
                       ctx->stream, static_cast<unsigned char *>(dst),
                       static_cast<const unsigned char *>(src), n);
  }
  TNH_HIP(hipGetLastError());
}

void axpby(tnh_ctx *ctx, int64_t n, double a, const double *x, double b, double *y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(ctx, n)), dim3(kThreads), 0, ctx->stream, n, a,
                     x, b, y);
  TNH_HIP(hipGetLastError());
}

void dot_f64(tnh_ctx *ctx, int64_t n, const double *x, const double *y, double *dst) {
  int g = grid_for(ctx, n, 4);
  if (g > 1024) g = 1024;
  double *part = static_cast<double *>(ctx->workspace(902, sizeof(double) * 1024));
  hipLaunchKernelGGL(dot_partial_kernel, dim3(g), dim3(kThreads), 0, ctx->stream, n, x, y, part);
  hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(kThreads), 0, ctx->stream, g, part, dst);
  TNH_HIP(hipGetLastError());
}

}  // namespace tnh
