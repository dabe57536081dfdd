// Split-BF16x3 GEMM: each operand is split once into hi + lo bf16 planes laid out
// k-contiguous ([row][k], k padded to a multiple of 32 with zeros), then the MODE_BF16P
// kernel accumulates hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 in FP32.
#include "gemm_kernel.hpp"
#include "tnh_internal.hpp"

namespace tnh {
namespace gk {

namespace {

// out[plane][b][r][k] (r < R, k < Kp; zero for k >= K) from X, batch stride sX:
//   RFAST: element (r, k) at X[r + ld k]  (transposed through a 32 x 32 LDS tile)
//   else : element (r, k) at X[k + ld r]
template <typename T, bool RFAST>
__global__ __launch_bounds__(256) void split_bf16_kernel(const T *__restrict__ X, int64_t ld, int64_t sX,
                                                         int R, int K, int Kp, int64_t plane,
                                                         unsigned short *__restrict__ out) {
  __shared__ float tile[32][33];
  const int k0 = blockIdx.x * 32, r0 = blockIdx.y * 32, b = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const T *Xb = X + int64_t(b) * sX;
  unsigned short *ob = out + int64_t(b) * 2 * plane;
  if (RFAST) {
    for (int kk = ty; kk < 32; kk += 8) {
      const int r = r0 + tx, k = k0 + kk;
      tile[kk][tx] = (r < R && k < K) ? float(Xb[int64_t(r) + ld * k]) : 0.0f;
    }
    __syncthreads();
    for (int rr = ty; rr < 32; rr += 8) {
      const int r = r0 + rr, k = k0 + tx;
      if (r >= R) continue;
      const float x = tile[tx][rr];
      const unsigned short h = f2bf(x);
      const int64_t o = int64_t(k) + int64_t(Kp) * r;
      ob[o] = h;
      ob[plane + o] = f2bf(x - bf2f(h));
    }
  } else {
    for (int rr = ty; rr < 32; rr += 8) {
      const int r = r0 + rr, k = k0 + tx;
      if (r >= R) continue;
      const float x = (k < K) ? float(Xb[int64_t(k) + ld * r]) : 0.0f;
      const unsigned short h = f2bf(x);
      const int64_t o = int64_t(k) + int64_t(Kp) * r;
      ob[o] = h;
      ob[plane + o] = f2bf(x - bf2f(h));
    }
  }
}

template <typename TC, int BM, int BN>
void launch_p(hipStream_t s, const Params &p, int batch) {
  dim3 grid(p.tiles_m * p.tiles_n, batch);
  hipLaunchKernelGGL((gemm_kernel<MODE_BF16P, unsigned short, unsigned short, TC, true, false, BM, BN, true>),
                     grid, dim3(256), 0, s, p);
}

}  // namespace

// Split R x K operand (batch entries at stride sX) into hi/lo planes at out
// (out holds batch * 2 * R * Kp bf16; plane = R * Kp).
void split_bf16(hipStream_t s, const void *X, int dtype, bool rfast, int64_t ld, int64_t sX, int R, int K,
                int Kp, int batch, unsigned short *out) {
  dim3 grid(Kp / 32, (R + 31) / 32, batch);
  const int64_t plane = int64_t(R) * Kp;
  if (dtype == TNH_F64) {
    if (rfast)
      hipLaunchKernelGGL((split_bf16_kernel<double, true>), grid, dim3(256), 0, s,
                         static_cast<const double *>(X), ld, sX, R, K, Kp, plane, out);
    else
      hipLaunchKernelGGL((split_bf16_kernel<double, false>), grid, dim3(256), 0, s,
                         static_cast<const double *>(X), ld, sX, R, K, Kp, plane, out);
  } else {
    if (rfast)
      hipLaunchKernelGGL((split_bf16_kernel<float, true>), grid, dim3(256), 0, s,
                         static_cast<const float *>(X), ld, sX, R, K, Kp, plane, out);
    else
      hipLaunchKernelGGL((split_bf16_kernel<float, false>), grid, dim3(256), 0, s,
                         static_cast<const float *>(X), ld, sX, R, K, Kp, plane, out);
  }
}

// p.A / p.B point at split planes ([row][k], row stride p.lda = p.ldb = Kp, p.K = Kp).
// tile: 0 = 64 x 64, 1 = 128 x 128 (kTileShapes in gemm.hip)
void launch_bf16p(hipStream_t s, const Params &p, int batch, int tC, int tile) {
  const bool c64 = tC == TNH_F64;
  switch (tile) {
    case 0:
      if (c64) launch_p<double, 64, 64>(s, p, batch);
      else launch_p<float, 64, 64>(s, p, batch);
      return;
    case 1:
      if (c64) launch_p<double, 128, 128>(s, p, batch);
      else launch_p<float, 128, 128>(s, p, batch);
      return;
    default: throw Error(TNH_ERR_ARG, "gemm: unknown split-BF16 tile " + std::to_string(tile));
  }
}

}  // namespace gk
}  // namespace tnh
