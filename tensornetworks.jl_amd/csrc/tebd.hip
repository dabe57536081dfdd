// Two-site gate application of TEBD: applygate(psi, site, gate, direction; ...),
// src/structures/mps/gatelist.jl:137-171 (rng = 2, rank 1).
//   theta = A1 A2; theta'[a,o1,o2,b] = sum G[o1,i1,o2,i2] theta[a,i1,i2,b]; split, no
//   normalisation (:157).
// The gate stage is an HBM-bound mixing pass over theta (the same kernel as the MPO
// mixing stage of the matvec, with the d^2 x d^2 gate as its matrix).
#include <vector>

#include "tnh_internal.hpp"

namespace tnh {

namespace {

// Wc[(i1,i2),(o1,o2)] = G[o1,i1,o2,i2] (planar complex)
template <typename T>
__global__ void gate_wc_kernel(const T *G, int64_t g_im, int d, double *Wc, bool cplx) {
  const int I = d * d;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= I * I) return;
  const int i = e % I, j = e / I;
  const int i1 = i % d, i2 = i / d, o1 = j % d, o2 = j / d;
  const int64_t q = o1 + int64_t(d) * (i1 + int64_t(d) * (o2 + int64_t(d) * i2));
  Wc[e] = double(G[q]);
  if (cplx) Wc[int64_t(I) * I + e] = g_im ? double(G[g_im + q]) : 0.0;
}

}  // namespace

// theta_out = G applied to theta (both (chi_l, d, d, chi_r)); theta_out (re)allocated.
void apply_gate_theta(tnh_ctx *ctx, const tnh_buf *theta, const tnh_buf *G, tnh_buf **out) {
  TNH_REQUIRE(G && G->ndim == 4, TNH_ERR_SHAPE, "gate must be (d, d, d, d)");
  const int64_t cl = theta->dims[0], d = theta->dims[1], cr = theta->dims[3];
  TNH_REQUIRE(theta->dims[2] == d && G->dims[0] == d && G->dims[1] == d && G->dims[2] == d &&
                  G->dims[3] == d,
              TNH_ERR_SHAPE, "gate dims must equal the physical dimension");
  TNH_REQUIRE(d * d <= 64, TNH_ERR_UNSUPPORTED, "d^2 > 64");
  const bool gc = is_complex(G->dtype);
  const bool tc = is_complex(theta->dtype);
  const bool oc = gc || tc;
  const int I = int(d * d);
  double *Wc = static_cast<double *>(ctx->workspace(840, sizeof(double) * I * I * 2));
  const int64_t gim = gc ? G->numel : 0;
  if (is_double(G->dtype))
    hipLaunchKernelGGL(gate_wc_kernel<double>, dim3((I * I + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const double *>(G->ptr), gim, int(d), Wc, gc);
  else
    hipLaunchKernelGGL(gate_wc_kernel<float>, dim3((I * I + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<const float *>(G->ptr), gim, int(d), Wc, gc);
  TNH_HIP(hipGetLastError());
  const int odt = is_double(theta->dtype) ? (oc ? TNH_C128 : TNH_F64) : (oc ? TNH_C64 : TNH_F32);
  buf_ensure(ctx, out, odt, 4, theta->dims);
  Small s;
  s.na = int(cl);
  s.nb = int(cr);
  s.I = I;
  s.J = I;
  s.in = theta->ptr;
  s.in_im = tc ? theta->numel : 0;
  s.tin = real_of(theta->dtype);
  s.out = (*out)->ptr;
  s.out_im = oc ? (*out)->numel : 0;
  s.tout = real_of(odt);
  s.Wc = Wc;
  s.Wc_cplx = gc;
  s.sbI = cl * I;
  s.sbJ = cl * I;
  s.offI.resize(I);
  s.offJ.resize(I);
  for (int i = 0; i < I; ++i) s.offI[i] = s.offJ[i] = cl * i;
  small_contract(ctx, s);
}

void applygate2(tnh_ctx *ctx, tnh_buf **A1, tnh_buf **A2, const tnh_buf *G, int direction,
                double cutoff, int maxdim, int mindim, int *k, double *disc) {
  TNH_REQUIRE(A1 && A2 && *A1 && *A2, TNH_ERR_ARG, "NULL site tensors");
  tnh_buf *theta = nullptr, *theta2 = nullptr;
  try {
    two_site_theta(ctx, *A1, *A2, &theta);  // :142-145
    apply_gate_theta(ctx, theta, G, &theta2);  // :149-154
    replacesites2(ctx, theta2, direction, 0, cutoff, maxdim, mindim, A1, A2, k, disc, nullptr, 0,
                  false);  // :157
  } catch (...) {
    buf_delete(ctx, theta);
    buf_delete(ctx, theta2);
    throw;
  }
  buf_delete(ctx, theta);
  buf_delete(ctx, theta2);
}

}  // namespace tnh

using namespace tnh;

extern "C" int tnh_applygate2(tnh_ctx *ctx, tnh_buf **A1, tnh_buf **A2, const tnh_buf *G,
                              int direction, double cutoff, int maxdim, int mindim, int *k,
                              double *discarded) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(maxdim >= 0 && mindim >= 0, TNH_ERR_ARG, "maxdim and mindim must be >= 0");
    applygate2(ctx, A1, A2, G, direction, cutoff, maxdim, mindim, k, discarded);
  });
}
