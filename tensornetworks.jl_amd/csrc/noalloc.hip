// The H_eff matvec for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): every temporary of the GEMM chain (mixing matrix, T1, T2,
// the FP32 copy of theta, split-BF16 operand planes, edge blocks) is carved from one
// file-scope device array instead of the context's allocator. Each GPU the library runs on
// holds its own copy of the array until the process ends.
#include <mutex>

#include "tnh_internal.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kScratchBytes];

// Device address of `scratch` per device. The name `scratch` is a host-side stand-in in
// host code: only these addresses are passed to kernels.
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

}  // namespace

namespace tnh {

void noalloc_init(int device) {
  TNH_REQUIRE(device >= 0 && device < kMaxDevices, TNH_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device]) return;
  void *ptr = nullptr;
  // the caller has made `device` current: this resolves that device's copy
  TNH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  TNH_REQUIRE(ptr != nullptr, TNH_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch");
  g_scratch[device] = ptr;
}

}  // namespace tnh

using namespace tnh;

extern "C" int tnh_heff_apply_noalloc(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2,
                                      const tnh_buf *R, const tnh_buf *x, tnh_buf *out, double coeff_re,
                                      double coeff_im, int precision) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(W1 && W2 && x && out, TNH_ERR_ARG, "heff_apply_noalloc: NULL tensor");
    TNH_REQUIRE(precision >= TNH_PREC_F64 && precision <= TNH_PREC_BF16X3, TNH_ERR_ARG,
                "unknown precision " + std::to_string(precision));
    void *base = nullptr;
    {
      std::lock_guard<std::mutex> lock(g_mu);
      base = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
    }
    TNH_REQUIRE(base != nullptr, TNH_ERR_HIP, "heff_apply_noalloc: the static scratch of this device is not resolved");
    const size_t need = heff_scratch_bound(L, W1, W2, R, x, precision);
    TNH_REQUIRE(need <= kScratchBytes, TNH_ERR_NOMEM,
                "heff_apply_noalloc: these shapes need up to " + std::to_string(need) +
                    " bytes of scratch, above the " + std::to_string(kScratchBytes) + "-byte static scratch");
    tnh_ctx::Arena arena;
    arena.base = static_cast<char *>(base);
    arena.cap = kScratchBytes;
    // no profiling events (their creation is a runtime allocation) and no allocator inside
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    ctx->arena = &arena;
    try {
      heff_apply(ctx, L, W1, W2, R, x, out, coeff_re, coeff_im, precision);
    } catch (...) {
      ctx->arena = nullptr;
      ctx->profiling = prof;
      throw;
    }
    ctx->arena = nullptr;
    ctx->profiling = prof;
  });
}
