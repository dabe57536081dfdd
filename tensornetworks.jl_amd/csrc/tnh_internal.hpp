// Internal declarations shared by the libtnhip sources: context, buffers, errors,
// the stream-ordered caching allocator and the launch helpers of every kernel family.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <exception>
#include <map>
#include <string>
#include <vector>

#include "tnhip.h"

namespace tnh {

struct Error : std::exception {
  int code;
  std::string msg;
  Error(int c, std::string m) : code(c), msg(std::move(m)) {}
  const char *what() const noexcept override { return msg.c_str(); }
};

#define TNH_HIP(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::tnh::Error(TNH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
  } while (0)

#define TNH_REQUIRE(cond, code, msg)                                                       \
  do {                                                                                     \
    if (!(cond)) throw ::tnh::Error((code), (msg));                                       \
  } while (0)

inline bool is_complex(int dt) { return dt == TNH_C128 || dt == TNH_C64; }
inline bool is_double(int dt) { return dt == TNH_F64 || dt == TNH_C128; }
// bytes of one real scalar of the dtype
inline size_t real_size(int dt) { return is_double(dt) ? 8 : 4; }
inline int real_of(int dt) { return is_double(dt) ? TNH_F64 : TNH_F32; }
inline int complex_of(int dt) { return is_double(dt) ? TNH_C128 : TNH_C64; }
inline bool valid_dtype(int dt) { return dt >= TNH_F64 && dt <= TNH_C64; }

}  // namespace tnh

struct tnh_ctx;

struct tnh_buf {
  int dtype = TNH_F64;
  int ndim = 0;
  int64_t dims[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  int64_t numel = 0;    // element count (a complex element counts once)
  void *ptr = nullptr;  // device memory; complex: planar [re | im]
  size_t bytes = 0;
  size_t cap = 0;       // allocated bytes (cache slot size)
  bool owned = true;

  int64_t dim(int i) const { return i < ndim ? dims[i] : 1; }
  // pointer to the real plane / imaginary plane (complex only)
  void *re() const { return ptr; }
  void *im() const {
    return static_cast<char *>(ptr) + numel * tnh::real_size(dtype);
  }
};

struct tnh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int cu_count = 256;

  // Stream-ordered caching allocator: freed blocks are kept per size class and
  // reused by later allocations on the same stream (all work of a context is on
  // `stream`, so reuse is ordered after every earlier use).
  std::multimap<size_t, void *> cache;
  size_t cached_bytes = 0;
  void *alloc(size_t bytes, size_t *cap_out);
  void release(void *p, size_t cap);
  void trim();

  // Named grow-only workspaces for temporaries of one call.
  std::map<int, std::pair<void *, size_t>> ws;
  void *workspace(int slot, size_t bytes);

  // The 1x1x1 ones block standing for environments beyond the chain ends.
  tnh_buf edge;

  // Pinned host staging for small device->host scalar transfers.
  double *host_scalars = nullptr;
  static constexpr int kHostScalars = 4096;

  ~tnh_ctx();
};

namespace tnh {

// ----------------------------------------------------------------- buffers
tnh_buf *buf_new(tnh_ctx *ctx, int dtype, int ndim, const int64_t *dims);
void buf_delete(tnh_ctx *ctx, tnh_buf *b);
// Reallocate *b to (dtype, dims) unless it already has that shape and dtype.
void buf_ensure(tnh_ctx *ctx, tnh_buf **b, int dtype, int ndim, const int64_t *dims);
std::string shape_str(const tnh_buf *b);
// Copy n real scalars of type dt from device to host (synchronises the stream).
void to_host(tnh_ctx *ctx, void *host, const void *dev, size_t bytes);

// ------------------------------------------------------------ GEMM (gemm.hip)
// Column-major C[M,N] = alpha * op(A)[M,K] op(B)[K,N] + beta * C, strided batch.
// A/B/C element types are F64 or F32 (real); `prec` picks the MFMA arithmetic.
struct Gemm {
  int M = 0, N = 0, K = 0;
  bool transA = false, transB = false;
  const void *A = nullptr; int64_t lda = 0, sA = 0; int tA = TNH_F64;
  const void *B = nullptr; int64_t ldb = 0, sB = 0; int tB = TNH_F64;
  void *C = nullptr; int64_t ldc = 0, sC = 0; int tC = TNH_F64;
  double alpha = 1.0, beta = 0.0;
  int batch = 1;
};
void gemm(tnh_ctx *ctx, const Gemm &g, int prec);

// Planar-complex operand description for complex GEMMs built from real GEMMs.
struct COperand {
  const void *re = nullptr;
  const void *im = nullptr;  // nullptr = purely real operand
  int dtype = TNH_F64;       // real dtype of each plane
  bool conj = false;
};
// C = alpha * op(A) op(B) + beta * C on planar complex data (C.im may be nullptr
// only if both A and B are real). alpha, beta real.
void cgemm(tnh_ctx *ctx, Gemm g, COperand A, COperand B, void *Cre, void *Cim, int tC,
           int prec);

// ---------------------------------------------------- contraction kernels
// out[a, j, b] = sum_i in[a, i, b] * Wc[i, j] (planar complex when the flags say so),
// element offsets in(a,i,b) = a + offI[i] + b*sbI, out(a,j,b) = a + offJ[j] + b*sbJ.
struct Small {
  int na = 0, nb = 0, I = 0, J = 0;
  const void *in = nullptr; int64_t in_im = 0;  // element offset of imaginary plane (0 = real)
  int tin = TNH_F64;
  void *out = nullptr; int64_t out_im = 0;
  int tout = TNH_F64;
  const double *Wc = nullptr;  // device, I*J real parts then I*J imaginary parts if Wc_cplx
  bool Wc_cplx = false;
  int64_t sbI = 0, sbJ = 0;
  std::vector<int64_t> offI, offJ;
  double beta = 0.0;  // out = in*Wc + beta*out
};
void small_contract(tnh_ctx *ctx, const Small &s);

// elementwise helpers (elementwise.hip)
void convert(tnh_ctx *ctx, const void *src, int tsrc, void *dst, int tdst, int64_t n);
void scale(tnh_ctx *ctx, void *x, int t, int64_t n, double a);
void fill(tnh_ctx *ctx, void *x, int t, int64_t n, double v);
void axpby(tnh_ctx *ctx, int64_t n, double a, const double *x, double b, double *y);

// FP64 reductions with deterministic two-stage sums; result stays on the device
// (dst) so that chains of Krylov steps need no host round trip.
void dot_f64(tnh_ctx *ctx, int64_t n, const double *x, const double *y, double *dst);

}  // namespace tnh

// Run `f` and convert exceptions to a status code + ctx->err.
template <class F>
static inline int tnh_guard(tnh_ctx *ctx, F &&f) {
  try {
    f();
    return TNH_OK;
  } catch (const tnh::Error &e) {
    if (ctx) ctx->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    if (ctx) ctx->err = "host allocation failed";
    return TNH_ERR_NOMEM;
  } catch (const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return TNH_ERR_ARG;
  } catch (...) {
    if (ctx) ctx->err = "unknown error";
    return TNH_ERR_ARG;
  }
}
