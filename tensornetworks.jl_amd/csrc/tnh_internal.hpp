// Internal declarations shared by the libtnhip sources: context, buffers, errors,
// the stream-ordered caching allocator and the launch helpers of every kernel family.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <complex>
#include <cstdint>
#include <cstring>
#include <exception>
#include <functional>
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
  // Helper stream (default priority) for work that overlaps the main stream inside one
  // call; every call that uses it joins it back into `stream` before returning, so the
  // context stays ordered on `stream` for its callers.
  hipStream_t aux = nullptr;
  hipStream_t aux_stream();
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

  // Caller-provided device scratch, set only for the duration of an entry point that must
  // not allocate: workspace() then carves its slots from it (a slot is reused while it is
  // large enough) and alloc() refuses, so no hipMalloc runs inside such a call.
  struct Arena {
    char *base = nullptr;
    size_t cap = 0, off = 0;
    std::map<int, std::pair<void *, size_t>> slots;
    tnh_buf edge;  // the ones(1,1,1) block, carved from the arena when needed
  };
  Arena *arena = nullptr;

  // The 1x1x1 ones block standing for environments beyond the chain ends.
  tnh_buf edge;

  // Optional kernel timing: HIP event pairs per kernel family, resolved on read.
  struct ProfRec {
    hipEvent_t a, b;
    double flops, bytes;
  };
  bool profiling = false;
  // Lowest stopping floor of the block-Jacobi Gram eigensolve, relative to max_i G_ii (0: the
  // FP64 rounding floor alone). An FP32 DMRG pass raises it for its own splits: its theta
  // carries far more error than the FP64 floor resolves.
  double gram_floor = 0.0;
  std::map<std::string, std::vector<ProfRec>> prof_pending;
  struct ProfSum {
    int64_t launches = 0;
    double ms = 0, flops = 0, bytes = 0;
  };
  std::map<std::string, ProfSum> prof_done;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t take_event();
  // bracket one launch: call prof_begin before and prof_end after the launch
  hipEvent_t prof_begin();
  void prof_end(const char *family, hipEvent_t start, double flops, double bytes);
  void prof_resolve();

  // Pinned host staging for small device->host scalar transfers.
  double *host_scalars = nullptr;
  static constexpr int kHostScalars = 4096;
  // Pinned staging of to_host (downloads) and to_device (uploads). Uploads take
  // consecutive pieces of an arena whose copies are queued on the stream; the arena is
  // reset whenever the stream has been synchronised (every to_host), and a full arena
  // first waits for the last queued upload.
  void *pin_down = nullptr;
  size_t pin_down_cap = 0;
  void *pin_up = nullptr;
  size_t pin_up_cap = 0, pin_up_off = 0;
  hipEvent_t up_done = nullptr;
  bool up_pending = false;
  // Device copies of the block-Jacobi round-robin schedules, keyed by column count.
  std::map<int, std::pair<void *, size_t>> schedules;

  // GEMM tile autotuner. A size class is (mode, operand types, transposes, vector paths,
  // M, N, K rounded to four classes per octave); its tile is chosen on the first product
  // of that class by timing every candidate tile on a scratch output, and reused after.
  bool gemm_autotune = true;
  struct TileChoice {
    int tile = 0;
    int rule = 0;
    float ms[4] = {0, 0, 0, 0};  // best of two timed runs per candidate
    int64_t M = 0, N = 0, K = 0;  // the product that was timed
  };
  std::map<std::array<int64_t, 12>, TileChoice> gemm_tiles;

  ~tnh_ctx();
};

namespace tnh {

// ----------------------------------------------------------------- buffers
tnh_buf *buf_new(tnh_ctx *ctx, int dtype, int ndim, const int64_t *dims);
void buf_delete(tnh_ctx *ctx, tnh_buf *b);
// Reallocate *b to (dtype, dims) unless it already has that shape and dtype.
void buf_ensure(tnh_ctx *ctx, tnh_buf **b, int dtype, int ndim, const int64_t *dims);
std::string shape_str(const tnh_buf *b);
// Copy bytes from device to host through pinned staging (synchronises the stream).
void to_host(tnh_ctx *ctx, void *host, const void *dev, size_t bytes);
// Queue a copy of bytes from host to device on the stream; the host data is copied into
// pinned staging before return, so it may be freed at once (no synchronisation needed).
void to_device(tnh_ctx *ctx, void *dev, const void *host, size_t bytes);

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
// Convert n elements between dtypes and complex layouts (planar or interleaved).
void reformat(tnh_ctx *ctx, const void *src, int tsrc, bool src_interleaved, void *dst,
              int tdst, bool dst_interleaved, int64_t n);
void axpby(tnh_ctx *ctx, int64_t n, double a, const double *x, double b, double *y);
// Device-to-device copy by a kernel on the context's stream.
void copy_device(tnh_ctx *ctx, void *dst, const void *src, size_t bytes);

// FP64 reductions with deterministic two-stage sums; result stays on the device
// (dst) so that chains of Krylov steps need no host round trip.
void dot_f64(tnh_ctx *ctx, int64_t n, const double *x, const double *y, double *dst);

// -------------------------------------------------------- hot-path operators
void heff_apply(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2,
                const tnh_buf *R, const tnh_buf *x, tnh_buf *out, double cr, double ci, int prec);
size_t heff_scratch_bound(const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2, const tnh_buf *R,
                          const tnh_buf *x, int prec);
// Resolve this device's copy of the static scratch of tnh_heff_apply_noalloc (the device
// must be current); called once per device when a context is created.
void noalloc_init(int device);
void env_left(tnh_ctx *ctx, const tnh_buf *Lp, const tnh_buf *A, const tnh_buf *W,
              tnh_buf **Lout, int prec);
// the same with a bra B and a ket K of possibly different bonds: L (b_B, w, b_K)
void env_left_pair(tnh_ctx *ctx, const tnh_buf *Lp, const tnh_buf *B, const tnh_buf *K,
                   const tnh_buf *W, tnh_buf **Lout, int prec);
void env_right(tnh_ctx *ctx, const tnh_buf *Rp, const tnh_buf *A, const tnh_buf *W,
               tnh_buf **Rout, int prec);
// one-site H_eff matvec (nsites = 1 DMRG): out = coeff H_eff x
void heff1_apply(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W, const tnh_buf *R,
                 const tnh_buf *x, tnh_buf *out, double coeff, int prec);
double env_calculate(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *A, const tnh_buf *W,
                     const tnh_buf *R, double *im_out);
void two_site_theta(tnh_ctx *ctx, const tnh_buf *A1, const tnh_buf *A2, tnh_buf **theta);
// rank-1 Hamiltonian terms (projector.hip): 2-index overlap blocks B[a_phi, a_psi] (FP64)
void proj_env_left(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *phi, const tnh_buf *psi,
                   tnh_buf **Lout);
void proj_env_right(tnh_ctx *ctx, const tnh_buf *R, const tnh_buf *phi, const tnh_buf *psi,
                    tnh_buf **Rout);
void proj_project(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *phi1, const tnh_buf *phi2,
                  const tnh_buf *R, tnh_buf **P);
void proj_project1(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *phi, const tnh_buf *R,
                   tnh_buf **P);
void proj_apply(tnh_ctx *ctx, const tnh_buf *P, const tnh_buf *x, tnh_buf *y, double coeff);
void proj_calculate(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *phi, const tnh_buf *psi,
                    const tnh_buf *R, double *re, double *im);
void cscale(tnh_ctx *ctx, void *x, int t, int64_t n, double cr, double ci);
void cdot_host(tnh_ctx *ctx, const tnh_buf *x, const tnh_buf *y, double *re, double *im);

// y = H_eff x for the Krylov solver
using Matvec = std::function<void(const tnh_buf *x, tnh_buf *y)>;
double eigsolve_ground_op(tnh_ctx *ctx, const Matvec &matvec, tnh_buf *theta, int kdim,
                          int maxiter, double tol, int *nmv_out);
// the same for a non-Hermitian operator: the Ritz value of smallest real part (complex theta)
std::complex<double> eigsolve_nonherm_op(tnh_ctx *ctx, const Matvec &matvec, tnh_buf *theta,
                                         int kdim, int maxiter, double tol, int *nmv_out);
// warm (optional, in/out): a p x p basis of the Gram space from an earlier split of the same
// bond, used as the Jacobi's starting basis when its size and dtype match, and replaced by
// this split's eigenvectors (block-Jacobi path, p > 64).
void replacesites2(tnh_ctx *ctx, const tnh_buf *theta, int direction, int normalize,
                   double cutoff, int maxdim, int mindim, tnh_buf **A1, tnh_buf **A2, int *k_out,
                   double *disc_out, double *schmidt, int schmidt_cap, bool force_fallback,
                   tnh_buf **warm = nullptr);
void mps_movecenter(tnh_ctx *ctx, int N, tnh_buf **s, int center, int idx, double cutoff,
                    int maxdim, int mindim);

// Outcome of one truncated decomposition: all singular values (descending), kept rank,
// discarded weight and whether the one-sided fallback ran.
struct SvdOut {
  int m = 0, n = 0, k = 0;
  std::vector<double> s;
  double discarded = 0;
  bool fallback = false;
};
// Factor the m x n matrix of `src` (any dtype; column-major, its leading dims are rows) as
// left * right with the truncation rule: side 0 -> left = U_k (orthonormal columns),
// right = U_k^H M; side 1 -> right = V_k^H (orthonormal rows), left = M V_k. left / right
// are (re)allocated with dims ldims / rdims whose last / first entry is replaced by k.
// force_fallback runs the one-sided Jacobi on M (accurate small singular values).
int split_matrix(tnh_ctx *ctx, const tnh_buf *src, int m, int n, int side, double cutoff,
                 int maxdim, int mindim, bool force_fallback, int nl, const int64_t *ldims,
                 int nr, const int64_t *rdims, tnh_buf **A1, tnh_buf **A2, SvdOut *o,
                 tnh_buf **warm = nullptr);
void mps_moveleft(tnh_ctx *ctx, tnh_buf **A, tnh_buf **B, double cutoff, int maxdim, int mindim);
void mps_moveright(tnh_ctx *ctx, tnh_buf **A, tnh_buf **B, double cutoff, int maxdim, int mindim);

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
