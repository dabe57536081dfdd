// Instantiations of the BF16X3 GEMM kernels and their launcher.
#include "gemm_kernel.hpp"
#include "tnh_internal.hpp"

namespace tnh {
namespace gk {

template <typename TA, typename TB, typename TC, bool TRA, bool TRB, int BM, int BN>
static void launch(hipStream_t s, const Params &p, int batch) {
  dim3 grid(p.tiles_m * p.tiles_n, batch);
  // the hi/lo split writes LDS element by element; the scalar load order keeps those
  // writes on consecutive banks, the vector order does not
  hipLaunchKernelGGL((gemm_kernel<MODE_BF16X3, TA, TB, TC, TRA, TRB, BM, BN, false>), grid, dim3(256), 0, s, p);
}

template <typename TA, typename TB, typename TC, int BM, int BN>
static void launch_tr(hipStream_t s, const Params &p, int batch, bool tra, bool trb) {
  if (!tra && !trb) launch<TA, TB, TC, false, false, BM, BN>(s, p, batch);
  else if (!tra && trb) launch<TA, TB, TC, false, true, BM, BN>(s, p, batch);
  else if (tra && !trb) launch<TA, TB, TC, true, false, BM, BN>(s, p, batch);
  else throw Error(TNH_ERR_UNSUPPORTED, "gemm: op(A)=T with op(B)=T is not instantiated");
}

template <int BM, int BN>
static void launch_types(hipStream_t s, const Params &p, int batch, bool tra, bool trb, int tA, int tB, int tC) {
  const bool a64 = tA == TNH_F64, b64 = tB == TNH_F64, c64 = tC == TNH_F64;
#define TNH_CASE(CA, CB, CC, TA_, TB_, TC_) \
  if (a64 == CA && b64 == CB && c64 == CC) return launch_tr<TA_, TB_, TC_, BM, BN>(s, p, batch, tra, trb);
  TNH_CASE(false, false, false, float, float, float)
  TNH_CASE(false, false, true, float, float, double)
  TNH_CASE(false, true, false, float, double, float)
  TNH_CASE(false, true, true, float, double, double)
  TNH_CASE(true, false, false, double, float, float)
  TNH_CASE(true, false, true, double, float, double)
  TNH_CASE(true, true, false, double, double, float)
  TNH_CASE(true, true, true, double, double, double)
#undef TNH_CASE
}

void launch_bf16x3(hipStream_t s, const Params &p, int batch, bool tra, bool trb, int tA, int tB, int tC, bool big) {
  if (big) launch_types<128, 128>(s, p, batch, tra, trb, tA, tB, tC);
  else launch_types<64, 64>(s, p, batch, tra, trb, tA, tB, tC);
}

}  // namespace gk
}  // namespace tnh
