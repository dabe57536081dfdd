// Instantiations of the F32 GEMM kernels and their launcher.
#include "gemm_kernel.hpp"
#include "tnh_internal.hpp"

namespace tnh {
namespace gk {

template <typename TA, typename TB, typename TC, bool TRA, bool TRB, int BM, int BN>
static void launch(hipStream_t s, const Params &p, int batch) {
  dim3 grid(p.tiles_m * p.tiles_n, batch);
  if (p.vecA && p.vecB)
    hipLaunchKernelGGL((gemm_kernel<MODE_F32, TA, TB, TC, TRA, TRB, BM, BN, true>), grid, dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((gemm_kernel<MODE_F32, TA, TB, TC, TRA, TRB, BM, BN, false>), grid, dim3(256), 0, s, p);
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

// tile: 0 = 64 x 64, 1 = 128 x 128, 2 = 128 x 64 (kTileShapes in gemm.hip)
void launch_f32(hipStream_t s, const Params &p, int batch, bool tra, bool trb, int tA, int tB, int tC, int tile) {
  switch (tile) {
    case 0: return launch_types<64, 64>(s, p, batch, tra, trb, tA, tB, tC);
    case 1: return launch_types<128, 128>(s, p, batch, tra, trb, tA, tB, tC);
    case 2: return launch_types<128, 64>(s, p, batch, tra, trb, tA, tB, tC);
    default: throw Error(TNH_ERR_ARG, "gemm: unknown FP32 tile " + std::to_string(tile));
  }
}

}  // namespace gk
}  // namespace tnh
