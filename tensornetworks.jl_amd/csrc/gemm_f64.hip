// Instantiations of the F64 GEMM kernels and their launcher.
#include "gemm_kernel.hpp"
#include "tnh_internal.hpp"

namespace tnh {
namespace gk {

template <typename TA, typename TB, typename TC, bool TRA, bool TRB, int BM, int BN>
static void launch(hipStream_t s, const Params &p, int batch) {
  dim3 grid(p.tiles_m * p.tiles_n, batch);
  if (p.vecA && p.vecB)
    hipLaunchKernelGGL((gemm_kernel<MODE_F64, TA, TB, TC, TRA, TRB, BM, BN, true>), grid, dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((gemm_kernel<MODE_F64, TA, TB, TC, TRA, TRB, BM, BN, false>), grid, dim3(256), 0, s, p);
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
  if (tA != TNH_F64 || tB != TNH_F64 || tC != TNH_F64)
    throw Error(TNH_ERR_UNSUPPORTED, "FP64 GEMM needs FP64 operands");
  launch_tr<double, double, double, BM, BN>(s, p, batch, tra, trb);
}

// tile: 0 = 64 x 64, 1 = 128 x 64, 2 = 128 x 128 (kTileShapes in gemm.hip)
void launch_f64(hipStream_t s, const Params &p, int batch, bool tra, bool trb, int tA, int tB, int tC, int tile) {
  switch (tile) {
    case 0: return launch_types<64, 64>(s, p, batch, tra, trb, tA, tB, tC);
    case 1: return launch_types<128, 64>(s, p, batch, tra, trb, tA, tB, tC);
    case 2: return launch_types<128, 128>(s, p, batch, tra, trb, tA, tB, tC);
    default: throw Error(TNH_ERR_ARG, "gemm: unknown FP64 tile " + std::to_string(tile));
  }
}

}  // namespace gk
}  // namespace tnh
