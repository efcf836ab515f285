// Split-BF16 instantiation of the fused MLP kernels (templates in mlp_impl.h): FP32 activations,
// stored tiles and accumulation, products on 16x16x32 BF16 MFMAs over BF16 operand planes.  The
// forward uses three planes (six plane products, FP32-level rounding, so ReLU decisions agree
// with an FP32 evaluation); the backward chain and the weight gradients use two (three products).
#include "mlp_impl.h"

namespace srf {
namespace mlp {

int fwd_bf16s(const FwdArgs& a, bool vd, bool tr, hipStream_t s) {
  if (vd) return tr ? launch_fwd<kBFS3, true, true>(a, s) : launch_fwd<kBFS3, true, false>(a, s);
  return tr ? launch_fwd<kBFS3, false, true>(a, s) : launch_fwd<kBFS3, false, false>(a, s);
}

int bwd_bf16s(const BwdArgs& a, bool vd, hipStream_t s) {
  return vd ? launch_bwd<kBFS2, true>(a, s) : launch_bwd<kBFS2, false>(a, s);
}

int dw_bf16s(const DwArgs& d, hipStream_t s) {
  hipLaunchKernelGGL(mlp_dw_split_kernel<kBFS2>, dim3(d.total_work), dim3(64 * kDwWaves), 0, s, d);
  return launch_status();
}

}  // namespace mlp
}  // namespace srf
