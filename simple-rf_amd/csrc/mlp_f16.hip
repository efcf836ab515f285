// FP16 instantiation of the fused MLP forward for inference (templates in mlp_impl.h).  Only the
// eval forward (no stored activations) is built for FP16 operands: it is what a full-image render
// needs, and the training kernels' FP16 instantiations are not part of this library.
#include "mlp_impl.h"

namespace srf {
namespace mlp {

int fwd_f16(const FwdArgs& a, bool vd, bool tr, hipStream_t s) {
  if (tr) return SRF_EUNSUPPORTED;
  return vd ? launch_fwd<kF16, true, false>(a, s) : launch_fwd<kF16, false, false>(a, s);
}

}  // namespace mlp
}  // namespace srf
