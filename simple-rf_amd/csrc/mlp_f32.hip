// f32 instantiation of the fused MLP kernels (templates in mlp_impl.h).
#include "mlp_impl.h"

namespace srf {
namespace mlp {

int fwd_f32(const FwdArgs& a, bool vd, bool tr, hipStream_t s) {
  if (vd) return tr ? launch_fwd<kF32, true, true>(a, s) : launch_fwd<kF32, true, false>(a, s);
  return tr ? launch_fwd<kF32, false, true>(a, s) : launch_fwd<kF32, false, false>(a, s);
}

int bwd_f32(const BwdArgs& a, bool vd, hipStream_t s) {
  return vd ? launch_bwd<kF32, true>(a, s) : launch_bwd<kF32, false>(a, s);
}

int dw_f32(const DwArgs& d, hipStream_t s) { return launch_dw<kF32>(d, s); }

int dw_f32_narrow(const DwArgs& d, hipStream_t s) { return launch_dw<kF32, 1, 10>(d, s); }

}  // namespace mlp
}  // namespace srf
