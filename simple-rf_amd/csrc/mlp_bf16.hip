// bf16 instantiation of the fused MLP kernels (templates in mlp_impl.h).
#include "mlp_impl.h"

namespace srf {
namespace mlp {

int fwd_bf16(const FwdArgs& a, bool vd, bool tr, hipStream_t s) {
  if (vd) return tr ? launch_fwd<kBF16, true, true>(a, s) : launch_fwd<kBF16, true, false>(a, s);
  return tr ? launch_fwd<kBF16, false, true>(a, s) : launch_fwd<kBF16, false, false>(a, s);
}

int bwd_bf16(const BwdArgs& a, bool vd, hipStream_t s) {
  return vd ? launch_bwd<kBF16, true>(a, s) : launch_bwd<kBF16, false>(a, s);
}

int dw_bf16(const DwArgs& d, hipStream_t s) { return launch_dw<kBF16>(d, s); }

}  // namespace mlp
}  // namespace srf
