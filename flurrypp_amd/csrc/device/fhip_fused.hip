// Fused tensor-product path (placeholder until the kernels land).
#include "fhip_internal.hpp"

namespace fhip {

bool fused_supported(const Ctx&) { return false; }
void fused_setup(Ctx&) {}
void fused_rk_step(Ctx&, bool) {}

}  // namespace fhip
