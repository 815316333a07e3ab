// Fused 3D FP64 stage kernels (placeholder until the fused kernel lands).
#include "fst_fused.hpp"

namespace fst {

bool fused_supported(const fst_config &cf) { (void)cf; return false; }

int fused_steps(const Geo &, const Phys &, const fst_config &, Bufs &, FusedBufs &, int, hipStream_t,
                std::string *err) {
  *err = "fused path not available";
  return FST_ERR_ARG;
}

void fused_free(FusedBufs &fb) {
  if (fb.stage) (void)hipFree(fb.stage);
  if (fb.prim) (void)hipFree(fb.prim);
  fb.stage = fb.prim = nullptr;
}

}  // namespace fst
