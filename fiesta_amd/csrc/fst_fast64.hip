// The fused 3D stage kernels of the fast FP64 form (precision =
// FST_FP64_FAST): fst_fused3d.hip compiled again with FST_FAST64_TU, which
// switches the WENO5 fluxes and the viscous / artificial-viscosity face terms
// to contracted arithmetic with reciprocal-based division and one-division
// WENO5 weights.  It exports fused_steps_fast64 only.
#define FST_FAST64_TU
#include "fst_fused3d.hip"
