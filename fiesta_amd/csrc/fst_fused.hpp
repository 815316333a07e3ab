// Fused stage path: one ghost fill + one RHS/RK kernel per Runge-Kutta stage.
#pragma once
#include <string>

#include "fst_common.hpp"
#include "fst_kernels.hpp"

namespace fst {

struct FusedBufs {
  double *stage = nullptr;  // U* (stage-1 output), nvt planes
  double *prim = nullptr;   // primitive planes rho, p, u, v, w
  size_t planes_stage = 0;
};

bool fused_supported(const fst_config &cf);
int fused_steps(const Geo &g, const Phys &ph, const fst_config &cf, Bufs &b, FusedBufs &fb, int nsteps,
                hipStream_t s, std::string *err);
void fused_free(FusedBufs &fb);

}  // namespace fst
