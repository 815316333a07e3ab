// Fused stage path: one ghost fill + one RHS/RK kernel per Runge-Kutta stage.
#pragma once
#include <string>

#include "fst_common.hpp"
#include "fst_kernels.hpp"

namespace fst {

struct FusedBufs {
  double *stage = nullptr;  // U* (stage-1 output), nvt planes
  double *prim = nullptr;   // two sets of primitive planes rho, p, u, v, w, nrg, Y_s
  Scalars *sc = nullptr;    // two sets of per-stage maxima
  int kchunk = 0;           // k planes per block (0: default)
  bool valid = false;       // prim[0] / sc[0] describe the current var (interior)
};

bool fused_supported(const fst_config &cf);
int fused_steps(const Geo &g, const Phys &ph, const fst_config &cf, Bufs &b, FusedBufs &fb, int nsteps,
                hipStream_t s, std::string *err);
void fused_free(FusedBufs &fb);

}  // namespace fst
