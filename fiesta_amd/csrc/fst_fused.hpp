// Fused stage path: one ghost fill + one RHS/RK kernel per Runge-Kutta stage.
#pragma once
#include <string>

#include "fst_common.hpp"
#include "fst_kernels.hpp"

namespace fst {

struct FusedBufs {
  double *stage = nullptr;   // U* (stage-1 output), nvt planes
  double *stage2 = nullptr;  // SSP-RK3: U2 (stage-2 output), nvt planes
  double *prim = nullptr;    // two sets of primitive planes rho, p, u, v, w, nrg, Y_s
  Scalars *sc = nullptr;     // two sets of per-stage maxima
  float *mt32 = nullptr;     // gen3d FP32 form with viscous terms: the 9 metric planes as floats
  bool mt32_ok = false;      // mt32 holds the current metrics
  int kchunk = 0;            // k planes per block (0: default)
  bool valid = false;        // prim[cur] / sc[cur] describe the current var (interior)
  int cur = 0;               // which set describes var (an SSP-RK3 step flips it)
};

struct LearnedBufs;

bool fused_supported(const fst_config &cf);
// The fused cart2d stage (fst_fused2d.hip): cartesian 2D, WENO5, FP64,
// ns <= 3; plain, learned or no C-equation.  fb.stage / fb.stage2 hold the
// stage states, fb.sc three maxima records used in rotation (fb.cur).
bool fused2d_supported(const fst_config &cf);
int fused2d_steps(const Geo &g, const Phys &ph, const fst_config &cf, Bufs &b, FusedBufs &fb, LearnedBufs *lb,
                  int nsteps, hipStream_t s, std::string *err);
// gen3d's advection and metric pressure gradient on the fused kernel (rho, p,
// contravariant velocity in the 5 planes of P, metric terms MT);
// gen_adv_supported says whether the configuration fits.
bool gen_adv_supported(const fst_config &cf, const Geo &g);
// buoy: add the gen_ext buoyancy after the pressure gradient in the same
// launch (else k_buoy_gen does, as the reference-sequence order has it).
void launch_gen_adv(const Geo &g, const Phys &ph, const fst_config &cf, const double *U, const double *P,
                    const double *MT, double *dvar, hipStream_t s, int buoy = 0);
// nsteps time steps of gen3d on the fused kernels (gen_adv_supported
// configurations): per stage one EOS / contravariant-velocity pass and one
// advection + stage-update launch (fst_gen3d.hip).
int fused_gen_steps(const Geo &g, const Phys &ph, const fst_config &cf, Bufs &b, FusedBufs &fb, int nsteps,
                    hipStream_t s, std::string *err);
// nsteps time steps of the fused path (RK2 midpoint, or SSP-RK3 with
// rk_order = 3); lb carries the learned C-indicator fields when cf.learned_c
// is set (else may be null).
int fused_steps(const Geo &g, const Phys &ph, const fst_config &cf, Bufs &b, FusedBufs &fb, LearnedBufs *lb,
                int nsteps, hipStream_t s, std::string *err);
// The same for precision = FST_FP64_FAST (fst_fast64.hip); fused_steps
// forwards to it.
int fused_steps_fast64(const Geo &g, const Phys &ph, const fst_config &cf, Bufs &b, FusedBufs &fb,
                       LearnedBufs *lb, int nsteps, hipStream_t s, std::string *err);
// Primitive planes (rho, p, u, v, w, nrg, Y_s) of U over every cell (3D).
int fused_primitives(const Geo &g, const Phys &ph, const fst_config &cf, const double *U, double *P, hipStream_t s);
void fused_free(FusedBufs &fb);

// Validation and derived sizes of a configuration (fst_config_dims), and the
// module constants fst_create builds from it: the effective config (2D and
// generalized-grid defaults applied), the launch geometry and the device
// constants.
int module_check(const fst_config &in, fst_dims *d, std::string *why);
void module_setup(const fst_config &in, const fst_dims &d, fst_config *cf, Geo *g, Phys *ph);

}  // namespace fst
