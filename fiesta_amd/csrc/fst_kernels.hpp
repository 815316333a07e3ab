// Device helpers shared by the reference-sequence and fused kernels, and the
// launcher declarations used by the C-ABI layer.
#pragma once
#include "fst_common.hpp"

#include <vector>

namespace fst {

// Event-bracketed timing of one kernel family on a stream.
// Live timing of the per-stage work with hipEvents on the context stream.
// A timed unit records `per` events: begin, (per - 2) phase marks, end.
struct Prof {
  static constexpr int MAXPER = 8;
  bool on = false;
  const char *name = "";
  int per = 2;      // events per timed unit
  int marked = 0;   // marks recorded in the open unit
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  bool grow(size_t n) {
    while (ev.size() < n) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) { on = false; return false; }
      ev.push_back(e);
    }
    return true;
  }
  void begin(hipStream_t s) {
    if (!on || !grow(used + per)) return;
    marked = 0;
    (void)hipEventRecord(ev[used], s);
  }
  void mark(hipStream_t s) {
    if (!on || marked + 2 >= per) return;
    ++marked;
    (void)hipEventRecord(ev[used + marked], s);
  }
  void end(hipStream_t s) {
    if (!on) return;
    while (marked + 2 < per) {  // unrecorded marks collapse onto the end
      ++marked;
      (void)hipEventRecord(ev[used + marked], s);
    }
    (void)hipEventRecord(ev[used + per - 1], s);
    used += per;
  }
  ~Prof() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

// Device buffers of one context.  Arrays are LayoutLeft planes of N cells.
struct Bufs {
  double *var = nullptr, *tmp1 = nullptr, *dvar = nullptr, *varx = nullptr;
  double *p = nullptr, *rho = nullptr, *T = nullptr, *vel = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr;
  double *stress = nullptr;  // 3D: 6 planes; 2D: sx00 sx11 sx01 sy00 sy11 sy01 qx qy
  double *gradRho = nullptr;
  double *mflux = nullptr;   // 3D: 27 planes ((face*3+dir)*3+w); 2D: 8 planes
  double *grid = nullptr, *metrics = nullptr;
  // noise filter (cf.noise): smoothing increments (1 plane), coarse-point
  // flags, and the maximum of the C variable that sets the cut-off
  double *ndn = nullptr;
  unsigned char *nflag = nullptr;
  unsigned long long *nmax = nullptr;
  double *prim = nullptr;  // gen3d: rho, p, vel[3] in one 5-plane block (rho, p, vel point into it)
  unsigned long long *mm = nullptr;  // status check: (max, min) order keys per variable
  Scalars *sc = nullptr;     // per-stage maxima (device)
  Prof *prof = nullptr;      // timing of the dominant kernel (host side)
};

// WENO5 of calculateFluxesG (src/flux.hpp:240-264): binary32 coefficients
// promoted to double, the reference association kept term by term.
__device__ __forceinline__ double weno5_f32c(double f1, double f2, double f3, double f4, double f5) {
  const double eps = 0.000001;
  double b1, b2, b3, w1, w2, w3, p1, p2, p3, a1, a2;
  a1 = f1 - 2.0f * f2 + f3;
  a2 = f1 - 4.0f * f2 + 3.0f * f3;
  b1 = (13.0f / 12.0f) * a1 * a1 + (0.25f) * a2 * a2;
  a1 = f2 - 2.0 * f3 + f4;
  a2 = f2 - f4;
  b2 = (13.0f / 12.0f) * a1 * a1 + (0.25f) * a2 * a2;
  a1 = f3 - 2.0 * f4 + f5;
  a2 = 3.0 * f3 - 4.0 * f4 + f5;
  b3 = (13.0f / 12.0f) * a1 * a1 + (0.25f) * a2 * a2;
  a1 = eps + b1;
  w1 = (0.1f) / (a1 * a1);
  a1 = eps + b2;
  w2 = (0.6f) / (a1 * a1);
  a1 = eps + b3;
  w3 = (0.3f) / (a1 * a1);
  p1 = (1.0f / 3.0f) * f1 + (-7.0f / 6.0f) * f2 + (11.0f / 6.0f) * f3;
  p2 = (-1.0f / 6.0f) * f2 + (5.0f / 6.0f) * f3 + (1.0f / 3.0f) * f4;
  p3 = (1.0f / 3.0f) * f3 + (5.0f / 6.0f) * f4 + (-1.0f / 6.0f) * f5;
  return (w1 * p1 + w2 * p2 + w3 * p3) / (w1 + w2 + w3);
}

// WENO5 of computeFluxWeno2D (src/flux.hpp:38-62): double coefficients.
__device__ __forceinline__ double weno5_f64c(double f1, double f2, double f3, double f4, double f5) {
  const double eps = 1e-6;
  double b1, b2, b3, w1, w2, w3, p1, p2, p3, a1, a2;
  a1 = f1 - 2.0 * f2 + f3;
  a2 = f1 - 4.0 * f2 + 3.0 * f3;
  b1 = (13.0 / 12.0) * a1 * a1 + (0.25) * a2 * a2;
  a1 = f2 - 2.0 * f3 + f4;
  a2 = f2 - f4;
  b2 = (13.0 / 12.0) * a1 * a1 + (0.25) * a2 * a2;
  a1 = f3 - 2.0 * f4 + f5;
  a2 = 3.0 * f3 - 4.0 * f4 + f5;
  b3 = (13.0 / 12.0) * a1 * a1 + (0.25) * a2 * a2;
  a1 = eps + b1;
  w1 = (0.1) / (a1 * a1);
  a1 = eps + b2;
  w2 = (0.6) / (a1 * a1);
  a1 = eps + b3;
  w3 = (0.3) / (a1 * a1);
  p1 = (1.0 / 3.0) * f1 + (-7.0 / 6.0) * f2 + (11.0 / 6.0) * f3;
  p2 = (-1.0 / 6.0) * f2 + (5.0 / 6.0) * f3 + (1.0 / 3.0) * f4;
  p3 = (1.0 / 3.0) * f3 + (5.0 / 6.0) * f4 + (-1.0 / 6.0) * f5;
  return (w1 * p1 + w2 * p2 + w3 * p3) / (w1 + w2 + w3);
}

void launch_eos_vel(const Geo &g, const Phys &ph, const Bufs &b, hipStream_t s);
// Learned C-indicator fields of the 2D form: C_hat [N], the tangent [2][N]
// and max C_hat (sc->maxCh), all on the device.
struct Ceq2dFields {
  const float *chat = nullptr, *dir = nullptr;
  const Scalars *sc = nullptr;
};
// cart3d/cart2d compute() (src/cart3d.cpp:220-298, src/cart2d.cpp:151-237).
// lf: the 2D learned form's fields (nullptr: the plain method); eos_done: the
// primitives of this state are already in b.
void launch_compute_ref(const Geo &g, const Phys &ph, const fst_config &cf, const Bufs &b, hipStream_t s,
                        const Ceq2dFields *lf = nullptr, bool eos_done = false);
void launch_apply_bcs(const Geo &g, const fst_config &cf, double *var, hipStream_t s);
void launch_apply_bcs(const Geo &g, const fst_config &cf, float *var, hipStream_t s);  // float stage state
// Stages per time step: 2 (RK2 midpoint, rkAdvance) or 3 (SSP-RK3, rk_order = 3).
inline int rk_stages(const fst_config &c) { return c.rk_order == 3 ? 3 : 2; }
// SSP-RK3 (Shu-Osher) stage s = 2, 3: var = A*tmp1 + B*(var + dt*dvar)
constexpr double RK3_A2 = 0.75, RK3_B2 = 0.25, RK3_A3 = 1.0 / 3.0, RK3_B3 = 2.0 / 3.0;
void launch_rk_stage(const Geo &g, int nstages, int stage, double dt, double *var, double *tmp, const double *dvar,
                     hipStream_t s);
// Output pack box (inclusive start..end, stride, extent per direction).
struct PackBox {
  long long start[3], stride[3], ext[3];
};
void launch_pack(const double *src, long long nx, long long ny, int ndim, int off, const PackBox &bx, int average,
                 bool f32, void *dst, hipStream_t s);
// Noise filter of postStep on var (cartesian modules; cf.noise = 1).
int noise_flag_count(const Geo &g, bool gen);
void launch_noise(const Geo &g, const Phys &ph, const fst_config &cf, const Bufs &b, hipStream_t s);
void launch_copy(const void *src, void *dst, size_t bytes, hipStream_t s);   // 32-byte elements
void launch_copy8(const void *src, void *dst, size_t bytes, hipStream_t s);  // 8-byte elements
// statusCheck reductions: max and min over the interior cells of each of the
// first nvar variables of `a` (planes of g.N values), into mm as order keys
// (mm[2v] max, mm[2v+1] min); decode with key_to_double.
void launch_minmax(const Geo &g, const double *a, int nvar, unsigned long long *mm, hipStream_t s);
inline double key_to_double(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
  return __builtin_bit_cast(double, b);
}
void launch_secondary(const Geo &g, const Phys &ph, const Bufs &b, hipStream_t s);
// generalized grids (gen2d/gen3d modules)
void launch_compute_gen(const Geo &g, const Phys &ph, const fst_config &cf, const Bufs &b, hipStream_t s);
// gen_ext viscous terms of gen3d in metric form (the faces of the three
// directions and their differences into b.dvar, x, y, z); mt32: the FP32
// form, with the metric planes as floats (launch_metrics_f32).
void launch_viscgen3d(const Geo &g, const Phys &ph, const Bufs &b, hipStream_t s, const float *mt32 = nullptr);
void launch_metrics_f32(const Geo &g, const double *mt, float *mt32, hipStream_t s);
void launch_metrics(const Geo &g, const Bufs &b, hipStream_t s);

}  // namespace fst
