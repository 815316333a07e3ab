// Fused 3D stage path (cart3d, src/cart3d.cpp:220-298 + src/rk.cpp:28-72).
//
// One Runge-Kutta stage = ghost fill (src/bc.cpp:179-249) + ghost primitives +
// three marching kernels:
//   k_adv3d   advective fluxes on every face (WENO5), flux differences,
//             pressure gradient and buoyancy -> partial RHS of the nv flow
//             variables (dvar planes);
//   k_vcr3d   viscous stresses and C-equation fluxes on every face, their
//             divergences, the stage update of the flow variables, and the
//             primitives + max wave speed of the new state for the next stage;
//   k_ceq3d   the C-equation sources and relaxation, the stage update of the
//             five C variables and max |C_hat| (plain C-equation only).
// In the FP64 form every term keeps the reference association and constants,
// so the result equals the reference kernel sequence bit for bit:
//   * divisions by run constants (dx, 4*dy, 12*dx, 12) use q0 = a*r,
//     e = fma(-b, q0, a), q = fma(e, r, q0) with r = RN(1/b) (Markstein's
//     theorem: q is the correctly rounded a/b); other divisions are IEEE.
//   * per-cell dvar is accumulated in the reference order: advection (x,y,z),
//     pressure gradient, buoyancy, viscous x, y, z, C-equation.
//
// Work decomposition of the advection and viscous kernels (block = TY+1 waves
// of 64 lanes, tile = 64 (i) x TY (j) cells, marching k over a chunk; TY = 7,
// and 11 for the binary32 advection kernel of the FP32 forms):
//   waves 0..TY-1  row r: x-face right of its cell, y-face below its row,
//                  z-face above its cell, then the cell itself;
//   wave TY        y-face above the tile, and the TY x-faces left of the tile.
// x and y face values are exchanged through LDS; the z-face of the previous
// plane stays in registers.  Each face is computed exactly once per kernel.
//
// Primitive planes P (rho, p, u, v, w, nrg, Y_0..Y_{ns-1}) of the stage input
// come from the previous stage's epilogue (interior) and a ghost pass; the
// per-stage maxima (max wave speed, max |C_hat|) likewise.  The FP32 forms
// store them as floats except the pressure (prim(), fst_common.hpp).
//
// Addressing: element offsets are 32-bit (checked on the host: every array
// stays below 2^32 elements) on uniform base pointers, scaled to 64-bit byte
// offsets (ld/st).
//
// Fast FP64 form (precision = FST_FP64_FAST): this file compiled a second time
// as fst_fast64.hip with FST_FAST64_TU defined.  There the WENO5 fluxes
// (weno5_onediv: one division per face value instead of four) and the
// viscous / artificial-viscosity face terms (face_vc, face_vc_ring) contract
// a*b+c into FMAs and divide through the hardware reciprocal (fast_div).
// Everything that sets the pressure and its gradient keeps the reference
// arithmetic: the EOS and primitives, the pressure gradient and buoyancy,
// the C-equation sensors and relaxation, and the RK update.  Near hydrostatic
// balance the momentum right-hand side is the small difference of the
// pressure gradient and gravity, so one ulp of p there would move it far
// more than the flux rounding does.  The form is held to relative L2
// <= 1e-12 against the reference arithmetic.  Only fused_steps is exported
// from that unit, as fused_steps_fast64.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdlib>

#include "fst_fused.hpp"
#include "fst_mlp.hpp"

namespace fst {

namespace {
#ifdef FST_FAST64_TU
// the fast FP64 form's kernels carry their own names (fst::(anonymous
// namespace)::fast64::k_...), apart from the reference-arithmetic ones
namespace fast64 {
#endif

// First statement of a function body whose arithmetic the fast FP64 form
// contracts (a no-op in the reference-arithmetic units).
#ifdef FST_FAST64_TU
#define FST_FAST_FP _Pragma("clang fp contract(fast)")
#else
#define FST_FAST_FP
#endif

constexpr int TX = 64;
constexpr int TY = 7;
constexpr int NWAVE = TY + 1;
constexpr int NTHREADS = 64 * NWAVE;
constexpr int NFV = 12;  // 3 stresses + 9 C-equation fluxes per face

struct FC {  // run constants
  double d[3], rd[3];      // cell sizes and RN(1/d)
  double d4[3], rd4[3];    // 4*d
  double d12[3], rd12[3];  // 12*d (C-equation gradients)
  double pg[3], rpg[3];    // pressure-gradient divisors 12*cd(1), 12*cd(2), 12*cd(2)
  double r12;              // RN(1/12), face-velocity interpolation
  double cpf[FST_MAX_SPECIES], cvf[FST_MAX_SPECIES], mus[FST_MAX_SPECIES];
  double g, rhoRef, kap, eps, alpha_cf, dxmag2, dxmag;
};

struct Grid32 {
  uint32_t N, s1, s2;  // plane size, j stride, k stride
  int ngi, ngj, ng, nci, ncj, nck, nv, nvt;
  __device__ __forceinline__ uint32_t id(int i, int j, int k) const {
    return (uint32_t)i + (uint32_t)j * s1 + (uint32_t)k * s2;
  }
};

__device__ __forceinline__ double ld(const double *__restrict__ base, uint32_t off) {
  return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + ((uint64_t)off << 3));
}
__device__ __forceinline__ void st(double *__restrict__ base, uint32_t off, double v) {
  *reinterpret_cast<double *>(reinterpret_cast<char *>(base) + ((uint64_t)off << 3)) = v;
}

// State value at element offset off of U (doubles), or of the float
// Runge-Kutta stage state of the learned FP32 form (UF).
template <bool UF>
__device__ __forceinline__ double uld(const double *__restrict__ U, uint32_t off) {
  if (UF) return (double)reinterpret_cast<const float *>(U)[off];
  return ld(U, off);
}
template <bool UF>
__device__ __forceinline__ void ust(double *__restrict__ U, uint32_t off, double v) {
  if (UF) reinterpret_cast<float *>(U)[off] = (float)v;
  else st(U, off, v);
}

// Primitive planes of cell c with 32-bit offsets, in the layout of prim()
// (fst_common.hpp): FP64 form all doubles; FP32 forms p as doubles, then
// rho, u, v, w, nrg, Y_s as floats.
template <bool F32>
__device__ __forceinline__ double pget(const double *__restrict__ P, uint32_t N, int q, uint32_t c) {
  if (!F32) return ld(P, c + N * (uint32_t)q);
  if (q == 1) return ld(P, c);
  return (double)reinterpret_cast<const float *>(P + N)[c + N * (uint32_t)(q == 0 ? 0 : q - 1)];
}
template <bool F32>
__device__ __forceinline__ void pset(double *__restrict__ P, uint32_t N, int q, uint32_t c, double v) {
  if (!F32) {
    st(P, c + N * (uint32_t)q, v);
  } else if (q == 1) {
    st(P, c, v);
  } else {
    reinterpret_cast<float *>(P + N)[c + N * (uint32_t)(q == 0 ? 0 : q - 1)] = (float)v;
  }
}

// Partial right-hand side handed from the advection kernel to the viscous
// kernel: FP64 in the FP64 form, FP32 in the FP32 forms (the pressure-
// gradient / buoyancy cancellation happens in FP64 before the store; the RK
// accumulation that adds it stays FP64).
template <bool F32>
using DvT = typename std::conditional<F32, float, double>::type;
template <bool F32>
__device__ __forceinline__ void st_dv(double *__restrict__ base, uint32_t off, double v) {
  if (F32) reinterpret_cast<float *>(base)[off] = (float)v;
  else st(base, off, v);
}
template <bool F32>
__device__ __forceinline__ double ld_dv(const double *__restrict__ base, uint32_t off) {
  if (F32) return (double)reinterpret_cast<const float *>(base)[off];
  return ld(base, off);
}

__device__ __forceinline__ double cdiv(double a, double b, double rb) {
  double q0 = a * rb;
  double e = fma(-b, q0, a);
  return fma(e, rb, q0);
}

// Stage update of the fused kernels' epilogues, one value: form 1: U + h dv
// (the first stage of either scheme), 2: Un + h dv (rkAdvance's second stage),
// 3: a Un + b (U + h dv) (SSP-RK3 stages 2 and 3); U is the stage input, Un
// the state at the start of the step.
struct Upd {
  double h, a, b;
  int form;
};
__device__ __forceinline__ double upd_val(const Upd &u, double us, double ub, double dv) {
  if (u.form == 3) return u.a * ub + u.b * (us + u.h * dv);
  return (u.form == 1 ? us : ub) + u.h * dv;
}

template <int D>
__device__ __forceinline__ uint32_t stride(const Grid32 &g) {
  return D == 0 ? 1u : (D == 1 ? g.s1 : g.s2);
}

// Tile placement shared by both kernels.  Hardware block b runs on XCD b % 8;
// the remap gives consecutive logical tiles (neighbours in i, then j) the same
// XCD, so halo rows are found in that XCD's L2.
struct Tile {
  int w, l, i0, j0, kb, ke, i, jrow;
  bool cellok, yfok, xbok;
};

template <int TYV = TY>
__device__ __forceinline__ Tile place(const Grid32 &g, int kchunk, int nbx, int nby, int nbz) {
  Tile t;
  const int tid = threadIdx.x;
  t.w = tid >> 6;
  t.l = tid & 63;
  const int nb = nbx * nby * nbz;
  const int hb = blockIdx.x;
  int lb = hb;
  const int full = (nb / 8) * 8;
  if (hb < full) lb = (hb % 8) * (full / 8) + hb / 8;
  const int bx = lb % nbx, by = (lb / nbx) % nby, bz = lb / (nbx * nby);
  t.i0 = g.ng + bx * TX;
  t.j0 = g.ng + by * TYV;
  t.kb = g.ng + bz * kchunk;
  t.ke = min(t.kb + kchunk, g.ng + g.nck);
  const int iend = g.ng + g.nci, jend = g.ng + g.ncj;
  t.i = t.i0 + t.l;
  t.jrow = t.j0 + t.w;
  t.cellok = (t.w < TYV) && (t.jrow < jend) && (t.i < iend);
  t.yfok = (t.i < iend) && (t.j0 + t.w - 1 < jend);
  t.xbok = (t.w == TYV) && (t.l < TYV) && (t.j0 + t.l < jend);
  return t;
}

// ---------------------------------------------------------------- arithmetic forms

// FP64 form: the reference arithmetic (divisions by run constants through
// cdiv, all others IEEE).  FP32 form (precision form): stencil, viscous and
// C-equation arithmetic in binary32 with hardware reciprocals; loads, flux
// assembly, pressure gradient, RK accumulation and EOS stay FP64.
template <bool F32>
struct Ar;
#ifdef FST_FAST64_TU
// a/b through the hardware reciprocal: r = rcp(b) refined by two
// Newton-Raphson steps r <- r + r (1 - b r), then q = a r with one residual
// correction q <- q + r (a - b q).  Within about an ulp of the IEEE quotient
// for the normal, finite operands of the face terms.
__device__ __forceinline__ double fast_div(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(e, r, r);
  e = fma(-b, r, 1.0);
  r = fma(e, r, r);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
#endif

template <>
struct Ar<false> {
  using R = double;
  static __device__ __forceinline__ R cdv(R a, double b, double rb) { return cdiv(a, b, rb); }
#ifdef FST_FAST64_TU
  static __device__ __forceinline__ R div(R a, R b) { return fast_div(a, b); }
#else
  static __device__ __forceinline__ R div(R a, R b) { return a / b; }
#endif
  static __device__ __forceinline__ R sq(R a) { return sqrt(a); }
};
template <>
struct Ar<true> {
  using R = float;
  static __device__ __forceinline__ R cdv(R a, double, double rb) { return a * (float)rb; }
  static __device__ __forceinline__ R div(R a, R b) { return a * __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ R sq(R a) { return __builtin_sqrtf(a); }
};

// WENO5 of the precision form: same stencils, weights and smoothness
// indicators as calculateFluxesG (src/flux.hpp:240-264) in binary32, with
// hardware reciprocals.  Callers pass deviations from the face-centre value
// (WENO5 is translation covariant), which keeps FP32 cancellation small.
__device__ __forceinline__ float weno5_fast(float f1, float f2, float f3, float f4, float f5) {
#pragma clang fp contract(fast)
  const float eps = 1.0e-6f;
  float a1 = f1 - 2.0f * f2 + f3, a2 = f1 - 4.0f * f2 + 3.0f * f3;
  const float b1 = (13.0f / 12.0f) * a1 * a1 + 0.25f * a2 * a2;
  a1 = f2 - 2.0f * f3 + f4;
  a2 = f2 - f4;
  const float b2 = (13.0f / 12.0f) * a1 * a1 + 0.25f * a2 * a2;
  a1 = f3 - 2.0f * f4 + f5;
  a2 = 3.0f * f3 - 4.0f * f4 + f5;
  const float b3 = (13.0f / 12.0f) * a1 * a1 + 0.25f * a2 * a2;
  const float w1 = 0.1f * __builtin_amdgcn_rcpf((eps + b1) * (eps + b1));
  const float w2 = 0.6f * __builtin_amdgcn_rcpf((eps + b2) * (eps + b2));
  const float w3 = 0.3f * __builtin_amdgcn_rcpf((eps + b3) * (eps + b3));
  const float p1 = (1.0f / 3.0f) * f1 + (-7.0f / 6.0f) * f2 + (11.0f / 6.0f) * f3;
  const float p2 = (-1.0f / 6.0f) * f2 + (5.0f / 6.0f) * f3 + (1.0f / 3.0f) * f4;
  const float p3 = (1.0f / 3.0f) * f3 + (5.0f / 6.0f) * f4 + (-1.0f / 6.0f) * f5;
  return (w1 * p1 + w2 * p2 + w3 * p3) * __builtin_amdgcn_rcpf(w1 + w2 + w3);
}

#ifdef FST_FAST64_TU
// WENO5 of calculateFluxesG (src/flux.hpp:240-264) with the reference's
// binary32-rounded constants, normalised with one division: for
// d_i = (eps + b_i)^2 and weights w_i = g_i / d_i,
//   (w1 p1 + w2 p2 + w3 p3) / (w1 + w2 + w3)
//     = (g1 p1 d2 d3 + g2 p2 d1 d3 + g3 p3 d1 d2) / (g1 d2 d3 + g2 d1 d3 + g3 d1 d2)
// (numerator and denominator times d1 d2 d3).  d_i >= 1e-12 and the flux
// variables stay far below 1e70, so the products neither underflow nor
// overflow.
__device__ __forceinline__ double weno5_onediv(double f1, double f2, double f3, double f4, double f5) {
  FST_FAST_FP
  const double eps = 0.000001;
  const double c13 = 13.0f / 12.0f, q = 0.25f;
  double a1 = f1 - 2.0 * f2 + f3, a2 = f1 - 4.0 * f2 + 3.0 * f3;
  const double b1 = c13 * a1 * a1 + q * a2 * a2;
  a1 = f2 - 2.0 * f3 + f4;
  a2 = f2 - f4;
  const double b2 = c13 * a1 * a1 + q * a2 * a2;
  a1 = f3 - 2.0 * f4 + f5;
  a2 = 3.0 * f3 - 4.0 * f4 + f5;
  const double b3 = c13 * a1 * a1 + q * a2 * a2;
  const double e1 = eps + b1, e2 = eps + b2, e3 = eps + b3;
  const double d1 = e1 * e1, d2 = e2 * e2, d3 = e3 * e3;
  const double d23 = d2 * d3, d13 = d1 * d3, d12 = d1 * d2;
  const double g1 = 0.1f, g2 = 0.6f, g3 = 0.3f;
  const double p1 = (1.0f / 3.0f) * f1 + (-7.0f / 6.0f) * f2 + (11.0f / 6.0f) * f3;
  const double p2 = (-1.0f / 6.0f) * f2 + (5.0f / 6.0f) * f3 + (1.0f / 3.0f) * f4;
  const double p3 = (1.0f / 3.0f) * f3 + (5.0f / 6.0f) * f4 + (-1.0f / 6.0f) * f5;
  const double W1 = g1 * d23, W2 = g2 * d13, W3 = g3 * d12;
  return fast_div(W1 * p1 + W2 * p2 + W3 * p3, W1 + W2 + W3);
}
#endif

// ---------------------------------------------------------------- faces

// Advective fluxes of the nv flow variables on one face (calculateFluxesG,
// src/flux.hpp:266-309) from its stencil: q[t][v] = var(v) + (v==3)*p at the
// six cells c-2e .. c+3e, vel_D at c-e .. c+2e.  Fluxes are FP64 in both forms.
template <int NS, bool F32>
__device__ __forceinline__ void flux_from_stencil(const FC &k, int D, const double *vd, const double (*q)[4 + NS],
                                                  double *out) {
  const double ud = cdiv(-vd[3] + 7.0f * vd[2] + 7.0f * vd[1] - vd[0], 12.0, k.r12);  // :278
  const bool neg = ud < 0.0;
#pragma unroll
  for (int v = 0; v < 4 + NS; ++v) {
    const double f1 = neg ? q[5][v] : q[0][v];
    const double f2 = neg ? q[4][v] : q[1][v];
    const double f3 = neg ? q[3][v] : q[2][v];
    const double f4 = neg ? q[2][v] : q[3][v];
    const double f5 = neg ? q[1][v] : q[4][v];
    if (!F32) {
#ifdef FST_FAST64_TU
      out[v] = (ud * weno5_onediv(f1, f2, f3, f4, f5)) * k.rd[D];
#else
      out[v] = cdiv(ud * weno5_f32c(f1, f2, f3, f4, f5), k.d[D], k.rd[D]);
#endif
    } else {
      const double c0 = q[2][v];
      const float w = weno5_fast((float)(f1 - c0), (float)(f2 - c0), (float)(f3 - c0), (float)(f4 - c0),
                                 (float)(f5 - c0));
      out[v] = cdiv(ud * (c0 + (double)w), k.d[D], k.rd[D]);
    }
  }
}

// q of one cell: var(v) + (v==3)*p for v < nv
template <int NS, bool F32>
__device__ __forceinline__ void load_q(const Grid32 &g, const double *__restrict__ U, const double *__restrict__ P,
                                       uint32_t cc, double *q) {
  const uint32_t N = g.N;
#pragma unroll
  for (int v = 0; v < 4 + NS; ++v) q[v] = (v == 3) ? ld(U, cc + N * v) + pget<F32>(P, N, 1, cc) : ld(U, cc + N * v);
}

// Advective fluxes on the face between c and c+e_D, stencil read from memory.
template <int D, int NS, bool F32>
__device__ __forceinline__ void face_adv(const FC &k, const Grid32 &g, const double *__restrict__ U,
                                         const double *__restrict__ P, uint32_t c, double *out) {
  const uint32_t N = g.N, e = stride<D>(g);
  const double vd[4] = {pget<F32>(P, N, 2 + D, c - e), pget<F32>(P, N, 2 + D, c), pget<F32>(P, N, 2 + D, c + e),
                        pget<F32>(P, N, 2 + D, c + 2 * e)};
  double q[6][4 + NS];
#pragma unroll
  for (int t = 0; t < 6; ++t) load_q<NS, F32>(g, U, P, c + (uint32_t)(t - 2) * e, q[t]);
  flux_from_stencil<NS, F32>(k, D, vd, q, out);
}

// Viscous stresses and C-equation fluxes on the face between c and c+e_D.
// out[0..2]: the stress components the divergence in direction D consumes
// (calculateStressTensor{x,y,z}3dv, src/viscosity.hpp:416-624);
// out[3+3*dir+w]: mFlux(D, dir, w) (calculateCeqFaces src/ceq3d.hpp:246-295
// times calculateCeqGrads src/ceq3d.hpp:306-331).  Both use the same nine
// face derivatives of the velocity.  LC (learned C-indicator): C_hat and the
// direction n come from the FP32 planes CH and DR instead of var(nv+1..nv+4).
template <int D, int NS, bool F32, bool LC>
__device__ __forceinline__ void face_vc(const FC &k, const Grid32 &g, const double *__restrict__ U,
                                        const double *__restrict__ P, const float *__restrict__ CH,
                                        const float *__restrict__ DR, double alpha, int visc, int ceq,
                                        uint32_t c, typename Ar<F32>::R *out) {
  FST_FAST_FP
  using A = Ar<F32>;
  using R = typename A::R;
  const uint32_t N = g.N, e = stride<D>(g);
  constexpr int T1 = D == 0 ? 1 : 0;
  constexpr int T2 = D == 2 ? 1 : 2;
  const uint32_t t1 = T1 == 1 ? g.s1 : 1u;
  const uint32_t t2 = T2 == 1 ? g.s1 : g.s2;
  auto V = [&](int q, uint32_t cell) { return (R)pget<F32>(P, N, q, cell); };
  R der[3][3];
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    const int vq = 2 + w;
    der[w][D] = A::cdv(V(vq, c + e) - V(vq, c), k.d[D], k.rd[D]);
    der[w][T1] = A::cdv((V(vq, c + t1) + V(vq, c + e + t1)) - (V(vq, c - t1) + V(vq, c + e - t1)), k.d4[T1],
                        k.rd4[T1]);
    der[w][T2] = A::cdv((V(vq, c + t2) + V(vq, c + e + t2)) - (V(vq, c - t2) + V(vq, c + e - t2)), k.d4[T2],
                        k.rd4[T2]);
  }
  if (visc) {
    R mu = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) mu += ((V(6 + s, c + e) + V(6 + s, c)) / (R)2.0) * (R)k.mus[s];
    const R dudx = der[0][0], dudy = der[0][1], dudz = der[0][2];
    const R dvdx = der[1][0], dvdy = der[1][1], dvdz = der[1][2];
    const R dwdx = der[2][0], dwdy = der[2][1], dwdz = der[2][2];
    const R c23 = (R)(2.0 / 3.0), two = (R)2.0;
    if (D == 0) {
      out[0] = c23 * mu * (two * dudx - dvdy - dwdz);
      out[1] = mu * (dudy + dvdx);
      out[2] = mu * (dudz + dwdx);
    } else if (D == 1) {
      out[0] = mu * (dudy + dvdx);
      out[1] = c23 * mu * (two * dvdy - dudx - dwdz);
      out[2] = mu * (dvdz + dwdy);
    } else {
      out[0] = mu * (dudz + dwdx);
      out[1] = mu * (dvdz + dwdy);
      out[2] = c23 * mu * (two * dwdz - dudx - dvdy);
    }
  }
  if (ceq) {
    const int nv = g.nv;
    auto Uv = [&](uint32_t off) { return (R)ld(U, off); };
    const R r = (V(0, c) + V(0, c + e)) / (R)2.0;
    R chat, cn[3];
    if (LC) {
      chat = ((R)CH[c] + (R)CH[c + e]) / (R)2.0;
#pragma unroll
      for (int q = 0; q < 3; ++q) cn[q] = ((R)DR[c + N * q] + (R)DR[c + e + N * q]) / (R)2.0;
    } else {
      chat = (Uv(c + N * (nv + 1)) + Uv(c + e + N * (nv + 1))) / (R)2.0;
      cn[0] = (Uv(c + N * (nv + 2)) + Uv(c + e + N * (nv + 2))) / (R)2.0;
      cn[1] = (Uv(c + N * (nv + 3)) + Uv(c + e + N * (nv + 3))) / (R)2.0;
      cn[2] = (Uv(c + N * (nv + 4)) + Uv(c + e + N * (nv + 4))) / (R)2.0;
    }
    R cmag = (R)1.0e-6;
    cmag += cn[0] * cn[0];
    cmag += cn[1] * cn[1];
    cmag += cn[2] * cn[2];
    const R arc = (R)alpha * r * chat;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
      const R dirac = (dir == D) ? (R)1.0 : (R)0.0;
      const R M = dirac - A::div(cn[D] * cn[dir], cmag);
      const R X = arc * M;
#pragma unroll
      for (int w = 0; w < 3; ++w) out[3 + 3 * dir + w] = X * der[w][dir];
    }
  }
}

// Primitives of one cell's state (calculateRhoPT3D + computeVelocity3D,
// src/secondary.hpp:109-151, src/velocity.hpp:68-83; nrg as
// src/ceq3d.hpp:138-141) and its max-wave-speed term (maxWaveSpeed,
// src/ceq3d.hpp:25-63).
template <int NS>
__device__ __forceinline__ void prims_of(const FC &k, const double *u, double *pr, double &ms) {
  double rho = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) rho = rho + u[4 + s];
  double Cp = 0, Cv = 0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double Y = u[4 + s] / rho;
    pr[6 + s] = Y;
    Cp = Cp + Y * k.cpf[s];
    Cv = Cv + Y * k.cvf[s];
  }
  const double gamma = Cp / Cv;
  const double p = (gamma - 1) * (u[3] - (0.5 / rho) * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
  const double vx = u[0] / rho, vy = u[1] / rho, vz = u[2] / rho;
  pr[0] = rho;
  pr[1] = p;
  pr[2] = vx;
  pr[3] = vy;
  pr[4] = vz;
  pr[5] = u[3] / rho - 0.5 * rho * (vx * vx + vy * vy + vz * vz);
  const double aa = sqrt(gamma * p / rho);
  ms = aa + sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / rho;
}

// ---------------------------------------------------------------- kernel 1

// advect3D (src/advect.hpp:52-55), applyPressureGradient3D (dz := cd(2),
// src/presgrad.hpp:53-67), computeBuoyancy3D (src/buoyancy.hpp:33-46):
// dvar[v] for v < nv.
template <int NS, bool F32>
__global__ void __launch_bounds__(NTHREADS, 1)
    k_adv3d(const Grid32 g, const FC k, const double *__restrict__ U, const double *__restrict__ P,
            double *__restrict__ dvar, int buoy, int kchunk, int nbx, int nby, int nbz) {
  constexpr int NV = 4 + NS;
  __shared__ double xsh[TY][NV][TX + 1];  // index 0: left of the tile; l+1: right of lane l
  __shared__ double ysh[NWAVE][NV][TX];   // row w: face below tile row w
  const Tile t = place(g, kchunk, nbx, nby, nbz);
  const int w = t.w, l = t.l;
  const uint32_t N = g.N;
  // z-stencil of this column carried in registers: qz[t] holds q at plane
  // kk-2+t for the face between kk and kk+1, vz[t] the z-velocity at kk-1+t.
  double qz[6][NV], vz[4];
  double Z[NV], Zp[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) Z[q] = Zp[q] = 0.0;
  const uint32_t c00 = g.id(t.i, t.jrow, t.kb - 1);
  if (t.cellok) {
#pragma unroll
    for (int s = 0; s < 6; ++s) load_q<NS, F32>(g, U, P, c00 + (uint32_t)(s - 2) * g.s2, qz[s]);
#pragma unroll
    for (int s = 0; s < 4; ++s) vz[s] = pget<F32>(P, N, 4, c00 + (uint32_t)(s - 1) * g.s2);
    flux_from_stencil<NS, F32>(k, 2, vz, qz, Zp);
  }
  for (int kk = t.kb; kk < t.ke; ++kk) {
    uint32_t c = g.id(t.i, t.jrow, kk);
    uint32_t cxb = g.id(t.i0 - 1, t.j0 + l, kk);
    asm volatile("" : "+v"(c), "+v"(cxb));  // keep address arithmetic per plane
    if (t.cellok) {
      double X[NV];
      face_adv<0, NS, F32>(k, g, U, P, c, X);
#pragma unroll
      for (int q = 0; q < NV; ++q) xsh[w][q][l + 1] = X[q];
    }
    if (t.xbok) {
      double X[NV];
      face_adv<0, NS, F32>(k, g, U, P, cxb, X);
#pragma unroll
      for (int q = 0; q < NV; ++q) xsh[l][q][0] = X[q];
    }
    if (t.yfok) {
      double Y[NV];
      face_adv<1, NS, F32>(k, g, U, P, c - g.s1, Y);
#pragma unroll
      for (int q = 0; q < NV; ++q) ysh[w][q][l] = Y[q];
    }
    if (t.cellok) {
#pragma unroll
      for (int s = 0; s < 5; ++s)
#pragma unroll
        for (int v = 0; v < NV; ++v) qz[s][v] = qz[s + 1][v];
      load_q<NS, F32>(g, U, P, c + 3 * g.s2, qz[5]);
#pragma unroll
      for (int s = 0; s < 3; ++s) vz[s] = vz[s + 1];
      vz[3] = pget<F32>(P, N, 4, c + 2 * g.s2);
      flux_from_stencil<NS, F32>(k, 2, vz, qz, Z);
    }
    __syncthreads();
    if (t.cellok) {
      double dv[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v)
        dv[v] = -((xsh[w][v][l + 1] - xsh[w][v][l]) + (ysh[w + 1][v][l] - ysh[w][v][l]) + (Z[v] - Zp[v]));
      const double *Pp = F32 ? P : P + N;  // pressure plane (FP64 in every form)
      const uint32_t sx = 1, sy = g.s1, sz = g.s2;
      const double dxp = cdiv(ld(Pp, c - 2 * sx) - 8.0 * ld(Pp, c - sx) + 8.0 * ld(Pp, c + sx) - ld(Pp, c + 2 * sx),
                              k.pg[0], k.rpg[0]);
      const double dyp = cdiv(ld(Pp, c - 2 * sy) - 8.0 * ld(Pp, c - sy) + 8.0 * ld(Pp, c + sy) - ld(Pp, c + 2 * sy),
                              k.pg[1], k.rpg[1]);
      const double dzp = cdiv(ld(Pp, c - 2 * sz) - 8.0 * ld(Pp, c - sz) + 8.0 * ld(Pp, c + sz) - ld(Pp, c + 2 * sz),
                              k.pg[2], k.rpg[2]);
      dv[0] = dv[0] - dxp;
      dv[1] = dv[1] - dyp;
      dv[2] = dv[2] - dzp;
      if (buoy) {
        const double rho = pget<F32>(P, N, 0, c);
        const double vv = ld(U, c + N) / rho;
        const double rhop = rho - k.rhoRef;
        const double f = -k.g * rhop;
        dv[1] += f;
        dv[3] += vv * f;
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) st_dv<F32>(dvar, c + N * v, dv[v]);
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) Zp[q] = Z[q];
    __syncthreads();
  }
}

// ---------------------------------------------------------------- kernel 1 with an LDS plane

// k_adv3d with the current plane staged in LDS: for rows j0-3..j0+TY+2 and
// columns i0-3..i0+TX+2 of the tile, q (var(v), plus p for the energy), u, v
// and p as doubles.  The x and y faces and the pressure gradient read it; the
// z stencil stays in the per-column register queue.  The next plane is loaded
// into registers during the current one.  Same values, same operations as
// k_adv3d.
constexpr int QY = TY + 6, QX = TX + 6, QPL = QY * QX;

template <int NS>
struct PlaneF {
  static constexpr int NV = 4 + NS;
  static constexpr int NQ = NV + 3;  // q_0..q_{nv-1}, u, v, p
  static constexpr int FU = NV, FV = NV + 1, FP = NV + 2;
  static constexpr int NPC = (QPL + NTHREADS - 1) / NTHREADS;  // plane cells per thread
  static constexpr int NPRE = NQ * NPC;
};

// Loads of plane p for the LDS plane, field-major: thread t takes the plane
// cells t, t + NTHREADS, ... of every field, so each load's field (and its
// array and precision) is a compile-time choice.
template <int NS, bool F32>
__device__ __forceinline__ void plane_fetch(const Grid32 &g, const double *__restrict__ U,
                                            const double *__restrict__ P, int i0, int j0, int p, double *pre) {
  using PF = PlaneF<NS>;
  const uint32_t N = g.N;
#pragma unroll
  for (int m = 0; m < PF::NPC; ++m) {
    const int e = threadIdx.x + m * NTHREADS;
    const int r = e / QX, q = e - r * QX;
    const int i = i0 - 3 + q, j = j0 - 3 + r;
    const bool ok = e < QPL && i < g.ngi && j < g.ngj;
    const uint32_t c = ok ? g.id(i, j, p) : 0u;
#pragma unroll
    for (int f = 0; f < PF::NQ; ++f) {
      double v = 0.0;
      if (ok) {
        if (f < PF::NV) v = (f == 3) ? ld(U, c + N * f) + pget<F32>(P, N, 1, c) : ld(U, c + N * f);
        else if (f == PF::FU) v = pget<F32>(P, N, 2, c);
        else if (f == PF::FV) v = pget<F32>(P, N, 3, c);
        else v = pget<F32>(P, N, 1, c);
      }
      pre[f * PF::NPC + m] = v;
    }
  }
}

template <int NS>
__device__ __forceinline__ void plane_store(double *qs, const double *pre) {
  using PF = PlaneF<NS>;
#pragma unroll
  for (int m = 0; m < PF::NPC; ++m) {
    const int e = threadIdx.x + m * NTHREADS;
    if (e < QPL) {
#pragma unroll
      for (int f = 0; f < PF::NQ; ++f) qs[f * QPL + e] = pre[f * PF::NPC + m];
    }
  }
}

// Advective fluxes on the face between the cell at plane row r, column q and
// its +e_D neighbour (D = 0: x, 1: y), stencil read from the LDS plane.  The
// upwind stencil is chosen by address (a0 + t dt, t = 0..4, one choice for all
// variables): the same values as flux_from_stencil's selection, so the same
// bits.
template <int D, int NS, bool F32>
__device__ __forceinline__ void face_adv_lds(const FC &k, const double *qs, int r, int q, double *out) {
  using PF = PlaneF<NS>;
  constexpr int NV = PF::NV;
  constexpr int st = D == 1 ? QX : 1;  // offset of +e_D in the plane
  const int base = r * QX + q;
  constexpr int fv = D == 0 ? PF::FU : PF::FV;
  const double *vp = qs + fv * QPL + base;
  const double ud = cdiv(-vp[2 * st] + 7.0f * vp[st] + 7.0f * vp[0] - vp[-st], 12.0, k.r12);  // src/flux.hpp:278
  const bool neg = ud < 0.0;
  const int a0 = neg ? base + 3 * st : base - 2 * st;
  const int dt = neg ? -st : st;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double *qv = qs + v * QPL + a0;
#ifdef FST_FAST64_TU
    out[v] = (ud * weno5_onediv(qv[0], qv[dt], qv[2 * dt], qv[3 * dt], qv[4 * dt])) * k.rd[D];
#else
    out[v] = cdiv(ud * weno5_f32c(qv[0], qv[dt], qv[2 * dt], qv[3 * dt], qv[4 * dt]), k.d[D], k.rd[D]);
#endif
  }
}

// PG = false (the generalized-grid module): the fluxes, their divergence and
// applyGenPressureGradient3D (src/presgrad.hpp:72-147) on the metric terms MT
// (9 planes, m(i,j,k,a,n) at MT + N (a + 3 n)); gen3d has no buoyancy of its
// own (buoy there: the gen_ext option's, after the pressure gradient as
// k_buoy_gen orders it).
template <int NS, bool F32, bool PG>
__device__ __forceinline__ void adv_tile(const Grid32 &g, const FC &k, const double *__restrict__ U,
                                         const double *__restrict__ P, double *__restrict__ dvar, int buoy,
                                         int kchunk, int nbx, int nby, int nbz, const double *__restrict__ MT) {
  using PF = PlaneF<NS>;
  constexpr int NV = 4 + NS;
  __shared__ double qs[PF::NQ * QPL];
  __shared__ double xsh[TY][NV][TX + 1];  // index 0: left of the tile; l+1: right of lane l
  __shared__ double ysh[NWAVE][NV][TX];   // row w: face below tile row w
  const Tile t = place(g, kchunk, nbx, nby, nbz);
  const int w = t.w, l = t.l;
  const uint32_t N = g.N;
  double qz[6][NV], vz[4];
  double Z[NV], Zp[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) Z[q] = Zp[q] = 0.0;
  const uint32_t c00 = g.id(t.i, t.jrow, t.kb - 1);
  if (t.cellok) {
#pragma unroll
    for (int s = 0; s < 6; ++s) load_q<NS, F32>(g, U, P, c00 + (uint32_t)(s - 2) * g.s2, qz[s]);
#pragma unroll
    for (int s = 0; s < 4; ++s) vz[s] = pget<F32>(P, N, 4, c00 + (uint32_t)(s - 1) * g.s2);
    flux_from_stencil<NS, F32>(k, 2, vz, qz, Zp);
  }
  // prefetch of the next plane when the registers allow it (ns <= 2)
  constexpr bool PREF = NS <= 2;
  double pre[PF::NPRE];
  if (PREF) plane_fetch<NS, F32>(g, U, P, t.i0, t.j0, t.kb, pre);
  for (int kk = t.kb; kk < t.ke; ++kk) {
    uint32_t c = g.id(t.i, t.jrow, kk);
    asm volatile("" : "+v"(c));
    if (!PREF) plane_fetch<NS, F32>(g, U, P, t.i0, t.j0, kk, pre);
    plane_store<NS>(qs, pre);
    __syncthreads();
    if (PREF && kk + 1 < t.ke) plane_fetch<NS, F32>(g, U, P, t.i0, t.j0, kk + 1, pre);
    const int rc = w + 3, qc = l + 3;  // plane coordinates of this lane's cell
    if (t.cellok) {  // x-face right of the cell and z-face above it, one region
      double X[NV];
      face_adv_lds<0, NS, F32>(k, qs, rc, qc, X);
#pragma unroll
      for (int q = 0; q < NV; ++q) xsh[w][q][l + 1] = X[q];
#pragma unroll
      for (int s = 0; s < 5; ++s)
#pragma unroll
        for (int v = 0; v < NV; ++v) qz[s][v] = qz[s + 1][v];
      load_q<NS, F32>(g, U, P, c + 3 * g.s2, qz[5]);
#pragma unroll
      for (int s = 0; s < 3; ++s) vz[s] = vz[s + 1];
      vz[3] = pget<F32>(P, N, 4, c + 2 * g.s2);
      flux_from_stencil<NS, F32>(k, 2, vz, qz, Z);
    }
    if (t.xbok) {  // x-face left of the tile, row l: cell (i0-1, j0+l)
      double X[NV];
      face_adv_lds<0, NS, F32>(k, qs, l + 3, 2, X);
#pragma unroll
      for (int q = 0; q < NV; ++q) xsh[l][q][0] = X[q];
    }
    if (t.yfok) {  // y-face below row w: cell (i0+l, j0+w-1)
      double Y[NV];
      face_adv_lds<1, NS, F32>(k, qs, w + 2, qc, Y);
#pragma unroll
      for (int q = 0; q < NV; ++q) ysh[w][q][l] = Y[q];
    }
    __syncthreads();
    if (t.cellok) {
      const double *pp = qs + PF::FP * QPL + rc * QX + qc;
      const double *Pp = F32 ? P : P + N;  // pressure plane (FP64 in every form)
      const uint32_t sz = g.s2;
      // the pressure gradient before the flux divergence holds its registers
      // (subtracted from dv[0..2] below, as the reference adds it after the
      // fluxes); on generalized grids one n at a time: unrolled, its metric
      // loads were hoisted (256 VGPRs with spills, 6.6 % slower)
      double gp[3];
      if constexpr (PG) {
        gp[0] = cdiv(pp[-2] - 8.0 * pp[-1] + 8.0 * pp[1] - pp[2], k.pg[0], k.rpg[0]);
        gp[1] = cdiv(pp[-2 * QX] - 8.0 * pp[-QX] + 8.0 * pp[QX] - pp[2 * QX], k.pg[1], k.rpg[1]);
        gp[2] = cdiv(ld(Pp, c - 2 * sz) - 8.0 * ld(Pp, c - sz) + 8.0 * ld(Pp, c + sz) - ld(Pp, c + 2 * sz),
                     k.pg[2], k.rpg[2]);
      } else {
        // metric pressure gradient: 4th-order differences of p m(.,a,n) along
        // xi, eta, zeta, divided by 12 (IEEE: cdiv with RN(1/12))
        const uint32_t sy = g.s1;
#pragma unroll 1
        for (int n = 0; n < 3; ++n) {
          const double *m0 = MT + (size_t)N * (size_t)(0 + 3 * n);
          const double *m1 = MT + (size_t)N * (size_t)(1 + 3 * n);
          const double *m2 = MT + (size_t)N * (size_t)(2 + 3 * n);
          const double dxi = cdiv(pp[-2] * ld(m0, c - 2) - 8.0 * pp[-1] * ld(m0, c - 1) + 8.0 * pp[1] * ld(m0, c + 1) -
                                      pp[2] * ld(m0, c + 2),
                                  12.0, k.r12);
          const double det = cdiv(pp[-2 * QX] * ld(m1, c - 2 * sy) - 8.0 * pp[-QX] * ld(m1, c - sy) +
                                      8.0 * pp[QX] * ld(m1, c + sy) - pp[2 * QX] * ld(m1, c + 2 * sy),
                                  12.0, k.r12);
          const double dzt = cdiv(ld(Pp, c - 2 * sz) * ld(m2, c - 2 * sz) - 8.0 * ld(Pp, c - sz) * ld(m2, c - sz) +
                                      8.0 * ld(Pp, c + sz) * ld(m2, c + sz) - ld(Pp, c + 2 * sz) * ld(m2, c + 2 * sz),
                                  12.0, k.r12);
          const double t3 = dxi + det + dzt;
          if (n == 0) gp[0] = t3;
          else if (n == 1) gp[1] = t3;
          else gp[2] = t3;
        }
      }
      double dv[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v)
        dv[v] = -((xsh[w][v][l + 1] - xsh[w][v][l]) + (ysh[w + 1][v][l] - ysh[w][v][l]) + (Z[v] - Zp[v]));
      dv[0] = dv[0] - gp[0];
      dv[1] = dv[1] - gp[1];
      dv[2] = dv[2] - gp[2];
      if (buoy) {
        const double rho = pget<F32>(P, N, 0, c);
        const double vv = qs[1 * QPL + rc * QX + qc] / rho;
        const double rhop = rho - k.rhoRef;
        const double f = -k.g * rhop;
        dv[1] += f;
        dv[3] += vv * f;
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) st_dv<F32>(dvar, c + N * v, dv[v]);
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) Zp[q] = Z[q];
    __syncthreads();
  }
}

template <int NS, bool F32, bool PG = true>
__global__ void __launch_bounds__(NTHREADS, 1)
    k_adv3d_lds(const Grid32 g, const FC k, const double *__restrict__ U, const double *__restrict__ P,
                double *__restrict__ dvar, int buoy, int kchunk, int nbx, int nby, int nbz) {
  adv_tile<NS, F32, PG>(g, k, U, P, dvar, buoy, kchunk, nbx, nby, nbz, nullptr);
}

// gen3d_func::compute's fluxes, their divergence and the metric pressure
// gradient (src/gen3d.cpp:352-366) in one launch
// (buoy: the gen_ext buoyancy in the same launch, for the fused gen3d steps)
template <int NS>
__global__ void __launch_bounds__(NTHREADS, 1)
    k_adv3d_gen(const Grid32 g, const FC k, const double *__restrict__ U, const double *__restrict__ P,
                double *__restrict__ dvar, const double *__restrict__ MT, int buoy, int kchunk, int nbx, int nby,
                int nbz) {
  adv_tile<NS, false, false>(g, k, U, P, dvar, buoy, kchunk, nbx, nby, nbz, MT);
}

// ---------------------------------------------------------------- kernel 1, FP32 forms, float plane

// Advection kernel of the FP32 forms with a binary32 LDS plane: q (var(v),
// plus p for the energy), u and v staged as floats for rows j0-3..j0+TYS+2 and
// columns i0-3..i0+TX+2, the z stencil as floats in a per-column register
// queue, WENO5 on deviations from the face-centre value, fluxes in binary32.
// The pressure gradient and buoyancy, whose cancellation is the hydrostatic
// balance, stay FP64 (p staged as doubles) and are added to the FP32 flux
// divergence in FP64; the partial RHS is stored as FP32.
constexpr int TYS = 11, NWS = TYS + 1, NTS = 64 * NWS;  // tile rows, waves and threads of the float-plane kernel
constexpr int QYS = TYS + 6, QPLS = QYS * QX;            // its staged plane: rows, cells

template <int NS>
struct PlaneS {
  static constexpr int NV = 4 + NS;
  static constexpr int NQ = NV + 2;  // q_0..q_{nv-1}, u, v
  static constexpr int FU = NV, FV = NV + 1;
  static constexpr int NPC = (QPLS + NTS - 1) / NTS;
};

template <int NS, bool FI>
__device__ __forceinline__ void plane_fetch_s(const Grid32 &g, const double *__restrict__ U,
                                              const double *__restrict__ P, int i0, int j0, int p, float *pre,
                                              double *prep) {
  using PS = PlaneS<NS>;
  const uint32_t N = g.N;
#pragma unroll
  for (int m = 0; m < PS::NPC; ++m) {
    const int e = threadIdx.x + m * NTS;
    const int r = e / QX, q = e - r * QX;
    const int i = i0 - 3 + q, j = j0 - 3 + r;
    const bool ok = e < QPLS && i < g.ngi && j < g.ngj;
    const uint32_t c = ok ? g.id(i, j, p) : 0u;
    const double pc = ok ? pget<true>(P, N, 1, c) : 0.0;
    prep[m] = pc;
#pragma unroll
    for (int f = 0; f < PS::NQ; ++f) {
      float v = 0.f;
      if (ok) {
        if (f < PS::NV) v = (float)((f == 3) ? uld<FI>(U, c + N * f) + pc : uld<FI>(U, c + N * f));
        else v = (float)pget<true>(P, N, f == PS::FU ? 2 : 3, c);
      }
      pre[f * PS::NPC + m] = v;
    }
  }
}

// WENO5 of the binary32 advection kernel on deviations from the upwind-centre
// value f3 (so f3 = 0 drops out of every smoothness indicator and candidate):
// the stencils, weights and indicators of calculateFluxesG
// (src/flux.hpp:240-264), with the linear weights (0.1, 0.6, 0.3) folded into
// the candidate polynomials and hardware reciprocals.
__device__ __forceinline__ float weno5_dev(float f1, float f2, float f4, float f5) {
#pragma clang fp contract(fast)
  const float eps = 1.0e-6f;
  float a1 = f1 - 2.0f * f2, a2 = f1 - 4.0f * f2;
  const float s1 = eps + ((13.0f / 12.0f) * a1 * a1 + 0.25f * a2 * a2);
  a1 = f2 + f4;
  a2 = f2 - f4;
  const float s2 = eps + ((13.0f / 12.0f) * a1 * a1 + 0.25f * a2 * a2);
  a1 = f5 - 2.0f * f4;
  a2 = f5 - 4.0f * f4;
  const float s3 = eps + ((13.0f / 12.0f) * a1 * a1 + 0.25f * a2 * a2);
  const float r1 = __builtin_amdgcn_rcpf(s1 * s1), r2 = __builtin_amdgcn_rcpf(s2 * s2),
              r3 = __builtin_amdgcn_rcpf(s3 * s3);
  const float p1 = (0.1f / 3.0f) * f1 - (0.7f / 6.0f) * f2;  // 0.1 (f1/3 - 7 f2/6 + 11 f3/6)
  const float p2 = 0.2f * f4 - 0.1f * f2;                    // 0.6 (-f2/6 + 5 f3/6 + f4/3)
  const float p3 = 0.25f * f4 - 0.05f * f5;                  // 0.3 (f3/3 + 5 f4/6 - f5/6)
  return (r1 * p1 + r2 * p2 + r3 * p3) * __builtin_amdgcn_rcpf(0.1f * r1 + 0.6f * r2 + 0.3f * r3);
}

// Fluxes of the nv flow variables on one face from its float stencil in
// registers (q[t][v] at cells c-2e .. c+3e, vd = vel_D at c-e .. c+2e),
// divided by d.
template <int NS>
__device__ __forceinline__ void flux_s(float rd, const float *vd, const float (*q)[4 + NS], float *out) {
  const float ud = (-vd[3] + 7.0f * vd[2] + 7.0f * vd[1] - vd[0]) * (1.0f / 12.0f);  // src/flux.hpp:278
  const bool neg = ud < 0.f;
  const float udr = ud * rd;
#pragma unroll
  for (int v = 0; v < 4 + NS; ++v) {
    const float c0 = neg ? q[3][v] : q[2][v];
    const float f1 = (neg ? q[5][v] : q[0][v]) - c0;
    const float f2 = (neg ? q[4][v] : q[1][v]) - c0;
    const float f4 = (neg ? q[2][v] : q[3][v]) - c0;
    const float f5 = (neg ? q[1][v] : q[4][v]) - c0;
    out[v] = udr * (c0 + weno5_dev(f1, f2, f4, f5));
  }
}

// The same on a face whose stencil lies in the LDS plane: the upwind stencil
// is chosen by address (a0 + t dt, t = 0..4), one choice for all variables.
template <int NS, int D>
__device__ __forceinline__ void face_s(const float *qs, float rd, int r, int q, float *out) {
  using PS = PlaneS<NS>;
  constexpr int NV = PS::NV;
  constexpr int st = D == 1 ? QX : 1;
  constexpr int fv = D == 0 ? PS::FU : PS::FV;
  const int base = r * QX + q;
  const float *vp = qs + fv * QPLS + base;
  const float ud = (-vp[2 * st] + 7.0f * vp[st] + 7.0f * vp[0] - vp[-st]) * (1.0f / 12.0f);  // src/flux.hpp:278
  const bool neg = ud < 0.f;
  const float udr = ud * rd;
  const int a0 = neg ? base + 3 * st : base - 2 * st;
  const int dt = neg ? -st : st;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const float *qv = qs + v * QPLS + a0;
    const float c0 = qv[2 * dt];
    const float f1 = qv[0] - c0, f2 = qv[dt] - c0, f4 = qv[3 * dt] - c0, f5 = qv[4 * dt] - c0;
    out[v] = udr * (c0 + weno5_dev(f1, f2, f4, f5));
  }
}

// FI: U is the float stage state of the learned FP32 form.  GEN: the
// generalized-grid module (P's float planes hold the contravariant velocity,
// MT the 9 metric planes): applyGenPressureGradient3D's metric pressure
// gradient (src/presgrad.hpp:72-147) in FP64 in place of the cartesian one.
// The partial RHS is stored as FP32 in both.
template <int NS, bool FI, bool GEN>
__device__ __forceinline__ void adv_s_tile(const Grid32 &g, const FC &k, const double *__restrict__ U,
                                           const double *__restrict__ P, double *__restrict__ dvar, int buoy,
                                           int kchunk, int nbx, int nby, int nbz, const double *__restrict__ MT) {
  using PS = PlaneS<NS>;
  constexpr int NV = 4 + NS;
  __shared__ float qs[PS::NQ * QPLS];
  __shared__ double ps[QPLS];
  __shared__ float xsh[TYS][NV][TX + 1];  // index 0: left of the tile; l+1: right of lane l
  __shared__ float ysh[NWS][NV][TX];   // row w: face below tile row w
  const Tile t = place<TYS>(g, kchunk, nbx, nby, nbz);
  const int w = t.w, l = t.l;
  const uint32_t N = g.N;
  const float rdx = (float)k.rd[0], rdy = (float)k.rd[1], rdz = (float)k.rd[2];
  float qz[6][NV], vz[4], Z[NV], Zp[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) Z[q] = Zp[q] = 0.f;
  auto load_qf = [&](uint32_t cc, float *q) {
    const double pc = pget<true>(P, N, 1, cc);
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = (float)((v == 3) ? uld<FI>(U, cc + N * v) + pc : uld<FI>(U, cc + N * v));
  };
  const uint32_t c00 = g.id(t.i, t.jrow, t.kb - 1);
  if (t.cellok) {
#pragma unroll
    for (int s = 0; s < 6; ++s) load_qf(c00 + (uint32_t)(s - 2) * g.s2, qz[s]);
#pragma unroll
    for (int s = 0; s < 4; ++s) vz[s] = (float)pget<true>(P, N, 4, c00 + (uint32_t)(s - 1) * g.s2);
    flux_s<NS>(rdz, vz, qz, Zp);
  }
  float pre[PS::NQ * PS::NPC];
  double prep[PS::NPC];
  plane_fetch_s<NS, FI>(g, U, P, t.i0, t.j0, t.kb, pre, prep);
  for (int kk = t.kb; kk < t.ke; ++kk) {
    uint32_t c = g.id(t.i, t.jrow, kk);
    asm volatile("" : "+v"(c));
#pragma unroll
    for (int m = 0; m < PS::NPC; ++m) {
      const int e = threadIdx.x + m * NTS;
      if (e < QPLS) {
#pragma unroll
        for (int f = 0; f < PS::NQ; ++f) qs[f * QPLS + e] = pre[f * PS::NPC + m];
        ps[e] = prep[m];
      }
    }
    __syncthreads();
    if (kk + 1 < t.ke) plane_fetch_s<NS, FI>(g, U, P, t.i0, t.j0, kk + 1, pre, prep);
    const int rc = w + 3, qc = l + 3;  // plane coordinates of this lane's cell
    if (t.cellok) {  // the three faces of a tile cell in one region
      float X[NV];
      face_s<NS, 0>(qs, rdx, rc, qc, X);
#pragma unroll
      for (int q = 0; q < NV; ++q) xsh[w][q][l + 1] = X[q];
#pragma unroll
      for (int s = 0; s < 5; ++s)
#pragma unroll
        for (int v = 0; v < NV; ++v) qz[s][v] = qz[s + 1][v];
      load_qf(c + 3 * g.s2, qz[5]);
#pragma unroll
      for (int s = 0; s < 3; ++s) vz[s] = vz[s + 1];
      vz[3] = (float)pget<true>(P, N, 4, c + 2 * g.s2);
      flux_s<NS>(rdz, vz, qz, Z);
      float Y[NV];  // y-face below row w (yfok holds for a tile cell)
      face_s<NS, 1>(qs, rdy, w + 2, qc, Y);
#pragma unroll
      for (int q = 0; q < NV; ++q) ysh[w][q][l] = Y[q];
    } else {
      if (t.xbok) {  // x-face left of the tile, row l: cell (i0-1, j0+l)
        float X[NV];
        face_s<NS, 0>(qs, rdx, l + 3, 2, X);
#pragma unroll
        for (int q = 0; q < NV; ++q) xsh[l][q][0] = X[q];
      }
      if (t.yfok) {  // y-face below row w: cell (i0+l, j0+w-1)
        float Y[NV];
        face_s<NS, 1>(qs, rdy, w + 2, qc, Y);
#pragma unroll
        for (int q = 0; q < NV; ++q) ysh[w][q][l] = Y[q];
      }
    }
    __syncthreads();
    if (t.cellok) {
      const double *pp = ps + rc * QX + qc;
      const uint32_t sz = g.s2;
      double gp[3];  // pressure gradient terms, subtracted from dv[0..2]
      if constexpr (!GEN) {
        gp[0] = cdiv(pp[-2] - 8.0 * pp[-1] + 8.0 * pp[1] - pp[2], k.pg[0], k.rpg[0]);
        gp[1] = cdiv(pp[-2 * QX] - 8.0 * pp[-QX] + 8.0 * pp[QX] - pp[2 * QX], k.pg[1], k.rpg[1]);
        gp[2] = cdiv(ld(P, c - 2 * sz) - 8.0 * ld(P, c - sz) + 8.0 * ld(P, c + sz) - ld(P, c + 2 * sz), k.pg[2],
                     k.rpg[2]);  // FP32 forms: P starts with the FP64 pressure plane
      } else {
        // the FP64 kernel's metric pressure gradient on the same FP64 p,
        // before the flux divergence holds its registers, one n at a time
        // (unrolled, the 36 metric loads were hoisted and spilled)
        const uint32_t sy = g.s1;
#pragma unroll 1
        for (int n = 0; n < 3; ++n) {
          const double *m0 = MT + (size_t)N * (size_t)(0 + 3 * n);
          const double *m1 = MT + (size_t)N * (size_t)(1 + 3 * n);
          const double *m2 = MT + (size_t)N * (size_t)(2 + 3 * n);
          const double dxi = cdiv(pp[-2] * ld(m0, c - 2) - 8.0 * pp[-1] * ld(m0, c - 1) + 8.0 * pp[1] * ld(m0, c + 1) -
                                      pp[2] * ld(m0, c + 2),
                                  12.0, k.r12);
          const double det = cdiv(pp[-2 * QX] * ld(m1, c - 2 * sy) - 8.0 * pp[-QX] * ld(m1, c - sy) +
                                      8.0 * pp[QX] * ld(m1, c + sy) - pp[2 * QX] * ld(m1, c + 2 * sy),
                                  12.0, k.r12);
          const double dzt = cdiv(ld(P, c - 2 * sz) * ld(m2, c - 2 * sz) - 8.0 * ld(P, c - sz) * ld(m2, c - sz) +
                                      8.0 * ld(P, c + sz) * ld(m2, c + sz) - ld(P, c + 2 * sz) * ld(m2, c + 2 * sz),
                                  12.0, k.r12);
          const double t3 = dxi + det + dzt;
          if (n == 0) gp[0] = t3;
          else if (n == 1) gp[1] = t3;
          else gp[2] = t3;
        }
      }
      double dv[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v)
        dv[v] = (double)(-((xsh[w][v][l + 1] - xsh[w][v][l]) + (ysh[w + 1][v][l] - ysh[w][v][l]) + (Z[v] - Zp[v])));
      dv[0] = dv[0] - gp[0];
      dv[1] = dv[1] - gp[1];
      dv[2] = dv[2] - gp[2];
      if (buoy) {
        const double rho = pget<true>(P, N, 0, c);
        const double vv = uld<FI>(U, c + N) / rho;
        const double rhop = rho - k.rhoRef;
        const double f = -k.g * rhop;
        dv[1] += f;
        dv[3] += vv * f;
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) st_dv<true>(dvar, c + N * v, dv[v]);
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) Zp[q] = Z[q];
    __syncthreads();
  }
}

template <int NS, bool FI>
__global__ void __launch_bounds__(NTS, 1)
    k_adv3d_s(const Grid32 g, const FC k, const double *__restrict__ U, const double *__restrict__ P,
              double *__restrict__ dvar, int buoy, int kchunk, int nbx, int nby, int nbz) {
  adv_s_tile<NS, FI, false>(g, k, U, P, dvar, buoy, kchunk, nbx, nby, nbz, nullptr);
}

// gen3d's fluxes in binary32 (FP32 form of the fused gen3d stage), their
// divergence, the FP64 metric pressure gradient and the gen_ext buoyancy
template <int NS>
__global__ void __launch_bounds__(NTS, 1)
    k_adv3d_sgen(const Grid32 g, const FC k, const double *__restrict__ U, const double *__restrict__ P,
                 double *__restrict__ dvar, const double *__restrict__ MT, int buoy, int kchunk, int nbx, int nby,
                 int nbz) {
  adv_s_tile<NS, false, true>(g, k, U, P, dvar, buoy, kchunk, nbx, nby, nbz, MT);
}

// ---------------------------------------------------------------- kernel 2

// applyViscousTerm{x,y,z}3dv (src/viscosity.hpp:472-497, 554-579, 637-662),
// calculateRhoGrad + updateCeq + applyCeq (src/ceq3d.hpp:143-186, 202-223,
// 343-365), the Runge-Kutta stage update (src/rk.cpp:116-119 / :137-139),
// then the primitives and the maxima of the new state.  dvar accumulates in
// FP64 in both forms.  LC: learned C-indicator (C_hat, n from CH, DR; no
// C-equation variables, nvt = nv; maxCh comes from the inference pass).
// FO: Uo is the float stage state of the learned FP32 form (stage 1).
template <int NS, bool F32, bool LC, bool FO = false, bool BL = false>
__global__ void __launch_bounds__(NTHREADS, 1)
    k_vcr3d(const Grid32 g, const FC k, const double *__restrict__ U, const double *__restrict__ Un,
            double *__restrict__ Uo, const double *__restrict__ P, double *__restrict__ Po,
            const double *__restrict__ dvar, const float *__restrict__ CH, const float *__restrict__ DR,
            const Scalars *scin, Scalars *scout, Upd up, int visc, int ceq, int kchunk, int nbx,
            int nby, int nbz) {
  using A = Ar<F32>;
  using R = typename A::R;
  constexpr int NV = 4 + NS;
  __shared__ R xsh[TY][NFV][TX + 1];
  __shared__ R ysh[NWAVE][NFV][TX];
  __shared__ double red[NWAVE];
  const Tile t = place(g, kchunk, nbx, nby, nbz);
  const int w = t.w, l = t.l;
  const uint32_t N = g.N;
  const double maxS = as_double(scin->maxS);
  const double alpha = (k.dxmag2 / (as_double(scin->maxCh) + 1.0e-6)) * k.alpha_cf;  // src/cart3d.cpp:285
  const bool faces = visc || ceq;
  R Z[NFV], Zp[NFV];
#pragma unroll
  for (int q = 0; q < NFV; ++q) Z[q] = Zp[q] = (R)0.0;
  if (faces && t.cellok) face_vc<2, NS, F32, LC>(k, g, U, P, CH, DR, alpha, visc, ceq, g.id(t.i, t.jrow, t.kb - 1), Zp);
  double mS = 0.0;
  for (int kk = t.kb; kk < t.ke; ++kk) {
    uint32_t c = g.id(t.i, t.jrow, kk);
    uint32_t cxb = g.id(t.i0 - 1, t.j0 + l, kk);
    asm volatile("" : "+v"(c), "+v"(cxb));
    if (faces) {
      if (t.cellok) {
        R X[NFV];
        face_vc<0, NS, F32, LC>(k, g, U, P, CH, DR, alpha, visc, ceq, c, X);
#pragma unroll
        for (int q = 0; q < NFV; ++q) xsh[w][q][l + 1] = X[q];
      }
      if (t.xbok) {
        R X[NFV];
        face_vc<0, NS, F32, LC>(k, g, U, P, CH, DR, alpha, visc, ceq, cxb, X);
#pragma unroll
        for (int q = 0; q < NFV; ++q) xsh[l][q][0] = X[q];
      }
      if (t.yfok) {
        R Y[NFV];
        face_vc<1, NS, F32, LC>(k, g, U, P, CH, DR, alpha, visc, ceq, c - g.s1, Y);
#pragma unroll
        for (int q = 0; q < NFV; ++q) ysh[w][q][l] = Y[q];
      }
      if (t.cellok) face_vc<2, NS, F32, LC>(k, g, U, P, CH, DR, alpha, visc, ceq, c, Z);
      __syncthreads();
    }
    if (t.cellok) {
      double dv[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) dv[v] = ld_dv<F32>(dvar, c + N * v);
      auto PV = [&](int q, uint32_t cell) { return (R)pget<F32>(P, N, q, cell); };
      if (visc) {
#pragma unroll
        for (int D = 0; D < 3; ++D) {
          const uint32_t e = D == 0 ? 1u : (D == 1 ? g.s1 : g.s2);
          R T0c, T0m, T1c, T1m, T2c, T2m;
          if (D == 0) {
            T0c = xsh[w][0][l + 1]; T0m = xsh[w][0][l];
            T1c = xsh[w][1][l + 1]; T1m = xsh[w][1][l];
            T2c = xsh[w][2][l + 1]; T2m = xsh[w][2][l];
          } else if (D == 1) {
            T0c = ysh[w + 1][0][l]; T0m = ysh[w][0][l];
            T1c = ysh[w + 1][1][l]; T1m = ysh[w][1][l];
            T2c = ysh[w + 1][2][l]; T2m = ysh[w][2][l];
          } else {
            T0c = Z[0]; T0m = Zp[0];
            T1c = Z[1]; T1m = Zp[1];
            T2c = Z[2]; T2m = Zp[2];
          }
          const R V0 = PV(2, c), V1 = PV(3, c), V2 = PV(4, c);
          const R ur = (PV(2, c + e) + V0) / (R)2.0, ul = (PV(2, c - e) + V0) / (R)2.0;
          const R vr = (PV(3, c + e) + V1) / (R)2.0, vl = (PV(3, c - e) + V1) / (R)2.0;
          const R wr = (PV(4, c + e) + V2) / (R)2.0, wl = (PV(4, c - e) + V2) / (R)2.0;
          const double dd = k.d[D], rdd = k.rd[D];
          const R ta = A::cdv(T0c - T0m, dd, rdd);
          const R tb = A::cdv(T1c - T1m, dd, rdd);
          const R tc = A::cdv(T2c - T2m, dd, rdd);
          const R td = A::cdv(ur * T0c - ul * T0m, dd, rdd) + A::cdv(vr * T1c - vl * T1m, dd, rdd) +
                       A::cdv(wr * T2c - wl * T2m, dd, rdd);
          dv[0] += ta;
          dv[1] += tb;
          dv[2] += tc;
          dv[3] += td;
        }
      }
      if (ceq) {
        // applyCeq, src/ceq3d.hpp:343-365
#pragma unroll
        for (int ww = 0; ww < 3; ++ww) {
          R diffu = (R)0.0;
#pragma unroll
          for (int dir = 0; dir < 3; ++dir)
            diffu += A::cdv(xsh[w][3 + 3 * dir + ww][l + 1] - xsh[w][3 + 3 * dir + ww][l], k.d[0], k.rd[0]);
#pragma unroll
          for (int dir = 0; dir < 3; ++dir)
            diffu += A::cdv(ysh[w + 1][3 + 3 * dir + ww][l] - ysh[w][3 + 3 * dir + ww][l], k.d[1], k.rd[1]);
#pragma unroll
          for (int dir = 0; dir < 3; ++dir)
            diffu += A::cdv(Z[3 + 3 * dir + ww] - Zp[3 + 3 * dir + ww], k.d[2], k.rd[2]);
          dv[ww] += diffu;
        }
      }
      // Runge-Kutta stage update of the flow variables and primitives of
      // the new state (FP64); the C variables are updated by k_ceq3d
      double un[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if constexpr (BL) un[v] = upd_val(up, ld(U, c + N * v), ld(Un, c + N * v), dv[v]);
        else un[v] = ((up.form == 1) ? ld(U, c + N * v) : ld(Un, c + N * v)) + up.h * dv[v];
        ust<FO>(Uo, c + N * v, un[v]);
        // float stage state: the primitives (and maxima) describe the stored,
        // rounded state, as the ghost-cell primitives computed from it do
        if (FO) un[v] = (double)(float)un[v];
      }
      double pr[6 + NS];
      double ms;
      prims_of<NS>(k, un, pr, ms);
#pragma unroll
      for (int q = 0; q < 6 + NS; ++q) pset<F32>(Po, N, q, c, pr[q]);
      mS = fmax(mS, ms);
    }
#pragma unroll
    for (int q = 0; q < NFV; ++q) Zp[q] = Z[q];
    if (faces) __syncthreads();
  }
  // block maxima -> one atomic per block and value
  mS = wave_max(mS);
  if (l == 0) red[w] = mS;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < NWAVE; ++q) s = fmax(s, red[q]);
    atomic_max_nonneg(&scout->maxS, s);
  }
}

// ---------------------------------------------------------------- kernel 2 with an LDS ring

// k_vcr3d with the face fields staged in LDS.  A ring of four planes holds,
// for rows j0-1..j0+TY and columns i0-1..i0+TX of the tile, u, v, w and, in
// the FP32 forms, also rho, Y_s and the C-equation fields (plain: the five C
// variables of U; learned: C_hat and the three direction planes) as FP32.
// The FP64 form stages u, v, w as doubles (the other fields do not fit next to
// its FP64 face exchange) and reads the rest from memory.  The block loads one
// new plane per k step (two steps ahead); faces and the viscous cell terms
// read the ring.  The arithmetic is k_vcr3d's: the same values in the same
// operations, so the FP64 form keeps its bitwise parity.
constexpr int RY = TY + 2, RX = TX + 2, RPL = RY * RX, NSLOT = 4;

template <int NS, bool F32, bool LC>
struct RingF {
  static constexpr int NFALL = 4 + NS + (LC ? 4 : 5);  // u v w rho Y_s C...
  static constexpr int NF = F32 ? NFALL : 3;           // fields held in the ring
  static constexpr int FY = 4, FC = 4 + NS;            // first Y, first C field
  static constexpr int NPC = (RPL + NTHREADS - 1) / NTHREADS;  // ring cells per thread
  static constexpr int NPRE = NF * NPC;
  using R = typename Ar<F32>::R;
};

// Global value of ring field f at cell c (as the form's type R).
template <int NS, bool F32, bool LC>
__device__ __forceinline__ typename Ar<F32>::R field_at(uint32_t N, int nv, const double *__restrict__ U,
                                                        const double *__restrict__ P, const float *__restrict__ CH,
                                                        const float *__restrict__ DR, int f, uint32_t c) {
  using RF = RingF<NS, F32, LC>;
  using R = typename RF::R;
  if (f < 3) return (R)pget<F32>(P, N, 2 + f, c);
  if (f == 3) return (R)pget<F32>(P, N, 0, c);
  if (f < RF::FC) return (R)pget<F32>(P, N, 6 + (f - RF::FY), c);
  if (LC) return (f == RF::FC) ? (R)CH[c] : (R)DR[c + N * (f - RF::FC - 1)];
  return (R)ld(U, c + N * (nv + (f - RF::FC)));
}

// Loads of one plane p of the ring, field-major: thread t takes the ring
// cells t, t + NTHREADS, ... of every field (compile-time field choice).
// Without the C-equation (ceq = 0) var has no C planes: the ring's C slots
// then take var's first planes instead, which exist (nv >= 5) and are never
// read (the faces read the C fields only with ceq set).  A select of the
// plane base keeps the loads unconditional: skipping them under a runtime
// test made the FP32 kernel 35 % slower.
template <int NS, bool F32, bool LC>
__device__ __forceinline__ void ring_fetch(const Grid32 &g, const double *__restrict__ U,
                                           const double *__restrict__ P, const float *__restrict__ CH,
                                           const float *__restrict__ DR, int i0, int j0, int p, int ceq,
                                           typename Ar<F32>::R *pre) {
  using RF = RingF<NS, F32, LC>;
  using R = typename RF::R;
  const int nvc = ceq ? g.nv : 0;  // first C plane
#pragma unroll
  for (int m = 0; m < RF::NPC; ++m) {
    const int e = threadIdx.x + m * NTHREADS;
    const int r = e / RX, q = e - r * RX;
    const int i = i0 - 1 + q, j = j0 - 1 + r;
    const bool ok = e < RPL && i < g.ngi && j < g.ngj;
    const uint32_t c = ok ? g.id(i, j, p) : 0u;
#pragma unroll
    for (int f = 0; f < RF::NF; ++f)
      pre[f * RF::NPC + m] = ok ? field_at<NS, F32, LC>(g.N, nvc, U, P, CH, DR, f, c) : (R)0;
  }
}

template <int NS, bool F32, bool LC>
__device__ __forceinline__ void ring_store(typename Ar<F32>::R *ring, int slot, const typename Ar<F32>::R *pre) {
  using RF = RingF<NS, F32, LC>;
#pragma unroll
  for (int m = 0; m < RF::NPC; ++m) {
    const int e = threadIdx.x + m * NTHREADS;
    if (e < RPL) {
#pragma unroll
      for (int f = 0; f < RF::NF; ++f) ring[(slot * RF::NF + f) * RPL + e] = pre[f * RF::NPC + m];
    }
  }
}

// Reader of the fields around a cell at ring row r, column q (global index
// c), with the ring slots of planes kk-1, kk, kk+1: ring fields from LDS,
// the others from memory.
template <int NS, bool F32, bool LC>
struct RingAt {
  using RF = RingF<NS, F32, LC>;
  using R = typename RF::R;
  const R *ring;
  const double *U, *P;
  const float *CH, *DR;
  uint32_t N, s1, s2;
  int nv;
  int sl[3];  // slots of relative planes -1, 0, +1
  int r, q;
  uint32_t c;
  __device__ __forceinline__ R operator()(int f, int dp, int dr, int dq) const {
    if (f < RF::NF) return ring[((sl[dp + 1] * RF::NF + f) * RY + (r + dr)) * RX + (q + dq)];
    const uint32_t cc = c + (uint32_t)dq + (uint32_t)dr * s1 + (uint32_t)dp * s2;
    return field_at<NS, F32, LC>(N, nv, U, P, CH, DR, f, cc);
  }
};

// face_vc on ring values; D: face direction; the face lies between the cell
// at (r, q, plane 0) and its +e_D neighbour.
template <int D, int NS, bool F32, bool LC>
__device__ __forceinline__ void face_vc_ring(const FC &k, const RingAt<NS, F32, LC> &at,
                                             typename Ar<F32>::R alpha, int visc, int ceq,
                                             typename Ar<F32>::R *out) {
  FST_FAST_FP
  using A = Ar<F32>;
  using RF = RingF<NS, F32, LC>;
  using R = typename A::R;
  constexpr int ep = D == 2, er = D == 1, eq = D == 0;  // offset of c+e: (plane, row, col)
  constexpr int T1 = D == 0 ? 1 : 0;
  constexpr int T2 = D == 2 ? 1 : 2;
  // offsets of +t for direction dir: x -> col, y -> row, z -> plane
  auto V = [&](int f, int dp, int dr, int dq) { return at(f, dp, dr, dq); };
  R der[3][3];
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    der[w][D] = A::cdv(V(w, ep, er, eq) - V(w, 0, 0, 0), k.d[D], k.rd[D]);
    {
      constexpr int tp = T1 == 2, tr = T1 == 1, tq = T1 == 0;
      der[w][T1] = A::cdv((V(w, tp, tr, tq) + V(w, ep + tp, er + tr, eq + tq)) -
                              (V(w, -tp, -tr, -tq) + V(w, ep - tp, er - tr, eq - tq)),
                          k.d4[T1], k.rd4[T1]);
    }
    {
      constexpr int tp = T2 == 2, tr = T2 == 1, tq = T2 == 0;
      der[w][T2] = A::cdv((V(w, tp, tr, tq) + V(w, ep + tp, er + tr, eq + tq)) -
                              (V(w, -tp, -tr, -tq) + V(w, ep - tp, er - tr, eq - tq)),
                          k.d4[T2], k.rd4[T2]);
    }
  }
  if (visc) {
    R mu = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s)
      mu += ((V(RF::FY + s, ep, er, eq) + V(RF::FY + s, 0, 0, 0)) / (R)2.0) * (R)k.mus[s];
    const R dudx = der[0][0], dudy = der[0][1], dudz = der[0][2];
    const R dvdx = der[1][0], dvdy = der[1][1], dvdz = der[1][2];
    const R dwdx = der[2][0], dwdy = der[2][1], dwdz = der[2][2];
    const R c23 = (R)(2.0 / 3.0), two = (R)2.0;
    if (D == 0) {
      out[0] = c23 * mu * (two * dudx - dvdy - dwdz);
      out[1] = mu * (dudy + dvdx);
      out[2] = mu * (dudz + dwdx);
    } else if (D == 1) {
      out[0] = mu * (dudy + dvdx);
      out[1] = c23 * mu * (two * dvdy - dudx - dwdz);
      out[2] = mu * (dvdz + dwdy);
    } else {
      out[0] = mu * (dudz + dwdx);
      out[1] = mu * (dvdz + dwdy);
      out[2] = c23 * mu * (two * dwdz - dudx - dvdy);
    }
  }
  if (ceq) {
    constexpr int fch = LC ? RF::FC : RF::FC + 1;  // C_hat
    constexpr int fcn = fch + 1;                  // direction n
    const R chat = (V(fch, 0, 0, 0) + V(fch, ep, er, eq)) / (R)2.0;
    if (LC && __ballot(chat != (R)0) == 0) {
      // learned form: C_hat vanishes on this face in every lane of the wave
      // (inactive cells), so the artificial-viscosity fluxes alpha rho C_hat M
      // du are zeros; their divergence adds +0 to dvar either way
#pragma unroll
      for (int q = 0; q < 9; ++q) out[3 + q] = (R)0;
      return;
    }
    const R r = (V(3, 0, 0, 0) + V(3, ep, er, eq)) / (R)2.0;
    R cn[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) cn[q] = (V(fcn + q, 0, 0, 0) + V(fcn + q, ep, er, eq)) / (R)2.0;
    R cmag = (R)1.0e-6;
    cmag += cn[0] * cn[0];
    cmag += cn[1] * cn[1];
    cmag += cn[2] * cn[2];
    const R arc = alpha * r * chat;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
      const R dirac = (dir == D) ? (R)1.0 : (R)0.0;
      const R M = dirac - A::div(cn[D] * cn[dir], cmag);
      const R X = arc * M;
#pragma unroll
      for (int w = 0; w < 3; ++w) out[3 + 3 * dir + w] = X * der[w][dir];
    }
  }
}

template <int NS, bool F32, bool LC, bool FO = false, bool BL = false>
__global__ void __launch_bounds__(NTHREADS, 1)
    k_vcr3d_ring(const Grid32 g, const FC k, const double *__restrict__ U, const double *__restrict__ Un,
                double *__restrict__ Uo, const double *__restrict__ P, double *__restrict__ Po,
                const double *__restrict__ dvar, const float *__restrict__ CH, const float *__restrict__ DR,
                const Scalars *scin, Scalars *scout, Upd up, int visc, int ceq, int kchunk, int nbx,
                int nby, int nbz) {
  using A = Ar<F32>;
  using RF = RingF<NS, F32, LC>;
  using R = typename A::R;
  constexpr int NV = 4 + NS;
  __shared__ R ring[NSLOT * RF::NF * RPL];
  __shared__ R xsh[TY][NFV][TX + 1];
  __shared__ R ysh[NWAVE][NFV][TX];
  __shared__ double red[NWAVE];
  const Tile t = place(g, kchunk, nbx, nby, nbz);
  const int w = t.w, l = t.l;
  const uint32_t N = g.N;
  const double maxS = as_double(scin->maxS);
  const R alpha = (R)((k.dxmag2 / (as_double(scin->maxCh) + 1.0e-6)) * k.alpha_cf);  // src/cart3d.cpp:285
  const bool faces = visc || ceq;
  R pre[RF::NPRE];
  // planes kb-1, kb, kb+1 of the ring
  for (int p = t.kb - 1; p <= t.kb + 1; ++p) {
    ring_fetch<NS, F32, LC>(g, U, P, CH, DR, t.i0, t.j0, p, ceq, pre);
    ring_store<NS, F32, LC>(ring, p & 3, pre);
  }
  __syncthreads();
  RingAt<NS, F32, LC> at;
  at.ring = ring;
  at.U = U; at.P = P; at.CH = CH; at.DR = DR;
  at.N = g.N; at.s1 = g.s1; at.s2 = g.s2; at.nv = g.nv;
  R Z[NFV], Zp[NFV];
#pragma unroll
  for (int q = 0; q < NFV; ++q) Z[q] = Zp[q] = (R)0.0;
  if (faces && t.cellok) {  // z-face between kb-1 and kb
    at.sl[0] = (t.kb - 2) & 3; at.sl[1] = (t.kb - 1) & 3; at.sl[2] = t.kb & 3;
    at.r = w + 1; at.q = l + 1;
    at.c = g.id(t.i, t.jrow, t.kb - 1);
    face_vc_ring<2, NS, F32, LC>(k, at, alpha, visc, ceq, Zp);
  }
  double mS = 0.0;
  for (int kk = t.kb; kk < t.ke; ++kk) {
    uint32_t c = g.id(t.i, t.jrow, kk);
    uint32_t cxb = g.id(t.i0 - 1, t.j0 + l, kk);
    asm volatile("" : "+v"(c), "+v"(cxb));
    const bool fetch = kk + 2 <= t.ke;  // plane kk+2 is read by iteration kk+1
    if (fetch) ring_fetch<NS, F32, LC>(g, U, P, CH, DR, t.i0, t.j0, kk + 2, ceq, pre);
    // FP32 forms (registers to spare): the cell's partial RHS and stage base
    // are loaded before the face phase so that their latency overlaps it (in
    // the FP64 form this measured slower)
    constexpr bool EARLY = F32;
    double dvp[EARLY ? NV : 1], bsp[EARLY ? NV : 1];
    if (EARLY && t.cellok) {
#pragma unroll
      for (int v = 0; v < NV; ++v) dvp[v] = ld_dv<F32>(dvar, c + N * v);
      if (!BL) {
#pragma unroll
        for (int v = 0; v < NV; ++v) bsp[v] = (up.form == 1) ? ld(U, c + N * v) : ld(Un, c + N * v);
      }
    }
    at.sl[0] = (kk - 1) & 3; at.sl[1] = kk & 3; at.sl[2] = (kk + 1) & 3;
    if (faces) {
      if (F32 && t.cellok) {
        // FP32 forms (registers to spare): one straight-line region for the
        // three faces of a tile cell, so that the loads of the later faces
        // are issued ahead of the earlier faces' arithmetic
        R X[NFV];
        at.r = w + 1; at.q = l + 1; at.c = c;
        face_vc_ring<0, NS, F32, LC>(k, at, alpha, visc, ceq, X);
#pragma unroll
        for (int q = 0; q < NFV; ++q) xsh[w][q][l + 1] = X[q];
        face_vc_ring<2, NS, F32, LC>(k, at, alpha, visc, ceq, Z);
        R Y[NFV];  // y-face below row w (yfok holds for a tile cell)
        at.r = w; at.c = c - g.s1;
        face_vc_ring<1, NS, F32, LC>(k, at, alpha, visc, ceq, Y);
#pragma unroll
        for (int q = 0; q < NFV; ++q) ysh[w][q][l] = Y[q];
      } else {
        if (!F32 && t.cellok) {
          R X[NFV];
          at.r = w + 1; at.q = l + 1; at.c = c;
          face_vc_ring<0, NS, F32, LC>(k, at, alpha, visc, ceq, X);
#pragma unroll
          for (int q = 0; q < NFV; ++q) xsh[w][q][l + 1] = X[q];
        }
        if (t.xbok) {  // x-face left of the tile, row l
          R X[NFV];
          at.r = l + 1; at.q = 0; at.c = cxb;
          face_vc_ring<0, NS, F32, LC>(k, at, alpha, visc, ceq, X);
#pragma unroll
          for (int q = 0; q < NFV; ++q) xsh[l][q][0] = X[q];
        }
        if (t.yfok) {  // y-face below row w (between rows w-1 and w)
          R Y[NFV];
          at.r = w; at.q = l + 1; at.c = c - g.s1;
          face_vc_ring<1, NS, F32, LC>(k, at, alpha, visc, ceq, Y);
#pragma unroll
          for (int q = 0; q < NFV; ++q) ysh[w][q][l] = Y[q];
        }
        if (!F32 && t.cellok) {
          at.r = w + 1; at.q = l + 1; at.c = c;
          face_vc_ring<2, NS, F32, LC>(k, at, alpha, visc, ceq, Z);
        }
      }
    }
    if (fetch) ring_store<NS, F32, LC>(ring, (kk + 2) & 3, pre);
    __syncthreads();
    if (t.cellok) {
      at.r = w + 1; at.q = l + 1; at.c = c;
      double dv[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) dv[v] = EARLY ? dvp[EARLY ? v : 0] : ld_dv<F32>(dvar, c + N * v);
      if (visc) {
#pragma unroll
        for (int D = 0; D < 3; ++D) {
          const int dp = D == 2, dr = D == 1, dq = D == 0;
          R T0c, T0m, T1c, T1m, T2c, T2m;
          if (D == 0) {
            T0c = xsh[w][0][l + 1]; T0m = xsh[w][0][l];
            T1c = xsh[w][1][l + 1]; T1m = xsh[w][1][l];
            T2c = xsh[w][2][l + 1]; T2m = xsh[w][2][l];
          } else if (D == 1) {
            T0c = ysh[w + 1][0][l]; T0m = ysh[w][0][l];
            T1c = ysh[w + 1][1][l]; T1m = ysh[w][1][l];
            T2c = ysh[w + 1][2][l]; T2m = ysh[w][2][l];
          } else {
            T0c = Z[0]; T0m = Zp[0];
            T1c = Z[1]; T1m = Zp[1];
            T2c = Z[2]; T2m = Zp[2];
          }
          const R V0 = at(0, 0, 0, 0), V1 = at(1, 0, 0, 0), V2 = at(2, 0, 0, 0);
          const R ur = (at(0, dp, dr, dq) + V0) / (R)2.0, ul = (at(0, -dp, -dr, -dq) + V0) / (R)2.0;
          const R vr = (at(1, dp, dr, dq) + V1) / (R)2.0, vl = (at(1, -dp, -dr, -dq) + V1) / (R)2.0;
          const R wr = (at(2, dp, dr, dq) + V2) / (R)2.0, wl = (at(2, -dp, -dr, -dq) + V2) / (R)2.0;
          const double dd = k.d[D], rdd = k.rd[D];
          const R ta = A::cdv(T0c - T0m, dd, rdd);
          const R tb = A::cdv(T1c - T1m, dd, rdd);
          const R tc = A::cdv(T2c - T2m, dd, rdd);
          const R td = A::cdv(ur * T0c - ul * T0m, dd, rdd) + A::cdv(vr * T1c - vl * T1m, dd, rdd) +
                           A::cdv(wr * T2c - wl * T2m, dd, rdd);
          dv[0] += ta;
          dv[1] += tb;
          dv[2] += tc;
          dv[3] += td;
        }
      }
      bool ceq_cell = ceq;
      if (LC && ceq) {  // learned form: skip when C_hat is 0 at the cell and its 6 neighbours in the whole wave
        constexpr int fch = RF::FC;
        const bool act = at(fch, 0, 0, 0) != (R)0 || at(fch, 0, 0, -1) != (R)0 || at(fch, 0, 0, 1) != (R)0 ||
                         at(fch, 0, -1, 0) != (R)0 || at(fch, 0, 1, 0) != (R)0 || at(fch, -1, 0, 0) != (R)0 ||
                         at(fch, 1, 0, 0) != (R)0;
        ceq_cell = __ballot(act) != 0;
      }
      if (ceq_cell) {
        // applyCeq, src/ceq3d.hpp:343-365
#pragma unroll
        for (int ww = 0; ww < 3; ++ww) {
          R diffu = (R)0.0;
#pragma unroll
          for (int dir = 0; dir < 3; ++dir)
            diffu += A::cdv(xsh[w][3 + 3 * dir + ww][l + 1] - xsh[w][3 + 3 * dir + ww][l], k.d[0], k.rd[0]);
#pragma unroll
          for (int dir = 0; dir < 3; ++dir)
            diffu += A::cdv(ysh[w + 1][3 + 3 * dir + ww][l] - ysh[w][3 + 3 * dir + ww][l], k.d[1], k.rd[1]);
#pragma unroll
          for (int dir = 0; dir < 3; ++dir)
            diffu += A::cdv(Z[3 + 3 * dir + ww] - Zp[3 + 3 * dir + ww], k.d[2], k.rd[2]);
          dv[ww] += diffu;
        }
      }
      // Runge-Kutta stage update of the flow variables and primitives of
      // the new state (FP64); the C variables are updated by k_ceq3d
      double un[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if constexpr (BL) {
          un[v] = upd_val(up, ld(U, c + N * v), ld(Un, c + N * v), dv[v]);
        } else {
          const double base = EARLY ? bsp[EARLY ? v : 0] : ((up.form == 1) ? ld(U, c + N * v) : ld(Un, c + N * v));
          un[v] = base + up.h * dv[v];
        }
        ust<FO>(Uo, c + N * v, un[v]);
        // float stage state: the primitives (and maxima) describe the stored,
        // rounded state, as the ghost-cell primitives computed from it do
        if (FO) un[v] = (double)(float)un[v];
      }
      double pr[6 + NS];
      double ms;
      prims_of<NS>(k, un, pr, ms);
#pragma unroll
      for (int q = 0; q < 6 + NS; ++q) pset<F32>(Po, N, q, c, pr[q]);
      mS = fmax(mS, ms);
    }
#pragma unroll
    for (int q = 0; q < NFV; ++q) Zp[q] = Z[q];
    __syncthreads();
  }
  mS = wave_max(mS);
  if (l == 0) red[w] = mS;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < NWAVE; ++q) s = fmax(s, red[q]);
    atomic_max_nonneg(&scout->maxS, s);
  }
}

// ---------------------------------------------------------------- kernel 3: C-equation sources

// calculateRhoGrad (src/ceq3d.hpp:143-186) and updateCeq (src/ceq3d.hpp:202-223)
// with the Runge-Kutta stage update of the five C variables (src/rk.cpp:42-50 /
// :63-69), and the maximum |C_hat| of the new state.  updateCeq alone assigns
// the C variables' right-hand side, so this pass is independent of the flow
// kernels of the stage (which read U's C variables, never Uo's).
//
// A block is 64 (i) x CY (j) columns marching k.  The x/y stencil fields of
// the current plane -- rho, nrg, u, v (4th-order, 2 cells of halo) and the C
// variables (Laplacian, 1 cell) -- are staged in LDS for rows j0-2..j0+CY+1
// and columns i0-2..i0+65, the next plane being loaded into registers during
// the current one.  The z stencils -- rho, nrg, w at k-2..k+2 and the C
// variables at k-1..k+1 -- stay in per-column register queues, loaded one
// plane ahead.  Same operations as the reference sequence, so the FP64 form
// stays bitwise equal.
constexpr int CY = 8, CNT = 64 * CY;
constexpr int CPR = CY + 4, CPX = 64 + 4, CPL = CPR * CPX;  // staged plane: rows, columns, cells
constexpr int CNF = 9;                                       // rho, nrg, u, v, C0..C4
constexpr int CNPC = (CPL + CNT - 1) / CNT;  // plane cells per thread
constexpr int CNPRE = CNF * CNPC;

template <bool F32>
__device__ __forceinline__ void ceq_fetch(const Grid32 &g, const double *__restrict__ U, const double *__restrict__ P,
                                          int i0, int j0, int kk, typename Ar<F32>::R *pre) {
  using R = typename Ar<F32>::R;
  const uint32_t N = g.N;
#pragma unroll
  for (int m = 0; m < CNPC; ++m) {  // field-major: compile-time field choice
    const int e = threadIdx.x + m * CNT;
    const int r = e / CPX, q = e - r * CPX;
    const int i = i0 - 2 + q, j = j0 - 2 + r;
    const bool ok = e < CPL && i < g.ngi && j < g.ngj;
    const uint32_t c = ok ? g.id(i, j, kk) : 0u;
#pragma unroll
    for (int f = 0; f < CNF; ++f) {
      // primitive planes rho 0, nrg 5, u 2, v 3; then the C variables of U
      const int pq = f == 0 ? 0 : f == 1 ? 5 : f == 2 ? 2 : 3;
      pre[f * CNPC + m] = !ok ? (R)0 : f < 4 ? (R)pget<F32>(P, N, pq, c) : (R)ld(U, c + N * (uint32_t)(g.nv + f - 4));
    }
  }
}

template <bool F32, bool BL = false>
__global__ void __launch_bounds__(CNT)
    k_ceq3d(const Grid32 g, const FC k, const double *__restrict__ U, const double *__restrict__ Un,
            double *__restrict__ Uo, const double *__restrict__ P, const Scalars *scin, Scalars *scout, Upd up,
            int kchunk, int nbx, int nby, int nbz) {
  using A = Ar<F32>;
  using R = typename A::R;
  __shared__ R pl[CNF * CPL];
  const int nb = nbx * nby * nbz;
  const int hb = blockIdx.x;
  int lb = hb;
  const int full = (nb / 8) * 8;
  if (hb < full) lb = (hb % 8) * (full / 8) + hb / 8;  // consecutive tiles on one XCD
  const int bx = lb % nbx, by = (lb / nbx) % nby, bz = lb / (nbx * nby);
  const int i0 = g.ng + bx * 64, j0 = g.ng + by * CY;
  const int l = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  const int i = i0 + l, j = j0 + w;
  const int kb = g.ng + bz * kchunk, ke = min(kb + kchunk, g.ng + g.nck);
  const bool ok = i < g.ng + g.nci && j < g.ng + g.ncj;
  const uint32_t N = g.N, s2 = g.s2;
  const uint32_t nvb = N * (uint32_t)g.nv;
  const double maxS = as_double(scin->maxS);
  const R c1 = (R)(maxS / (k.eps * k.dxmag));
  const R c2 = (R)(k.kap * maxS * k.dxmag);
  auto PV = [&](int q, uint32_t cell) { return (R)pget<F32>(P, N, q, cell); };
  auto CU = [&](uint32_t off) { return (R)ld(U, off); };
  constexpr int rb = 0, eb = 5, wb = 4;  // primitive planes rho, nrg, w
  // queues: [0..4] = planes k-2..k+2 of rho, nrg, w; qc[0..2] = planes k-1..k+1 of C
  R qr[5], qe[5], qw[5], qc[3][5], nr = 0, ne = 0, nw = 0, nc[5];
  uint32_t c = g.id(min(i, g.ngi - 1), min(j, g.ngj - 1), kb);
  if (ok) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t cc = c + (uint32_t)q * s2 - 2u * s2;
      qr[q] = PV(rb, cc);
      qe[q] = PV(eb, cc);
      qw[q] = PV(wb, cc);
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      qc[0][n] = CU(c - s2 + nvb + N * n);
      qc[1][n] = CU(c + nvb + N * n);
    }
    nr = PV(rb, c + 2 * s2);
    ne = PV(eb, c + 2 * s2);
    nw = PV(wb, c + 2 * s2);
#pragma unroll
    for (int n = 0; n < 5; ++n) nc[n] = CU(c + s2 + nvb + N * n);
  }
  R pre[CNPRE];
  ceq_fetch<F32>(g, U, P, i0, j0, kb, pre);
  double mC = 0.0;
  for (int kk = kb; kk < ke; ++kk, c += s2) {
    __syncthreads();  // the previous plane's readers are done
#pragma unroll
    for (int m = 0; m < CNPC; ++m) {
      const int e = threadIdx.x + m * CNT;
      if (e < CPL) {
#pragma unroll
        for (int f = 0; f < CNF; ++f) pl[f * CPL + e] = pre[f * CNPC + m];
      }
    }
    __syncthreads();
    if (kk + 1 < ke) ceq_fetch<F32>(g, U, P, i0, j0, kk + 1, pre);
    if (ok) {
      qr[4] = nr;
      qe[4] = ne;
      qw[4] = nw;
#pragma unroll
      for (int n = 0; n < 5; ++n) qc[2][n] = nc[n];
      // the cell's C variables in U (stage input) and, for the SSP-RK3 blend, Un
      double base[5], ub[BL ? 5 : 1];
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        if constexpr (BL) {
          base[n] = ld(U, c + nvb + N * n);
          ub[n] = ld(Un, c + nvb + N * n);
        } else {
          base[n] = (up.form == 1) ? ld(U, c + nvb + N * n) : ld(Un, c + nvb + N * n);
        }
      }
      if (kk + 1 < ke) {  // z-queue entries of the next plane
        nr = PV(rb, c + 3 * s2);
        ne = PV(eb, c + 3 * s2);
        nw = PV(wb, c + 3 * s2);
#pragma unroll
        for (int n = 0; n < 5; ++n) nc[n] = CU(c + 2 * s2 + nvb + N * n);
      }
      const R *pc = pl + (w + 2) * CPX + (l + 2);  // this column's cell in field 0
      auto L = [&](int f, int dr, int dq) { return pc[f * CPL + dr * CPX + dq]; };
      // calculateRhoGrad, src/ceq3d.hpp:143-186: 4th-order gradients of rho
      // and nrg, divergence of u
      R dr[3], de[3], du[3];
      dr[0] = A::cdv(L(0, 0, -2) - (R)8.0 * L(0, 0, -1) + (R)8.0 * L(0, 0, 1) - L(0, 0, 2), k.d12[0], k.rd12[0]);
      de[0] = A::cdv(L(1, 0, -2) - (R)8.0 * L(1, 0, -1) + (R)8.0 * L(1, 0, 1) - L(1, 0, 2), k.d12[0], k.rd12[0]);
      du[0] = A::cdv(L(2, 0, -2) - (R)8.0 * L(2, 0, -1) + (R)8.0 * L(2, 0, 1) - L(2, 0, 2), k.d12[0], k.rd12[0]);
      dr[1] = A::cdv(L(0, -2, 0) - (R)8.0 * L(0, -1, 0) + (R)8.0 * L(0, 1, 0) - L(0, 2, 0), k.d12[1], k.rd12[1]);
      de[1] = A::cdv(L(1, -2, 0) - (R)8.0 * L(1, -1, 0) + (R)8.0 * L(1, 1, 0) - L(1, 2, 0), k.d12[1], k.rd12[1]);
      du[1] = A::cdv(L(3, -2, 0) - (R)8.0 * L(3, -1, 0) + (R)8.0 * L(3, 1, 0) - L(3, 2, 0), k.d12[1], k.rd12[1]);
      dr[2] = A::cdv(qr[0] - (R)8.0 * qr[1] + (R)8.0 * qr[3] - qr[4], k.d12[2], k.rd12[2]);
      de[2] = A::cdv(qe[0] - (R)8.0 * qe[1] + (R)8.0 * qe[3] - qe[4], k.d12[2], k.rd12[2]);
      du[2] = A::cdv(qw[0] - (R)8.0 * qw[1] + (R)8.0 * qw[3] - qw[4], k.d12[2], k.rd12[2]);
      const R rgrad = A::sq(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      const R divu = du[0] + du[1] + du[2];
      const R dnednr = (dr[0] * de[0] + dr[1] * de[1] + dr[2] * de[2]) * (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      const int ind = dnednr < (R)0.0 ? 1 : 0;
      R G[5];
      G[0] = (divu < 0) ? (R)(1 - ind) * rgrad : (R)0;
      G[1] = (R)ind * rgrad;
      G[2] = dr[0];
      G[3] = dr[1];
      G[4] = dr[2];
      // updateCeq, src/ceq3d.hpp:202-223, then the stage update of the C variables
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        const R cc = qc[1][n];
        const R lap = A::cdv(L(4 + n, 0, -1) - (R)2 * cc + L(4 + n, 0, 1), k.d[0], k.rd[0]) +
                      A::cdv(L(4 + n, -1, 0) - (R)2 * cc + L(4 + n, 1, 0), k.d[1], k.rd[1]) +
                      A::cdv(qc[0][n] - (R)2 * cc + qc[2][n], k.d[2], k.rd[2]);
        const double dv = c1 * (G[n] - cc) + c2 * lap;
        double un;
        if constexpr (BL) un = upd_val(up, base[n], ub[n], dv);
        else un = base[n] + up.h * dv;
        st(Uo, c + nvb + N * n, un);
        if (n == 1) mC = fmax(mC, fabs(un));
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        qr[q] = qr[q + 1];
        qe[q] = qe[q + 1];
        qw[q] = qw[q + 1];
      }
#pragma unroll
      for (int n = 0; n < 5; ++n) {
        qc[0][n] = qc[1][n];
        qc[1][n] = qc[2][n];
      }
    }
  }
  double v1[1] = {mC};
  unsigned long long *const d1[1] = {&scout->maxCh};
  block_max_atomic<1>(v1, d1);
}

// ---------------------------------------------------------------- primitives

// Cells of the ghost shell: thread t enumerates the k slabs (k < ng or
// k >= ng+nck), then the j slabs of the interior k range, then the i slabs of
// the interior j and k ranges.  Returns false past the last ghost cell.
__device__ __forceinline__ bool ghost_cell(const Grid32 &g, uint32_t t, uint32_t *cell) {
  const int ng = g.ng;
  const uint32_t nk = 2u * ng * g.s2;                     // k < ng or k >= ng+nck
  const uint32_t nj = 2u * ng * (uint32_t)g.ngi * g.nck;  // interior k, j ghost
  const uint32_t ni = 2u * ng * (uint32_t)g.ncj * g.nck;  // interior k and j, i ghost
  if (t < nk) {
    const uint32_t kq = t / g.s2, r = t % g.s2;
    const uint32_t kk = kq < (uint32_t)ng ? kq : kq + g.nck;
    *cell = r + kk * g.s2;
  } else if (t < nk + nj) {
    t -= nk;
    const uint32_t i = t % g.ngi, q = t / g.ngi;
    const uint32_t jq = q % (2u * ng), kq = q / (2u * ng);
    const uint32_t j = jq < (uint32_t)ng ? jq : jq + g.ncj;
    *cell = i + j * g.s1 + (kq + ng) * g.s2;
  } else if (t < nk + nj + ni) {
    t -= nk + nj;
    const uint32_t iq = t % (2u * ng), q = t / (2u * ng);
    const uint32_t j = q % g.ncj, kq = q / g.ncj;
    const uint32_t i = iq < (uint32_t)ng ? iq : iq + g.nci;
    *cell = i + (j + ng) * g.s1 + (kq + ng) * g.s2;
  } else {
    return false;
  }
  return true;
}

// Ghost cells of the last SSP-RK3 stage.  The stage update covers every cell
// (as rkAdvance's does) with dvar = 0 on the ghosts: U^{n+1} = a Un + b (U + h 0)
// with Un the step's ghost-filled initial state (Uo, in place) and U the
// ghost-filled stage input.  Earlier stages need no such pass: the next
// stage's ghost fill rewrites every ghost cell.
__global__ void k_rk_ghost(const Grid32 g, const double *__restrict__ U, double *__restrict__ Uo, Upd up) {
  uint32_t cell;
  if (!ghost_cell(g, blockIdx.x * blockDim.x + threadIdx.x, &cell)) return;
  for (int v = 0; v < g.nvt; ++v) {
    const uint32_t c = cell + g.N * v;
    Uo[c] = upd_val(up, U[c], Uo[c], 0.0);
  }
}

// Primitives of U over the whole ghost range, with the maxima of the interior
// (ghost_only = 0), or over the ghost cells only (ghost_only = 1: threads
// enumerate the k slabs, then the j slabs, then the i slabs of the ghost
// shell).  Threads 0..3 of block 0 clear `clear` when given.
// FU: U is the float stage state of the learned FP32 form.
template <int NS, bool F32, bool FU = false>
__global__ void k_prims3d(const Grid32 g, const FC k, const double *__restrict__ U, double *__restrict__ P,
                          Scalars *sc, int ghost_only, Scalars *clear) {
  if (clear && blockIdx.x == 0 && threadIdx.x < 4) reinterpret_cast<unsigned long long *>(clear)[threadIdx.x] = 0ull;
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ng = g.ng;
  bool inrange, interior = false;
  uint32_t cell = 0;
  if (ghost_only) {
    inrange = ghost_cell(g, t, &cell);
  } else {
    inrange = t < g.N;
    cell = t;
    const int i = (int)(t % (uint32_t)g.ngi);
    const int j = (int)((t / (uint32_t)g.ngi) % (uint32_t)g.ngj);
    const int kk = (int)(t / g.s2);
    interior = i >= ng && i < ng + g.nci && j >= ng && j < ng + g.ncj && kk >= ng && kk < ng + g.nck;
  }
  double ms = 0.0, mc = 0.0;
  if (inrange) {
    const uint32_t N = g.N;
    double u[4 + NS];
#pragma unroll
    for (int v = 0; v < 4 + NS; ++v) u[v] = uld<FU>(U, cell + N * v);
    double pr[6 + NS];
    prims_of<NS>(k, u, pr, ms);
#pragma unroll
    for (int q = 0; q < 6 + NS; ++q) pset<F32>(P, N, q, cell, pr[q]);
    if (!interior) ms = 0.0;
    else if (g.nvt > g.nv) mc = fabs(uld<FU>(U, cell + N * (g.nv + 1)));
  }
  if (!ghost_only && sc) {  // uniform over the block
    double v2[2] = {ms, mc};
    unsigned long long *const d2[2] = {&sc->maxS, &sc->maxCh};
    block_max_atomic<2>(v2, d2);
  }
}

__global__ void k_clear(Scalars *sc) {
  if (threadIdx.x < 4) reinterpret_cast<unsigned long long *>(sc)[threadIdx.x] = 0ull;
}

FC make_fc(const Phys &ph, const fst_config &cf) {
  FC k;
  const double d[3] = {cf.dx, cf.dy, cf.dz};
  for (int q = 0; q < 3; ++q) {
    k.d[q] = d[q];
    k.rd[q] = 1.0 / d[q];
    k.d4[q] = 4 * d[q];
    k.rd4[q] = 1.0 / k.d4[q];
    k.d12[q] = 12.0 * d[q];
    k.rd12[q] = 1.0 / k.d12[q];
  }
  k.pg[0] = 12.0 * ph.cd[1];
  k.pg[1] = 12.0 * ph.cd[2];
  k.pg[2] = 12.0 * ph.cd[2];
  for (int q = 0; q < 3; ++q) k.rpg[q] = 1.0 / k.pg[q];
  k.r12 = 1.0 / 12.0;
  for (int s = 0; s < FST_MAX_SPECIES; ++s) {
    k.cpf[s] = ph.cpf[s];
    k.cvf[s] = ph.cvf[s];
    k.mus[s] = ph.mus[s];
  }
  k.g = ph.g;
  k.rhoRef = ph.rhoRef;
  k.kap = ph.kap;
  k.eps = ph.eps;
  k.alpha_cf = ph.alpha;
  k.dxmag2 = ph.dxmag2;
  k.dxmag = std::sqrt(ph.cd[1] * ph.cd[1] + ph.cd[2] * ph.cd[2] + ph.cd[3] * ph.cd[3]);
  return k;
}

Grid32 make_grid(const Geo &g) {
  Grid32 q;
  q.N = (uint32_t)g.N;
  q.s1 = (uint32_t)g.ngi;
  q.s2 = (uint32_t)g.ngi * (uint32_t)g.ngj;
  q.ngi = g.ngi; q.ngj = g.ngj; q.ng = g.ng;
  q.nci = g.nci; q.ncj = g.ncj; q.nck = g.nck;
  q.nv = g.nv; q.nvt = g.nvt;
  return q;
}

struct Launch {
  int nbx, nby, nbz, nb;
};

Launch shape(const Geo &g, int kchunk) {
  Launch L;
  L.nbx = (g.nci + TX - 1) / TX;
  L.nby = (g.ncj + TY - 1) / TY;
  L.nbz = (g.nck + kchunk - 1) / kchunk;
  L.nb = L.nbx * L.nby * L.nbz;
  return L;
}

#ifndef FST_GEN3D_TU  // the cartesian stage drivers

// fs: the stage state U* is float (learned FP32 form): stage 1 writes it,
// stage 2 reads it; U^n and U^{n+1} stay FP64.
template <int NS, bool F32, bool LC>
void run_stage(const Grid32 &q, const Geo &g, const FC &k, const fst_config &cf, const double *U,
               const double *Un, double *Uo, const double *P, double *Po, double *dvar, const float *CH,
               const float *DR, const Scalars *scin, Scalars *scout, Upd up, int stage, int kchunk,
               Prof *prof, hipStream_t s, bool fs) {
  const Launch L = shape(g, kchunk);
  if (F32) {  // FP32 forms: binary32 plane, 11-row tiles (one implementation)
    const int nbys = (g.ncj + TYS - 1) / TYS, nbs = L.nbx * nbys * L.nbz;
    if (fs && stage == 2)
      k_adv3d_s<NS, true><<<nbs, NTS, 0, s>>>(q, k, U, P, dvar, cf.buoyancy, kchunk, L.nbx, nbys, L.nbz);
    else
      k_adv3d_s<NS, false><<<nbs, NTS, 0, s>>>(q, k, U, P, dvar, cf.buoyancy, kchunk, L.nbx, nbys, L.nbz);
  } else if (!getenv("FST_ADV_GLOBAL")) {
    k_adv3d_lds<NS, false><<<L.nb, NTHREADS, 0, s>>>(q, k, U, P, dvar, cf.buoyancy, kchunk, L.nbx, L.nby, L.nbz);
  } else {
    k_adv3d<NS, false><<<L.nb, NTHREADS, 0, s>>>(q, k, U, P, dvar, cf.buoyancy, kchunk, L.nbx, L.nby, L.nbz);
  }
  if (prof) prof->mark(s);  // end of the advection kernel
  const bool ring = !getenv("FST_VCR_GLOBAL");
  if constexpr (F32 && LC) {
    if (fs && stage == 1) {
      if (ring)
        k_vcr3d_ring<NS, F32, LC, true><<<L.nb, NTHREADS, 0, s>>>(q, k, U, Un, Uo, P, Po, dvar, CH, DR, scin, scout,
                                                                  up, cf.visc, cf.ceq, kchunk, L.nbx, L.nby,
                                                                  L.nbz);
      else
        k_vcr3d<NS, F32, LC, true><<<L.nb, NTHREADS, 0, s>>>(q, k, U, Un, Uo, P, Po, dvar, CH, DR, scin, scout, up,
                                                             cf.visc, cf.ceq, kchunk, L.nbx, L.nby, L.nbz);
      if (prof) prof->mark(s);
      return;
    }
  }
  const bool bl = up.form == 3;  // SSP-RK3 blend of U^n and the stage result
  if (ring && bl)
    k_vcr3d_ring<NS, F32, LC, false, true><<<L.nb, NTHREADS, 0, s>>>(q, k, U, Un, Uo, P, Po, dvar, CH, DR, scin,
                                                                     scout, up, cf.visc, cf.ceq, kchunk, L.nbx,
                                                                     L.nby, L.nbz);
  else if (ring)
    k_vcr3d_ring<NS, F32, LC><<<L.nb, NTHREADS, 0, s>>>(q, k, U, Un, Uo, P, Po, dvar, CH, DR, scin, scout, up,
                                                       cf.visc, cf.ceq, kchunk, L.nbx, L.nby, L.nbz);
  else if (bl)
    k_vcr3d<NS, F32, LC, false, true><<<L.nb, NTHREADS, 0, s>>>(q, k, U, Un, Uo, P, Po, dvar, CH, DR, scin, scout,
                                                                up, cf.visc, cf.ceq, kchunk, L.nbx, L.nby, L.nbz);
  else
    k_vcr3d<NS, F32, LC><<<L.nb, NTHREADS, 0, s>>>(q, k, U, Un, Uo, P, Po, dvar, CH, DR, scin, scout, up,
                                                  cf.visc, cf.ceq, kchunk, L.nbx, L.nby, L.nbz);
  if (prof) prof->mark(s);  // end of the viscous / stage-update kernel
  if (cf.ceq && !LC) {  // the C variables: sources, relaxation and stage update
    const int kc = 32;
    const int nbx = (g.nci + 63) / 64, nby = (g.ncj + CY - 1) / CY, nbz = (g.nck + kc - 1) / kc;
    if (bl)
      k_ceq3d<F32, true><<<nbx * nby * nbz, CNT, 0, s>>>(q, k, U, Un, Uo, P, scin, scout, up, kc, nbx, nby, nbz);
    else
      k_ceq3d<F32><<<nbx * nby * nbz, CNT, 0, s>>>(q, k, U, Un, Uo, P, scin, scout, up, kc, nbx, nby, nbz);
  }
}

template <int NS, bool F32>
void run_prims(const Grid32 &q, const FC &k, const double *U, double *P, Scalars *sc, int ghost_only,
               Scalars *clear, hipStream_t s, bool uf) {
  uint32_t n = q.N;
  if (ghost_only)
    n = 2u * q.ng * q.s2 + 2u * q.ng * (uint32_t)q.ngi * q.nck + 2u * q.ng * (uint32_t)q.ncj * q.nck;
  if (F32 && uf)  // the float stage state of the learned FP32 form
    k_prims3d<NS, F32, true><<<(n + 255) / 256, 256, 0, s>>>(q, k, U, P, sc, ghost_only, clear);
  else
    k_prims3d<NS, F32><<<(n + 255) / 256, 256, 0, s>>>(q, k, U, P, sc, ghost_only, clear);
}

using StageFn = void (*)(const Grid32 &, const Geo &, const FC &, const fst_config &, const double *,
                         const double *, double *, const double *, double *, double *, const float *,
                         const float *, const Scalars *, Scalars *, Upd, int, int, Prof *, hipStream_t, bool);
using PrimFn = void (*)(const Grid32 &, const FC &, const double *, double *, Scalars *, int, Scalars *,
                        hipStream_t, bool);

#ifdef FST_FAST64_TU  // the fast FP64 form: FP64 plain stage kernels only
bool pick(int ns, bool f32, bool lc, StageFn *sf, PrimFn *pf) {
  if (f32 || lc) return false;
  switch (ns) {
    case 1: *sf = run_stage<1, false, false>; *pf = run_prims<1, false>; return true;
    case 2: *sf = run_stage<2, false, false>; *pf = run_prims<2, false>; return true;
    case 3: *sf = run_stage<3, false, false>; *pf = run_prims<3, false>; return true;
    default: return false;
  }
}
#else
template <int NS>
StageFn pick_ns(bool f32, bool lc) {
  return f32 ? (lc ? run_stage<NS, true, true> : run_stage<NS, true, false>)
             : (lc ? run_stage<NS, false, true> : run_stage<NS, false, false>);
}

bool pick(int ns, bool f32, bool lc, StageFn *sf, PrimFn *pf) {
  switch (ns) {
    case 1: *sf = pick_ns<1>(f32, lc); *pf = f32 ? run_prims<1, true> : run_prims<1, false>; return true;
    case 2: *sf = pick_ns<2>(f32, lc); *pf = f32 ? run_prims<2, true> : run_prims<2, false>; return true;
    case 3: *sf = pick_ns<3>(f32, lc); *pf = f32 ? run_prims<3, true> : run_prims<3, false>; return true;
    default: return false;
  }
}
#endif  // FST_FAST64_TU

#endif  // FST_GEN3D_TU

#ifdef FST_FAST64_TU
}  // namespace fast64
using namespace fast64;
#endif
}  // namespace

#ifdef FST_GEN3D_TU
// gen3d_func::compute's fluxes, their divergence and the metric pressure
// gradient (calculateFluxesG + advect3D for every flow variable, then
// applyGenPressureGradient3D; src/gen3d.cpp:352-366) in one launch: P holds
// rho, p and the contravariant velocity (planes 0, 1, 2..4), MT the 9 metric
// planes, the cell sizes are 1.  The same operations as k_flux3d +
// k_advect3d + k_genpresgrad3d, so the result is bit for bit the reference
// sequence's.
void launch_gen_adv(const Geo &g, const Phys &ph, const fst_config &cf, const double *U, const double *P,
                    const double *MT, double *dvar, hipStream_t s, int buoy) {
  const Grid32 q = make_grid(g);
  const FC k = make_fc(ph, cf);
  constexpr int kchunk = 8;
  const Launch L = shape(g, kchunk);
  switch (cf.ns) {
    case 1: k_adv3d_gen<1><<<L.nb, NTHREADS, 0, s>>>(q, k, U, P, dvar, MT, buoy, kchunk, L.nbx, L.nby, L.nbz); break;
    case 2: k_adv3d_gen<2><<<L.nb, NTHREADS, 0, s>>>(q, k, U, P, dvar, MT, buoy, kchunk, L.nbx, L.nby, L.nbz); break;
    default: k_adv3d_gen<3><<<L.nb, NTHREADS, 0, s>>>(q, k, U, P, dvar, MT, buoy, kchunk, L.nbx, L.nby, L.nbz); break;
  }
}

namespace {

// calculateRhoPT3D (src/secondary.hpp:109-151) and computeGenVelocity3D
// (src/velocity.hpp:23-45) of cell c of U: rho and p as k_eos_vel3d computes
// them and the contravariant velocity as k_genvel does (the cartesian
// velocity k_eos_vel3d writes first is only an intermediate of the reference
// sequence), into the 5 planes of P.  u: the cell's flow variables.  F32
// (FP32 form): the layout of the cartesian FP32 forms, p as doubles at P, then
// rho and the contravariant velocity rounded to floats.
template <int NS, bool F32, typename MTT>
__device__ __forceinline__ void eos_genvel(const FC &k, long long N, long long c, const double *u,
                                           const MTT *__restrict__ MT, double *__restrict__ P) {
  double rr = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) rr = rr + u[4 + s];
  double Cp = 0, Cv = 0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double Y = u[4 + s] / rr;
    Cp = Cp + Y * k.cpf[s];
    Cv = Cv + Y * k.cvf[s];
  }
  const double gamma = Cp / Cv;
  const double m0 = u[0], m1 = u[1], m2 = u[2];
  const double p = (gamma - 1) * (u[3] - (0.5 / rr) * (m0 * m0 + m1 * m1 + m2 * m2));
  double U3[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    U3[d] = (double)MT[c + N * (d + 0)] * m0 / rr + (double)MT[c + N * (d + 3)] * m1 / rr +
            (double)MT[c + N * (d + 6)] * m2 / rr;
  if constexpr (F32) {
    float *Pf = reinterpret_cast<float *>(P + N);
    P[c] = p;
    Pf[c] = (float)rr;
#pragma unroll
    for (int d = 0; d < 3; ++d) Pf[c + N * (1 + d)] = (float)U3[d];
  } else {
    P[c] = rr;
    P[c + N] = p;
#pragma unroll
    for (int d = 0; d < 3; ++d) P[c + N * (2 + d)] = U3[d];
  }
}

// eos_genvel over every cell (ghost_only = 0) or over the ghost shell only
// (after the gen_ext ghost fill rewrote it).  F32: float metrics (MT32).
template <int NS, bool F32>
__global__ void k_eos_genvel3d(const Grid32 q, const FC k, const double *__restrict__ U,
                               const double *__restrict__ MT, const float *__restrict__ MT32,
                               double *__restrict__ P, int ghost_only) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t cell = t;
  if (ghost_only) {
    if (!ghost_cell(q, t, &cell)) return;
  } else if (t >= q.N) {
    return;
  }
  const long long N = q.N, c = cell;
  double u[4 + NS];
#pragma unroll
  for (int v = 0; v < 4 + NS; ++v) u[v] = U[c + N * v];
  if constexpr (F32) eos_genvel<NS, F32>(k, N, c, u, MT32, P);
  else eos_genvel<NS, F32>(k, N, c, u, MT, P);
}

// The Runge-Kutta stage update of a fused gen3d stage over every cell
// (rkAdvance, src/rk.cpp:28-72: dvar is zero on the ghosts, where the gen
// modules never write it), Uo = upd(U, Un, dvar), then eos_genvel of the new
// state: P describes Uo for the next stage.
// F32: the partial RHS as floats, float metrics (MT32).
template <int NS, bool F32>
__global__ void k_rk_eos_gen(const Grid32 q, const FC k, const double *__restrict__ U,
                             const double *__restrict__ Un, const double *__restrict__ dvar, double *__restrict__ Uo,
                             const double *__restrict__ MT, const float *__restrict__ MT32, double *__restrict__ P,
                             Upd up) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= q.N) return;
  const int i = (int)(t % (uint32_t)q.ngi), j = (int)((t / (uint32_t)q.ngi) % (uint32_t)q.ngj), kk = (int)(t / q.s2);
  const bool interior = i >= q.ng && i < q.ng + q.nci && j >= q.ng && j < q.ng + q.ncj && kk >= q.ng &&
                        kk < q.ng + q.nck;
  const long long N = q.N, c = t;
  double u[4 + NS];
#pragma unroll
  for (int v = 0; v < 4 + NS; ++v) {
    const long long e = c + N * v;
    const double us = up.form != 2 ? U[e] : 0.0;
    const double ub = up.form != 1 ? Un[e] : 0.0;
    u[v] = upd_val(up, us, ub, interior ? ld_dv<F32>(dvar, (uint32_t)e) : 0.0);
    Uo[e] = u[v];
  }
  if constexpr (F32) eos_genvel<NS, F32>(k, N, c, u, MT32, P);
  else eos_genvel<NS, F32>(k, N, c, u, MT, P);
}

// visc: the gen_ext viscous terms (launch_viscgen3d on the stage input) after
// the advection kernel; they read rho from P's plane 0 in the FP64 form and
// sum it from the species densities (the EOS's sum) in the FP32 form, whose
// rho plane is rounded.  F32: binary32 fluxes (k_adv3d_sgen, 11-row tiles),
// float primitive planes and binary32 viscous faces on float velocities,
// mass fractions and metrics (mt32, which also gives the contravariant
// velocity), the partial RHS as floats; the pressure gradient, the EOS and
// the RK accumulation stay FP64.
template <int NS, bool F32>
void gen_stage(const Grid32 &q, const FC &k, int kchunk, const double *Uin, const double *Un, double *Uout,
               const double *MT, double *P, double *dvar, Upd up, int eos, int buoy, const Geo &g, const Phys &ph,
               const Bufs &b, bool visc, const float *mt32, uint32_t ngh, Prof *prof, hipStream_t s) {
  if (eos == 2) k_eos_genvel3d<NS, F32><<<(q.N + 255) / 256, 256, 0, s>>>(q, k, Uin, MT, mt32, P, 0);
  else if (eos == 1) k_eos_genvel3d<NS, F32><<<(ngh + 255) / 256, 256, 0, s>>>(q, k, Uin, MT, mt32, P, 1);
  if (prof) prof->mark(s);
  const Launch L = shape(g, kchunk);
  if constexpr (F32) {
    const int nbys = (g.ncj + TYS - 1) / TYS, nbs = L.nbx * nbys * L.nbz;
    k_adv3d_sgen<NS><<<nbs, NTS, 0, s>>>(q, k, Uin, P, dvar, MT, buoy, kchunk, L.nbx, nbys, L.nbz);
  } else {
    k_adv3d_gen<NS><<<L.nb, NTHREADS, 0, s>>>(q, k, Uin, P, dvar, MT, buoy, kchunk, L.nbx, L.nby, L.nbz);
  }
  if (visc) {
    Bufs vb = b;
    vb.var = const_cast<double *>(Uin);
    vb.rho = F32 ? nullptr : P;
    vb.dvar = dvar;
    launch_viscgen3d(g, ph, vb, s, F32 ? mt32 : nullptr);
  }
  if (prof) prof->mark(s);
  k_rk_eos_gen<NS, F32><<<(q.N + 255) / 256, 256, 0, s>>>(q, k, Uin, Un, dvar, Uout, MT, mt32, P, up);
}

}  // namespace

// Fused gen3d steps.  Per Runge-Kutta stage: the gen_ext ghost fill (when
// set) and the primitives of the cells it rewrote, k_adv3d_gen (fluxes,
// metric pressure gradient and the gen_ext buoyancy in one launch), the
// gen_ext viscous terms (when set), then k_rk_eos_gen (the stage update of every cell and the primitives of the new
// state, which the next stage reads).  Stage states as the cartesian fused
// path keeps them: RK2 U* in fb.stage, U^{n+1} in place over U^n (read
// pointwise); SSP-RK3 U1, U2 in fb.stage / fb.stage2.  Bit for bit the
// reference sequence (compute + rkAdvance) on every cell.  fb.valid: b.prim
// holds the primitives of var.
int fused_gen_steps(const Geo &g, const Phys &ph, const fst_config &cf, Bufs &b, FusedBufs &fb, int nsteps,
                    hipStream_t s, std::string *err) {
  if (!gen_adv_supported(cf, g) || !b.prim || !b.metrics) {
    *err = "fused gen3d path: configuration not supported";
    return FST_ERR_ARG;
  }
  const bool rk3 = rk_stages(cf) == 3;
  const size_t bytes = (size_t)g.N * g.nvt * sizeof(double);
  if (!fb.stage && hipMalloc((void **)&fb.stage, bytes) != hipSuccess) {
    *err = "fused gen3d path: device allocation failed";
    return FST_ERR_NOMEM;
  }
  if (rk3 && !fb.stage2 && hipMalloc((void **)&fb.stage2, bytes) != hipSuccess) {
    *err = "fused gen3d path: device allocation failed";
    return FST_ERR_NOMEM;
  }
  const Grid32 q = make_grid(g);
  const FC k = make_fc(ph, cf);
  const bool f32 = cf.precision == FST_FP32_BULK;
  // planes per block: 8 (FP64), 32 (FP32 form; 11.73, 11.44, 11.38 ms per
  // stage at 8, 16, 32 on the 512 x 512 x 256 terrain grid, 3 species)
  int kchunk = f32 ? 32 : 8;
  if (const char *ev = getenv("FST_KCHUNK")) kchunk = std::max(1, atoi(ev));  // tuning override
  const int buoy = cf.gen_ext && cf.buoyancy;
  const bool visc = cf.gen_ext && cf.visc && b.stress;
  if (f32 && !fb.mt32_ok) {  // float metrics of the FP32 form (cleared by preSim and uploads)
    if (!fb.mt32 && hipMalloc((void **)&fb.mt32, (size_t)g.N * 9 * sizeof(float)) != hipSuccess) {
      *err = "fused gen3d path: device allocation failed";
      return FST_ERR_NOMEM;
    }
    launch_metrics_f32(g, b.metrics, fb.mt32, s);
    fb.mt32_ok = true;
  }
  const uint32_t ngh = 2u * q.ng * q.s2 + 2u * q.ng * (uint32_t)q.ngi * q.nck + 2u * q.ng * (uint32_t)q.ncj * q.nck;
  Prof *prof = b.prof && b.prof->on ? b.prof : nullptr;
  if (prof) {
    prof->name = "gen3d stage: [ghost fill + ghost primitives] + k_adv3d_gen [+ viscous] + k_rk_eos_gen";
    prof->per = 4;  // begin | ghost fill + primitives | advection | update + primitives
  }
  auto stage_run = [&](double *Uin, double *Uout, Upd up) {
    if (prof) prof->begin(s);
    if (cf.gen_ext) launch_apply_bcs(g, cf, Uin, s);  // src/cart3d.cpp:221, as gen_ext applies it
    // P describes Uin unless the fill rewrote its ghosts (or var changed)
    const int eos = !fb.valid ? 2 : (cf.gen_ext ? 1 : 0);
#define FST_GEN_STAGE(NS_, F_)                                                                                  \
  gen_stage<NS_, F_>(q, k, kchunk, Uin, b.var, Uout, b.metrics, b.prim, b.dvar, up, eos, buoy, g, ph, b, visc,   \
                     fb.mt32, ngh, prof, s)
    switch (cf.ns * 2 + (f32 ? 1 : 0)) {
      case 2: FST_GEN_STAGE(1, false); break;
      case 3: FST_GEN_STAGE(1, true); break;
      case 4: FST_GEN_STAGE(2, false); break;
      case 5: FST_GEN_STAGE(2, true); break;
      case 6: FST_GEN_STAGE(3, false); break;
      default: FST_GEN_STAGE(3, true); break;
    }
#undef FST_GEN_STAGE
    fb.valid = true;
    if (prof) prof->end(s);
  };
  for (int t = 0; t < nsteps; ++t) {
    if (!rk3) {
      stage_run(b.var, fb.stage, Upd{0.5f * cf.dt, 0.0, 0.0, 1});
      stage_run(fb.stage, b.var, Upd{cf.dt, 0.0, 0.0, 2});
    } else {
      stage_run(b.var, fb.stage, Upd{cf.dt, 0.0, 0.0, 1});
      stage_run(fb.stage, fb.stage2, Upd{cf.dt, RK3_A2, RK3_B2, 3});
      stage_run(fb.stage2, b.var, Upd{cf.dt, RK3_A3, RK3_B3, 3});
    }
  }
  return hipGetLastError() == hipSuccess ? FST_OK : FST_ERR_HIP;
}

#else
#ifndef FST_FAST64_TU
bool gen_adv_supported(const fst_config &cf, const Geo &g) {
  // 32-bit element offsets: the largest array (nvt planes, or the 5 primitive
  // planes) below 2^32 elements
  return cf.ndim == 3 && cf.grid == 1 && cf.ns >= 1 && cf.ns <= 3 && cf.scheme == FST_WENO5 &&
         (double)g.N * std::max(g.nvt, 5) < 4294967296.0;
}

bool fused_supported(const fst_config &cf) {
  if (!(cf.ndim == 3 && cf.grid == 0 &&
        (cf.precision == FST_FP64 || cf.precision == FST_FP32_BULK || cf.precision == FST_FP64_FAST) &&
        cf.scheme == FST_WENO5 && cf.ns >= 1 && cf.ns <= 3 && (!cf.learned_c || cf.ceq)))
    return false;
  // 32-bit element offsets: the largest array (nvt planes) below 2^32 elements
  const double N = (double)(cf.nci + 2 * cf.ng) * (cf.ncj + 2 * cf.ng) * (cf.nck + 2 * cf.ng);
  const int nvt = 4 + cf.ns + (cf.ceq && !cf.learned_c ? 5 : 0);
  return N * std::max(nvt, 6 + cf.ns) < 4294967296.0;
}

#endif  // FST_FAST64_TU

#ifdef FST_FAST64_TU
int fused_steps_fast64(const Geo &g, const Phys &ph, const fst_config &cf, Bufs &b, FusedBufs &fb,
                       LearnedBufs *lb, int nsteps, hipStream_t s, std::string *err) {
#else
int fused_steps(const Geo &g, const Phys &ph, const fst_config &cf, Bufs &b, FusedBufs &fb, LearnedBufs *lb,
                int nsteps, hipStream_t s, std::string *err) {
  if (cf.precision == FST_FP64_FAST) return fused_steps_fast64(g, ph, cf, b, fb, lb, nsteps, s, err);
#endif
  StageFn sf;
  PrimFn pf;
  const bool lc = cf.learned_c != 0;
  if (!pick(cf.ns, cf.precision == FST_FP32_BULK, lc, &sf, &pf) || !fused_supported(cf)) {
    *err = "fused path: configuration not supported";
    return FST_ERR_ARG;
  }
  if (lc && (!lb || !lb->have_weights)) {
    *err = "learned C-indicator: no network weights set";
    return FST_ERR_STATE;
  }
  const int np = 6 + cf.ns;
  const bool rk3 = rk_stages(cf) == 3;
  const size_t bytes = (size_t)g.N * sizeof(double);
  if (!fb.stage) {
    if (hipMalloc((void **)&fb.stage, bytes * g.nvt) != hipSuccess ||
        hipMalloc((void **)&fb.prim, bytes * np * 2) != hipSuccess ||
        hipMalloc((void **)&fb.sc, sizeof(Scalars) * 2) != hipSuccess) {
      *err = "fused path: device allocation failed";
      return FST_ERR_NOMEM;
    }
    (void)hipMemsetAsync(fb.stage, 0, bytes * g.nvt, s);
    (void)hipMemsetAsync(fb.prim, 0, bytes * np * 2, s);
  }
  if (rk3 && !fb.stage2) {
    if (hipMalloc((void **)&fb.stage2, bytes * g.nvt) != hipSuccess) {
      *err = "fused path: device allocation failed";
      return FST_ERR_NOMEM;
    }
    (void)hipMemsetAsync(fb.stage2, 0, bytes * g.nvt, s);
  }
  const FC k = make_fc(ph, cf);
  const Grid32 q = make_grid(g);
  // planes per block: 8 for the FP64 kernels, 32 for the FP32 forms (k-chunk
  // sweeps at 256^3 over 4..32 planes)
  int kchunk = fb.kchunk > 0 ? fb.kchunk : (cf.precision == FST_FP32_BULK ? 32 : 8);
  if (const char *ev = getenv("FST_KCHUNK")) kchunk = std::max(1, atoi(ev));  // tuning override
  double *const Pset[2] = {fb.prim, fb.prim + (size_t)g.N * np};
  Scalars *const Sset[2] = {fb.sc, fb.sc + 1};
  Prof *prof = b.prof && b.prof->on ? b.prof : nullptr;
  if (prof) {
    prof->name = lc ? "stage: ghost fill + ghost primitives + [learned C fields, stage 1] + k_adv3d + k_vcr3d"
                    : "stage: ghost fill + ghost primitives + k_adv3d + k_vcr3d + k_ceq3d";
    prof->per = 6;  // begin | ghosts | learned fields | advection | viscous + RK | C-equation
  }
  const float *CH = lc ? lb->chat : nullptr, *DR = lc ? lb->dir : nullptr;
  if (!fb.valid) {  // primitives and maxima of var, all cells
    fb.cur = 0;
    k_clear<<<1, 64, 0, s>>>(Sset[0]);
    pf(q, k, b.var, Pset[0], Sset[0], 0, nullptr, s, false);
    fb.valid = true;
  }
  // learned FP32 form (RK2): the stage state U* is stored as floats (it only
  // feeds the binary32 right-hand side of stage 2; U^n and U^{n+1} stay FP64)
  const bool fs = lc && cf.precision == FST_FP32_BULK && !rk3;
  // stage s: input Uin (ghosts filled here), update form up, output Uout;
  // P[cur]/S[cur] describe Uin on entry, P[cur^1]/S[cur^1] receive Uout's
  auto stage_run = [&](int st, const double *Uin, double *Uout, Upd up, bool fin, bool fout) -> int {
    double *Pi = Pset[fb.cur], *Po = Pset[fb.cur ^ 1];
    Scalars *Si = Sset[fb.cur], *So = Sset[fb.cur ^ 1];
    if (prof) prof->begin(s);
    if (fin) launch_apply_bcs(g, cf, reinterpret_cast<float *>(const_cast<double *>(Uin)), s);
    else launch_apply_bcs(g, cf, const_cast<double *>(Uin), s);  // src/cart3d.cpp:221
    pf(q, k, Uin, Pi, nullptr, 1, So, s, fin);                     // ghost primitives; clears So
    if (prof) prof->mark(s);
    if (lc && st == 1) {  // C_hat and n of U^n, used by every stage of the step
      const int rc = learned_fields(g, cf, ph, Pi, cf.precision == FST_FP32_BULK, *lb, s);
      if (rc) {
        *err = "learned C-indicator: launch failed";
        return rc;
      }
    }
    if (prof) prof->mark(s);
    sf(q, g, k, cf, Uin, b.var, Uout, Pi, Po, b.dvar, CH, DR, lc ? lb->sc : Si, So, up, st, kchunk, prof, s,
       fs && (fout || fin));
    fb.cur ^= 1;
    return FST_OK;
  };
  for (int t = 0; t < nsteps; ++t) {
    int rc;
    if (!rk3) {
      // rkAdvance: U* = U^n + dt/2 R(U^n) (fb.stage), then U^{n+1} = U^n + dt R(U*)
      // written in place into b.var (read pointwise)
      if ((rc = stage_run(1, b.var, fb.stage, Upd{0.5f * cf.dt, 0.0, 0.0, 1}, false, fs))) return rc;
      if (prof) prof->end(s);
      if ((rc = stage_run(2, fb.stage, b.var, Upd{cf.dt, 0.0, 0.0, 2}, fs, false))) return rc;
      if (prof) prof->end(s);
    } else {
      // SSP-RK3: U1 (fb.stage), U2 (fb.stage2), U^{n+1} in place into b.var
      if ((rc = stage_run(1, b.var, fb.stage, Upd{cf.dt, 0.0, 0.0, 1}, false, false))) return rc;
      if (prof) prof->end(s);
      if ((rc = stage_run(2, fb.stage, fb.stage2, Upd{cf.dt, RK3_A2, RK3_B2, 3}, false, false))) return rc;
      if (prof) prof->end(s);
      const Upd u3{cf.dt, RK3_A3, RK3_B3, 3};
      if ((rc = stage_run(3, fb.stage2, b.var, u3, false, false))) return rc;
      const uint32_t ngh = 2u * q.ng * q.s2 + 2u * q.ng * (uint32_t)q.ngi * q.nck + 2u * q.ng * (uint32_t)q.ncj * q.nck;
      k_rk_ghost<<<(ngh + 255) / 256, 256, 0, s>>>(q, fb.stage2, b.var, u3);
      if (prof) prof->end(s);
    }
  }
  return FST_OK;
}

#ifndef FST_FAST64_TU
int fused_primitives(const Geo &g, const Phys &ph, const fst_config &cf, const double *U, double *P, hipStream_t s) {
  StageFn sf;
  PrimFn pf;
  // 32-bit element offsets in k_prims3d: 6 + ns planes of N cells below 2^32
  if (!pick(cf.ns, false, false, &sf, &pf) || g.ndim != 3 || cf.grid != 0 ||
      (double)g.N * (6 + cf.ns) >= 4294967296.0)
    return FST_ERR_ARG;
  const FC k = make_fc(ph, cf);
  const Grid32 q = make_grid(g);
  pf(q, k, U, P, nullptr, 0, nullptr, s, false);
  return hipGetLastError() == hipSuccess ? FST_OK : FST_ERR_HIP;
}

void fused_free(FusedBufs &fb) {
  if (fb.stage) (void)hipFree(fb.stage);
  if (fb.stage2) (void)hipFree(fb.stage2);
  if (fb.prim) (void)hipFree(fb.prim);
  if (fb.sc) (void)hipFree(fb.sc);
  if (fb.mt32) (void)hipFree(fb.mt32);
  fb.stage = fb.stage2 = fb.prim = nullptr;
  fb.sc = nullptr;
  fb.mt32 = nullptr;
  fb.mt32_ok = false;
}
#endif  // FST_FAST64_TU

#endif  // FST_GEN3D_TU

}  // namespace fst
