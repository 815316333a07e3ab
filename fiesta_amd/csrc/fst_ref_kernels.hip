// Reference-sequence kernels: one HIP kernel per reference Kokkos kernel, same
// ranges, same arithmetic order (compiled with -ffp-contract=off).  They back
// fst_compute()/fst_rk_stage() (the rk_func::compute contract, dvar written
// to HBM) and every 2D configuration.  The fused 3D stage kernel lives in
// fst_fused3d.hip and must agree with these bit for bit.
#include <algorithm>
#include <cfloat>

#include "fst_common.hpp"
#include "fst_fused.hpp"
#include "fst_kernels.hpp"

namespace fst {

namespace {

struct Range {
  int i0, j0, k0, ni, nj, nk;
};

__device__ __forceinline__ bool decode(const Range &r, int &i, int &j, int &k) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long n = (long long)r.ni * r.nj * r.nk;
  if (t >= n) return false;
  i = r.i0 + (int)(t % r.ni);
  t /= r.ni;
  j = r.j0 + (int)(t % r.nj);
  k = r.k0 + (int)(t / r.nj);
  return true;
}

inline Range ghost_range(const Geo &g) { return {0, 0, 0, g.ngi, g.ngj, g.ngk}; }
inline Range cell_range(const Geo &g) {
  if (g.ndim == 2) return {g.ng, g.ng, 0, g.nci, g.ncj, 1};
  return {g.ng, g.ng, g.ng, g.nci, g.ncj, g.nck};
}
// [ng-1, n-ng) in every active dimension (weno/face/tile range)
inline Range face_range(const Geo &g) {
  if (g.ndim == 2) return {g.ng - 1, g.ng - 1, 0, g.nci + 1, g.ncj + 1, 1};
  return {g.ng - 1, g.ng - 1, g.ng - 1, g.nci + 1, g.ncj + 1, g.nck + 1};
}

inline dim3 grid_for(const Range &r, int bs = 256) {
  long long n = (long long)r.ni * r.nj * r.nk;
  return dim3((unsigned)((n + bs - 1) / bs));
}

// ------------------------------------------------------------- 3D

// calculateRhoPT3D + computeVelocity3D (src/secondary.hpp:109-151,
// src/velocity.hpp:68-83), fused: both are pointwise over the ghost range.
__global__ void k_eos_vel3d(Geo g, Phys ph, Range r, const double *__restrict__ var, double *__restrict__ p,
                            double *__restrict__ rho, double *__restrict__ T, double *__restrict__ vel) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long c = g.id(i, j, k), N = g.N;
  double rr = 0.0;
  for (int s = 0; s < g.ns; ++s) rr = rr + var[c + N * (4 + s)];
  double Cp = 0, Cv = 0;
  for (int s = 0; s < g.ns; ++s) {
    double Y = var[c + N * (4 + s)] / rr;
    Cp = Cp + Y * ph.cpf[s];
    Cv = Cv + Y * ph.cvf[s];
  }
  double gamma = Cp / Cv;
  double m0 = var[c], m1 = var[c + N], m2 = var[c + 2 * N];
  double pp = (gamma - 1) * (var[c + 3 * N] - (0.5 / rr) * (m0 * m0 + m1 * m1 + m2 * m2));
  rho[c] = rr;
  p[c] = pp;
  T[c] = pp / ((Cp - Cv) * rr);
  vel[c] = m0 / rr;
  vel[c + N] = m1 / rr;
  vel[c + 2 * N] = m2 / rr;
}

// calculateFluxesG, src/flux.hpp:266-309 over the tile range
__global__ void k_flux3d(Geo g, Phys ph, Range r, int v, const double *__restrict__ var, const double *__restrict__ p,
                         const double *__restrict__ vel, double *__restrict__ fx, double *__restrict__ fy,
                         double *__restrict__ fz) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const double *q = var + N * v;
  const double isE = (double)(v == 3);
  auto Q = [&](int a, int b, int cc) {
    long long c = g.id(a, b, cc);
    return q[c] + isE * p[c];
  };
  auto U = [&](int a, int b, int cc, int d) { return vel[g.id(a, b, cc) + N * d]; };
  const long long c = g.id(i, j, k);
  double f1, f2, f3, f4, f5, f6, ur;
  f1 = Q(i + 3, j, k); f2 = Q(i + 2, j, k); f3 = Q(i + 1, j, k);
  f4 = Q(i, j, k);     f5 = Q(i - 1, j, k); f6 = Q(i - 2, j, k);
  ur = (-U(i + 2, j, k, 0) + 7.0f * U(i + 1, j, k, 0) + 7.0f * U(i, j, k, 0) - U(i - 1, j, k, 0)) / 12.0f;
  fx[c] = (ur < 0.0) ? ur * weno5_f32c(f1, f2, f3, f4, f5) / ph.dx : ur * weno5_f32c(f6, f5, f4, f3, f2) / ph.dx;
  f1 = Q(i, j + 3, k); f2 = Q(i, j + 2, k); f3 = Q(i, j + 1, k);
  f4 = Q(i, j, k);     f5 = Q(i, j - 1, k); f6 = Q(i, j - 2, k);
  ur = (-U(i, j + 2, k, 1) + 7.0f * U(i, j + 1, k, 1) + 7.0f * U(i, j, k, 1) - U(i, j - 1, k, 1)) / 12.0f;
  fy[c] = (ur < 0.0) ? ur * weno5_f32c(f1, f2, f3, f4, f5) / ph.dy : ur * weno5_f32c(f6, f5, f4, f3, f2) / ph.dy;
  f1 = Q(i, j, k + 3); f2 = Q(i, j, k + 2); f3 = Q(i, j, k + 1);
  f4 = Q(i, j, k);     f5 = Q(i, j, k - 1); f6 = Q(i, j, k - 2);
  ur = (-U(i, j, k + 2, 2) + 7.0f * U(i, j, k + 1, 2) + 7.0f * U(i, j, k, 2) - U(i, j, k - 1, 2)) / 12.0f;
  fz[c] = (ur < 0.0) ? ur * weno5_f32c(f1, f2, f3, f4, f5) / ph.dz : ur * weno5_f32c(f6, f5, f4, f3, f2) / ph.dz;
}

// advect3D, src/advect.hpp:39-61
__global__ void k_advect3d(Geo g, Range r, int v, const double *__restrict__ fx, const double *__restrict__ fy,
                           const double *__restrict__ fz, double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long c = g.id(i, j, k);
  double res = -((fx[c] - fx[g.id(i - 1, j, k)]) + (fy[c] - fy[g.id(i, j - 1, k)]) + (fz[c] - fz[g.id(i, j, k - 1)]));
  dvar[c + g.N * v] = res;
}

// applyPressureGradient3D (dz := cd(2)), src/presgrad.hpp:44-71, then
// computeBuoyancy3D, src/buoyancy.hpp:22-47 (both pointwise in dvar)
__global__ void k_presgrad_buoy3d(Geo g, Phys ph, Range r, int buoyancy, const double *__restrict__ var,
                                  const double *__restrict__ p, const double *__restrict__ rho,
                                  double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long c = g.id(i, j, k), N = g.N;
  const double dx = ph.cd[1], dy = ph.cd[2], dz = ph.cd[2];
  auto P = [&](int a, int b, int cc) { return p[g.id(a, b, cc)]; };
  double dxp = (P(i - 2, j, k) - 8.0 * P(i - 1, j, k) + 8.0 * P(i + 1, j, k) - P(i + 2, j, k)) / (12.0 * dx);
  double dyp = (P(i, j - 2, k) - 8.0 * P(i, j - 1, k) + 8.0 * P(i, j + 1, k) - P(i, j + 2, k)) / (12.0 * dy);
  double dzp = (P(i, j, k - 2) - 8.0 * P(i, j, k - 1) + 8.0 * P(i, j, k + 1) - P(i, j, k + 2)) / (12.0 * dz);
  dvar[c] = dvar[c] - dxp;
  dvar[c + N] = dvar[c + N] - dyp;
  dvar[c + 2 * N] = dvar[c + 2 * N] - dzp;
  if (buoyancy) {
    double vv = var[c + N] / rho[c];
    double rhop = rho[c] - ph.rhoRef;
    double f = -ph.g * rhop;
    dvar[c + N] += f;
    dvar[c + 3 * N] += vv * f;
  }
}

// calculateStressTensor{x,y,z}3dv, src/viscosity.hpp:416-624, weno range
__global__ void k_stress3d(Geo g, Phys ph, Range r, int dir, const double *__restrict__ var,
                           const double *__restrict__ rho, const double *__restrict__ vel,
                           double *__restrict__ stress) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const double dx = ph.cd[1], dy = ph.cd[2], dz = ph.cd[3];
  auto u = [&](int a, int b, int cc, int d) { return vel[g.id(a, b, cc) + N * d]; };
  const int ip = dir == 0, jp = dir == 1, kp = dir == 2;
  const long long c = g.id(i, j, k), cn = g.id(i + ip, j + jp, k + kp);
  double mu = 0.0;
  for (int s = 0; s < g.ns; ++s)
    mu += ((var[cn + N * (4 + s)] / rho[cn] + var[c + N * (4 + s)] / rho[c]) / 2.0) * ph.mus[s];
  double dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz;
  if (dir == 0) {
    dudx = (u(i + 1, j, k, 0) - u(i, j, k, 0)) / dx;
    dvdx = (u(i + 1, j, k, 1) - u(i, j, k, 1)) / dx;
    dwdx = (u(i + 1, j, k, 2) - u(i, j, k, 2)) / dx;
    dudy = ((u(i + 1, j + 1, k, 0) + u(i, j + 1, k, 0)) - (u(i + 1, j - 1, k, 0) + u(i, j - 1, k, 0))) / (4 * dy);
    dvdy = ((u(i + 1, j + 1, k, 1) + u(i, j + 1, k, 1)) - (u(i + 1, j - 1, k, 1) + u(i, j - 1, k, 1))) / (4 * dy);
    dwdy = ((u(i + 1, j + 1, k, 2) + u(i, j + 1, k, 2)) - (u(i + 1, j - 1, k, 2) + u(i, j - 1, k, 2))) / (4 * dy);
    dudz = ((u(i + 1, j, k + 1, 0) + u(i, j, k + 1, 0)) - (u(i + 1, j, k - 1, 0) + u(i, j, k - 1, 0))) / (4 * dz);
    dvdz = ((u(i + 1, j, k + 1, 1) + u(i, j, k + 1, 1)) - (u(i + 1, j, k - 1, 1) + u(i, j, k - 1, 1))) / (4 * dz);
    dwdz = ((u(i + 1, j, k + 1, 2) + u(i, j, k + 1, 2)) - (u(i + 1, j, k - 1, 2) + u(i, j, k - 1, 2))) / (4 * dz);
  } else if (dir == 1) {
    dudx = ((u(i + 1, j + 1, k, 0) + u(i + 1, j, k, 0)) - (u(i - 1, j + 1, k, 0) + u(i - 1, j, k, 0))) / (4 * dx);
    dvdx = ((u(i + 1, j + 1, k, 1) + u(i + 1, j, k, 1)) - (u(i - 1, j + 1, k, 1) + u(i - 1, j, k, 1))) / (4 * dx);
    dwdx = ((u(i + 1, j + 1, k, 2) + u(i + 1, j, k, 2)) - (u(i - 1, j + 1, k, 2) + u(i - 1, j, k, 2))) / (4 * dx);
    dudy = (u(i, j + 1, k, 0) - u(i, j, k, 0)) / dy;
    dvdy = (u(i, j + 1, k, 1) - u(i, j, k, 1)) / dy;
    dwdy = (u(i, j + 1, k, 2) - u(i, j, k, 2)) / dy;
    dudz = ((u(i, j + 1, k + 1, 0) + u(i, j, k + 1, 0)) - (u(i, j + 1, k - 1, 0) + u(i, j, k - 1, 0))) / (4 * dz);
    dvdz = ((u(i, j + 1, k + 1, 1) + u(i, j, k + 1, 1)) - (u(i, j + 1, k - 1, 1) + u(i, j, k - 1, 1))) / (4 * dz);
    dwdz = ((u(i, j + 1, k + 1, 2) + u(i, j, k + 1, 2)) - (u(i, j + 1, k - 1, 2) + u(i, j, k - 1, 2))) / (4 * dz);
  } else {
    dudx = ((u(i + 1, j, k + 1, 0) + u(i + 1, j, k, 0)) - (u(i - 1, j, k + 1, 0) + u(i - 1, j, k, 0))) / (4 * dx);
    dvdx = ((u(i + 1, j, k + 1, 1) + u(i + 1, j, k, 1)) - (u(i - 1, j, k + 1, 1) + u(i - 1, j, k, 1))) / (4 * dx);
    dwdx = ((u(i + 1, j, k + 1, 2) + u(i + 1, j, k, 2)) - (u(i - 1, j, k + 1, 2) + u(i - 1, j, k, 2))) / (4 * dx);
    dudy = ((u(i, j + 1, k + 1, 0) + u(i, j + 1, k, 0)) - (u(i, j - 1, k + 1, 0) + u(i, j - 1, k, 0))) / (4 * dy);
    dvdy = ((u(i, j + 1, k + 1, 1) + u(i, j + 1, k, 1)) - (u(i, j - 1, k + 1, 1) + u(i, j - 1, k, 1))) / (4 * dy);
    dwdy = ((u(i, j + 1, k + 1, 2) + u(i, j + 1, k, 2)) - (u(i, j - 1, k + 1, 2) + u(i, j - 1, k, 2))) / (4 * dy);
    dudz = (u(i, j, k + 1, 0) - u(i, j, k, 0)) / dz;
    dvdz = (u(i, j, k + 1, 1) - u(i, j, k, 1)) / dz;
    dwdz = (u(i, j, k + 1, 2) - u(i, j, k, 2)) / dz;
  }
  stress[c + N * 0] = (2.0 / 3.0) * mu * (2.0 * dudx - dvdy - dwdz);
  stress[c + N * 1] = (2.0 / 3.0) * mu * (2.0 * dvdy - dudx - dwdz);
  stress[c + N * 2] = (2.0 / 3.0) * mu * (2.0 * dwdz - dudx - dvdy);
  stress[c + N * 3] = mu * (dudy + dvdx);
  stress[c + N * 4] = mu * (dudz + dwdx);
  stress[c + N * 5] = mu * (dvdz + dwdy);
}

// applyViscousTerm{x,y,z}3dv, src/viscosity.hpp:460-663
__global__ void k_applyvisc3d(Geo g, Phys ph, Range r, int dir, const double *__restrict__ vel,
                              const double *__restrict__ stress, double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const int s0 = dir == 0 ? 0 : (dir == 1 ? 3 : 4);
  const int s1 = dir == 0 ? 3 : (dir == 1 ? 1 : 5);
  const int s2 = dir == 0 ? 4 : (dir == 1 ? 5 : 2);
  const double d = ph.cd[1 + dir];
  const int ih = dir == 0, jh = dir == 1, kh = dir == 2;
  const long long c = g.id(i, j, k), cp = g.id(i + ih, j + jh, k + kh), cm = g.id(i - ih, j - jh, k - kh);
  double ur = (vel[cp] + vel[c]) / 2.0;
  double ul = (vel[cm] + vel[c]) / 2.0;
  double vr = (vel[cp + N] + vel[c + N]) / 2.0;
  double vl = (vel[cm + N] + vel[c + N]) / 2.0;
  double wr = (vel[cp + 2 * N] + vel[c + 2 * N]) / 2.0;
  double wl = (vel[cm + 2 * N] + vel[c + 2 * N]) / 2.0;
  auto S = [&](long long q, int s) { return stress[q + N * s]; };
  double a = (S(c, s0) - S(cm, s0)) / d;
  double b = (S(c, s1) - S(cm, s1)) / d;
  double cc = (S(c, s2) - S(cm, s2)) / d;
  double dd = (ur * S(c, s0) - ul * S(cm, s0)) / d + (vr * S(c, s1) - vl * S(cm, s1)) / d +
              (wr * S(c, s2) - wl * S(cm, s2)) / d;
  dvar[c] += a;
  dvar[c + N] += b;
  dvar[c + 2 * N] += cc;
  dvar[c + 3 * N] += dd;
}

// calculateRhoGrad (src/ceq3d.hpp:95-187) fused with the two per-stage maxima
// maxWaveSpeed (:25-63) and maxGradFunctor(var, nv+1) (:82-93); all cell range
__global__ void k_rhograd_max3d(Geo g, Phys ph, Range r, const double *__restrict__ var,
                                const double *__restrict__ p, const double *__restrict__ rho,
                                const double *__restrict__ vel, double *__restrict__ gradRho, Scalars *sc) {
  int i, j, k;
  double ms = 0.0, mc = 0.0;
  if (decode(r, i, j, k)) {
    const long long N = g.N;
    const double dx = ph.dx, dy = ph.dy, dz = ph.dz;
    auto R = [&](int a, int b, int cc) { return rho[g.id(a, b, cc)]; };
    auto V = [&](int a, int b, int cc, int w) { return vel[g.id(a, b, cc) + N * w]; };
    auto E = [&](int a, int b, int cc) {
      long long q = g.id(a, b, cc);
      return var[q + 3 * N] / rho[q] -
             0.5 * rho[q] * (vel[q] * vel[q] + vel[q + N] * vel[q + N] + vel[q + 2 * N] * vel[q + 2 * N]);
    };
    double dxr = (R(i - 2, j, k) - 8.0 * R(i - 1, j, k) + 8.0 * R(i + 1, j, k) - R(i + 2, j, k)) / (12.0 * dx);
    double dyr = (R(i, j - 2, k) - 8.0 * R(i, j - 1, k) + 8.0 * R(i, j + 1, k) - R(i, j + 2, k)) / (12.0 * dy);
    double dzr = (R(i, j, k - 2) - 8.0 * R(i, j, k - 1) + 8.0 * R(i, j, k + 1) - R(i, j, k + 2)) / (12.0 * dz);
    double dxe = (E(i - 2, j, k) - 8.0 * E(i - 1, j, k) + 8.0 * E(i + 1, j, k) - E(i + 2, j, k)) / (12.0 * dx);
    double dye = (E(i, j - 2, k) - 8.0 * E(i, j - 1, k) + 8.0 * E(i, j + 1, k) - E(i, j + 2, k)) / (12.0 * dy);
    double dze = (E(i, j, k - 2) - 8.0 * E(i, j, k - 1) + 8.0 * E(i, j, k + 1) - E(i, j, k + 2)) / (12.0 * dz);
    double dxu = (V(i - 2, j, k, 0) - 8.0 * V(i - 1, j, k, 0) + 8.0 * V(i + 1, j, k, 0) - V(i + 2, j, k, 0)) / (12.0 * dx);
    double dyv = (V(i, j - 2, k, 1) - 8.0 * V(i, j - 1, k, 1) + 8.0 * V(i, j + 1, k, 1) - V(i, j + 2, k, 1)) / (12.0 * dy);
    double dzw = (V(i, j, k - 2, 2) - 8.0 * V(i, j, k - 1, 2) + 8.0 * V(i, j, k + 1, 2) - V(i, j, k + 2, 2)) / (12.0 * dz);
    double rgrad = sqrt(dxr * dxr + dyr * dyr + dzr * dzr);
    double divu = dxu + dyv + dzw;
    double dnednr = (dxr * dxe + dyr * dye + dzr * dze) * (dxr * dxr + dyr * dyr + dzr * dzr);
    int ind = dnednr < 0.0 ? 1 : 0;
    const long long c = g.id(i, j, k);
    gradRho[c] = (divu < 0) ? (1 - ind) * rgrad : 0;
    gradRho[c + N] = ind * rgrad;
    gradRho[c + 2 * N] = dxr;
    gradRho[c + 3 * N] = dyr;
    gradRho[c + 4 * N] = dzr;
    // maxWaveSpeed
    double Cp = 0, Cv = 0;
    for (int s = 0; s < g.ns; ++s) {
      double Y = var[c + N * (4 + s)] / rho[c];
      Cp = Cp + Y * ph.cpf[s];
      Cv = Cv + Y * ph.cvf[s];
    }
    double gamma = Cp / Cv;
    double a = sqrt(gamma * p[c] / rho[c]);
    ms = a + sqrt(var[c] * var[c] + var[c + N] * var[c + N] + var[c + 2 * N] * var[c + 2 * N]) / rho[c];
    mc = fabs(var[c + N * (g.nv + 1)]);
  }
  double v2[2] = {ms, mc};
  unsigned long long *const d2[2] = {&sc->maxS, &sc->maxCh};
  block_max_atomic<2>(v2, d2);
}

// updateCeq, src/ceq3d.hpp:189-224 (maxS from the device reduction)
__global__ void k_updateceq3d(Geo g, Phys ph, Range r, const double *__restrict__ var,
                              const double *__restrict__ gradRho, const Scalars *sc, double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const double dx = ph.cd[1], dy = ph.cd[2], dz = ph.cd[3];
  const int nc = g.nv;
  const double maxS = as_double(sc->maxS);
  const double dxmag = sqrt(dx * dx + dy * dy + dz * dz);
  const long long c = g.id(i, j, k);
  for (int n = 0; n < 5; ++n) {
    const double *C = var + N * (nc + n);
    double lap = (C[g.id(i - 1, j, k)] - 2 * C[c] + C[g.id(i + 1, j, k)]) / dx +
                 (C[g.id(i, j - 1, k)] - 2 * C[c] + C[g.id(i, j + 1, k)]) / dy +
                 (C[g.id(i, j, k - 1)] - 2 * C[c] + C[g.id(i, j, k + 1)]) / dz;
    dvar[c + N * (nc + n)] = (maxS / (ph.eps * dxmag)) * (gradRho[c + N * n] - C[c]) + ph.kap * maxS * dxmag * lap;
  }
}

// calculateCeqFaces (src/ceq3d.hpp:226-296) and calculateCeqGrads (:298-332),
// fused per face (the product X*deriv is formed exactly as the two passes do)
__global__ void k_ceqfaces3d(Geo g, Phys ph, Range r, const double *__restrict__ var, const double *__restrict__ rho,
                             const double *__restrict__ vel, const Scalars *sc, double *__restrict__ mflux) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const int nv = g.nv;
  const double alpha = (ph.dxmag2 / (as_double(sc->maxCh) + 1.0e-6)) * ph.alpha;
  const double dx = ph.dx, dy = ph.dy, dz = ph.dz;
  auto vv = [&](int a, int b, int cc, int w) { return vel[g.id(a, b, cc) + N * w]; };
  const long long c = g.id(i, j, k);
  double X[3][3];
  for (int face = 0; face < 3; ++face) {
    const long long q = g.id(i + (face == 0), j + (face == 1), k + (face == 2));
    double rr = (rho[c] + rho[q]) / 2.0;
    double chat = (var[c + N * (nv + 1)] + var[q + N * (nv + 1)]) / 2.0;
    double cn[3];
    cn[0] = (var[c + N * (nv + 2)] + var[q + N * (nv + 2)]) / 2.0;
    cn[1] = (var[c + N * (nv + 3)] + var[q + N * (nv + 3)]) / 2.0;
    cn[2] = (var[c + N * (nv + 4)] + var[q + N * (nv + 4)]) / 2.0;
    double cmag = 1.0e-6;
    cmag += cn[0] * cn[0];
    cmag += cn[1] * cn[1];
    cmag += cn[2] * cn[2];
    for (int dir = 0; dir < 3; ++dir) {
      double dirac = (face == dir) ? 1.0 : 0.0;
      double M = dirac - (cn[face] * cn[dir]) / cmag;
      X[face][dir] = alpha * rr * chat * M;
    }
  }
  // mflux layout: ((face*3+dir)*3+w) planes of N
  auto MF = [&](int f, int d, int w) -> double & { return mflux[c + N * ((f * 3 + d) * 3 + w)]; };
  for (int w = 0; w < 3; ++w) {
    MF(0, 0, w) = X[0][0] * ((vv(i + 1, j, k, w) - vv(i, j, k, w)) / dx);
    MF(0, 1, w) = X[0][1] * (((vv(i, j + 1, k, w) + vv(i + 1, j + 1, k, w)) - (vv(i, j - 1, k, w) + vv(i + 1, j - 1, k, w))) / (4 * dy));
    MF(0, 2, w) = X[0][2] * (((vv(i, j, k + 1, w) + vv(i + 1, j, k + 1, w)) - (vv(i, j, k - 1, w) + vv(i + 1, j, k - 1, w))) / (4 * dz));
    MF(1, 0, w) = X[1][0] * (((vv(i + 1, j + 1, k, w) + vv(i + 1, j, k, w)) - (vv(i - 1, j + 1, k, w) + vv(i - 1, j, k, w))) / (4 * dx));
    MF(1, 1, w) = X[1][1] * ((vv(i, j + 1, k, w) - vv(i, j, k, w)) / dy);
    MF(1, 2, w) = X[1][2] * (((vv(i, j + 1, k + 1, w) + vv(i, j, k + 1, w)) - (vv(i, j + 1, k - 1, w) + vv(i, j, k - 1, w))) / (4 * dz));
    MF(2, 0, w) = X[2][0] * (((vv(i + 1, j, k + 1, w) + vv(i + 1, j, k, w)) - (vv(i - 1, j, k + 1, w) + vv(i - 1, j, k, w))) / (4 * dx));
    MF(2, 1, w) = X[2][1] * (((vv(i, j + 1, k + 1, w) + vv(i, j + 1, k, w)) - (vv(i, j - 1, k + 1, w) + vv(i, j - 1, k, w))) / (4 * dy));
    MF(2, 2, w) = X[2][2] * ((vv(i, j, k + 1, w) - vv(i, j, k, w)) / dz);
  }
}

// applyCeq, src/ceq3d.hpp:334-366
__global__ void k_applyceq3d(Geo g, Phys ph, Range r, const double *__restrict__ mflux, double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const double d[3] = {ph.dx, ph.dy, ph.dz};
  const long long c = g.id(i, j, k);
  for (int w = 0; w < 3; ++w) {
    double diffu = 0.0;
    for (int face = 0; face < 3; ++face) {
      const long long cm = g.id(i - (face == 0), j - (face == 1), k - (face == 2));
      for (int dir = 0; dir < 3; ++dir) {
        const long long pl = N * ((face * 3 + dir) * 3 + w);
        diffu += (mflux[c + pl] - mflux[cm + pl]) / d[face];
      }
    }
    dvar[c + N * w] += diffu;
  }
}

// ------------------------------------------------------------- 2D

// calculateRhoPT2D + computeVelocity2D, src/secondary.hpp:60-107, src/velocity.hpp:85-99
__global__ void k_eos_vel2d(Geo g, Phys ph, Range r, const double *__restrict__ var, double *__restrict__ p,
                            double *__restrict__ rho, double *__restrict__ T, double *__restrict__ vel) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long c = g.id(i, j, 0), N = g.N;
  double rr = 0.0;
  for (int s = 0; s < g.ns; ++s) rr = rr + var[c + N * (3 + s)];
  double Cp = 0, Cv = 0;
  for (int s = 0; s < g.ns; ++s) {
    double Y = var[c + N * (3 + s)] / rr;
    Cp = Cp + Y * ph.cpf[s];
    Cv = Cv + Y * ph.cvf[s];
  }
  double gamma = Cp / Cv;
  double m0 = var[c], m1 = var[c + N];
  double pp = (gamma - 1) * (var[c + 2 * N] - (0.5 / rr) * (m0 * m0 + m1 * m1));
  rho[c] = rr;
  p[c] = pp;
  T[c] = pp / ((Cp - Cv) * rr);
  vel[c] = m0 / rr;
  vel[c + N] = m1 / rr;
}

// computeFluxWeno2D / computeFluxCentered2D / computeFluxQuick2D
// (src/flux.hpp:23-221), face range
__global__ void k_flux2d(Geo g, Phys ph, Range r, int v, int scheme, const double *__restrict__ var,
                         const double *__restrict__ p, const double *__restrict__ rho,
                         const double *__restrict__ vel, double *__restrict__ fx, double *__restrict__ fy) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const double isE = (double)(v == 2);
  auto VAR = [&](int a, int b, int w) { return var[g.id(a, b, 0) + N * w]; };
  auto P = [&](int a, int b) { return p[g.id(a, b, 0)]; };
  auto R = [&](int a, int b) { return rho[g.id(a, b, 0)]; };
  auto U = [&](int a, int b, int w) { return vel[g.id(a, b, 0) + N * w]; };
  auto Q = [&](int a, int b) { return VAR(a, b, v) + isE * P(a, b); };
  const long long c = g.id(i, j, 0);
  if (scheme == FST_CENTERED4) {
    const double dx = ph.cd[1], dy = ph.cd[2];
    double ur = (0.0 - VAR(i + 2, j, 0) / R(i + 2, j) + 7.0 * VAR(i + 1, j, 0) / R(i + 1, j) +
                 7.0 * VAR(i, j, 0) / R(i, j) - VAR(i - 1, j, 0) / R(i - 1, j)) / 12.0;
    double vr = (0.0 - VAR(i, j + 2, 1) / R(i, j + 2) + 7.0 * VAR(i, j + 1, 1) / R(i, j + 1) +
                 7.0 * VAR(i, j, 1) / R(i, j) - VAR(i, j - 1, 1) / R(i, j - 1)) / 12.0;
    double px = (0.0 - P(i + 2, j) + 7.0 * P(i + 1, j) + 7.0 * P(i, j) - P(i - 1, j)) / 12.0;
    double py = (0.0 - P(i, j + 2) + 7.0 * P(i, j + 1) + 7.0 * P(i, j) - P(i, j - 1)) / 12.0;
    double x1 = (0.0 - VAR(i + 2, j, v) + 7.0 * VAR(i + 1, j, v) + 7.0 * VAR(i, j, v) - VAR(i - 1, j, v)) / 12.0;
    double y1 = (0.0 - VAR(i, j + 2, v) + 7.0 * VAR(i, j + 1, v) + 7.0 * VAR(i, j, v) - VAR(i, j - 1, v)) / 12.0;
    if (v == 2) {
      x1 = x1 + px;
      y1 = y1 + py;
    }
    fx[c] = ur * x1 / dx;
    fy[c] = vr * y1 / dy;
  } else if (scheme == FST_QUICK) {
    const double dx = ph.cd[1], dy = ph.cd[2];
    double f1, f2, f3;
    double ur = (-U(i + 2, j, 0) + 7.0 * U(i + 1, j, 0) + 7.0 * U(i, j, 0) - U(i - 1, j, 0)) / 12.0;
    if (ur < 0.0) { f1 = Q(i + 2, j); f2 = Q(i + 1, j); f3 = Q(i, j); }
    else          { f1 = Q(i - 1, j); f2 = Q(i, j);     f3 = Q(i + 1, j); }
    fx[c] = ur * (0.75 * f2 + 0.375 * f3 - 0.125 * f1) / dx;
    double vr = (-U(i, j + 2, 1) + 7.0 * U(i, j + 1, 1) + 7.0 * U(i, j, 1) - U(i, j - 1, 1)) / 12.0;
    if (vr < 0.0) { f1 = Q(i, j + 2); f2 = Q(i, j + 1); f3 = Q(i, j); }
    else          { f1 = Q(i, j - 1); f2 = Q(i, j);     f3 = Q(i, j + 1); }
    fy[c] = vr * (0.75 * f2 + 0.375 * f3 - 0.125 * f1) / dy;
  } else {
    double f1, f2, f3, f4, f5;
    double ur = (-U(i + 2, j, 0) + 7.0 * U(i + 1, j, 0) + 7.0 * U(i, j, 0) - U(i - 1, j, 0)) / 12.0;
    double vr = (-U(i, j + 2, 1) + 7.0 * U(i, j + 1, 1) + 7.0 * U(i, j, 1) - U(i, j - 1, 1)) / 12.0;
    if (ur < 0.0) { f1 = Q(i + 3, j); f2 = Q(i + 2, j); f3 = Q(i + 1, j); f4 = Q(i, j); f5 = Q(i - 1, j); }
    else          { f1 = Q(i - 2, j); f2 = Q(i - 1, j); f3 = Q(i, j); f4 = Q(i + 1, j); f5 = Q(i + 2, j); }
    fx[c] = ur * weno5_f64c(f1, f2, f3, f4, f5) / ph.dx;
    if (vr < 0.0) { f1 = Q(i, j + 3); f2 = Q(i, j + 2); f3 = Q(i, j + 1); f4 = Q(i, j); f5 = Q(i, j - 1); }
    else          { f1 = Q(i, j - 2); f2 = Q(i, j - 1); f3 = Q(i, j); f4 = Q(i, j + 1); f5 = Q(i, j + 2); }
    fy[c] = vr * weno5_f64c(f1, f2, f3, f4, f5) / ph.dy;
  }
}

// advect2D, src/advect.hpp:23-37
__global__ void k_advect2d(Geo g, Range r, int v, const double *__restrict__ fx, const double *__restrict__ fy,
                           double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long c = g.id(i, j, 0);
  dvar[c + g.N * v] = -((fx[c] - fx[g.id(i - 1, j, 0)]) + (fy[c] - fy[g.id(i, j - 1, 0)]));
}

// applyPressureGradient2D (src/presgrad.hpp:22-42) then computeBuoyancy
// (src/buoyancy.hpp:49-73)
__global__ void k_presgrad_buoy2d(Geo g, Phys ph, Range r, int buoyancy, const double *__restrict__ var,
                                  const double *__restrict__ p, const double *__restrict__ rho,
                                  double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long c = g.id(i, j, 0), N = g.N;
  const double dx = ph.cd[1], dy = ph.cd[2];
  auto P = [&](int a, int b) { return p[g.id(a, b, 0)]; };
  double dxp = (P(i - 2, j) - 8.0 * P(i - 1, j) + 8.0 * P(i + 1, j) - P(i + 2, j)) / (12.0 * dx);
  double dyp = (P(i, j - 2) - 8.0 * P(i, j - 1) + 8.0 * P(i, j + 1) - P(i, j + 2)) / (12.0 * dy);
  dvar[c] -= dxp;
  dvar[c + N] -= dyp;
  if (buoyancy) {
    double vv = var[c + N] / var[c + 3 * N];
    double rhop = rho[c] - ph.rhoRef;
    double eps = 1e-6;
    if (rhop >= eps || rhop <= -eps) {
      double f = -ph.g * rhop;
      dvar[c + N] += f;
      dvar[c + 2 * N] += fabs(vv) * f;
    }
  }
}

// calculateStressTensor2dv + calculateHeatFlux2dv, src/viscosity.hpp:25-136,
// face range; stress planes: sx00, sx11, sx01, sy00, sy11, sy01, qx, qy
__global__ void k_stress2d(Geo g, Phys ph, Range r, const double *__restrict__ vel, const double *__restrict__ T,
                           double *__restrict__ st) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const double dx = ph.cd[1], dy = ph.cd[2];
  auto U = [&](int a, int b, int w) { return vel[g.id(a, b, 0) + N * w]; };
  const double mur = 2.928e-5, mut = 2.928e-5;
  const long long c = g.id(i, j, 0);
  double dudx = (U(i + 1, j, 0) - U(i, j, 0)) / dx;
  double dvdx = (U(i + 1, j, 1) - U(i, j, 1)) / dx;
  double dudy = ((U(i, j + 1, 0) + U(i + 1, j + 1, 0)) - (U(i, j - 1, 1) + U(i + 1, j - 1, 0))) / (4.0 * dy);
  double dvdy = ((U(i, j + 1, 1) + U(i + 1, j + 1, 1)) - (U(i, j - 1, 1) + U(i + 1, j - 1, 1))) / (4.0 * dy);
  st[c + N * 0] = (2.0 / 3.0) * mur * (2.0 * dudx - dvdy);
  st[c + N * 1] = (2.0 / 3.0) * mur * (2.0 * dvdy - dudx);
  st[c + N * 2] = mur * (dudy + dvdx);
  dudx = ((U(i + 1, j, 0) + U(i + 1, j + 1, 0)) - (U(i - 1, j, 0) + U(i - 1, j + 1, 0))) / (4.0 * dx);
  dvdx = ((U(i + 1, j, 1) + U(i + 1, j + 1, 1)) - (U(i - 1, j, 1) + U(i - 1, j + 1, 1))) / (4.0 * dx);
  dudy = (U(i, j + 1, 0) - U(i, j, 0)) / dy;
  dvdy = (U(i, j + 1, 1) - U(i, j, 1)) / dy;
  st[c + N * 3] = (2.0 / 3.0) * mut * (2.0 * dudx - dvdy);
  st[c + N * 4] = (2.0 / 3.0) * mut * (2.0 * dvdy - dudx);
  st[c + N * 5] = mut * (dudy + dvdx);
  const double kk = 0.026;
  st[c + N * 6] = -kk * (T[g.id(i + 1, j, 0)] - T[c]) / dx;
  st[c + N * 7] = -kk * (T[g.id(i, j + 1, 0)] - T[c]) / dy;
}

// applyViscousTerm2dv, src/viscosity.hpp:138-199
__global__ void k_applyvisc2d(Geo g, Phys ph, Range r, const double *__restrict__ vel, const double *__restrict__ st,
                              double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const double dx = ph.cd[1], dy = ph.cd[2];
  auto U = [&](int a, int b, int w) { return vel[g.id(a, b, 0) + N * w]; };
  const long long c = g.id(i, j, 0), cl = g.id(i - 1, j, 0), cb = g.id(i, j - 1, 0);
  double ur = (U(i + 1, j, 0) + U(i, j, 0)) / 2.0;
  double ul = (U(i - 1, j, 0) + U(i, j, 0)) / 2.0;
  double vr = (U(i + 1, j, 1) + U(i, j, 1)) / 2.0;
  double vl = (U(i - 1, j, 1) + U(i, j, 1)) / 2.0;
  double ut = (U(i, j + 1, 0) + U(i, j, 0)) / 2.0;
  double ub = (U(i, j - 1, 0) + U(i, j, 0)) / 2.0;
  double vt = (U(i, j + 1, 1) + U(i, j, 1)) / 2.0;
  double vb = (U(i, j - 1, 1) + U(i, j, 1)) / 2.0;
  auto S = [&](long long q, int s) { return st[q + N * s]; };
  // sx00=0 sx11=1 sx01=sx10=2 sy00=3 sy11=4 sy01=sy10=5 qx=6 qy=7
  double a = (S(c, 0) - S(cl, 0)) / dx + (S(c, 5) - S(cb, 5)) / dy;
  double b = (S(c, 2) - S(cl, 2)) / dx + (S(c, 4) - S(cb, 4)) / dy;
  double c1 = (ur * S(c, 0) - ul * S(cl, 0)) / dx + (vr * S(c, 2) - vl * S(cl, 2)) / dx +
              (ut * S(c, 5) - ub * S(cb, 5)) / dy + (vt * S(c, 4) - vb * S(cb, 4)) / dy;
  double c2 = (S(c, 6) - S(cl, 6)) / dx + (S(c, 7) - S(cb, 7)) / dy;
  dvar[c] = dvar[c] + a;
  dvar[c + N] = dvar[c + N] + b;
  dvar[c + 2 * N] = dvar[c + 2 * N] + c1 - c2;
}

// calculateRhoGrad2D (src/ceq.hpp:80-151) with maxWaveSpeed2D (:43-78, the
// reference's stride-2 species lookup kept) and maxCvar2D for C and C_hat
__global__ void k_rhograd_max2d(Geo g, Phys ph, Range r, const double *__restrict__ var,
                                const double *__restrict__ p, const double *__restrict__ rho,
                                const double *__restrict__ vel, double *__restrict__ gradRho, Scalars *sc) {
  int i, j, k;
  double ms = 0.0, mc = 0.0, mch = 0.0;
  if (decode(r, i, j, k)) {
    const long long N = g.N;
    const double dx = ph.dx, dy = ph.dy;
    auto R = [&](int a, int b) { return rho[g.id(a, b, 0)]; };
    auto U = [&](int a, int b, int w) { return vel[g.id(a, b, 0) + N * w]; };
    auto E = [&](int a, int b) {
      long long q = g.id(a, b, 0);
      return var[q + 2 * N] / rho[q] - 0.5 * rho[q] * (vel[q] * vel[q] + vel[q + N] * vel[q + N]);
    };
    double dxr = (R(i - 2, j) - 8.0 * R(i - 1, j) + 8.0 * R(i + 1, j) - R(i + 2, j)) / (12.0 * dx);
    double dyr = (R(i, j - 2) - 8.0 * R(i, j - 1) + 8.0 * R(i, j + 1) - R(i, j + 2)) / (12.0 * dy);
    double dxe = (E(i - 2, j) - 8.0 * E(i - 1, j) + 8.0 * E(i + 1, j) - E(i + 2, j)) / (12.0 * dx);
    double dye = (E(i, j - 2) - 8.0 * E(i, j - 1) + 8.0 * E(i, j + 1) - E(i, j + 2)) / (12.0 * dy);
    double dxu = (U(i - 2, j, 0) - 8.0 * U(i - 1, j, 0) + 8.0 * U(i + 1, j, 0) - U(i + 2, j, 0)) / (12.0 * dx);
    double dyv = (U(i, j - 2, 1) - 8.0 * U(i, j - 1, 1) + 8.0 * U(i, j + 1, 1) - U(i, j + 2, 1)) / (12.0 * dy);
    double rgrad = sqrt(dxr * dxr + dyr * dyr);
    double divu = dxu + dyv;
    double dnednr = (dxr * dxe + dyr * dye) * (dxr * dxr + dyr * dyr);
    double i1 = dnednr < 0.0 ? 1.0 : 0.0;
    double i2 = divu < 0.0 ? 1.0 : 0.0;
    const long long c = g.id(i, j, 0);
    gradRho[c] = (1 - i1) * i2 * rgrad;
    gradRho[c + N] = i1 * rgrad;
    gradRho[c + 2 * N] = -dyr;
    gradRho[c + 3 * N] = dxr;
    double Cp = 0, Cv = 0;
    for (int s = 0; s < g.ns; ++s) {
      double gammas = ph.cd[6 + 2 * s], Rs = ph.cd[6 + 2 * s + 1];
      Cp = Cp + (var[c + N * (3 + s)] / rho[c]) * (gammas * Rs / (gammas - 1));
      Cv = Cv + (var[c + N * (3 + s)] / rho[c]) * (Rs / (gammas - 1));
    }
    double gamma = Cp / Cv;
    double a = sqrt(gamma * p[c] / rho[c]);
    ms = a + sqrt(var[c] * var[c] + var[c + N] * var[c + N]);
    mc = fabs(var[c + N * g.nv]);
    mch = fabs(var[c + N * (g.nv + 1)]);
  }
  double v3[3] = {ms, mc, mch};
  unsigned long long *const d3[3] = {&sc->maxS, &sc->maxC, &sc->maxCh};
  block_max_atomic<3>(v3, d3);
}

// updateCeq2D, src/ceq.hpp:153-195
__global__ void k_updateceq2d(Geo g, Phys ph, Range r, const double *__restrict__ var,
                              const double *__restrict__ gradRho, const Scalars *sc, double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const double dx = ph.cd[1], dy = ph.cd[2];
  const int nc = g.nv;
  const double maxS = as_double(sc->maxS);
  const double dxmag = sqrt(dx * dx + dy * dy);
  const long long c = g.id(i, j, 0);
  for (int n = 0; n < 4; ++n) {
    const double *C = var + N * (nc + n);
    double dx_right = (C[g.id(i + 1, j, 0)] - C[c]) / dx;
    double dx_left = (C[c] - C[g.id(i - 1, j, 0)]) / dx;
    double dy_top = (C[g.id(i, j + 1, 0)] - C[c]) / dy;
    double dy_bot = (C[c] - C[g.id(i, j - 1, 0)]) / dy;
    double lap = dx_right - dx_left + dy_top - dy_bot;
    dvar[c + N * (nc + n)] = (maxS / (ph.eps * dxmag)) * (gradRho[c + N * n] - C[c]) + ph.kap * maxS * dxmag * lap;
  }
}

// computeCeqFlux2D (src/ceq.hpp:197-245) and computeCeqFaces2D (:247-267)
// fused per face; m planes ((f*2+d)*2+w).  LRN: the learned form, the same
// face composition on the given C_hat and tangent planes (float) in place of
// var(nv+1..nv+3).
template <bool LRN>
__global__ void k_ceqfaces2d(Geo g, Phys ph, Range r, const double *__restrict__ var, const float *__restrict__ lchat,
                             const float *__restrict__ ldir, const double *__restrict__ rho,
                             const double *__restrict__ vel, const Scalars *sc, double *__restrict__ m) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const int nv = g.nv;
  const double alpha = (ph.dx * ph.dx + ph.dy * ph.dy) / (as_double(sc->maxCh) + 1.0e-6) * ph.alpha;
  const double dx = ph.cd[1], dy = ph.cd[2];
  auto U = [&](int a, int b, int w) { return vel[g.id(a, b, 0) + N * w]; };
  const long long c = g.id(i, j, 0);
  double X[2][2];
  for (int f = 0; f < 2; ++f) {
    const long long q = g.id(i + (f == 0), j + (f == 1), 0);
    double rr = (rho[c] + rho[q]) / 2.0;
    double chat, ct[2];
    if (LRN) {
      chat = ((double)lchat[c] + (double)lchat[q]) / 2.0;
      ct[0] = ((double)ldir[c] + (double)ldir[q]) / 2.0;
      ct[1] = ((double)ldir[c + N] + (double)ldir[q + N]) / 2.0;
    } else {
      chat = (var[c + N * (nv + 1)] + var[q + N * (nv + 1)]) / 2.0;
      ct[0] = (var[c + N * (nv + 2)] + var[q + N * (nv + 2)]) / 2.0;
      ct[1] = (var[c + N * (nv + 3)] + var[q + N * (nv + 3)]) / 2.0;
    }
    double ctmag = (ct[0] * ct[0] + ct[1] * ct[1]) + 1.0e-6;
    for (int d = 0; d < 2; ++d) X[f][d] = alpha * rr * chat * (ct[f] * ct[d]) / ctmag;
  }
  for (int w = 0; w < 2; ++w) {
    m[c + N * ((0 * 2 + 0) * 2 + w)] = X[0][0] * ((U(i + 1, j, w) - U(i, j, w)) / dx);
    m[c + N * ((0 * 2 + 1) * 2 + w)] = X[0][1] * (((U(i, j + 1, w) + U(i + 1, j + 1, w)) - (U(i, j - 1, w) + U(i + 1, j - 1, w))) / (4.0 * dy));
    m[c + N * ((1 * 2 + 0) * 2 + w)] = X[1][0] * (((U(i + 1, j, w) + U(i + 1, j + 1, w)) - (U(i - 1, j, w) + U(i - 1, j + 1, w))) / (4.0 * dx));
    m[c + N * ((1 * 2 + 1) * 2 + w)] = X[1][1] * ((U(i, j + 1, w) - U(i, j, w)) / dy);
  }
}

// applyCeq2D, src/ceq.hpp:269-293
__global__ void k_applyceq2d(Geo g, Phys ph, Range r, const double *__restrict__ m, double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long N = g.N;
  const long long c = g.id(i, j, 0), cl = g.id(i - 1, j, 0), cb = g.id(i, j - 1, 0);
  for (int w = 0; w < 2; ++w) {
    auto M = [&](long long q, int f, int d) { return m[q + N * ((f * 2 + d) * 2 + w)]; };
    double diffu = 0;
    diffu += (M(c, 0, 0) - M(cl, 0, 0)) / ph.dx;
    diffu += (M(c, 0, 1) - M(cl, 0, 1)) / ph.dx;
    diffu += (M(c, 1, 0) - M(cb, 1, 0)) / ph.dy;
    diffu += (M(c, 1, 1) - M(cb, 1, 1)) / ph.dy;
    dvar[c + N * w] += diffu;
  }
}

// ------------------------------------------------------------- ghosts / RK

// bc_xPer / bc_yPer / bc_zPer, src/bc.cpp:133-176: one thread per (a, b, v)
// Ghost fills work on double state arrays and, in the learned FP32 form, on
// its float Runge-Kutta stage state (T = float; arithmetic in double).
template <typename T>
__global__ void k_bc_periodic(Geo g, int dir, T *__restrict__ var) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int na = dir == 0 ? g.ngj : g.ngi;
  int nb = dir == 2 ? g.ngj : g.ngk;
  long long n = (long long)na * nb * g.nvt;
  if (t >= n) return;
  int a = (int)(t % na);
  t /= na;
  int b = (int)(t % nb);
  int v = (int)(t / nb);
  // sources (interior cells) and targets (ghosts) never overlap: a periodic
  // direction has at least ng interior cells
  const T *__restrict__ src = var + g.N * v;
  T *__restrict__ dst = var + g.N * v;
  const int ng = g.ng;
  for (int gg = 0; gg < ng; ++gg) {
    if (dir == 0) {
      dst[g.id(gg, a, b)] = src[g.id(g.nci + gg, a, b)];
      dst[g.id(g.nci + ng + gg, a, b)] = src[g.id(ng + gg, a, b)];
    } else if (dir == 1) {
      dst[g.id(a, gg, b)] = src[g.id(a, g.ncj + gg, b)];
      dst[g.id(a, g.ncj + ng + gg, b)] = src[g.id(a, ng + gg, b)];
    } else {
      dst[g.id(a, b, gg)] = src[g.id(a, b, g.nck + gg)];
      dst[g.id(a, b, g.nck + ng + gg)] = src[g.id(a, b, ng + gg)];
    }
  }
}

// bc_xPer (src/bc.cpp:133-146) with 8 lanes per (j, k, v) row: lanes 0..ng-1
// fill the low ghosts i = r from i = nci + r, lanes ng..2ng-1 the high ghosts
// i = nci + ng + r from i = ng + r (sources and targets never overlap: a
// periodic direction has at least ng interior cells), so a wave touches 8
// rows with adjacent addresses.
template <typename T>
__global__ void k_bc_xper(Geo g, T *__restrict__ var) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long row = t >> 3;
  const int r = (int)(t & 7);
  const long long nrow = (long long)g.ngj * g.ngk * g.nvt;
  const int ng = g.ng;
  if (row >= nrow || r >= 2 * ng) return;
  T *u = var + row * g.ngi;  // rows of one variable plane are contiguous, planes too
  if (r < ng) u[r] = u[g.nci + r];
  else u[g.nci + r] = u[r];  // i = nci + ng + (r - ng) from i = ng + (r - ng)
}

// bc_gen, src/bc.cpp:33-131, on one boundary plane (i0 / j0 / k0 fixed).
// Threads (column, v): blockDim.x columns of the plane times blockDim.y = nvt
// variables.  Each thread mirrors its variable's ng cells (sources are the
// interior cells plane - n, targets the ghosts plane + n + 1: they never
// overlap, so the loads go out before the stores).  For hydrostatic, one
// thread per column then redoes the energy ghosts in the reference's order
// after a block barrier; the reflected energy it replaces is never read, so
// that thread's columns skip it.  Same values and operations as the
// reference's per-column loops.
template <typename T>
__global__ void k_bc_gen(Geo g, int type, int ih, int jh, int kh, int plane, T *__restrict__ var) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = threadIdx.y;
  int i = 0, j = 0, k = 0;
  bool ok;
  if (ih != 0) {
    ok = t < (long long)g.ngj * g.ngk;
    i = plane; j = (int)(t % g.ngj); k = (int)(t / g.ngj);
  } else if (jh != 0) {
    ok = t < (long long)g.ngi * g.ngk;
    j = plane; i = (int)(t % g.ngi); k = (int)(t / g.ngi);
  } else {
    ok = t < (long long)g.ngi * g.ngj;
    k = plane; i = (int)(t % g.ngi); j = (int)(t / g.ngi);
  }
  const int ng = g.ng;
  const bool hydro = type == FST_BC_HYDROSTATIC;
  const int edx = g.ndim, rdx = g.ndim + 1;
  if (ok && !(hydro && v == edx)) {
    const T *__restrict__ src = var + g.N * v;
    T *__restrict__ dst = var + g.N * v;
    const long long c0 = g.id(i, j, k);
    const long long st = (long long)ih + (long long)jh * g.ngi + (long long)kh * g.ngi * g.ngj;  // one cell outward
    if (type == FST_BC_OUTFLOW) {
      for (int n = 0; n < ng; ++n) dst[c0 + st * (n + 1)] = src[c0 - st * n];
    } else {
      double reflect = 1.0;
      if (type == FST_BC_NOSLIP) {
        if (v == 0 || v == 1 || v == 2) reflect = -1.0;
      } else {  // reflective, and the reflection pass of hydrostatic
        if ((ih != 0 && v == 0) || (jh != 0 && v == 1) || (kh != 0 && v == 2)) reflect = -1.0;
      }
      for (int n = 0; n < ng; ++n) dst[c0 + st * (n + 1)] = (T)(reflect * (double)src[c0 - st * n]);
    }
  }
  if (!hydro) return;
  __syncthreads();
  if (!ok || v != 0) return;
  auto U = [&](int a, int b, int c, int w) -> T & { return var[g.id(a, b, c) + g.N * w]; };
  auto D = [&](int a, int b, int c, int w) { return (double)U(a, b, c, w); };
  for (int n = 0; n < ng; ++n) {
    int ig = i + ih * (n + 1), jg = j + jh * (n + 1), kg = k + kh * (n + 1);
    int i1 = ig - ih, j1 = jg - jh, k1 = kg - kh;
    int i2 = ig - 2 * ih, j2 = jg - 2 * jh, k2 = kg - 2 * kh;
    double ke = 0, ke1 = 0, ke2 = 0;
    for (int d = 0; d < g.ndim; ++d) {
      ke += 0.5 * (D(ig, jg, kg, d) * D(ig, jg, kg, d));
      ke1 += 0.5 * (D(i1, j1, k1, d) * D(i1, j1, k1, d));
      ke2 += 0.5 * (D(i2, j2, k2, d) * D(i2, j2, k2, d));
    }
    ke = ke / D(ig, jg, kg, rdx);
    ke1 = ke1 / D(i1, j1, k1, rdx);
    ke2 = ke2 / D(i2, j2, k2, rdx);
    U(ig, jg, kg, edx) = (T)(2 * (D(i1, j1, k1, edx) - ke1) - (D(i2, j2, k2, edx) - ke2) + ke);
  }
}

// clears the per-stage maxima (a kernel, not a memset node, so the sequence
// stays capturable in a HIP graph)
__global__ void k_zero_scalars(Scalars *sc) {
  unsigned long long *w = reinterpret_cast<unsigned long long *>(sc);
  if (threadIdx.x < 4) w[threadIdx.x] = 0ull;  // per-lane (vector) stores
}

// rkAdvance Loop1 / Loop2, src/rk.cpp:42-50 and :63-69, over every cell
// rkAdvance stage updates over every cell (src/rk.cpp:42-50, :63-69); with
// three stages the SSP-RK3 updates (no counterpart in the reference)
__global__ void k_rk_stage(long long n, int nstages, int stage, double dt, double *__restrict__ var,
                           double *__restrict__ tmp, const double *__restrict__ dvar) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (nstages == 2) {
    if (stage == 1) {
      double u = var[t];
      tmp[t] = u;
      var[t] = u + 0.5f * dt * dvar[t];
    } else {
      var[t] = tmp[t] + dt * dvar[t];
    }
  } else if (stage == 1) {
    double u = var[t];
    tmp[t] = u;
    var[t] = u + dt * dvar[t];
  } else {
    const double a = stage == 2 ? RK3_A2 : RK3_A3, b = stage == 2 ? RK3_B2 : RK3_B3;
    var[t] = a * tmp[t] + b * (var[t] + dt * dvar[t]);
  }
}

// ------------------------------------------------------------- noise filter
// postStep's noise filter (src/cart3d.cpp:319-374, src/cart2d.cpp:257-307,
// detectNoise/removeNoise src/noise.hpp:26-261) with an order-free definition
// of its overlapping block writes and in-place sweeps: a cell is flagged when
// any coarse point whose 3x3(x3) block covers it flags it (and, with coff > 0,
// its C value is below coff), and each sweep reads the values from before it.

// coarse points per direction: 0..noise_half(nc), src/cart3d.cpp:325-341
__host__ __device__ __forceinline__ int noise_half(int nc) { return (nc + 1) % 2 == 0 ? (nc + 1) / 2 : nc / 2; }

// detectNoise3D's stencil in the reference's summation order (noise.hpp:63-82):
// (di, dj, dk, coefficient)
__constant__ signed char NOISE3D_ST[27][4] = {
    {-1, 1, 1, 2}, {0, 1, 1, 2}, {1, 1, 1, 2}, {-1, 0, 1, 2}, {0, 0, 1, 8}, {1, 0, 1, 2},
    {-1, -1, 1, 2}, {0, -1, 1, 2}, {1, -1, 1, 2}, {-1, 1, 0, 2}, {0, 1, 0, 8}, {1, 1, 0, 2},
    {-1, 0, 0, 8}, {0, 0, 0, -88}, {1, 0, 0, 8}, {-1, -1, 0, 2}, {0, -1, 0, 8}, {1, -1, 0, 2},
    {-1, 1, -1, 2}, {0, 1, -1, 2}, {1, 1, -1, 2}, {-1, 0, -1, 2}, {0, 0, -1, 8}, {1, 0, -1, 2},
    {-1, -1, -1, 2}, {0, -1, -1, 2}, {1, -1, -1, 2}};

__global__ void k_fill_u64(unsigned long long *a, int n, unsigned long long v) {
  if ((int)threadIdx.x < n) a[threadIdx.x] = v;
}

// max |var(.., cv)| over the cells (maxCvar3D(var, 6), src/cart3d.cpp:343-352;
// maxCvar2D(var, 1) = C_hat, src/cart2d.cpp:274-283), one rank
__global__ void k_noise_cmax(Geo g, const double *__restrict__ var, int cv, unsigned long long *out) {
  const long long n = (long long)g.nci * g.ncj * g.nck;
  double m = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % g.nci) + g.ng, j = (int)((t / g.nci) % g.ncj) + g.ng;
    const int k = g.ndim == 3 ? (int)(t / ((long long)g.nci * g.ncj)) + g.ng : 0;
    m = fmax(m, fabs(var[g.id(i, j, k) + g.N * cv]));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0 && m > 0.0) atomic_max_nonneg(out, m);
}

// detectNoise3D / detectNoise2D's test on each coarse point (ii, jj, kk)
// px, py, pz: coarse points per direction (pz = 1 in 2D)
__global__ void k_noise_flags(Geo g, Phys ph, const fst_config cf, const double *__restrict__ var, int v, int px,
                              int py, int pz, unsigned char *__restrict__ flag) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)px * py * pz) return;
  const int ii = (int)(t % px), jj = (int)((t / px) % py), kk = (int)(t / ((long long)px * py));
  const int i = 2 * ii + g.ng, j = 2 * jj + g.ng;
  const double dx = ph.cd[1], dy = ph.cd[2], dz = ph.cd[3];
  const double *U = var + g.N * v;
  bool f;
  if (g.ndim == 3) {
    const int k = 2 * kk + g.ng;
    const double a = sqrt(165.0 * dx * dy * dz) / 15840.0;
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const double term =
          (double)NOISE3D_ST[q][3] * U[g.id(i + NOISE3D_ST[q][0], j + NOISE3D_ST[q][1], k + NOISE3D_ST[q][2])];
      sum = q == 0 ? term : sum + term;
    }
    const double c = -a * sum;
    const double cref = cf.n_dh * sqrt(dx * dy * dz / 12.0);
    f = fabs(c) >= cref;
  } else {
    auto W = [&](int a, int b) { return U[g.id(a, b, 0)]; };
    const double a = sqrt(6.0 * dx * dy);
    const double c = -(a / 192.0) * (W(i - 1, j + 1) + 2 * W(i, j + 1) + W(i + 1, j + 1) + 2 * W(i - 1, j) -
                                     12 * W(i, j) + 2 * W(i + 1, j) + W(i - 1, j - 1) + 2 * W(i, j - 1) +
                                     W(i + 1, j - 1));
    const double cref = cf.n_dh * a / 16.0;
    f = fabs(c) >= cref;
  }
  flag[t] = f ? 1 : 0;
}

// removeNoise3D / removeNoise2D increments of one sweep on the cells: dn
__global__ void k_noise_dn(Geo g, Phys ph, const fst_config cf, const double *__restrict__ var, int v, int px,
                           int py, int pz, const unsigned char *__restrict__ flag, const unsigned long long *cmax,
                           int use_coff, double *__restrict__ dn) {
  const long long n = (long long)g.nci * g.ncj * g.nck;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.nci) + g.ng, j = (int)((t / g.nci) % g.ncj) + g.ng;
  const int k = g.ndim == 3 ? (int)(t / ((long long)g.nci * g.ncj)) + g.ng : 0;
  const int ng = g.ng;
  // the coarse points covering (i, j, k): point q sits at ng + 2q and covers
  // ng + 2q - 1 .. ng + 2q + 1, so per axis those of x-1, x, x+1 on the lattice
  bool cov = false;
  for (int ci = i - 1; ci <= i + 1; ++ci) {
    const int di = ci - ng;
    if (di < 0 || (di & 1) || di / 2 >= px) continue;
    for (int cj = j - 1; cj <= j + 1; ++cj) {
      const int dj = cj - ng;
      if (dj < 0 || (dj & 1) || dj / 2 >= py) continue;
      for (int ck = (g.ndim == 3 ? k - 1 : 0); ck <= (g.ndim == 3 ? k + 1 : 0); ++ck) {
        const int dk = g.ndim == 3 ? ck - ng : 0;
        if (dk < 0 || (dk & 1) || dk / 2 >= pz) continue;
        cov = cov || flag[di / 2 + (long long)px * (dj / 2 + (long long)py * (dk / 2))] != 0;
      }
    }
  }
  const double coff = use_coff ? cf.n_coff * as_double(*cmax) : 0.0;
  const double dx = ph.cd[1], dy = ph.cd[2], dz = ph.cd[3];
  const double *U = var + g.N * v;
  const long long c = g.id(i, j, k);
  if (g.ndim == 3) {
    const double ch = fabs(var[c + g.N * (g.nv + 1)]);
    const int noise = cov && (coff == 0.0 || ch < coff) ? 1 : 0;
    const long long sj = g.ngi, sk = (long long)g.ngi * g.ngj;
    const double lap = (U[c - 1] - 2 * U[c] + U[c + 1]) / (dx * dx) + (U[c - sj] - 2 * U[c] + U[c + sj]) / (dy * dy) +
                       (U[c - sk] - 2 * U[c] + U[c + sk]) / (dz * dz);
    dn[c] = cf.n_eta * (dx * dx + dy * dy + dz * dz) * noise * lap;
  } else {
    const double ch = var[c + g.N * (g.nv + 1)];  // no abs in the reference (noise.hpp:222)
    const int noise = cov && (coff == 0.0 || ch < coff) ? 1 : 0;
    const long long sj = g.ngi;
    const double lgrad = (U[c] - U[c - 1]) / dx;
    const double rgrad = (U[c + 1] - U[c]) / dx;
    const double bgrad = (U[c] - U[c - sj]) / dy;
    const double tgrad = (U[c + sj] - U[c]) / dy;
    dn[c] = cf.n_eta * (dx * dx + dy * dy) * noise * ((rgrad - lgrad) / dx + (tgrad - bgrad) / dy);
  }
}

__global__ void k_noise_add(Geo g, double *__restrict__ var, int v, const double *__restrict__ dn) {
  const long long n = (long long)g.nci * g.ncj * g.nck;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.nci) + g.ng, j = (int)((t / g.nci) % g.ncj) + g.ng;
  const int k = g.ndim == 3 ? (int)(t / ((long long)g.nci * g.ncj)) + g.ng : 0;
  const long long c = g.id(i, j, k);
  var[c + g.N * v] += dn[c];
}

// ------------------------------------------------------------- output pack
// blockWriter::dataPack (src/block.cpp:404-454): one thread per packed value;
// the mean sums the stride block in the reference's loop order (ia, ja, ka)
// and divides by the block size.  src: plane vn of the array, with cell
// offset `off` (ng for var/varx, 0 for the grid) and extents nx, ny.
template <typename T>
__global__ void k_pack(const double *__restrict__ src, long long nx, long long ny, int ndim, int off, PackBox bx,
                       int average, T *__restrict__ dst) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long e0 = bx.ext[0], e1 = bx.ext[1], e2 = ndim == 3 ? bx.ext[2] : 1;
  if (t >= e0 * e1 * e2) return;
  const long long ii = t % e0, jj = (t / e0) % e1, kk = t / (e0 * e1);
  const long long i = bx.start[0] + ii * bx.stride[0], j = bx.start[1] + jj * bx.stride[1];
  const long long k = ndim == 3 ? bx.start[2] + kk * bx.stride[2] : 0;
  auto at = [&](long long a, long long b, long long c) {
    return src[(a + off) + nx * ((b + off) + ny * (ndim == 3 ? c + off : 0))];
  };
  if (average) {
    double mean = 0.0;
    const long long s2 = ndim == 3 ? bx.stride[2] : 1;
    for (long long ia = 0; ia < bx.stride[0]; ++ia)
      for (long long ja = 0; ja < bx.stride[1]; ++ja)
        for (long long ka = 0; ka < s2; ++ka) mean += at(i + ia, j + ja, k + ka);
    dst[t] = (T)(mean / (double)(ndim == 3 ? bx.stride[0] * bx.stride[1] * bx.stride[2] : bx.stride[0] * bx.stride[1]));
  } else {
    dst[t] = (T)at(i, j, k);
  }
}

// plain copy, 16 bytes per lane per iteration (bandwidth probe)
// Streaming copies for the bandwidth ceiling.  k_copy32: 32-byte elements,
// four in flight per thread per iteration (consecutive threads on
// consecutive elements); the fastest of the shapes swept on MI355X (8- and
// 32-byte elements, 1 or 4 per thread, 1024..16384 blocks of 256, with and
// without nontemporal hints).  k_copy8: one double per thread per iteration,
// the access width of the stage kernels, for scaling the byte counters.
typedef double dv4 __attribute__((ext_vector_type(4)));
__global__ void k_copy32(const dv4 *__restrict__ src, dv4 *__restrict__ dst, long long n) {
  constexpr int U = 4;
  long long base = (long long)blockIdx.x * blockDim.x * U + threadIdx.x;
  const long long step = (long long)gridDim.x * blockDim.x * U;
  for (; base < n; base += step) {
    dv4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = base + (long long)u * blockDim.x;
      if (i < n) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = base + (long long)u * blockDim.x;
      if (i < n) dst[i] = v[u];
    }
  }
}
__global__ void k_copy8(const double *__restrict__ src, double *__restrict__ dst, long long n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; t < n; t += stride) dst[t] = src[t];
}

// copyExtraVars2D/3D (src/secondary.hpp:25-59): varx = (vel..., p, T, rho)
__global__ void k_copy_extra(long long N, int ndim, const double *__restrict__ vel, const double *__restrict__ p,
                             const double *__restrict__ T, const double *__restrict__ rho, double *__restrict__ varx) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  for (int d = 0; d < ndim; ++d) varx[c + N * d] = vel[c + N * d];
  varx[c + N * ndim] = p[c];
  varx[c + N * (ndim + 1)] = T[c];
  varx[c + N * (ndim + 2)] = rho[c];
}

// ------------------------------------------------------------- generalized grids
//
// grid(i,j,k,v): node coordinates, LayoutLeft over (ni,nj,nk,3) with
// ni = nci+1, nj = ncj+1, nk = nck+1 (1 in 2D) (src/gen3d.cpp:41,
// src/gen2d.cpp:25, src/input.cpp:401-406).  metrics(i,j,k,m,n) = d xi_m / d x_n:
// plane c + N*(m + ndim*n) (LayoutLeft over (ngi,ngj,ngk,ndim,ndim)).

struct Nodes {
  int ni, nj, nk;
  __device__ __forceinline__ long long at(int i, int j, int k, int v) const {
    return (long long)i + (long long)ni * ((long long)j + (long long)nj * ((long long)k + (long long)nk * v));
  }
};

// computeMetrics3D, src/metric.hpp:59-164, over the cell range.  Node indices
// are cell index - 3 as in the reference (the literal 3, not ng).
__global__ void k_metrics3d(Geo g, Range r, Nodes nd, const double *__restrict__ grid, double *__restrict__ mt) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const int ii = i - 3, jj = j - 3, kk = k - 3;
  auto G = [&](int a, int b, int c, int v) { return grid[nd.at(a, b, c, v)]; };
  double xm[3], xp[3], ym[3], yp[3], zm[3], zp[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    xm[v] = (G(ii, jj, kk, v) + G(ii, jj + 1, kk, v) + G(ii, jj, kk + 1, v) + G(ii, jj + 1, kk + 1, v)) / 4.0;
    xp[v] = (G(ii + 1, jj, kk, v) + G(ii + 1, jj + 1, kk, v) + G(ii + 1, jj, kk + 1, v) +
             G(ii + 1, jj + 1, kk + 1, v)) / 4.0;
    zm[v] = (G(ii, jj, kk, v) + G(ii + 1, jj, kk, v) + G(ii, jj + 1, kk, v) + G(ii + 1, jj + 1, kk, v)) / 4.0;
    zp[v] = (G(ii, jj, kk + 1, v) + G(ii + 1, jj, kk + 1, v) + G(ii, jj + 1, kk + 1, v) +
             G(ii + 1, jj + 1, kk + 1, v)) / 4.0;
    ym[v] = (G(ii, jj, kk, v) + G(ii + 1, jj, kk, v) + G(ii, jj, kk + 1, v) + G(ii + 1, jj, kk + 1, v)) / 4.0;
    yp[v] = (G(ii, jj + 1, kk, v) + G(ii + 1, jj + 1, kk, v) + G(ii, jj + 1, kk + 1, v) +
             G(ii + 1, jj + 1, kk + 1, v)) / 4.0;
  }
  const double xxi = xp[0] - xm[0], xet = yp[0] - ym[0], xzt = zp[0] - zm[0];
  const double yxi = xp[1] - xm[1], yet = yp[1] - ym[1], yzt = zp[1] - zm[1];
  const double zxi = xp[2] - xm[2], zet = yp[2] - ym[2], zzt = zp[2] - zm[2];
  const double jac = 1.0 / (xxi * (yet * zzt - yzt * zet) + xet * (yzt * zxi - yxi * zzt) + xzt * (yxi * zet - yet * zxi));
  const long long c = g.id(i, j, k), N = g.N;
  mt[c + N * 0] = jac * (yet * zzt - yzt * zet);  // (0,0)
  mt[c + N * 3] = jac * (xzt * zet - xet * zzt);  // (0,1)
  mt[c + N * 6] = jac * (xet * yzt - xzt * yet);  // (0,2)
  mt[c + N * 1] = jac * (yzt * zxi - yxi * zzt);  // (1,0)
  mt[c + N * 4] = jac * (xxi * zzt - xzt * zxi);  // (1,1)
  mt[c + N * 7] = jac * (xzt * yxi - xxi * yzt);  // (1,2)
  mt[c + N * 2] = jac * (yxi * zet - yet * zxi);  // (2,0)
  mt[c + N * 5] = jac * (xet * zxi - xxi * zet);  // (2,1)
  mt[c + N * 8] = jac * (xxi * yet - xet * yxi);  // (2,2)
}

// computeMetrics2D, src/metric.hpp:21-57, over the cell range
__global__ void k_metrics2d(Geo g, Range r, Nodes nd, const double *__restrict__ grid, double *__restrict__ mt) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const int ii = i - 3, jj = j - 3;
  auto G = [&](int a, int b, int v) { return grid[nd.at(a, b, 0, v)]; };
  const double x_xi = (G(ii + 1, jj, 0) + G(ii + 1, jj + 1, 0)) / 2.0 - (G(ii, jj, 0) + G(ii, jj + 1, 0)) / 2.0;
  const double y_xi = (G(ii + 1, jj, 1) + G(ii + 1, jj + 1, 1)) / 2.0 - (G(ii, jj, 1) + G(ii, jj + 1, 1)) / 2.0;
  const double x_et = (G(ii, jj + 1, 0) + G(ii + 1, jj + 1, 0)) / 2.0 - (G(ii, jj, 0) + G(ii + 1, jj, 0)) / 2.0;
  const double y_et = (G(ii, jj + 1, 1) + G(ii + 1, jj + 1, 1)) / 2.0 - (G(ii, jj, 1) + G(ii + 1, jj, 1)) / 2.0;
  const double jac = 1.0 / (x_xi * y_et - x_et * y_xi);
  const long long c = g.id(i, j, 0), N = g.N;
  mt[c + N * 0] = jac * y_et;    // (0,0)
  mt[c + N * 2] = -jac * x_et;   // (0,1)
  mt[c + N * 1] = -jac * y_xi;   // (1,0)
  mt[c + N * 3] = jac * x_xi;    // (1,1)
}

// Metric ghosts of a rank without neighbours on one face: mirror copy of the
// ng layers next to it (src/gen3d.cpp:241-300, src/gen2d.cpp:285-329), over the
// full range of the other two indices.  dir 0/1/2, side -1 (min) or +1 (max).
__global__ void k_metric_ghost(Geo g, int dir, int side, double *__restrict__ mt) {
  const int e0 = dir == 0 ? g.ng : g.ngi, e1 = dir == 1 ? g.ng : g.ngj, e2 = dir == 2 ? g.ng : g.ngk;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)e0 * e1 * e2) return;
  const int a = (int)(t % e0);
  t /= e0;
  const int b = (int)(t % e1);
  const int c = (int)(t / e1);
  int di = a, dj = b, dk = c, si = a, sj = b, sk = c;
  if (dir == 0) {
    di = side < 0 ? g.ng - a - 1 : g.ngi - 1 - a;
    si = side < 0 ? g.ng + a : g.nci + a;
  } else if (dir == 1) {
    dj = side < 0 ? g.ng - b - 1 : g.ngj - 1 - b;
    sj = side < 0 ? g.ng + b : g.ncj + b;
  } else {
    dk = side < 0 ? g.ng - c - 1 : g.ngk - 1 - c;
    sk = side < 0 ? g.ng + c : g.nck + c;
  }
  const long long d = g.id(di, dj, dk), s = g.id(si, sj, sk);
  const int nm = g.ndim * g.ndim;
  for (int q = 0; q < nm; ++q) mt[d + g.N * q] = mt[s + g.N * q];
}

// computeGenVelocity3D / 2D (src/velocity.hpp:23-66): contravariant velocity
// vel(d) = sum_m metrics(d,m) * var(m) / rho over every cell, after the
// primitive pass has written rho.
__global__ void k_genvel(Geo g, Range r, const double *__restrict__ var, const double *__restrict__ rho,
                         const double *__restrict__ mt, double *__restrict__ vel) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long c = g.id(i, j, k), N = g.N;
  const double rr = rho[c];
  if (g.ndim == 3) {
    const double m0 = var[c], m1 = var[c + N], m2 = var[c + 2 * N];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      vel[c + N * d] = (mt[c + N * (d + 0)] * m0 / rr + mt[c + N * (d + 3)] * m1 / rr + mt[c + N * (d + 6)] * m2 / rr);
  } else {
    const double m0 = var[c], m1 = var[c + N];
#pragma unroll
    for (int d = 0; d < 2; ++d) vel[c + N * d] = (mt[c + N * (d + 0)] * m0 / rr + mt[c + N * (d + 2)] * m1 / rr);
  }
}

// applyGenPressureGradient3D, src/presgrad.hpp:72-147, over the cell range:
// dvar(n) -= sum_d d/dxi_d (p * metrics(d,n)), 4th-order central in each
// computational direction.
__global__ void k_genpresgrad3d(Geo g, Range r, const double *__restrict__ p, const double *__restrict__ mt,
                                double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long c = g.id(i, j, k), N = g.N;
  auto P = [&](int a, int b, int cc) { return p[g.id(a, b, cc)]; };
  auto M = [&](int a, int b, int cc, int m, int n) { return mt[g.id(a, b, cc) + N * (m + 3 * n)]; };
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const double dxi = (P(i - 2, j, k) * M(i - 2, j, k, 0, n) - 8.0 * P(i - 1, j, k) * M(i - 1, j, k, 0, n) +
                        8.0 * P(i + 1, j, k) * M(i + 1, j, k, 0, n) - P(i + 2, j, k) * M(i + 2, j, k, 0, n)) /
                       (12.0);
    const double det = (P(i, j - 2, k) * M(i, j - 2, k, 1, n) - 8.0 * P(i, j - 1, k) * M(i, j - 1, k, 1, n) +
                        8.0 * P(i, j + 1, k) * M(i, j + 1, k, 1, n) - P(i, j + 2, k) * M(i, j + 2, k, 1, n)) /
                       (12.0);
    const double dzt = (P(i, j, k - 2) * M(i, j, k - 2, 2, n) - 8.0 * P(i, j, k - 1) * M(i, j, k - 1, 2, n) +
                        8.0 * P(i, j, k + 1) * M(i, j, k + 1, 2, n) - P(i, j, k + 2) * M(i, j, k + 2, 2, n)) /
                       (12.0);
    dvar[c + N * n] = (dvar[c + N * n] - (dxi + det + dzt));
  }
}

// applyGenPressureGradient2D, src/presgrad.hpp:149-190, over the cell range
__global__ void k_genpresgrad2d(Geo g, Range r, const double *__restrict__ p, const double *__restrict__ mt,
                                double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long c = g.id(i, j, 0), N = g.N;
  auto P = [&](int a, int b) { return p[g.id(a, b, 0)]; };
  auto M = [&](int a, int b, int m, int n) { return mt[g.id(a, b, 0) + N * (m + 2 * n)]; };
#pragma unroll
  for (int n = 0; n < 2; ++n) {
    const double dxi = (P(i - 2, j) * M(i - 2, j, 0, n) - 8.0 * P(i - 1, j) * M(i - 1, j, 0, n) +
                        8.0 * P(i + 1, j) * M(i + 1, j, 0, n) - P(i + 2, j) * M(i + 2, j, 0, n)) /
                       (12.0);
    const double det = (P(i, j - 2) * M(i, j - 2, 1, n) - 8.0 * P(i, j - 1) * M(i, j - 1, 1, n) +
                        8.0 * P(i, j + 1) * M(i, j + 1, 1, n) - P(i, j + 2) * M(i, j + 2, 1, n)) /
                       (12.0);
    dvar[c + N * n] = (dvar[c + N * n] - (dxi + det));
  }
}


// ---------------------------------------------------------------- status check

// Order key of a double: unsigned order of the key is numeric order (negative
// values reversed; -0 orders just below +0).
__device__ __forceinline__ unsigned long long order_key(double d) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(d);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}

__global__ void k_minmax_init(int nvar, unsigned long long *__restrict__ mm) {
  const int v = threadIdx.x;
  if (v < nvar) {
    mm[2 * v] = order_key(-DBL_MAX);  // the identities of Kokkos::Max / Kokkos::Min
    mm[2 * v + 1] = order_key(DBL_MAX);
  }
}

// maxVarFunctor / minVarFunctor (src/status.cpp:40-98) over the cell range:
// one block per group of (j, k) rows of variable blockIdx.y, threads along i.
// `s > max` and `s < min` are false for NaN, so NaN cells leave both
// reductions unchanged, as in the reference's join.
__global__ void __launch_bounds__(256) k_minmax(const double *__restrict__ a, long long N, int ndim, int ng, int ngi,
                                                int ngj, int nci, int ncj, int rows, unsigned long long *__restrict__ mm) {
  const int v = blockIdx.y;
  const double *p = a + N * v;
  const int ko = ndim == 3 ? ng : 0;
  double mx = -DBL_MAX, mn = DBL_MAX;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    const int j = r % ncj, k = r / ncj;
    const double *row = p + (long long)ngi * ((j + ng) + (long long)ngj * (k + ko)) + ng;
    for (int i = threadIdx.x; i < nci; i += blockDim.x) {
      const double x = row[i];
      if (x > mx) mx = x;
      if (x < mn) mn = x;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double x = __shfl_xor(mx, o), y = __shfl_xor(mn, o);
    if (x > mx) mx = x;
    if (y < mn) mn = y;
  }
  __shared__ double smx[4], smn[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    smx[w] = mx;
    smn[w] = mn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
      if (smx[q] > mx) mx = smx[q];
      if (smn[q] < mn) mn = smn[q];
    }
    atomicMax(&mm[2 * v], order_key(mx));
    atomicMin(&mm[2 * v + 1], order_key(mn));
  }
}

}  // namespace

// ------------------------------------------------------------- launchers

void launch_eos_vel(const Geo &g, const Phys &ph, const Bufs &b, hipStream_t s) {
  Range r = ghost_range(g);
  if (g.ndim == 3) k_eos_vel3d<<<grid_for(r), 256, 0, s>>>(g, ph, r, b.var, b.p, b.rho, b.T, b.vel);
  else k_eos_vel2d<<<grid_for(r), 256, 0, s>>>(g, ph, r, b.var, b.p, b.rho, b.T, b.vel);
}

void launch_compute_ref(const Geo &g, const Phys &ph, const fst_config &cf, const Bufs &b, hipStream_t s,
                        const Ceq2dFields *lf, bool eos_done) {
  if (!eos_done) launch_eos_vel(g, ph, b, s);
  Range cr = cell_range(g), fr = face_range(g);
  if (g.ndim == 3) {
    for (int v = 0; v < g.nv; ++v) {
      if (b.prof) { b.prof->name = "k_flux3d"; b.prof->per = 2; b.prof->begin(s); }
      k_flux3d<<<grid_for(fr), 256, 0, s>>>(g, ph, fr, v, b.var, b.p, b.vel, b.fx, b.fy, b.fz);
      if (b.prof) b.prof->end(s);
      k_advect3d<<<grid_for(cr), 256, 0, s>>>(g, cr, v, b.fx, b.fy, b.fz, b.dvar);
    }
    k_presgrad_buoy3d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, cf.buoyancy, b.var, b.p, b.rho, b.dvar);
    if (cf.visc) {
      for (int d = 0; d < 3; ++d) {
        k_stress3d<<<grid_for(fr), 256, 0, s>>>(g, ph, fr, d, b.var, b.rho, b.vel, b.stress);
        k_applyvisc3d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, d, b.vel, b.stress, b.dvar);
      }
    }
    if (cf.ceq) {
      k_zero_scalars<<<1, 64, 0, s>>>(b.sc);
      k_rhograd_max3d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, b.var, b.p, b.rho, b.vel, b.gradRho, b.sc);
      k_updateceq3d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, b.var, b.gradRho, b.sc, b.dvar);
      k_ceqfaces3d<<<grid_for(fr), 256, 0, s>>>(g, ph, fr, b.var, b.rho, b.vel, b.sc, b.mflux);
      k_applyceq3d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, b.mflux, b.dvar);
    }
  } else {
    for (int v = 0; v < g.nv; ++v) {
      k_flux2d<<<grid_for(fr), 256, 0, s>>>(g, ph, fr, v, cf.scheme, b.var, b.p, b.rho, b.vel, b.fx, b.fy);
      k_advect2d<<<grid_for(cr), 256, 0, s>>>(g, cr, v, b.fx, b.fy, b.dvar);
    }
    k_presgrad_buoy2d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, cf.buoyancy, b.var, b.p, b.rho, b.dvar);
    if (cf.visc) {
      k_stress2d<<<grid_for(fr), 256, 0, s>>>(g, ph, fr, b.vel, b.T, b.stress);
      k_applyvisc2d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, b.vel, b.stress, b.dvar);
    }
    if (cf.ceq && lf) {
      // learned form: no relaxation solve; the artificial-viscosity terms
      // (src/cart2d.cpp:223-236) on the given C_hat, tangent and max C_hat
      k_ceqfaces2d<true><<<grid_for(fr), 256, 0, s>>>(g, ph, fr, b.var, lf->chat, lf->dir, b.rho, b.vel, lf->sc,
                                                      b.mflux);
      k_applyceq2d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, b.mflux, b.dvar);
    } else if (cf.ceq) {
      k_zero_scalars<<<1, 64, 0, s>>>(b.sc);
      k_rhograd_max2d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, b.var, b.p, b.rho, b.vel, b.gradRho, b.sc);
      k_updateceq2d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, b.var, b.gradRho, b.sc, b.dvar);
      k_ceqfaces2d<false><<<grid_for(fr), 256, 0, s>>>(g, ph, fr, b.var, nullptr, nullptr, b.rho, b.vel, b.sc,
                                                       b.mflux);
      k_applyceq2d<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, b.mflux, b.dvar);
    }
  }
}

// Buoyancy of cart3d / cart2d on the generalized grids (gen_ext): the
// source terms of computeBuoyancy3D (src/buoyancy.hpp:22-47) and
// computeBuoyancy (:49-73) on the stored physical momenta, cell range.
__global__ void k_buoy_gen(Geo g, Phys ph, Range r, const double *__restrict__ var, const double *__restrict__ rho,
                           double *__restrict__ dvar) {
  int i, j, k;
  if (!decode(r, i, j, k)) return;
  const long long c = g.id(i, j, k), N = g.N;
  if (g.ndim == 3) {
    double vv = var[c + N] / rho[c];
    double rhop = rho[c] - ph.rhoRef;
    double f = -ph.g * rhop;
    dvar[c + N] += f;
    dvar[c + 3 * N] += vv * f;
  } else {
    double vv = var[c + N] / var[c + 3 * N];
    double rhop = rho[c] - ph.rhoRef;
    double eps = 1e-6;
    if (rhop >= eps || rhop <= -eps) {
      double f = -ph.g * rhop;
      dvar[c + N] += f;
      dvar[c + 2 * N] += fabs(vv) * f;
    }
  }
}

namespace {

// The cartesian velocity var(w) / rho and the mass fractions var(4+s) / rho of
// every cell into u0, u1, u2 and Y (ns planes): the inputs of the gen_ext
// viscous terms, the same quotients they would form.  rho == nullptr: rho is
// summed from the species densities as calculateRhoPT3D sums it
// (src/secondary.hpp:126-130), so the same value.  T = float (FP32 form):
// the FP64 quotients rounded once.
template <typename T>
__global__ void k_cartvel3d(long long N, int ns, const double *__restrict__ var, const double *__restrict__ rho,
                            T *__restrict__ u0, T *__restrict__ u1, T *__restrict__ u2, T *__restrict__ Y) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double r = 0.0;
  if (rho) {
    r = rho[c];
  } else {
    for (int q = 0; q < ns; ++q) r = r + var[c + N * (4 + q)];
  }
  u0[c] = (T)(var[c] / r);
  u1[c] = (T)(var[c + N] / r);
  u2[c] = (T)(var[c + 2 * N] / r);
  for (int q = 0; q < ns; ++q) Y[c + N * q] = (T)(var[c + N * (4 + q)] / r);
}

// the 9 metric planes rounded to floats (the FP32 form's viscous terms)
__global__ void k_metrics_f32(long long n, const double *__restrict__ mt, float *__restrict__ m32) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) m32[c] = (float)mt[c];
}

// gen_ext viscous terms on a generalized 3D grid (not in the reference: the
// cartesian split of src/viscosity.hpp:416-663 in index space, with the chain
// rule through the metric terms m(a,n) = d xi_a / d x_n as
// applyGenPressureGradient3D applies it, src/presgrad.hpp:72-147).  On the
// face between c and c + e_dir: index-space derivatives g[w][a] of the
// cartesian velocity u_w = var(w) / rho (normal: u(c+e) - u(c); tangential: the 4-point
// difference / 4), face metrics mf(a,n) = (m(c) + m(c+e)) / 2,
// du_w/dx_n = (mf(0,n) g[w][0] + mf(1,n) g[w][1]) + mf(2,n) g[w][2], the
// stress of calculateStressTensor, and the contravariant fluxes through the
// face: momentum F_w = sum_n mf(dir,n) tau_wn and energy
// H_w = sum_n mf(dir,n) (uf_w tau_wn), uf the face velocity (out[0..5]).
// On a uniform grid of power-of-two spacing each value is the cartesian
// one scaled by a power of two, so the module equals cart3d bit for bit.
// Cell offsets: the face lies between cell c and c + e_DIR; sy, sz the j and
// k strides.  T = float: the FP32 form (float inputs, binary32 arithmetic).
template <int DIR, typename T>
__device__ __forceinline__ void viscgen3d_face(long long N, long long sy, long long sz, int ns, const Phys &ph,
                                               long long c, const T *__restrict__ Y, const T *__restrict__ u0,
                                               const T *__restrict__ u1, const T *__restrict__ u2,
                                               const T *__restrict__ mt, T *out) {
  const long long st[3] = {1, sy, sz};
  const long long e = st[DIR], cn = c + e;
  const T *const uw[3] = {u0, u1, u2};
  T mu = 0;
  for (int s = 0; s < ns; ++s) mu += ((Y[cn + N * s] + Y[c + N * s]) / (T)2) * (T)ph.mus[s];
  T gr[3][3], mf[3][3], du[3][3], uf[3];
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    const T *u = uw[w];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a == DIR) {
        gr[w][a] = u[cn] - u[c];
      } else {
        const long long t = st[a];
        gr[w][a] = ((u[c + t] + u[cn + t]) - (u[c - t] + u[cn - t])) / (T)4;
      }
    }
    uf[w] = (u[cn] + u[c]) / (T)2;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int n = 0; n < 3; ++n) mf[a][n] = (mt[c + N * (a + 3 * n)] + mt[cn + N * (a + 3 * n)]) / (T)2;
#pragma unroll
  for (int w = 0; w < 3; ++w)
#pragma unroll
    for (int n = 0; n < 3; ++n) du[w][n] = (mf[0][n] * gr[w][0] + mf[1][n] * gr[w][1]) + mf[2][n] * gr[w][2];
  const T dudx = du[0][0], dudy = du[0][1], dudz = du[0][2];
  const T dvdx = du[1][0], dvdy = du[1][1], dvdz = du[1][2];
  const T dwdx = du[2][0], dwdy = du[2][1], dwdz = du[2][2];
  const T c23 = (T)(2.0 / 3.0);
  T t[3][3];
  t[0][0] = c23 * mu * ((T)2 * dudx - dvdy - dwdz);
  t[1][1] = c23 * mu * ((T)2 * dvdy - dudx - dwdz);
  t[2][2] = c23 * mu * ((T)2 * dwdz - dudx - dvdy);
  t[0][1] = t[1][0] = mu * (dudy + dvdx);
  t[0][2] = t[2][0] = mu * (dudz + dwdx);
  t[1][2] = t[2][1] = mu * (dvdz + dwdy);
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    out[w] = (mf[DIR][0] * t[w][0] + mf[DIR][1] * t[w][1]) + mf[DIR][2] * t[w][2];
    out[3 + w] = (mf[DIR][0] * (uf[w] * t[w][0]) + mf[DIR][1] * (uf[w] * t[w][1])) + mf[DIR][2] * (uf[w] * t[w][2]);
  }
}

// One launch for the three directions.  A block is VW + 1 waves over a tile
// of 63 (i) x VW (j) cells marching k: lane l holds column i0 - 1 + l (lane 0
// the tile's left neighbour), wave w < VW row j0 + w, wave VW the row above
// the tile.  Each face is computed once: the x face right of every lane's
// column (its left face comes from the lane below through a cross-lane
// move), the y face below every wave's row (through LDS), the z face above
// each cell (the one below is the previous plane's, in registers).  Then
// each cell adds x, y, z in that order, as the per-direction kernels of the
// reference sequence do (applyViscousTerm, src/viscosity.hpp:460-498), so
// the values are those of the face-by-face evaluation.
// T = float: the FP32 form; the face differences are formed in binary32 and
// added in FP64 to the float partial RHS, which is stored back as a float.
constexpr int VW = 7, VTX = 63, VKC = 16;
template <typename T>
__global__ void __launch_bounds__(64 * (VW + 1)) k_viscgen3d_march(Geo g, Phys ph, const T *__restrict__ Y,
                                                                     const T *__restrict__ u0,
                                                                     const T *__restrict__ u1,
                                                                     const T *__restrict__ u2,
                                                                     const T *__restrict__ mt,
                                                                     T *__restrict__ dvar, int nbx, int nby) {
  __shared__ T ysh[VW + 1][6][64];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, bz = blockIdx.x / (nbx * nby);
  const int i0 = g.ng + bx * VTX, j0 = g.ng + by * VW;
  const int kb = g.ng + bz * VKC, ke = min(kb + VKC, g.ng + g.nck);
  const int i = i0 - 1 + l, j = j0 + w;
  const bool colok = i < g.ng + g.nci;                 // column inside the interior (or the left neighbour)
  const bool own = l > 0 && colok && w < VW && j < g.ng + g.ncj;  // a cell this block updates
  const bool rowok = w < VW && j < g.ng + g.ncj;
  const bool yface = l > 0 && colok && j - 1 < g.ng + g.ncj;      // y face below row j0 + w
  const long long N = g.N, sy = g.ngi, sz = (long long)g.ngi * g.ngj;
  T Zp[6] = {0, 0, 0, 0, 0, 0};
  if (own) viscgen3d_face<2, T>(N, sy, sz, g.ns, ph, g.id(i, j, kb - 1), Y, u0, u1, u2, mt, Zp);
  for (int k = kb; k < ke; ++k) {
    const long long c = g.id(i, j, k);
    T X[6] = {0, 0, 0, 0, 0, 0}, Yf[6], Z[6];
    if (rowok && colok) viscgen3d_face<0, T>(N, sy, sz, g.ns, ph, c, Y, u0, u1, u2, mt, X);
    if (yface) {
      viscgen3d_face<1, T>(N, sy, sz, g.ns, ph, c - sy, Y, u0, u1, u2, mt, Yf);
#pragma unroll
      for (int q = 0; q < 6; ++q) ysh[w][q][l] = Yf[q];
    }
    if (own) viscgen3d_face<2, T>(N, sy, sz, g.ns, ph, c, Y, u0, u1, u2, mt, Z);
    T Xm[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) Xm[q] = __shfl_up(X[q], 1, 64);
    __syncthreads();
    if (own) {
      double d0 = (double)dvar[c], d1 = (double)dvar[c + N], d2 = (double)dvar[c + 2 * N],
             d3 = (double)dvar[c + 3 * N];
      d0 += X[0] - Xm[0];
      d1 += X[1] - Xm[1];
      d2 += X[2] - Xm[2];
      d3 += ((X[3] - Xm[3]) + (X[4] - Xm[4])) + (X[5] - Xm[5]);
      auto YS = [&](int r, int q) { return ysh[r][q][l]; };
      d0 += YS(w + 1, 0) - YS(w, 0);
      d1 += YS(w + 1, 1) - YS(w, 1);
      d2 += YS(w + 1, 2) - YS(w, 2);
      d3 += ((YS(w + 1, 3) - YS(w, 3)) + (YS(w + 1, 4) - YS(w, 4))) + (YS(w + 1, 5) - YS(w, 5));
      d0 += Z[0] - Zp[0];
      d1 += Z[1] - Zp[1];
      d2 += Z[2] - Zp[2];
      d3 += ((Z[3] - Zp[3]) + (Z[4] - Zp[4])) + (Z[5] - Zp[5]);
      dvar[c] = (T)d0;
      dvar[c + N] = (T)d1;
      dvar[c + 2 * N] = (T)d2;
      dvar[c + 3 * N] = (T)d3;
#pragma unroll
      for (int q = 0; q < 6; ++q) Zp[q] = Z[q];
    }
    __syncthreads();
  }
}

}  // namespace

// Directions x, y, z in that order (src/cart3d.cpp:265-270); fx, fy, fz (the
// per-variable flux planes, free once the fluxes are differenced) hold the
// cartesian velocity, the stress planes (max(6, ns) on generalized grids) the
// mass fractions.  mt32 (FP32 form): the metric planes as floats; the
// velocity and mass fractions are then stored as floats in the same planes,
// and b.dvar holds the partial RHS as floats.
void launch_viscgen3d(const Geo &g, const Phys &ph, const Bufs &b, hipStream_t s, const float *mt32) {
  const unsigned nb = (unsigned)((g.N + 255) / 256);
  const int nbx = (g.nci + VTX - 1) / VTX, nby = (g.ncj + VW - 1) / VW, nbz = (g.nck + VKC - 1) / VKC;
  if (mt32) {
    auto F = [](double *p) { return reinterpret_cast<float *>(p); };
    k_cartvel3d<float><<<nb, 256, 0, s>>>(g.N, g.ns, b.var, b.rho, F(b.fx), F(b.fy), F(b.fz), F(b.stress));
    k_viscgen3d_march<float><<<nbx * nby * nbz, 64 * (VW + 1), 0, s>>>(g, ph, F(b.stress), F(b.fx), F(b.fy),
                                                                       F(b.fz), mt32, F(b.dvar), nbx, nby);
    return;
  }
  k_cartvel3d<double><<<nb, 256, 0, s>>>(g.N, g.ns, b.var, b.rho, b.fx, b.fy, b.fz, b.stress);
  k_viscgen3d_march<double><<<nbx * nby * nbz, 64 * (VW + 1), 0, s>>>(g, ph, b.stress, b.fx, b.fy, b.fz,
                                                                      b.metrics, b.dvar, nbx, nby);
}

void launch_metrics_f32(const Geo &g, const double *mt, float *mt32, hipStream_t s) {
  const long long n = 9LL * g.N;
  k_metrics_f32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, mt, mt32);
}

// gen3d_func::compute (src/gen3d.cpp:341-369) / gen2d_func::compute
// (src/gen2d.cpp:100-131): primitives, contravariant velocity, per-variable
// fluxes and flux differences, metric pressure gradient.  No boundary
// conditions and no viscous, buoyancy or C-equation terms (gen_ext adds the
// ghost fill, in compute_seq / preStep, the buoyancy and, in 3D, the viscous
// terms in metric form).  The 3D flux
// kernel's dz is cf.dx (src/gen3d.cpp:359); both are 1 on these grids.
void launch_compute_gen(const Geo &g, const Phys &ph, const fst_config &cf, const Bufs &b, hipStream_t s) {
  launch_eos_vel(g, ph, b, s);
  Range gr = ghost_range(g), cr = cell_range(g), fr = face_range(g);
  k_genvel<<<grid_for(gr), 256, 0, s>>>(g, gr, b.var, b.rho, b.metrics, b.vel);
  if (g.ndim == 3 && b.prim) {  // fluxes + divergence + metric pressure gradient in one fused launch
    if (b.prof) { b.prof->name = "k_adv3d_gen"; b.prof->per = 2; b.prof->begin(s); }
    launch_gen_adv(g, ph, cf, b.var, b.prim, b.metrics, b.dvar, s);
    if (b.prof) b.prof->end(s);
  } else if (g.ndim == 3) {
    Phys pg = ph;
    pg.dz = ph.dx;
    for (int v = 0; v < g.nv; ++v) {
      if (b.prof) { b.prof->name = "k_flux3d"; b.prof->per = 2; b.prof->begin(s); }
      k_flux3d<<<grid_for(fr), 256, 0, s>>>(g, pg, fr, v, b.var, b.p, b.vel, b.fx, b.fy, b.fz);
      if (b.prof) b.prof->end(s);
      k_advect3d<<<grid_for(cr), 256, 0, s>>>(g, cr, v, b.fx, b.fy, b.fz, b.dvar);
    }
    k_genpresgrad3d<<<grid_for(cr), 256, 0, s>>>(g, cr, b.p, b.metrics, b.dvar);
  } else {
    for (int v = 0; v < g.nv; ++v) {
      k_flux2d<<<grid_for(fr), 256, 0, s>>>(g, ph, fr, v, cf.scheme, b.var, b.p, b.rho, b.vel, b.fx, b.fy);
      k_advect2d<<<grid_for(cr), 256, 0, s>>>(g, cr, v, b.fx, b.fy, b.dvar);
    }
    k_genpresgrad2d<<<grid_for(cr), 256, 0, s>>>(g, cr, b.p, b.metrics, b.dvar);
  }
  // gen_ext: the cartesian modules' buoyancy after the pressure gradient, as
  // they order it (src/cart3d.cpp:268-270, src/cart2d.cpp:186-188)
  if (cf.gen_ext && cf.buoyancy) k_buoy_gen<<<grid_for(cr), 256, 0, s>>>(g, ph, cr, b.var, b.rho, b.dvar);
  // gen_ext: the viscous terms in metric form, after the buoyancy as cart3d
  // orders them (src/cart3d.cpp:265-270)
  if (cf.gen_ext && cf.visc && g.ndim == 3) launch_viscgen3d(g, ph, b, s);
}

// gen3d_func::preSim / gen2d_func::preSim: metrics of the interior cells, then
// the mirror-copied metric ghosts of a rank without neighbours, 3D in the order
// x-, x+, y-, y+, z-, z+ (src/gen3d.cpp:131-304), 2D y-, y+, x-, x+
// (src/gen2d.cpp:187-331).
void launch_metrics(const Geo &g, const Bufs &b, hipStream_t s) {
  const Nodes nd{g.nci + 1, g.ncj + 1, g.ndim == 3 ? g.nck + 1 : 1};
  Range cr = cell_range(g);
  if (g.ndim == 3) k_metrics3d<<<grid_for(cr), 256, 0, s>>>(g, cr, nd, b.grid, b.metrics);
  else k_metrics2d<<<grid_for(cr), 256, 0, s>>>(g, cr, nd, b.grid, b.metrics);
  static const int order3[6][2] = {{0, -1}, {0, 1}, {1, -1}, {1, 1}, {2, -1}, {2, 1}};
  static const int order2[4][2] = {{1, -1}, {1, 1}, {0, -1}, {0, 1}};
  const int nf = g.ndim == 3 ? 6 : 4;
  for (int f = 0; f < nf; ++f) {
    const int dir = g.ndim == 3 ? order3[f][0] : order2[f][0];
    const int side = g.ndim == 3 ? order3[f][1] : order2[f][1];
    const long long n = (long long)(dir == 0 ? g.ng : g.ngi) * (dir == 1 ? g.ng : g.ngj) * (dir == 2 ? g.ng : g.ngk);
    k_metric_ghost<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(g, dir, side, b.metrics);
  }
}

template <typename T>
void apply_bcs_t(const Geo &g, const fst_config &cf, T *var, hipStream_t s) {
  const int bs = 256;
  if (cf.xPer) {
    if (2 * g.ng <= 8) {
      long long n = 8LL * g.ngj * g.ngk * g.nvt;
      k_bc_xper<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(g, var);
    } else {
      long long n = (long long)g.ngj * g.ngk * g.nvt;
      k_bc_periodic<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(g, 0, var);
    }
  }
  if (cf.yPer) {
    long long n = (long long)g.ngi * g.ngk * g.nvt;
    k_bc_periodic<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(g, 1, var);
  }
  if (g.ndim == 3 && cf.zPer) {
    long long n = (long long)g.ngi * g.ngj * g.nvt;
    k_bc_periodic<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(g, 2, var);
  }
  const int ng = g.ng;
  // k_bc_gen: 32 columns x nvt variables per block (nvt <= 17)
  const dim3 gb(32, g.nvt);
  auto gen = [&](long long ncol, int type, int ih, int jh, int kh, int plane) {
    k_bc_gen<<<(unsigned)((ncol + 31) / 32), gb, 0, s>>>(g, type, ih, jh, kh, plane, var);
  };
  if (!cf.xPer) {
    const long long n = (long long)g.ngj * g.ngk;
    gen(n, cf.bcL, -1, 0, 0, ng);
    gen(n, cf.bcR, 1, 0, 0, ng + g.nci - 1);
  }
  if (!cf.yPer) {
    const long long n = (long long)g.ngi * g.ngk;
    gen(n, cf.bcB, 0, -1, 0, ng);
    gen(n, cf.bcT, 0, 1, 0, ng + g.ncj - 1);
  }
  if (g.ndim == 3 && !cf.zPer) {
    const long long n = (long long)g.ngi * g.ngj;
    gen(n, cf.bcH, 0, 0, -1, ng);
    gen(n, cf.bcF, 0, 0, 1, ng + g.nck - 1);
  }
}

void launch_apply_bcs(const Geo &g, const fst_config &cf, double *var, hipStream_t s) { apply_bcs_t(g, cf, var, s); }
void launch_apply_bcs(const Geo &g, const fst_config &cf, float *var, hipStream_t s) { apply_bcs_t(g, cf, var, s); }

void launch_rk_stage(const Geo &g, int nstages, int stage, double dt, double *var, double *tmp, const double *dvar,
                     hipStream_t s) {
  long long n = g.N * g.nvt;
  k_rk_stage<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, nstages, stage, dt, var, tmp, dvar);
}

// Coarse points per direction: cartesian modules 0..noise_half(nc)
// (src/cart3d.cpp:325-341, src/cart2d.cpp:261-272); gen2d 0..M-1 with
// M = (nc-1)/2 or nc/2 (src/gen2d.cpp:155-164).
void noise_points(const Geo &g, bool gen, int p[3]) {
  const int nc[3] = {g.nci, g.ncj, g.nck};
  for (int a = 0; a < 3; ++a) {
    if (a == 2 && g.ndim == 2) { p[a] = 1; continue; }
    p[a] = gen ? ((nc[a] - 1) % 2 == 0 ? (nc[a] - 1) / 2 : nc[a] / 2) : noise_half(nc[a]) + 1;
  }
}

int noise_flag_count(const Geo &g, bool gen) {
  int p[3];
  noise_points(g, gen, p);
  return p[0] * p[1] * p[2];
}

void launch_noise(const Geo &g, const Phys &ph, const fst_config &cf, const Bufs &b, hipStream_t s) {
  const bool gen = cf.grid == 1;
  const long long ncell = (long long)g.nci * g.ncj * g.nck;
  const unsigned nb = (unsigned)((ncell + 255) / 256);
  // the C_hat cut-off: cartesian modules with the C-equation on; gen2d uses 0
  const bool use_coff = !gen && cf.ceq;
  if (use_coff) {
    k_fill_u64<<<1, 64, 0, s>>>(b.nmax, 1, 0ull);
    k_noise_cmax<<<std::min(nb, 1024u), 256, 0, s>>>(g, b.var, g.ndim == 3 ? 6 : g.nv + 1, b.nmax);
  }
  int p[3];
  noise_points(g, gen, p);
  const int nf = p[0] * p[1] * p[2];
  // variables: mode 1 the momentum components, else the energy (cartesian);
  // gen2d always the two momentum components
  int vars[3], nvars;
  if (gen || cf.n_mode == 1) {
    nvars = g.ndim;
    for (int q = 0; q < nvars; ++q) vars[q] = q;
  } else {
    nvars = 1;
    vars[0] = g.ndim;
  }
  if (nf == 0 || ncell == 0) return;
  for (int q = 0; q < nvars; ++q) {
    k_noise_flags<<<(unsigned)((nf + 255) / 256), 256, 0, s>>>(g, ph, cf, b.var, vars[q], p[0], p[1], p[2],
                                                                b.nflag);
    for (int tau = 0; tau < cf.n_nt; ++tau) {
      k_noise_dn<<<nb, 256, 0, s>>>(g, ph, cf, b.var, vars[q], p[0], p[1], p[2], b.nflag, b.nmax,
                                    use_coff ? 1 : 0, b.ndn);
      k_noise_add<<<nb, 256, 0, s>>>(g, b.var, vars[q], b.ndn);
    }
  }
}

void launch_pack(const double *src, long long nx, long long ny, int ndim, int off, const PackBox &bx, int average,
                 bool f32, void *dst, hipStream_t s) {
  const long long n = bx.ext[0] * bx.ext[1] * (ndim == 3 ? bx.ext[2] : 1);
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (f32) k_pack<float><<<nb, 256, 0, s>>>(src, nx, ny, ndim, off, bx, average, (float *)dst);
  else k_pack<double><<<nb, 256, 0, s>>>(src, nx, ny, ndim, off, bx, average, (double *)dst);
}

void launch_copy(const void *src, void *dst, size_t bytes, hipStream_t s) {
  k_copy32<<<8192, 256, 0, s>>>((const dv4 *)src, (dv4 *)dst, (long long)(bytes / 32));
}

void launch_copy8(const void *src, void *dst, size_t bytes, hipStream_t s) {
  k_copy8<<<8192, 256, 0, s>>>((const double *)src, (double *)dst, (long long)(bytes / 8));
}

void launch_minmax(const Geo &g, const double *a, int nvar, unsigned long long *mm, hipStream_t s) {
  k_minmax_init<<<1, 64, 0, s>>>(nvar, mm);
  const int rows = g.ncj * (g.ndim == 3 ? g.nck : 1);
  const unsigned nb = (unsigned)std::min(rows, 2048);
  k_minmax<<<dim3(nb, (unsigned)nvar), 256, 0, s>>>(a, g.N, g.ndim, g.ng, g.ngi, g.ngj, g.nci, g.ncj, rows, mm);
}

void launch_secondary(const Geo &g, const Phys &ph, const Bufs &b, hipStream_t s) {
  launch_eos_vel(g, ph, b, s);
  k_copy_extra<<<(unsigned)((g.N + 255) / 256), 256, 0, s>>>(g.N, g.ndim, b.vel, b.p, b.T, b.rho, b.varx);
}

}  // namespace fst
