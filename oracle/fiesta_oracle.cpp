/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.
 *
 * CPU restatement of FIESTA's right-hand-side evaluation, Runge-Kutta update
 * and ghost-cell fill, written from the reference's kernels in the same
 * arithmetic order (same association, same constants, no FMA contraction:
 * build with -ffp-contract=off).  Only tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py may load it; the product never does.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the FIESTA source tree).  Loops are OpenMP-parallel over the slowest index;
 * every value is a per-cell function and the only reductions are maxima, so
 * the result is independent of the thread count.
 *
 * Pinning (what checks each function independently of this restatement):
 *   apply_bcs       the reference's own known-answer tests
 *                   (tests/bc_outflow.cpp, bc_reflective.cpp, bc_hydrostatic.cpp,
 *                   tests/test.cpp:327-456) restated in tests/test_oracle_bc.py
 *   eos / velocity  closed forms (ideal-gas single species, mass-fraction mixture)
 *   weno3d/weno2d   constant data (binary32 coefficient sums), linear data,
 *                   brute-force high-precision evaluation
 *   advect+presgrad telescoping conservation on periodic domains
 *   rk              amplification factor 1+z+z^2/2 on dU/dt = lambda*U
 *   viscous         linear shear (uniform stress: zero divergence), quadratic
 *                   profile with analytic divergence
 *   C-equation      uniform state (zero sources), planar-shock switch logic,
 *                   relaxation of a uniform C field; artificial-viscosity
 *                   faces (calculateCeqFaces/Grads/applyCeq and the 2D set)
 *                   on quadratic velocities: the anisotropic Laplacian
 *                   alpha rho C_hat sum M_fd H_fd exactly
 *   full 2D/3D step Sod and planar 3D shock tubes against the exact Riemann
 *                   solution
 *   2D quirks       fixed mu = 2.928e-5 and k = 0.026 (quadratic u, quadratic
 *                   T), the x-face dudy index (linear v gives -mu s/(4 dy)),
 *                   maxWaveSpeed2D's stride-2 table reads and |m| (uniform
 *                   state: the C relaxation exposes maxS)
 */
#include "fiesta_hip.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <cfloat>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Orc {
  fst_config c;
  int ndim, ns, ng, nci, ncj, nck, ngi, ngj, ngk, nv, nvt, nvx;
  size_t N;
  std::vector<double> cd;  // device config array: src/cart3d.cpp:152-171
  std::vector<double> var, tmp1, dvar, varx, p, rho, T, vel;
  std::vector<double> fluxx, fluxy, fluxz;
  std::vector<double> stress;            // 3D (ngi,ngj,ngk,6)
  std::vector<double> stressx, stressy;  // 2D (ngi,ngj,2,2)
  std::vector<double> qx, qy;            // 2D heat flux
  std::vector<double> gradRho;           // (.,5) 3D, (.,4) 2D
  std::vector<double> mFlux;             // 3D (3,3,ngi,ngj,ngk,3), 2D (2,2,ngi,ngj,2)
  // generalized grids (gen2d/gen3d modules, cf.grid = 1)
  bool gen = false;
  int ni = 0, nj = 0, nk = 0;            // grid nodes: nc + 1 (nk = 1 in 2D), src/input.cpp:401-406
  std::vector<double> grid;              // (ni,nj,nk,3)
  std::vector<double> metrics;           // (ngi,ngj,ngk,3,3) 3D, (ngi,ngj,2,2) 2D
  double dxmag;                          // src/cart3d.cpp:173 (no sqrt)
  double maxS = 0, maxC = 0, maxCh = 0, alpha = 0;
  // learned C-indicator form: C_hat and direction fields given from outside
  bool learned = false;
  std::vector<double> lchat, ldir;  // [N], [3][N]
  double lmaxCh = 0;

  size_t id(int i, int j, int k) const {
    return (size_t)i + (size_t)ngi * ((size_t)j + (size_t)ngj * (size_t)k);
  }
  double &V(std::vector<double> &a, int i, int j, int k, int v) {
    return a[id(i, j, k) + N * (size_t)v];
  }
  // mFlux(face,dir,i,j,k,w): LayoutLeft over (3,3,ngi,ngj,ngk,3)
  double &MF(int f, int d, int i, int j, int k, int w) {
    return mFlux[(size_t)f + 3 * ((size_t)d + 3 * (id(i, j, k) + N * (size_t)w))];
  }
  // 2D m(face,dir,i,j,w): (2,2,ngi,ngj,2)
  double &M2(int f, int d, int i, int j, int w) {
    return mFlux[(size_t)f + 2 * ((size_t)d + 2 * (id(i, j, 0) + N * (size_t)w))];
  }
};

// ---------------------------------------------------------------- 3D kernels

// calculateRhoPT3D, src/secondary.hpp:109-151
void eos3d(Orc &o) {
  const int ns = o.ns;
#pragma omp parallel for
  for (int k = 0; k < o.ngk; ++k)
    for (int j = 0; j < o.ngj; ++j)
      for (int i = 0; i < o.ngi; ++i) {
        size_t c = o.id(i, j, k);
        double r = 0.0;
        for (int s = 0; s < ns; ++s) r = r + o.var[c + o.N * (4 + s)];
        double Cp = 0, Cv = 0;
        for (int s = 0; s < ns; ++s) {
          double gammas = o.cd[6 + 3 * s];
          double Rs = o.cd[6 + 3 * s + 1];
          Cp = Cp + (o.var[c + o.N * (4 + s)] / r) * (gammas * Rs / (gammas - 1));
          Cv = Cv + (o.var[c + o.N * (4 + s)] / r) * (Rs / (gammas - 1));
        }
        double gamma = Cp / Cv;
        double m0 = o.var[c], m1 = o.var[c + o.N], m2 = o.var[c + 2 * o.N];
        double pp = (gamma - 1) * (o.var[c + 3 * o.N] - (0.5 / r) * (m0 * m0 + m1 * m1 + m2 * m2));
        o.rho[c] = r;
        o.p[c] = pp;
        o.T[c] = pp / ((Cp - Cv) * r);
      }
}

// computeVelocity3D, src/velocity.hpp:68-83
void vel3d(Orc &o) {
#pragma omp parallel for
  for (int k = 0; k < o.ngk; ++k)
    for (int j = 0; j < o.ngj; ++j)
      for (int i = 0; i < o.ngi; ++i) {
        size_t c = o.id(i, j, k);
        for (int d = 0; d < 3; ++d) o.vel[c + o.N * d] = o.var[c + o.N * d] / o.rho[c];
      }
}

// calculateFluxesG::weno, src/flux.hpp:240-264 (binary32 constants)
double weno3d(double f1, double f2, double f3, double f4, double f5) {
  const double eps = 0.000001;  // src/flux.hpp:233
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

// calculateFluxesG::operator(), src/flux.hpp:266-309, launched over the tile
// range [ng-1, n-ng)^3 (src/cart3d.cpp:228-231, 245)
void flux3d(Orc &o, int v, double dz) {
  const int ng = o.ng;
  const double dx = o.c.dx, dy = o.c.dy;
  double *var = o.var.data() + o.N * v;
  const double *p = o.p.data();
  const double *vel = o.vel.data();
  const size_t N = o.N;
  const double isE = (double)(v == 3);
#pragma omp parallel for
  for (int k = ng - 1; k < o.ngk - ng; ++k)
    for (int j = ng - 1; j < o.ngj - ng; ++j)
      for (int i = ng - 1; i < o.ngi - ng; ++i) {
        double ur, vr, wr, f1, f2, f3, f4, f5, f6;
        auto q = [&](int ii, int jj, int kk) {
          size_t c = o.id(ii, jj, kk);
          return var[c] + isE * p[c];
        };
        auto u = [&](int ii, int jj, int kk, int d) { return vel[o.id(ii, jj, kk) + N * d]; };
        size_t c = o.id(i, j, k);
        f1 = q(i + 3, j, k); f2 = q(i + 2, j, k); f3 = q(i + 1, j, k);
        f4 = q(i, j, k);     f5 = q(i - 1, j, k); f6 = q(i - 2, j, k);
        ur = (-u(i + 2, j, k, 0) + 7.0f * u(i + 1, j, k, 0) + 7.0f * u(i, j, k, 0) - u(i - 1, j, k, 0)) / 12.0f;
        if (ur < 0.0) o.fluxx[c] = ur * weno3d(f1, f2, f3, f4, f5) / dx;
        else          o.fluxx[c] = ur * weno3d(f6, f5, f4, f3, f2) / dx;

        f1 = q(i, j + 3, k); f2 = q(i, j + 2, k); f3 = q(i, j + 1, k);
        f4 = q(i, j, k);     f5 = q(i, j - 1, k); f6 = q(i, j - 2, k);
        vr = (-u(i, j + 2, k, 1) + 7.0f * u(i, j + 1, k, 1) + 7.0f * u(i, j, k, 1) - u(i, j - 1, k, 1)) / 12.0f;
        if (vr < 0.0) o.fluxy[c] = vr * weno3d(f1, f2, f3, f4, f5) / dy;
        else          o.fluxy[c] = vr * weno3d(f6, f5, f4, f3, f2) / dy;

        f1 = q(i, j, k + 3); f2 = q(i, j, k + 2); f3 = q(i, j, k + 1);
        f4 = q(i, j, k);     f5 = q(i, j, k - 1); f6 = q(i, j, k - 2);
        wr = (-u(i, j, k + 2, 2) + 7.0f * u(i, j, k + 1, 2) + 7.0f * u(i, j, k, 2) - u(i, j, k - 1, 2)) / 12.0f;
        if (wr < 0.0) o.fluxz[c] = wr * weno3d(f1, f2, f3, f4, f5) / dz;
        else          o.fluxz[c] = wr * weno3d(f6, f5, f4, f3, f2) / dz;
      }
}

// advect3D, src/advect.hpp:39-61 (assignment into dvar)
void advect3d(Orc &o, int v) {
  const int ng = o.ng;
#pragma omp parallel for
  for (int k = ng; k < o.ngk - ng; ++k)
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        double res = -((o.fluxx[o.id(i, j, k)] - o.fluxx[o.id(i - 1, j, k)]) +
                       (o.fluxy[o.id(i, j, k)] - o.fluxy[o.id(i, j - 1, k)]) +
                       (o.fluxz[o.id(i, j, k)] - o.fluxz[o.id(i, j, k - 1)]));
        o.V(o.dvar, i, j, k, v) = res;
      }
}

// applyPressureGradient3D, src/presgrad.hpp:44-71 (dz := cd(2), :56)
void presgrad3d(Orc &o) {
  const int ng = o.ng;
  const double dx = o.cd[1], dy = o.cd[2], dz = o.cd[2];
#pragma omp parallel for
  for (int k = ng; k < o.ngk - ng; ++k)
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        auto P = [&](int a, int b, int cc) { return o.p[o.id(a, b, cc)]; };
        double dxp = (P(i - 2, j, k) - 8.0 * P(i - 1, j, k) + 8.0 * P(i + 1, j, k) - P(i + 2, j, k)) / (12.0 * dx);
        double dyp = (P(i, j - 2, k) - 8.0 * P(i, j - 1, k) + 8.0 * P(i, j + 1, k) - P(i, j + 2, k)) / (12.0 * dy);
        double dzp = (P(i, j, k - 2) - 8.0 * P(i, j, k - 1) + 8.0 * P(i, j, k + 1) - P(i, j, k + 2)) / (12.0 * dz);
        o.V(o.dvar, i, j, k, 0) = o.V(o.dvar, i, j, k, 0) - dxp;
        o.V(o.dvar, i, j, k, 1) = o.V(o.dvar, i, j, k, 1) - dyp;
        o.V(o.dvar, i, j, k, 2) = o.V(o.dvar, i, j, k, 2) - dzp;
      }
}

// computeBuoyancy3D, src/buoyancy.hpp:22-47
void buoyancy3d(Orc &o) {
  const int ng = o.ng;
  const double g = o.c.gAccel, rhoRef = o.c.rhoRef;
#pragma omp parallel for
  for (int k = ng; k < o.ngk - ng; ++k)
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        size_t c = o.id(i, j, k);
        double v = o.var[c + o.N] / o.rho[c];
        double rhop = o.rho[c] - rhoRef;
        double f = -g * rhop;
        o.dvar[c + o.N] += f;
        o.dvar[c + 3 * o.N] += v * f;
      }
}

// calculateStressTensor{x,y,z}3dv, src/viscosity.hpp:416-459, 499-541, 581-624,
// launched over the weno range (src/cart3d.cpp:265-270)
void stress3d(Orc &o, int dir) {
  const int ng = o.ng, ns = o.ns;
  const double dx = o.cd[1], dy = o.cd[2], dz = o.cd[3];
  const size_t N = o.N;
#pragma omp parallel for
  for (int k = ng - 1; k < o.ngk - ng; ++k)
    for (int j = ng - 1; j < o.ngj - ng; ++j)
      for (int i = ng - 1; i < o.ngi - ng; ++i) {
        auto u = [&](int a, int b, int cc, int d) { return o.vel[o.id(a, b, cc) + N * d]; };
        auto Y = [&](int a, int b, int cc, int s) {
          size_t q = o.id(a, b, cc);
          return o.var[q + N * (4 + s)] / o.rho[q];
        };
        double dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz, mu;
        const int ip = dir == 0, jp = dir == 1, kp = dir == 2;
        mu = 0.0;
        for (int s = 0; s < ns; ++s)
          mu += ((Y(i + ip, j + jp, k + kp, s) + Y(i, j, k, s)) / 2.0) * o.cd[6 + 3 * s + 2];
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
        size_t c = o.id(i, j, k);
        o.stress[c + N * 0] = (2.0 / 3.0) * mu * (2.0 * dudx - dvdy - dwdz);
        o.stress[c + N * 1] = (2.0 / 3.0) * mu * (2.0 * dvdy - dudx - dwdz);
        o.stress[c + N * 2] = (2.0 / 3.0) * mu * (2.0 * dwdz - dudx - dvdy);
        o.stress[c + N * 3] = mu * (dudy + dvdx);
        o.stress[c + N * 4] = mu * (dudz + dwdx);
        o.stress[c + N * 5] = mu * (dvdz + dwdy);
      }
}

// applyViscousTerm{x,y,z}3dv, src/viscosity.hpp:460-498, 542-580, 625-663
void applyvisc3d(Orc &o, int dir) {
  const int ng = o.ng;
  const size_t N = o.N;
  // stress components entering a, b, c, d for each direction
  static const int comp[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
  const double d = o.cd[1 + dir];
  const int ih = dir == 0, jh = dir == 1, kh = dir == 2;
#pragma omp parallel for
  for (int k = ng; k < o.ngk - ng; ++k)
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        size_t c = o.id(i, j, k), cp = o.id(i + ih, j + jh, k + kh), cm = o.id(i - ih, j - jh, k - kh);
        auto u = [&](size_t q, int w) { return o.vel[q + N * w]; };
        double ur = (u(cp, 0) + u(c, 0)) / 2.0;
        double ul = (u(cm, 0) + u(c, 0)) / 2.0;
        double vr = (u(cp, 1) + u(c, 1)) / 2.0;
        double vl = (u(cm, 1) + u(c, 1)) / 2.0;
        double wr = (u(cp, 2) + u(c, 2)) / 2.0;
        double wl = (u(cm, 2) + u(c, 2)) / 2.0;
        auto S = [&](size_t q, int s) { return o.stress[q + N * s]; };
        const int s0 = comp[dir][0], s1 = comp[dir][1], s2 = comp[dir][2];
        double a = (S(c, s0) - S(cm, s0)) / d;
        double b = (S(c, s1) - S(cm, s1)) / d;
        double cc = (S(c, s2) - S(cm, s2)) / d;
        double dd = (ur * S(c, s0) - ul * S(cm, s0)) / d +
                    (vr * S(c, s1) - vl * S(cm, s1)) / d +
                    (wr * S(c, s2) - wl * S(cm, s2)) / d;
        o.dvar[c + N * 0] += a;
        o.dvar[c + N * 1] += b;
        o.dvar[c + N * 2] += cc;
        o.dvar[c + N * 3] += dd;
      }
}

// calculateRhoGrad, src/ceq3d.hpp:95-187
void rhograd3d(Orc &o) {
  const int ng = o.ng;
  const size_t N = o.N;
  const double dx = o.c.dx, dy = o.c.dy, dz = o.c.dz;
#pragma omp parallel for
  for (int k = ng; k < o.ngk - ng; ++k)
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        auto rho = [&](int a, int b, int cc) { return o.rho[o.id(a, b, cc)]; };
        auto vel = [&](int a, int b, int cc, int w) { return o.vel[o.id(a, b, cc) + N * w]; };
        auto nrg = [&](int a, int b, int cc) {
          size_t q = o.id(a, b, cc);
          return o.var[q + 3 * N] / o.rho[q] -
                 0.5 * o.rho[q] * (o.vel[q] * o.vel[q] + o.vel[q + N] * o.vel[q + N] + o.vel[q + 2 * N] * o.vel[q + 2 * N]);
        };
        auto dRho = [&](int ih, int jh, int kh, double d) {
          return (rho(i - 2 * ih, j - 2 * jh, k - 2 * kh) - 8.0 * rho(i - ih, j - jh, k - kh) +
                  8.0 * rho(i + ih, j + jh, k + kh) - rho(i + 2 * ih, j + 2 * jh, k + 2 * kh)) / (12.0 * d);
        };
        auto dE = [&](int ih, int jh, int kh, double d) {
          return (nrg(i - 2 * ih, j - 2 * jh, k - 2 * kh) - 8.0 * nrg(i - ih, j - jh, k - kh) +
                  8.0 * nrg(i + ih, j + jh, k + kh) - nrg(i + 2 * ih, j + 2 * jh, k + 2 * kh)) / (12.0 * d);
        };
        auto dV = [&](int ih, int jh, int kh, int w, double d) {
          return (vel(i - 2 * ih, j - 2 * jh, k - 2 * kh, w) - 8.0 * vel(i - ih, j - jh, k - kh, w) +
                  8.0 * vel(i + ih, j + jh, k + kh, w) - vel(i + 2 * ih, j + 2 * jh, k + 2 * kh, w)) / (12.0 * d);
        };
        double dxr = dRho(1, 0, 0, dx), dyr = dRho(0, 1, 0, dy), dzr = dRho(0, 0, 1, dz);
        double dxe = dE(1, 0, 0, dx), dye = dE(0, 1, 0, dy), dze = dE(0, 0, 1, dz);
        double dxu = dV(1, 0, 0, 0, dx), dyv = dV(0, 1, 0, 1, dy), dzw = dV(0, 0, 1, 2, dz);
        double n1 = dxr, n2 = dyr, n3 = dzr;
        double rgrad = std::sqrt(dxr * dxr + dyr * dyr + dzr * dzr);
        double divu = dxu + dyv + dzw;
        double dnednr = (n1 * dxe + n2 * dye + n3 * dze) * (n1 * dxr + n2 * dyr + n3 * dzr);
        int indicator = dnednr < 0.0 ? 1 : 0;
        size_t c = o.id(i, j, k);
        if (divu < 0) o.gradRho[c] = (1 - indicator) * rgrad;
        else          o.gradRho[c] = 0;
        o.gradRho[c + N * 1] = indicator * rgrad;
        o.gradRho[c + N * 2] = dxr;
        o.gradRho[c + N * 3] = dyr;
        o.gradRho[c + N * 4] = dzr;
      }
}

// maxWaveSpeed reduce (src/ceq3d.hpp:25-63) and maxGradFunctor on var(nv+1)
// (src/ceq3d.hpp:82-93, called src/cart3d.cpp:281-283); fsMax returns the
// local maximum (src/cart3d.cpp:190-198, the one-rank value)
void maxima3d(Orc &o) {
  const int ng = o.ng, ns = o.ns;
  const size_t N = o.N;
  double ms = -DBL_MAX, mc = -DBL_MAX;
#pragma omp parallel for reduction(max : ms, mc)
  for (int k = ng; k < o.ngk - ng; ++k)
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        size_t c = o.id(i, j, k);
        double Cp = 0, Cv = 0;
        for (int s = 0; s < ns; ++s) {
          double gammas = o.cd[6 + 3 * s], Rs = o.cd[6 + 3 * s + 1];
          Cp = Cp + (o.var[c + N * (4 + s)] / o.rho[c]) * (gammas * Rs / (gammas - 1));
          Cv = Cv + (o.var[c + N * (4 + s)] / o.rho[c]) * (Rs / (gammas - 1));
        }
        double gamma = Cp / Cv;
        double a = std::sqrt(gamma * o.p[c] / o.rho[c]);
        double s = a + std::sqrt(o.var[c] * o.var[c] + o.var[c + N] * o.var[c + N] + o.var[c + 2 * N] * o.var[c + 2 * N]) / o.rho[c];
        if (s > ms) ms = s;
        double ch = std::abs(o.var[c + N * (o.nv + 1)]);
        if (ch > mc) mc = ch;
      }
  o.maxS = ms;
  o.maxCh = mc;
}

// updateCeq, src/ceq3d.hpp:189-224
void updateceq3d(Orc &o) {
  const int ng = o.ng;
  const size_t N = o.N;
  const double dx = o.cd[1], dy = o.cd[2], dz = o.cd[3];
  const int nc = (int)o.cd[0] + 4;
  const double maxS = o.maxS, kap = o.c.kap, eps = o.c.eps;
  const double dxmag = std::sqrt(dx * dx + dy * dy + dz * dz);
#pragma omp parallel for
  for (int k = ng; k < o.ngk - ng; ++k)
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        for (int n = 0; n < 5; ++n) {
          auto C = [&](int a, int b, int cc) { return o.var[o.id(a, b, cc) + N * (nc + n)]; };
          double lap = (C(i - 1, j, k) - 2 * C(i, j, k) + C(i + 1, j, k)) / dx +
                       (C(i, j - 1, k) - 2 * C(i, j, k) + C(i, j + 1, k)) / dy +
                       (C(i, j, k - 1) - 2 * C(i, j, k) + C(i, j, k + 1)) / dz;
          size_t c = o.id(i, j, k);
          o.dvar[c + N * (nc + n)] =
              (maxS / (eps * dxmag)) * (o.gradRho[c + N * n] - o.var[c + N * (nc + n)]) + kap * maxS * dxmag * lap;
        }
      }
}

// calculateCeqFaces (src/ceq3d.hpp:226-296) then calculateCeqGrads
// (src/ceq3d.hpp:298-332), both over the weno range (src/cart3d.cpp:288-289)
void ceqfaces3d(Orc &o) {
  const int ng = o.ng, nv = o.nv;
  const size_t N = o.N;
  const double alpha = o.alpha;
#pragma omp parallel for
  for (int k = ng - 1; k < o.ngk - ng; ++k)
    for (int j = ng - 1; j < o.ngj - ng; ++j)
      for (int i = ng - 1; i < o.ngi - ng; ++i) {
        double r, chat, cn[3], M, cmag;
        for (int face = 0; face < 3; ++face) {
          int ip = face == 0, jp = face == 1, kp = face == 2;
          size_t c = o.id(i, j, k), q = o.id(i + ip, j + jp, k + kp);
          r = (o.rho[c] + o.rho[q]) / 2.0;
          if (o.learned) {  // learned form: the same face composition on the given fields
            chat = (o.lchat[c] + o.lchat[q]) / 2.0;
            for (int d = 0; d < 3; ++d) cn[d] = (o.ldir[c + N * d] + o.ldir[q + N * d]) / 2.0;
          } else {
            chat = (o.var[c + N * (nv + 1)] + o.var[q + N * (nv + 1)]) / 2.0;
            cn[0] = (o.var[c + N * (nv + 2)] + o.var[q + N * (nv + 2)]) / 2.0;
            cn[1] = (o.var[c + N * (nv + 3)] + o.var[q + N * (nv + 3)]) / 2.0;
            cn[2] = (o.var[c + N * (nv + 4)] + o.var[q + N * (nv + 4)]) / 2.0;
          }
          cmag = 1.0e-6;
          cmag += cn[0] * cn[0];
          cmag += cn[1] * cn[1];
          cmag += cn[2] * cn[2];
          for (int dir = 0; dir < 3; ++dir) {
            double dirac = 0.0;
            if (face == dir) dirac = 1.0;
            M = dirac - (cn[face] * cn[dir]) / cmag;
            for (int w = 0; w < 3; ++w) o.MF(face, dir, i, j, k, w) = alpha * r * chat * M;
          }
        }
      }
  const double dx = o.c.dx, dy = o.c.dy, dz = o.c.dz;
#pragma omp parallel for
  for (int k = ng - 1; k < o.ngk - ng; ++k)
    for (int j = ng - 1; j < o.ngj - ng; ++j)
      for (int i = ng - 1; i < o.ngi - ng; ++i) {
        auto vel = [&](int a, int b, int cc, int w) { return o.vel[o.id(a, b, cc) + N * w]; };
        for (int w = 0; w < 3; ++w) {
          o.MF(0, 0, i, j, k, w) *= (vel(i + 1, j, k, w) - vel(i, j, k, w)) / dx;
          o.MF(0, 1, i, j, k, w) *= ((vel(i, j + 1, k, w) + vel(i + 1, j + 1, k, w)) - (vel(i, j - 1, k, w) + vel(i + 1, j - 1, k, w))) / (4 * dy);
          o.MF(0, 2, i, j, k, w) *= ((vel(i, j, k + 1, w) + vel(i + 1, j, k + 1, w)) - (vel(i, j, k - 1, w) + vel(i + 1, j, k - 1, w))) / (4 * dz);
          o.MF(1, 0, i, j, k, w) *= ((vel(i + 1, j + 1, k, w) + vel(i + 1, j, k, w)) - (vel(i - 1, j + 1, k, w) + vel(i - 1, j, k, w))) / (4 * dx);
          o.MF(1, 1, i, j, k, w) *= (vel(i, j + 1, k, w) - vel(i, j, k, w)) / dy;
          o.MF(1, 2, i, j, k, w) *= ((vel(i, j + 1, k + 1, w) + vel(i, j, k + 1, w)) - (vel(i, j + 1, k - 1, w) + vel(i, j, k - 1, w))) / (4 * dz);
          o.MF(2, 0, i, j, k, w) *= ((vel(i + 1, j, k + 1, w) + vel(i + 1, j, k, w)) - (vel(i - 1, j, k + 1, w) + vel(i - 1, j, k, w))) / (4 * dx);
          o.MF(2, 1, i, j, k, w) *= ((vel(i, j + 1, k + 1, w) + vel(i, j + 1, k, w)) - (vel(i, j - 1, k + 1, w) + vel(i, j - 1, k, w))) / (4 * dy);
          o.MF(2, 2, i, j, k, w) *= (vel(i, j, k + 1, w) - vel(i, j, k, w)) / dz;
        }
      }
}

// applyCeq, src/ceq3d.hpp:334-366
void applyceq3d(Orc &o) {
  const int ng = o.ng;
  const size_t N = o.N;
  const double d[3] = {o.c.dx, o.c.dy, o.c.dz};
#pragma omp parallel for
  for (int k = ng; k < o.ngk - ng; ++k)
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        for (int w = 0; w < 3; ++w) {
          double diffu = 0.0;
          for (int face = 0; face < 3; ++face) {
            int ih = face == 0, jh = face == 1, kh = face == 2;
            for (int dir = 0; dir < 3; ++dir)
              diffu += (o.MF(face, dir, i, j, k, w) - o.MF(face, dir, i - ih, j - jh, k - kh, w)) / d[face];
          }
          o.dvar[o.id(i, j, k) + N * w] += diffu;
        }
      }
}

// ---------------------------------------------------------------- 2D kernels

// calculateRhoPT2D, src/secondary.hpp:60-107
void eos2d(Orc &o) {
  const int ns = o.ns;
#pragma omp parallel for
  for (int j = 0; j < o.ngj; ++j)
    for (int i = 0; i < o.ngi; ++i) {
      size_t c = o.id(i, j, 0);
      double r = 0.0;
      for (int s = 0; s < ns; ++s) r = r + o.var[c + o.N * (3 + s)];
      double Cp = 0, Cv = 0;
      for (int s = 0; s < ns; ++s) {
        double gammas = o.cd[6 + 3 * s], Rs = o.cd[6 + 3 * s + 1];
        Cp = Cp + (o.var[c + o.N * (3 + s)] / r) * (gammas * Rs / (gammas - 1));
        Cv = Cv + (o.var[c + o.N * (3 + s)] / r) * (Rs / (gammas - 1));
      }
      double gamma = Cp / Cv;
      double m0 = o.var[c], m1 = o.var[c + o.N];
      double pp = (gamma - 1) * (o.var[c + 2 * o.N] - (0.5 / r) * (m0 * m0 + m1 * m1));
      o.rho[c] = r;
      o.p[c] = pp;
      o.T[c] = pp / ((Cp - Cv) * r);
    }
}

// computeVelocity2D, src/velocity.hpp:85-99
void vel2d(Orc &o) {
#pragma omp parallel for
  for (int j = 0; j < o.ngj; ++j)
    for (int i = 0; i < o.ngi; ++i) {
      size_t c = o.id(i, j, 0);
      o.vel[c] = o.var[c] / o.rho[c];
      o.vel[c + o.N] = o.var[c + o.N] / o.rho[c];
    }
}

// computeFluxWeno2D::weno, src/flux.hpp:38-62 (double constants)
double weno2d(double f1, double f2, double f3, double f4, double f5) {
  const double eps = 1e-6;  // src/flux.hpp:32
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

// computeFluxWeno2D (src/flux.hpp:64-105), computeFluxCentered2D (:121-166),
// computeFluxQuick2D (:184-220), all over the face range (src/cart2d.cpp:154,164-172)
void flux2d(Orc &o, int v) {
  const int ng = o.ng;
  const size_t N = o.N;
  const double isE = (double)(v == 2);
  const int scheme = o.c.scheme;
#pragma omp parallel for
  for (int j = ng - 1; j < o.ngj - ng; ++j)
    for (int i = ng - 1; i < o.ngi - ng; ++i) {
      auto var = [&](int a, int b, int w) { return o.var[o.id(a, b, 0) + N * w]; };
      auto P = [&](int a, int b) { return o.p[o.id(a, b, 0)]; };
      auto R = [&](int a, int b) { return o.rho[o.id(a, b, 0)]; };
      auto u = [&](int a, int b, int w) { return o.vel[o.id(a, b, 0) + N * w]; };
      auto q = [&](int a, int b) { return var(a, b, v) + isE * P(a, b); };
      size_t c = o.id(i, j, 0);
      if (scheme == FST_CENTERED4) {
        const double dx = o.cd[1], dy = o.cd[2];
        double ur, vr, x1, y1, px, py;
        ur = (0.0 - var(i + 2, j, 0) / R(i + 2, j) + 7.0 * var(i + 1, j, 0) / R(i + 1, j) +
              7.0 * var(i, j, 0) / R(i, j) - var(i - 1, j, 0) / R(i - 1, j)) / 12.0;
        vr = (0.0 - var(i, j + 2, 1) / R(i, j + 2) + 7.0 * var(i, j + 1, 1) / R(i, j + 1) +
              7.0 * var(i, j, 1) / R(i, j) - var(i, j - 1, 1) / R(i, j - 1)) / 12.0;
        px = (0.0 - P(i + 2, j) + 7.0 * P(i + 1, j) + 7.0 * P(i, j) - P(i - 1, j)) / 12.0;
        py = (0.0 - P(i, j + 2) + 7.0 * P(i, j + 1) + 7.0 * P(i, j) - P(i, j - 1)) / 12.0;
        x1 = (0.0 - var(i + 2, j, v) + 7.0 * var(i + 1, j, v) + 7.0 * var(i, j, v) - var(i - 1, j, v)) / 12.0;
        y1 = (0.0 - var(i, j + 2, v) + 7.0 * var(i, j + 1, v) + 7.0 * var(i, j, v) - var(i, j - 1, v)) / 12.0;
        if (v == 2) {
          x1 = x1 + px;
          y1 = y1 + py;
        }
        o.fluxx[c] = ur * x1 / dx;
        o.fluxy[c] = vr * y1 / dy;
      } else if (scheme == FST_QUICK) {
        const double dx = o.cd[1], dy = o.cd[2];
        auto quick = [](double f1, double f2, double f3) { return 0.75 * f2 + 0.375 * f3 - 0.125 * f1; };
        double ur, vr, f1, f2, f3;
        ur = (-u(i + 2, j, 0) + 7.0 * u(i + 1, j, 0) + 7.0 * u(i, j, 0) - u(i - 1, j, 0)) / 12.0;
        if (ur < 0.0) { f1 = q(i + 2, j); f2 = q(i + 1, j); f3 = q(i, j); }
        else          { f1 = q(i - 1, j); f2 = q(i, j);     f3 = q(i + 1, j); }
        o.fluxx[c] = ur * quick(f1, f2, f3) / dx;
        vr = (-u(i, j + 2, 1) + 7.0 * u(i, j + 1, 1) + 7.0 * u(i, j, 1) - u(i, j - 1, 1)) / 12.0;
        if (vr < 0.0) { f1 = q(i, j + 2); f2 = q(i, j + 1); f3 = q(i, j); }
        else          { f1 = q(i, j - 1); f2 = q(i, j);     f3 = q(i, j + 1); }
        o.fluxy[c] = vr * quick(f1, f2, f3) / dy;
      } else {
        const double dx = o.c.dx, dy = o.c.dy;
        double ur, vr, f1, f2, f3, f4, f5;
        ur = (-u(i + 2, j, 0) + 7.0 * u(i + 1, j, 0) + 7.0 * u(i, j, 0) - u(i - 1, j, 0)) / 12.0;
        vr = (-u(i, j + 2, 1) + 7.0 * u(i, j + 1, 1) + 7.0 * u(i, j, 1) - u(i, j - 1, 1)) / 12.0;
        if (ur < 0.0) { f1 = q(i + 3, j); f2 = q(i + 2, j); f3 = q(i + 1, j); f4 = q(i, j); f5 = q(i - 1, j); }
        else          { f1 = q(i - 2, j); f2 = q(i - 1, j); f3 = q(i, j); f4 = q(i + 1, j); f5 = q(i + 2, j); }
        o.fluxx[c] = ur * weno2d(f1, f2, f3, f4, f5) / dx;
        if (vr < 0.0) { f1 = q(i, j + 3); f2 = q(i, j + 2); f3 = q(i, j + 1); f4 = q(i, j); f5 = q(i, j - 1); }
        else          { f1 = q(i, j - 2); f2 = q(i, j - 1); f3 = q(i, j); f4 = q(i, j + 1); f5 = q(i, j + 2); }
        o.fluxy[c] = vr * weno2d(f1, f2, f3, f4, f5) / dy;
      }
    }
}

// advect2D, src/advect.hpp:23-37
void advect2d(Orc &o, int v) {
  const int ng = o.ng;
#pragma omp parallel for
  for (int j = ng; j < o.ngj - ng; ++j)
    for (int i = ng; i < o.ngi - ng; ++i)
      o.V(o.dvar, i, j, 0, v) = -((o.fluxx[o.id(i, j, 0)] - o.fluxx[o.id(i - 1, j, 0)]) +
                                  (o.fluxy[o.id(i, j, 0)] - o.fluxy[o.id(i, j - 1, 0)]));
}

// applyPressureGradient2D, src/presgrad.hpp:22-42
void presgrad2d(Orc &o) {
  const int ng = o.ng;
  const double dx = o.cd[1], dy = o.cd[2];
#pragma omp parallel for
  for (int j = ng; j < o.ngj - ng; ++j)
    for (int i = ng; i < o.ngi - ng; ++i) {
      auto P = [&](int a, int b) { return o.p[o.id(a, b, 0)]; };
      double dxp = (P(i - 2, j) - 8.0 * P(i - 1, j) + 8.0 * P(i + 1, j) - P(i + 2, j)) / (12.0 * dx);
      double dyp = (P(i, j - 2) - 8.0 * P(i, j - 1) + 8.0 * P(i, j + 1) - P(i, j + 2)) / (12.0 * dy);
      o.V(o.dvar, i, j, 0, 0) -= dxp;
      o.V(o.dvar, i, j, 0, 1) -= dyp;
    }
}

// computeBuoyancy (2D), src/buoyancy.hpp:49-73
void buoyancy2d(Orc &o) {
  const int ng = o.ng;
  const double g = o.c.gAccel, rhoRef = o.c.rhoRef;
  const size_t N = o.N;
#pragma omp parallel for
  for (int j = ng; j < o.ngj - ng; ++j)
    for (int i = ng; i < o.ngi - ng; ++i) {
      size_t c = o.id(i, j, 0);
      double v = o.var[c + N] / o.var[c + 3 * N];
      double rhop = o.rho[c] - rhoRef;
      double eps = 1e-6;
      if (rhop >= eps || rhop <= -eps) {
        double f = -g * rhop;
        o.dvar[c + N] += f;
        o.dvar[c + 2 * N] += std::abs(v) * f;
      }
    }
}

// 2D stress on x and y faces (2,2) layout helper
inline double &S2(std::vector<double> &a, const Orc &o, size_t c, int r, int s) {
  return a[c + o.N * (size_t)(r + 2 * s)];
}

// calculateStressTensor2dv (src/viscosity.hpp:25-109) and calculateHeatFlux2dv
// (:111-136) over the face range, applyViscousTerm2dv (:138-199) over cells
void visc2d(Orc &o) {
  const int ng = o.ng;
  const size_t N = o.N;
  const double dx = o.cd[1], dy = o.cd[2];
#pragma omp parallel for
  for (int j = ng - 1; j < o.ngj - ng; ++j)
    for (int i = ng - 1; i < o.ngi - ng; ++i) {
      auto vel = [&](int a, int b, int w) { return o.vel[o.id(a, b, 0) + N * w]; };
      double dudx, dvdy, dudy, dvdx;
      double mur = 2.928e-5, mut = 2.928e-5;
      size_t c = o.id(i, j, 0);
      dudx = (vel(i + 1, j, 0) - vel(i, j, 0)) / dx;
      dvdx = (vel(i + 1, j, 1) - vel(i, j, 1)) / dx;
      dudy = ((vel(i, j + 1, 0) + vel(i + 1, j + 1, 0)) - (vel(i, j - 1, 1) + vel(i + 1, j - 1, 0))) / (4.0 * dy);
      dvdy = ((vel(i, j + 1, 1) + vel(i + 1, j + 1, 1)) - (vel(i, j - 1, 1) + vel(i + 1, j - 1, 1))) / (4.0 * dy);
      S2(o.stressx, o, c, 0, 0) = (2.0 / 3.0) * mur * (2.0 * dudx - dvdy);
      S2(o.stressx, o, c, 1, 1) = (2.0 / 3.0) * mur * (2.0 * dvdy - dudx);
      S2(o.stressx, o, c, 0, 1) = mur * (dudy + dvdx);
      S2(o.stressx, o, c, 1, 0) = S2(o.stressx, o, c, 0, 1);
      dudx = ((vel(i + 1, j, 0) + vel(i + 1, j + 1, 0)) - (vel(i - 1, j, 0) + vel(i - 1, j + 1, 0))) / (4.0 * dx);
      dvdx = ((vel(i + 1, j, 1) + vel(i + 1, j + 1, 1)) - (vel(i - 1, j, 1) + vel(i - 1, j + 1, 1))) / (4.0 * dx);
      dudy = (vel(i, j + 1, 0) - vel(i, j, 0)) / dy;
      dvdy = (vel(i, j + 1, 1) - vel(i, j, 1)) / dy;
      S2(o.stressy, o, c, 0, 0) = (2.0 / 3.0) * mut * (2.0 * dudx - dvdy);
      S2(o.stressy, o, c, 1, 1) = (2.0 / 3.0) * mut * (2.0 * dvdy - dudx);
      S2(o.stressy, o, c, 0, 1) = mut * (dudy + dvdx);
      S2(o.stressy, o, c, 1, 0) = S2(o.stressy, o, c, 0, 1);
    }
#pragma omp parallel for
  for (int j = ng - 1; j < o.ngj - ng; ++j)
    for (int i = ng - 1; i < o.ngi - ng; ++i) {
      const double kk = 0.026;
      size_t c = o.id(i, j, 0);
      o.qx[c] = -kk * (o.T[o.id(i + 1, j, 0)] - o.T[c]) / dx;
      o.qy[c] = -kk * (o.T[o.id(i, j + 1, 0)] - o.T[c]) / dy;
    }
#pragma omp parallel for
  for (int j = ng; j < o.ngj - ng; ++j)
    for (int i = ng; i < o.ngi - ng; ++i) {
      auto vel = [&](int a, int b, int w) { return o.vel[o.id(a, b, 0) + N * w]; };
      size_t c = o.id(i, j, 0), cl = o.id(i - 1, j, 0), cb = o.id(i, j - 1, 0);
      double ur = (vel(i + 1, j, 0) + vel(i, j, 0)) / 2.0;
      double ul = (vel(i - 1, j, 0) + vel(i, j, 0)) / 2.0;
      double vr = (vel(i + 1, j, 1) + vel(i, j, 1)) / 2.0;
      double vl = (vel(i - 1, j, 1) + vel(i, j, 1)) / 2.0;
      double ut = (vel(i, j + 1, 0) + vel(i, j, 0)) / 2.0;
      double ub = (vel(i, j - 1, 0) + vel(i, j, 0)) / 2.0;
      double vt = (vel(i, j + 1, 1) + vel(i, j, 1)) / 2.0;
      double vb = (vel(i, j - 1, 1) + vel(i, j, 1)) / 2.0;
      auto sx = [&](size_t q, int r, int s) { return S2(o.stressx, o, q, r, s); };
      auto sy = [&](size_t q, int r, int s) { return S2(o.stressy, o, q, r, s); };
      double a = (sx(c, 0, 0) - sx(cl, 0, 0)) / dx + (sy(c, 0, 1) - sy(cb, 0, 1)) / dy;
      double b = (sx(c, 1, 0) - sx(cl, 1, 0)) / dx + (sy(c, 1, 1) - sy(cb, 1, 1)) / dy;
      double c1 = (ur * sx(c, 0, 0) - ul * sx(cl, 0, 0)) / dx + (vr * sx(c, 0, 1) - vl * sx(cl, 0, 1)) / dx +
                  (ut * sy(c, 1, 0) - ub * sy(cb, 1, 0)) / dy + (vt * sy(c, 1, 1) - vb * sy(cb, 1, 1)) / dy;
      double c2 = (o.qx[c] - o.qx[cl]) / dx + (o.qy[c] - o.qy[cb]) / dy;
      o.dvar[c] = o.dvar[c] + a;
      o.dvar[c + N] = o.dvar[c + N] + b;
      o.dvar[c + 2 * N] = o.dvar[c + 2 * N] + c1 - c2;
    }
}

// 2D C-equation: calculateRhoGrad2D (src/ceq.hpp:80-151), maxWaveSpeed2D
// (:43-78, stride-2 species lookup and |m| kept), maxCvar2D (:23-41),
// updateCeq2D (:153-195), computeCeqFlux2D (:197-245), computeCeqFaces2D
// (:247-267), applyCeq2D (:269-293); sequence src/cart2d.cpp:201-236
// computeCeqFlux2D, computeCeqFaces2D and applyCeq2D (src/ceq.hpp:197-293)
// over the face range; the learned form reads C_hat and the tangent from
// lchat / ldir in place of var(nv+1..nv+3).
void ceqfaces2d(Orc &o) {
  const int ng = o.ng, nv = o.nv;
  const size_t N = o.N;
  const double dx = o.c.dx, dy = o.c.dy;
  const double alpha = o.alpha;
#pragma omp parallel for
  for (int j = ng - 1; j < o.ngj - ng; ++j)
    for (int i = ng - 1; i < o.ngi - ng; ++i) {
      for (int f = 0; f < 2; ++f) {
        int ih = f == 0, jh = f == 1;
        size_t c = o.id(i, j, 0), q = o.id(i + ih, j + jh, 0);
        double r = (o.rho[c] + o.rho[q]) / 2.0;
        double chat, ct[2];
        if (o.learned) {
          chat = (o.lchat[c] + o.lchat[q]) / 2.0;
          ct[0] = (o.ldir[c] + o.ldir[q]) / 2.0;
          ct[1] = (o.ldir[c + N] + o.ldir[q + N]) / 2.0;
        } else {
          chat = (o.var[c + N * (nv + 1)] + o.var[q + N * (nv + 1)]) / 2.0;
          ct[0] = (o.var[c + N * (nv + 2)] + o.var[q + N * (nv + 2)]) / 2.0;
          ct[1] = (o.var[c + N * (nv + 3)] + o.var[q + N * (nv + 3)]) / 2.0;
        }
        double ctmag = (ct[0] * ct[0] + ct[1] * ct[1]) + 1.0e-6;
        for (int d = 0; d < 2; ++d)
          for (int w = 0; w < 2; ++w) o.M2(f, d, i, j, w) = alpha * r * chat * (ct[f] * ct[d]) / ctmag;
      }
    }
  {
    const double cdx = o.cd[1], cdy = o.cd[2];
#pragma omp parallel for
    for (int j = ng - 1; j < o.ngj - ng; ++j)
      for (int i = ng - 1; i < o.ngi - ng; ++i) {
        auto v = [&](int a, int b, int w) { return o.vel[o.id(a, b, 0) + N * w]; };
        for (int w = 0; w < 2; ++w) {
          o.M2(0, 0, i, j, w) *= (v(i + 1, j, w) - v(i, j, w)) / cdx;
          o.M2(0, 1, i, j, w) *= ((v(i, j + 1, w) + v(i + 1, j + 1, w)) - (v(i, j - 1, w) + v(i + 1, j - 1, w))) / (4.0 * cdy);
          o.M2(1, 0, i, j, w) *= ((v(i + 1, j, w) + v(i + 1, j + 1, w)) - (v(i - 1, j, w) + v(i - 1, j + 1, w))) / (4.0 * cdx);
          o.M2(1, 1, i, j, w) *= (v(i, j + 1, w) - v(i, j, w)) / cdy;
        }
      }
  }
#pragma omp parallel for
  for (int j = ng; j < o.ngj - ng; ++j)
    for (int i = ng; i < o.ngi - ng; ++i) {
      for (int w = 0; w < 2; ++w) {
        double diffu = 0;
        diffu += (o.M2(0, 0, i, j, w) - o.M2(0, 0, i - 1, j, w)) / dx;
        diffu += (o.M2(0, 1, i, j, w) - o.M2(0, 1, i - 1, j, w)) / dx;
        diffu += (o.M2(1, 0, i, j, w) - o.M2(1, 0, i, j - 1, w)) / dy;
        diffu += (o.M2(1, 1, i, j, w) - o.M2(1, 1, i, j - 1, w)) / dy;
        o.dvar[o.id(i, j, 0) + N * w] += diffu;
      }
    }
}

void ceq2d(Orc &o) {
  const int ng = o.ng, ns = o.ns;
  const size_t N = o.N;
  const double dx = o.c.dx, dy = o.c.dy;
  if (o.learned) {
    // learned form: no relaxation solve; the artificial-viscosity terms of
    // src/cart2d.cpp:223-236 on the given C_hat, tangent and max C_hat
    o.alpha = (o.c.dx * o.c.dx + o.c.dy * o.c.dy) / (o.lmaxCh + 1.0e-6) * o.c.alpha;
    ceqfaces2d(o);
    return;
  }
#pragma omp parallel for
  for (int j = ng; j < o.ngj - ng; ++j)
    for (int i = ng; i < o.ngi - ng; ++i) {
      auto R = [&](int a, int b) { return o.rho[o.id(a, b, 0)]; };
      auto U = [&](int a, int b, int w) { return o.vel[o.id(a, b, 0) + N * w]; };
      auto nrg = [&](int a, int b) {
        size_t q = o.id(a, b, 0);
        return o.var[q + 2 * N] / o.rho[q] - 0.5 * o.rho[q] * (o.vel[q] * o.vel[q] + o.vel[q + N] * o.vel[q + N]);
      };
      auto dR = [&](int ih, int jh, double d) {
        return (R(i - 2 * ih, j - 2 * jh) - 8.0 * R(i - ih, j - jh) + 8.0 * R(i + ih, j + jh) - R(i + 2 * ih, j + 2 * jh)) / (12.0 * d);
      };
      auto dE = [&](int ih, int jh, double d) {
        return (nrg(i - 2 * ih, j - 2 * jh) - 8.0 * nrg(i - ih, j - jh) + 8.0 * nrg(i + ih, j + jh) - nrg(i + 2 * ih, j + 2 * jh)) / (12.0 * d);
      };
      auto dV = [&](int ih, int jh, int w, double d) {
        return (U(i - 2 * ih, j - 2 * jh, w) - 8.0 * U(i - ih, j - jh, w) + 8.0 * U(i + ih, j + jh, w) - U(i + 2 * ih, j + 2 * jh, w)) / (12.0 * d);
      };
      double dxr = dR(1, 0, dx), dyr = dR(0, 1, dy);
      double dxe = dE(1, 0, dx), dye = dE(0, 1, dy);
      double dxu = dV(1, 0, 0, dx), dyv = dV(0, 1, 1, dy);
      double n1 = dxr, n2 = dyr;
      double rgrad = std::sqrt(dxr * dxr + dyr * dyr);
      double divu = dxu + dyv;
      double dnednr = (n1 * dxe + n2 * dye) * (n1 * dxr + n2 * dyr);
      double i1 = 0;
      if (dnednr < 0.0) i1 = 1.0; else i1 = 0.0;
      double i2 = 0;
      if (divu < 0.0) i2 = 1.0;
      size_t c = o.id(i, j, 0);
      o.gradRho[c] = (1 - i1) * i2 * rgrad;
      o.gradRho[c + N] = i1 * rgrad;
      o.gradRho[c + 2 * N] = -dyr;
      o.gradRho[c + 3 * N] = dxr;
    }
  const int nv = o.nv;
  double ms = -DBL_MAX, mc = -DBL_MAX, mch = -DBL_MAX;
#pragma omp parallel for reduction(max : ms, mc, mch)
  for (int j = ng; j < o.ngj - ng; ++j)
    for (int i = ng; i < o.ngi - ng; ++i) {
      size_t c = o.id(i, j, 0);
      double Cp = 0, Cv = 0;
      for (int s = 0; s < ns; ++s) {
        double gammas = o.cd[6 + 2 * s], Rs = o.cd[6 + 2 * s + 1];
        Cp = Cp + (o.var[c + N * (3 + s)] / o.rho[c]) * (gammas * Rs / (gammas - 1));
        Cv = Cv + (o.var[c + N * (3 + s)] / o.rho[c]) * (Rs / (gammas - 1));
      }
      double gamma = Cp / Cv;
      double a = std::sqrt(gamma * o.p[c] / o.rho[c]);
      double s = a + std::sqrt(o.var[c] * o.var[c] + o.var[c + N] * o.var[c + N]);
      if (s > ms) ms = s;
      double c0 = std::abs(o.var[c + N * (nv + 0)]);
      if (c0 > mc) mc = c0;
      double c1 = std::abs(o.var[c + N * (nv + 1)]);
      if (c1 > mch) mch = c1;
    }
  o.maxS = ms;
  o.maxC = mc;
  o.maxCh = mch;
  o.alpha = (o.c.dx * o.c.dx + o.c.dy * o.c.dy) / (o.maxCh + 1.0e-6) * o.c.alpha;
  {
    const double cdx = o.cd[1], cdy = o.cd[2];
    const int nc = (int)o.cd[0] + 3;
    const double maxS = o.maxS, kap = o.c.kap, eps = o.c.eps;
    const double dxmag = std::sqrt(cdx * cdx + cdy * cdy);
#pragma omp parallel for
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        for (int n = 0; n < 4; ++n) {
          auto C = [&](int a, int b) { return o.var[o.id(a, b, 0) + N * (nc + n)]; };
          double dx_right = (C(i + 1, j) - C(i, j)) / cdx;
          double dx_left = (C(i, j) - C(i - 1, j)) / cdx;
          double dy_top = (C(i, j + 1) - C(i, j)) / cdy;
          double dy_bot = (C(i, j) - C(i, j - 1)) / cdy;
          double lap = dx_right - dx_left + dy_top - dy_bot;
          size_t c = o.id(i, j, 0);
          o.dvar[c + N * (nc + n)] =
              (maxS / (eps * dxmag)) * (o.gradRho[c + N * n] - C(i, j)) + kap * maxS * dxmag * lap;
        }
      }
  }
  ceqfaces2d(o);
}

// ---------------------------------------------------------------- generalized grids

// grid(i,j,k,v) over (ni,nj,nk,3), LayoutLeft (src/gen3d.cpp:41, src/gen2d.cpp:25)
inline double GR(const Orc &o, int i, int j, int k, int v) {
  return o.grid[(size_t)i + (size_t)o.ni * ((size_t)j + (size_t)o.nj * ((size_t)k + (size_t)o.nk * (size_t)v))];
}
// metrics(i,j,k,m,n) over (ngi,ngj,ngk,3,3) in 3D (src/gen3d.cpp:43) and
// metrics(i,j,m,n) over (ngi,ngj,2,2) in 2D (src/gen2d.cpp:26)
inline double &MT(Orc &o, int i, int j, int k, int m, int n) {
  return o.metrics[o.id(i, j, k) + o.N * ((size_t)m + (size_t)o.ndim * (size_t)n)];
}

// computeMetrics3D, src/metric.hpp:59-164, over the cell range
// (src/gen3d.cpp:138-144).  Node offsets use the literal 3 of the reference
// (ii = i - 3), not ng.
void metrics3d(Orc &o) {
#pragma omp parallel for
  for (int k = o.ng; k < o.ngk - o.ng; ++k)
    for (int j = o.ng; j < o.ngj - o.ng; ++j)
      for (int i = o.ng; i < o.ngi - o.ng; ++i) {
        const int ii = i - 3, jj = j - 3, kk = k - 3;
        auto g = [&](int a, int b, int c, int v) { return GR(o, a, b, c, v); };
        double xmx = (g(ii, jj, kk, 0) + g(ii, jj + 1, kk, 0) + g(ii, jj, kk + 1, 0) + g(ii, jj + 1, kk + 1, 0)) / 4.0;
        double xpx = (g(ii + 1, jj, kk, 0) + g(ii + 1, jj + 1, kk, 0) + g(ii + 1, jj, kk + 1, 0) +
                      g(ii + 1, jj + 1, kk + 1, 0)) / 4.0;
        double xmy = (g(ii, jj, kk, 1) + g(ii, jj + 1, kk, 1) + g(ii, jj, kk + 1, 1) + g(ii, jj + 1, kk + 1, 1)) / 4.0;
        double xpy = (g(ii + 1, jj, kk, 1) + g(ii + 1, jj + 1, kk, 1) + g(ii + 1, jj, kk + 1, 1) +
                      g(ii + 1, jj + 1, kk + 1, 1)) / 4.0;
        double xmz = (g(ii, jj, kk, 2) + g(ii, jj + 1, kk, 2) + g(ii, jj, kk + 1, 2) + g(ii, jj + 1, kk + 1, 2)) / 4.0;
        double xpz = (g(ii + 1, jj, kk, 2) + g(ii + 1, jj + 1, kk, 2) + g(ii + 1, jj, kk + 1, 2) +
                      g(ii + 1, jj + 1, kk + 1, 2)) / 4.0;

        double zmx = (g(ii, jj, kk, 0) + g(ii + 1, jj, kk, 0) + g(ii, jj + 1, kk, 0) + g(ii + 1, jj + 1, kk, 0)) / 4.0;
        double zpx = (g(ii, jj, kk + 1, 0) + g(ii + 1, jj, kk + 1, 0) + g(ii, jj + 1, kk + 1, 0) +
                      g(ii + 1, jj + 1, kk + 1, 0)) / 4.0;
        double zmy = (g(ii, jj, kk, 1) + g(ii + 1, jj, kk, 1) + g(ii, jj + 1, kk, 1) + g(ii + 1, jj + 1, kk, 1)) / 4.0;
        double zpy = (g(ii, jj, kk + 1, 1) + g(ii + 1, jj, kk + 1, 1) + g(ii, jj + 1, kk + 1, 1) +
                      g(ii + 1, jj + 1, kk + 1, 1)) / 4.0;
        double zmz = (g(ii, jj, kk, 2) + g(ii + 1, jj, kk, 2) + g(ii, jj + 1, kk, 2) + g(ii + 1, jj + 1, kk, 2)) / 4.0;
        double zpz = (g(ii, jj, kk + 1, 2) + g(ii + 1, jj, kk + 1, 2) + g(ii, jj + 1, kk + 1, 2) +
                      g(ii + 1, jj + 1, kk + 1, 2)) / 4.0;

        double ymx = (g(ii, jj, kk, 0) + g(ii + 1, jj, kk, 0) + g(ii, jj, kk + 1, 0) + g(ii + 1, jj, kk + 1, 0)) / 4.0;
        double ypx = (g(ii, jj + 1, kk, 0) + g(ii + 1, jj + 1, kk, 0) + g(ii, jj + 1, kk + 1, 0) +
                      g(ii + 1, jj + 1, kk + 1, 0)) / 4.0;
        double ymy = (g(ii, jj, kk, 1) + g(ii + 1, jj, kk, 1) + g(ii, jj, kk + 1, 1) + g(ii + 1, jj, kk + 1, 1)) / 4.0;
        double ypy = (g(ii, jj + 1, kk, 1) + g(ii + 1, jj + 1, kk, 1) + g(ii, jj + 1, kk + 1, 1) +
                      g(ii + 1, jj + 1, kk + 1, 1)) / 4.0;
        double ymz = (g(ii, jj, kk, 2) + g(ii + 1, jj, kk, 2) + g(ii, jj, kk + 1, 2) + g(ii + 1, jj, kk + 1, 2)) / 4.0;
        double ypz = (g(ii, jj + 1, kk, 2) + g(ii + 1, jj + 1, kk, 2) + g(ii, jj + 1, kk + 1, 2) +
                      g(ii + 1, jj + 1, kk + 1, 2)) / 4.0;

        double xxi = xpx - xmx, xet = ypx - ymx, xzt = zpx - zmx;
        double yxi = xpy - xmy, yet = ypy - ymy, yzt = zpy - zmy;
        double zxi = xpz - xmz, zet = ypz - ymz, zzt = zpz - zmz;
        double jac = 1.0 / (xxi * (yet * zzt - yzt * zet) + xet * (yzt * zxi - yxi * zzt) +
                            xzt * (yxi * zet - yet * zxi));
        MT(o, i, j, k, 0, 0) = jac * (yet * zzt - yzt * zet);
        MT(o, i, j, k, 0, 1) = jac * (xzt * zet - xet * zzt);
        MT(o, i, j, k, 0, 2) = jac * (xet * yzt - xzt * yet);
        MT(o, i, j, k, 1, 0) = jac * (yzt * zxi - yxi * zzt);
        MT(o, i, j, k, 1, 1) = jac * (xxi * zzt - xzt * zxi);
        MT(o, i, j, k, 1, 2) = jac * (xzt * yxi - xxi * yzt);
        MT(o, i, j, k, 2, 0) = jac * (yxi * zet - yet * zxi);
        MT(o, i, j, k, 2, 1) = jac * (xet * zxi - xxi * zet);
        MT(o, i, j, k, 2, 2) = jac * (xxi * yet - xet * yxi);
      }
}

// computeMetrics2D, src/metric.hpp:21-57, over the cell range (src/gen2d.cpp:192)
void metrics2d(Orc &o) {
  for (int j = o.ng; j < o.ngj - o.ng; ++j)
    for (int i = o.ng; i < o.ngi - o.ng; ++i) {
      const int ii = i - 3, jj = j - 3;
      auto g = [&](int a, int b, int v) { return GR(o, a, b, 0, v); };
      double x_xi = (g(ii + 1, jj, 0) + g(ii + 1, jj + 1, 0)) / 2.0 - (g(ii, jj, 0) + g(ii, jj + 1, 0)) / 2.0;
      double y_xi = (g(ii + 1, jj, 1) + g(ii + 1, jj + 1, 1)) / 2.0 - (g(ii, jj, 1) + g(ii, jj + 1, 1)) / 2.0;
      double x_et = (g(ii, jj + 1, 0) + g(ii + 1, jj + 1, 0)) / 2.0 - (g(ii, jj, 0) + g(ii + 1, jj, 0)) / 2.0;
      double y_et = (g(ii, jj + 1, 1) + g(ii + 1, jj + 1, 1)) / 2.0 - (g(ii, jj, 1) + g(ii + 1, jj, 1)) / 2.0;
      double jac = 1.0 / (x_xi * y_et - x_et * y_xi);
      MT(o, i, j, 0, 0, 0) = jac * y_et;
      MT(o, i, j, 0, 0, 1) = -jac * x_et;
      MT(o, i, j, 0, 1, 0) = -jac * y_xi;
      MT(o, i, j, 0, 1, 1) = jac * x_xi;
    }
}

// Metric ghost cells of one rank without neighbours: mirror copies, in the
// reference order -- 3D x-, x+, y-, y+, z-, z+ (src/gen3d.cpp:241-300), 2D
// y-, y+, x-, x+ (src/gen2d.cpp:285-329) -- each over the full range of the
// other indices.
void metric_ghosts(Orc &o) {
  const int ng = o.ng, nm = o.ndim;
  auto copy = [&](int id, int jd, int kd, int is, int js, int ks) {
    for (int m = 0; m < nm; ++m)
      for (int n = 0; n < nm; ++n) MT(o, id, jd, kd, m, n) = MT(o, is, js, ks, m, n);
  };
  auto xm = [&]() {
    for (int k = 0; k < o.ngk; ++k)
      for (int j = 0; j < o.ngj; ++j)
        for (int i = 0; i < ng; ++i) copy(ng - i - 1, j, k, ng + i, j, k);
  };
  auto xp = [&]() {
    for (int k = 0; k < o.ngk; ++k)
      for (int j = 0; j < o.ngj; ++j)
        for (int i = 0; i < ng; ++i) copy(o.ngi - 1 - i, j, k, o.nci + i, j, k);
  };
  auto ym = [&]() {
    for (int k = 0; k < o.ngk; ++k)
      for (int j = 0; j < ng; ++j)
        for (int i = 0; i < o.ngi; ++i) copy(i, ng - j - 1, k, i, ng + j, k);
  };
  auto yp = [&]() {
    for (int k = 0; k < o.ngk; ++k)
      for (int j = 0; j < ng; ++j)
        for (int i = 0; i < o.ngi; ++i) copy(i, o.ngj - 1 - j, k, i, o.ncj + j, k);
  };
  if (o.ndim == 3) {
    xm(); xp(); ym(); yp();
    for (int k = 0; k < ng; ++k)
      for (int j = 0; j < o.ngj; ++j)
        for (int i = 0; i < o.ngi; ++i) copy(i, j, ng - k - 1, i, j, ng + k);
    for (int k = 0; k < ng; ++k)
      for (int j = 0; j < o.ngj; ++j)
        for (int i = 0; i < o.ngi; ++i) copy(i, j, o.ngk - 1 - k, i, j, o.nck + k);
  } else {
    ym(); yp(); xm(); xp();
  }
}

// computeGenVelocity3D, src/velocity.hpp:23-48: contravariant velocity
// vel(d) = sum_m metrics(d,m) * m_m / rho, over all cells
void genvel3d(Orc &o) {
  const size_t N = o.N;
#pragma omp parallel for
  for (int k = 0; k < o.ngk; ++k)
    for (int j = 0; j < o.ngj; ++j)
      for (int i = 0; i < o.ngi; ++i) {
        const size_t c = o.id(i, j, k);
        const double r = o.rho[c], m0 = o.var[c], m1 = o.var[c + N], m2 = o.var[c + 2 * N];
        for (int d = 0; d < 3; ++d)
          o.vel[c + N * d] = (MT(o, i, j, k, d, 0) * m0 / r + MT(o, i, j, k, d, 1) * m1 / r +
                              MT(o, i, j, k, d, 2) * m2 / r);
      }
}

// computeGenVelocity2D, src/velocity.hpp:50-66
void genvel2d(Orc &o) {
  const size_t N = o.N;
  for (int j = 0; j < o.ngj; ++j)
    for (int i = 0; i < o.ngi; ++i) {
      const size_t c = o.id(i, j, 0);
      const double r = o.rho[c], m0 = o.var[c], m1 = o.var[c + N];
      for (int d = 0; d < 2; ++d)
        o.vel[c + N * d] = (MT(o, i, j, 0, d, 0) * m0 / r + MT(o, i, j, 0, d, 1) * m1 / r);
    }
}

// applyGenPressureGradient3D, src/presgrad.hpp:72-147, over the cell range
void genpresgrad3d(Orc &o) {
  const size_t N = o.N;
  const int ng = o.ng;
#pragma omp parallel for
  for (int k = ng; k < o.ngk - ng; ++k)
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        auto P = [&](int a, int b, int c) { return o.p[o.id(a, b, c)]; };
        auto M = [&](int a, int b, int c, int m, int n) { return MT(o, a, b, c, m, n); };
        double d[3];
        for (int n = 0; n < 3; ++n) {
          double dxi = (P(i - 2, j, k) * M(i - 2, j, k, 0, n) - 8.0 * P(i - 1, j, k) * M(i - 1, j, k, 0, n) +
                        8.0 * P(i + 1, j, k) * M(i + 1, j, k, 0, n) - P(i + 2, j, k) * M(i + 2, j, k, 0, n)) /
                       (12.0);
          double det = (P(i, j - 2, k) * M(i, j - 2, k, 1, n) - 8.0 * P(i, j - 1, k) * M(i, j - 1, k, 1, n) +
                        8.0 * P(i, j + 1, k) * M(i, j + 1, k, 1, n) - P(i, j + 2, k) * M(i, j + 2, k, 1, n)) /
                       (12.0);
          double dzt = (P(i, j, k - 2) * M(i, j, k - 2, 2, n) - 8.0 * P(i, j, k - 1) * M(i, j, k - 1, 2, n) +
                        8.0 * P(i, j, k + 1) * M(i, j, k + 1, 2, n) - P(i, j, k + 2) * M(i, j, k + 2, 2, n)) /
                       (12.0);
          d[n] = dxi + det + dzt;
        }
        const size_t c = o.id(i, j, k);
        for (int n = 0; n < 3; ++n) o.dvar[c + N * n] = (o.dvar[c + N * n] - d[n]);
      }
}

// applyGenPressureGradient2D, src/presgrad.hpp:149-190, over the cell range
void genpresgrad2d(Orc &o) {
  const size_t N = o.N;
  const int ng = o.ng;
  for (int j = ng; j < o.ngj - ng; ++j)
    for (int i = ng; i < o.ngi - ng; ++i) {
      auto P = [&](int a, int b) { return o.p[o.id(a, b, 0)]; };
      auto M = [&](int a, int b, int m, int n) { return MT(o, a, b, 0, m, n); };
      double d[2];
      for (int n = 0; n < 2; ++n) {
        double dxi = (P(i - 2, j) * M(i - 2, j, 0, n) - 8.0 * P(i - 1, j) * M(i - 1, j, 0, n) +
                      8.0 * P(i + 1, j) * M(i + 1, j, 0, n) - P(i + 2, j) * M(i + 2, j, 0, n)) /
                     (12.0);
        double det = (P(i, j - 2) * M(i, j - 2, 1, n) - 8.0 * P(i, j - 1) * M(i, j - 1, 1, n) +
                      8.0 * P(i, j + 1) * M(i, j + 1, 1, n) - P(i, j + 2) * M(i, j + 2, 1, n)) /
                     (12.0);
        d[n] = dxi + det;
      }
      const size_t c = o.id(i, j, 0);
      for (int n = 0; n < 2; ++n) o.dvar[c + N * n] = (o.dvar[c + N * n] - d[n]);
    }
}

// ---------------------------------------------------------------- ghosts

// bc_gen, src/bc.cpp:33-131, launched on one plane (src/bc.cpp:217-247)
void bc_gen(Orc &o, int type, int ih, int jh, int kh, int i0, int i1, int j0, int j1, int k0, int k1) {
  const int ng = o.ng, nv = o.nvt, ndim = o.ndim;
  for (int k = k0; k < k1; ++k)
    for (int j = j0; j < j1; ++j)
      for (int i = i0; i < i1; ++i) {
        auto U = [&](int a, int b, int c, int v) -> double & { return o.var[o.id(a, b, c) + o.N * v]; };
        if (type == FST_BC_OUTFLOW) {
          for (int v = 0; v < nv; ++v)
            for (int n = 0; n < ng; ++n)
              U(i + ih * (n + 1), j + jh * (n + 1), k + kh * (n + 1), v) = U(i - ih * n, j - jh * n, k - kh * n, v);
        }
        if (type == FST_BC_REFLECTIVE || type == FST_BC_HYDROSTATIC) {
          for (int v = 0; v < nv; ++v) {
            double reflect = 1.0;
            if (ih != 0 && v == 0) reflect = -1.0;
            if (jh != 0 && v == 1) reflect = -1.0;
            if (kh != 0 && v == 2) reflect = -1.0;
            for (int n = 0; n < ng; ++n)
              U(i + ih * (n + 1), j + jh * (n + 1), k + kh * (n + 1), v) = reflect * U(i - ih * n, j - jh * n, k - kh * n, v);
          }
        }
        if (type == FST_BC_NOSLIP) {
          for (int v = 0; v < nv; ++v) {
            double reflect = 1.0;
            if (ih != 0 && (v == 0 || v == 1 || v == 2)) reflect = -1.0;
            if (jh != 0 && (v == 0 || v == 1 || v == 2)) reflect = -1.0;
            if (kh != 0 && (v == 0 || v == 1 || v == 2)) reflect = -1.0;
            for (int n = 0; n < ng; ++n)
              U(i + ih * (n + 1), j + jh * (n + 1), k + kh * (n + 1), v) = reflect * U(i - ih * n, j - jh * n, k - kh * n, v);
          }
        }
        if (type == FST_BC_HYDROSTATIC) {
          const int edx = ndim, rdx = ndim + 1;
          for (int n = 0; n < ng; ++n) {
            int ig = i + ih * (n + 1), jg = j + jh * (n + 1), kg = k + kh * (n + 1);
            int i1_ = ig - ih, j1_ = jg - jh, k1_ = kg - kh;
            int i2_ = ig - 2 * ih, j2_ = jg - 2 * jh, k2_ = kg - 2 * kh;
            double ke = 0, ke1 = 0, ke2 = 0;
            for (int d = 0; d < ndim; ++d) {
              ke += 0.5 * (U(ig, jg, kg, d) * U(ig, jg, kg, d));
              ke1 += 0.5 * (U(i1_, j1_, k1_, d) * U(i1_, j1_, k1_, d));
              ke2 += 0.5 * (U(i2_, j2_, k2_, d) * U(i2_, j2_, k2_, d));
            }
            ke = ke / U(ig, jg, kg, rdx);
            ke1 = ke1 / U(i1_, j1_, k1_, rdx);
            ke2 = ke2 / U(i2_, j2_, k2_, rdx);
            U(ig, jg, kg, edx) = 2 * (U(i1_, j1_, k1_, edx) - ke1) - (U(i2_, j2_, k2_, edx) - ke2) + ke;
          }
        }
      }
}

// applyBCs, src/bc.cpp:179-249 (no-MPI branch); bc_xPer/yPer/zPer :133-176
void apply_bcs(Orc &o) {
  const int ng = o.ng, nvt = o.nvt;
  auto U = [&](int a, int b, int c, int v) -> double & { return o.var[o.id(a, b, c) + o.N * v]; };
  if (o.c.xPer == 1)
    for (int v = 0; v < nvt; ++v)
      for (int k = 0; k < o.ngk; ++k)
        for (int j = 0; j < o.ngj; ++j)
          for (int g = 0; g < ng; ++g) {
            U(g, j, k, v) = U(o.nci + g, j, k, v);
            U(o.nci + ng + g, j, k, v) = U(ng + g, j, k, v);
          }
  if (o.c.yPer == 1)
    for (int v = 0; v < nvt; ++v)
      for (int k = 0; k < o.ngk; ++k)
        for (int i = 0; i < o.ngi; ++i)
          for (int g = 0; g < ng; ++g) {
            U(i, g, k, v) = U(i, o.ncj + g, k, v);
            U(i, o.ncj + ng + g, k, v) = U(i, ng + g, k, v);
          }
  if (o.ndim == 3 && o.c.zPer == 1)
    for (int v = 0; v < nvt; ++v)
      for (int j = 0; j < o.ngj; ++j)
        for (int i = 0; i < o.ngi; ++i)
          for (int g = 0; g < ng; ++g) {
            U(i, j, g, v) = U(i, j, o.nck + g, v);
            U(i, j, o.nck + ng + g, v) = U(i, j, ng + g, v);
          }
  // xMinus = -1 + xPer etc. (src/input.cpp:408-413)
  if (!o.c.xPer) {
    bc_gen(o, o.c.bcL, -1, 0, 0, ng, ng + 1, 0, o.ngj, 0, o.ngk);
    bc_gen(o, o.c.bcR, 1, 0, 0, ng + o.nci - 1, ng + o.nci, 0, o.ngj, 0, o.ngk);
  }
  if (!o.c.yPer) {
    bc_gen(o, o.c.bcB, 0, -1, 0, 0, o.ngi, ng, ng + 1, 0, o.ngk);
    bc_gen(o, o.c.bcT, 0, 1, 0, 0, o.ngi, ng + o.ncj - 1, ng + o.ncj, 0, o.ngk);
  }
  if (o.ndim == 3 && !o.c.zPer) {
    bc_gen(o, o.c.bcH, 0, 0, -1, 0, o.ngi, 0, o.ngj, ng, ng + 1);
    bc_gen(o, o.c.bcF, 0, 0, 1, 0, o.ngi, 0, o.ngj, ng + o.nck - 1, ng + o.nck);
  }
}

// ---------------------------------------------------------------- modules

// cart3d_func::compute, src/cart3d.cpp:220-298
void compute3d(Orc &o) {
  apply_bcs(o);
  eos3d(o);
  vel3d(o);
  for (int v = 0; v < o.nv; ++v) {
    flux3d(o, v, o.c.dz);
    advect3d(o, v);
  }
  presgrad3d(o);
  if (o.c.buoyancy) buoyancy3d(o);
  if (o.c.visc) {
    for (int d = 0; d < 3; ++d) {
      stress3d(o, d);
      applyvisc3d(o, d);
    }
  }
  if (o.c.ceq && o.learned) {
    // learned form: no relaxation solve; the artificial-viscosity terms of
    // src/cart3d.cpp:285-292 on the given C_hat, n and max C_hat
    o.alpha = (o.dxmag / (o.lmaxCh + 1.0e-6)) * o.c.alpha;
    ceqfaces3d(o);
    applyceq3d(o);
  } else if (o.c.ceq) {
    rhograd3d(o);
    maxima3d(o);
    o.alpha = (o.dxmag / (o.maxCh + 1.0e-6)) * o.c.alpha;
    updateceq3d(o);
    ceqfaces3d(o);
    applyceq3d(o);
  }
}

// Features of the learned C-indicator on the current primitives (eos3d and
// vel3d must have run), in the order and precision of the GPU feature pass:
// FP64 arithmetic, one rounding to FP32 per output.
void features3d(Orc &o, float *F, float *M, float *Dir, float *G1) {
  const size_t N = o.N;
  const double d[3] = {o.c.dx, o.c.dy, o.c.dz};
  const double hmin = std::min(d[0], std::min(d[1], d[2]));
  const long long st[3] = {1, o.ngi, (long long)o.ngi * o.ngj};
#pragma omp parallel for
  for (int k = 0; k < o.ngk; ++k)
    for (int j = 0; j < o.ngj; ++j)
      for (int i = 0; i < o.ngi; ++i) {
        const size_t t = o.id(i, j, k);
        const bool in1 = i >= 1 && j >= 1 && k >= 1 && i < o.ngi - 1 && j < o.ngj - 1 && k < o.ngk - 1;
        const bool in2 = i >= 2 && j >= 2 && k >= 2 && i < o.ngi - 2 && j < o.ngj - 2 && k < o.ngk - 2;
        float f0 = 0, f1 = 0, f2 = 0, d0 = 0, d1 = 0, d2 = 0, gg = 0;
        if (in1) {
          auto nrg = [&](size_t q) {
            const double r = o.rho[q], u = o.vel[q], v = o.vel[q + N], w = o.vel[q + 2 * N];
            return o.var[q + 3 * N] / r - 0.5 * r * (u * u + v * v + w * w);
          };
          double gr2 = 0.0, gp2 = 0.0, divu = 0.0, dr[3] = {0, 0, 0}, de[3] = {0, 0, 0};
          for (int a = 0; a < 3; ++a) {
            const size_t sp = t + st[a], sm = t - st[a];
            const double h2 = 2.0 * d[a];
            const double r1 = (o.rho[sp] - o.rho[sm]) / h2;
            const double p1 = (o.p[sp] - o.p[sm]) / h2;
            gr2 = gr2 + r1 * r1;
            gp2 = gp2 + p1 * p1;
            divu = divu + (o.vel[sp + N * a] - o.vel[sm + N * a]) / h2;
            if (in2) {
              const size_t sp2 = t + 2 * st[a], sm2 = t - 2 * st[a];
              const double h12 = 12.0 * d[a];
              dr[a] = (o.rho[sm2] - 8.0 * o.rho[sm] + 8.0 * o.rho[sp] - o.rho[sp2]) / h12;
              de[a] = (nrg(sm2) - 8.0 * nrg(sm) + 8.0 * nrg(sp) - nrg(sp2)) / h12;
            }
          }
          const double rho = o.rho[t], p = o.p[t];
          double Cp = 0.0, Cv = 0.0;
          for (int s = 0; s < o.ns; ++s) {
            const double gs = o.cd[6 + 3 * s], Rs = o.cd[6 + 3 * s + 1];
            const double Y = o.var[t + N * (4 + s)] / rho;
            Cp = Cp + Y * (gs * Rs / (gs - 1));
            Cv = Cv + Y * (Rs / (gs - 1));
          }
          const double a_snd = std::sqrt((Cp / Cv) * p / rho);
          f0 = (float)(std::sqrt(gr2) * hmin / rho);
          f1 = (float)(std::sqrt(gp2) * hmin / p);
          f2 = (float)(divu * hmin / a_snd);
          if (in2) {
            d0 = (float)dr[0];
            d1 = (float)dr[1];
            d2 = (float)dr[2];
            const double rg2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            const double dnednr = (dr[0] * de[0] + dr[1] * de[1] + dr[2] * de[2]) * rg2;
            gg = dnednr < 0.0 ? (float)(std::sqrt(rg2) * hmin / rho) : 0.f;
          }
        }
        F[t] = f0;
        F[t + N] = f1;
        F[t + 2 * N] = f2;
        M[t] = std::max(std::fabs(f0), std::max(std::fabs(f1), std::fabs(f2)));
        Dir[t] = d0;
        Dir[t + N] = d1;
        Dir[t + 2 * N] = d2;
        G1[t] = gg;
      }
}

// Features of the 2D learned C-indicator on the current primitives (eos2d and
// vel2d must have run), in the order and precision of the GPU feature pass.
// The tangent (-d rho/dy, d rho/dx) and the contact switch follow
// calculateRhoGrad2D (src/ceq.hpp:80-151): 4th-order gradients of rho and of
// nrg = E / rho - rho |u|^2 / 2; G1 = |grad rho| h / rho where
// (grad rho . grad nrg) |grad rho|^2 < 0.
void features2d(Orc &o, float *F, float *M, float *Dir, float *G1) {
  const size_t N = o.N;
  const double d[2] = {o.c.dx, o.c.dy};
  const double hmin = std::min(d[0], d[1]);
  const long long st[2] = {1, o.ngi};
#pragma omp parallel for
  for (int j = 0; j < o.ngj; ++j)
    for (int i = 0; i < o.ngi; ++i) {
      const size_t t = o.id(i, j, 0);
      const bool in1 = i >= 1 && j >= 1 && i < o.ngi - 1 && j < o.ngj - 1;
      const bool in2 = i >= 2 && j >= 2 && i < o.ngi - 2 && j < o.ngj - 2;
      float f0 = 0, f1 = 0, f2 = 0, d0 = 0, d1 = 0, gg = 0;
      if (in1) {
        auto nrg = [&](size_t q) {
          const double r = o.rho[q], u = o.vel[q], v = o.vel[q + N];
          return o.var[q + 2 * N] / r - 0.5 * r * (u * u + v * v);
        };
        double gr2 = 0.0, gp2 = 0.0, divu = 0.0, dr[2] = {0, 0}, de[2] = {0, 0};
        for (int a = 0; a < 2; ++a) {
          const size_t sp = t + st[a], sm = t - st[a];
          const double h2 = 2.0 * d[a];
          const double r1 = (o.rho[sp] - o.rho[sm]) / h2;
          const double p1 = (o.p[sp] - o.p[sm]) / h2;
          gr2 = gr2 + r1 * r1;
          gp2 = gp2 + p1 * p1;
          divu = divu + (o.vel[sp + N * a] - o.vel[sm + N * a]) / h2;
          if (in2) {
            const size_t sp2 = t + 2 * st[a], sm2 = t - 2 * st[a];
            const double h12 = 12.0 * d[a];
            dr[a] = (o.rho[sm2] - 8.0 * o.rho[sm] + 8.0 * o.rho[sp] - o.rho[sp2]) / h12;
            de[a] = (nrg(sm2) - 8.0 * nrg(sm) + 8.0 * nrg(sp) - nrg(sp2)) / h12;
          }
        }
        const double rho = o.rho[t], p = o.p[t];
        double Cp = 0.0, Cv = 0.0;
        for (int s = 0; s < o.ns; ++s) {
          const double gs = o.cd[6 + 3 * s], Rs = o.cd[6 + 3 * s + 1];
          const double Y = o.var[t + N * (3 + s)] / rho;
          Cp = Cp + Y * (gs * Rs / (gs - 1));
          Cv = Cv + Y * (Rs / (gs - 1));
        }
        const double a_snd = std::sqrt((Cp / Cv) * p / rho);
        f0 = (float)(std::sqrt(gr2) * hmin / rho);
        f1 = (float)(std::sqrt(gp2) * hmin / p);
        f2 = (float)(divu * hmin / a_snd);
        if (in2) {
          d0 = (float)(-dr[1]);
          d1 = (float)dr[0];
          const double rg2 = dr[0] * dr[0] + dr[1] * dr[1];
          const double dnednr = (dr[0] * de[0] + dr[1] * de[1]) * rg2;
          gg = dnednr < 0.0 ? (float)(std::sqrt(rg2) * hmin / rho) : 0.f;
        }
      }
      F[t] = f0;
      F[t + N] = f1;
      F[t + 2 * N] = f2;
      M[t] = std::max(std::fabs(f0), std::max(std::fabs(f1), std::fabs(f2)));
      Dir[t] = d0;
      Dir[t + N] = d1;
      Dir[t + 2 * N] = 0.f;
      G1[t] = gg;
    }
}

// gen_ext viscous terms on a generalized grid (the reference's gen3d has none;
// the cartesian split of src/viscosity.hpp:416-663 in index space, with the
// chain rule through the metric terms m(a,n) = d xi_a / d x_n as
// applyGenPressureGradient3D applies it, src/presgrad.hpp:72-147).  On the
// face between c and c + e_D (stored at c, over the cartesian stress range):
//   g[w][a]  index-space derivatives of the cartesian velocity u_w = var(w)/rho:
//            a = D: u(c+e) - u(c); a != D: the 4-point difference / 4
//   mf(a,n)  (m(c,a,n) + m(c+e,a,n)) / 2
//   du_w/dx_n = (mf(0,n) g[w][0] + mf(1,n) g[w][1]) + mf(2,n) g[w][2]
//   tau      from du/dx as calculateStressTensor (mass-fraction mu)
//   F_w      = (mf(D,0) tau_w0 + mf(D,1) tau_w1) + mf(D,2) tau_w2     (momentum)
//   H_w      = (mf(D,0) (uf_w tau_w0) + mf(D,1) (uf_w tau_w1)) + mf(D,2) (uf_w tau_w2)
//            with uf the face velocity (u(c+e) + u(c)) / 2           (energy)
// then dvar(w) += F_w(c) - F_w(c-e), dvar(3) += sum_w (H_w(c) - H_w(c-e)).
// On a uniform grid of power-of-two spacing (mf = diag(1/d)) every step is
// the cartesian operation scaled by a power of two, so the result equals the
// cartesian module's bit for bit.  F and H go through o.stress (6 planes).
void viscgen3d(Orc &o, int dir) {
  const int ng = o.ng, ns = o.ns;
  const size_t N = o.N;
  const int ip = dir == 0, jp = dir == 1, kp = dir == 2;
  auto U = [&](int a, int b, int cc, int w) {
    const size_t q = o.id(a, b, cc);
    return o.var[q + N * w] / o.rho[q];
  };
#pragma omp parallel for
  for (int k = ng - 1; k < o.ngk - ng; ++k)
    for (int j = ng - 1; j < o.ngj - ng; ++j)
      for (int i = ng - 1; i < o.ngi - ng; ++i) {
        double mu = 0.0;
        for (int s = 0; s < ns; ++s) {
          const size_t q1 = o.id(i + ip, j + jp, k + kp), q0 = o.id(i, j, k);
          mu += ((o.var[q1 + N * (4 + s)] / o.rho[q1] + o.var[q0 + N * (4 + s)] / o.rho[q0]) / 2.0) *
                o.cd[6 + 3 * s + 2];
        }
        double g[3][3], mf[3][3], du[3][3], uf[3];
        for (int w = 0; w < 3; ++w) {
          for (int a = 0; a < 3; ++a) {
            if (a == dir) {
              g[w][a] = U(i + ip, j + jp, k + kp, w) - U(i, j, k, w);
            } else {
              const int ta = a == 0, tb = a == 1, tc = a == 2;
              g[w][a] = ((U(i + ta, j + tb, k + tc, w) + U(i + ip + ta, j + jp + tb, k + kp + tc, w)) -
                         (U(i - ta, j - tb, k - tc, w) + U(i + ip - ta, j + jp - tb, k + kp - tc, w))) /
                        4.0;
            }
          }
          uf[w] = (U(i + ip, j + jp, k + kp, w) + U(i, j, k, w)) / 2.0;
        }
        for (int a = 0; a < 3; ++a)
          for (int n = 0; n < 3; ++n) mf[a][n] = (MT(o, i, j, k, a, n) + MT(o, i + ip, j + jp, k + kp, a, n)) / 2.0;
        for (int w = 0; w < 3; ++w)
          for (int n = 0; n < 3; ++n) du[w][n] = (mf[0][n] * g[w][0] + mf[1][n] * g[w][1]) + mf[2][n] * g[w][2];
        const double dudx = du[0][0], dudy = du[0][1], dudz = du[0][2];
        const double dvdx = du[1][0], dvdy = du[1][1], dvdz = du[1][2];
        const double dwdx = du[2][0], dwdy = du[2][1], dwdz = du[2][2];
        double t[3][3];
        t[0][0] = (2.0 / 3.0) * mu * (2.0 * dudx - dvdy - dwdz);
        t[1][1] = (2.0 / 3.0) * mu * (2.0 * dvdy - dudx - dwdz);
        t[2][2] = (2.0 / 3.0) * mu * (2.0 * dwdz - dudx - dvdy);
        t[0][1] = t[1][0] = mu * (dudy + dvdx);
        t[0][2] = t[2][0] = mu * (dudz + dwdx);
        t[1][2] = t[2][1] = mu * (dvdz + dwdy);
        const size_t c = o.id(i, j, k);
        for (int w = 0; w < 3; ++w) {
          o.stress[c + N * w] = (mf[dir][0] * t[w][0] + mf[dir][1] * t[w][1]) + mf[dir][2] * t[w][2];
          o.stress[c + N * (3 + w)] =
              (mf[dir][0] * (uf[w] * t[w][0]) + mf[dir][1] * (uf[w] * t[w][1])) + mf[dir][2] * (uf[w] * t[w][2]);
        }
      }
#pragma omp parallel for
  for (int k = ng; k < o.ngk - ng; ++k)
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        const size_t c = o.id(i, j, k), cm = o.id(i - ip, j - jp, k - kp);
        auto S = [&](size_t q, int s) { return o.stress[q + N * s]; };
        o.dvar[c + N * 0] += S(c, 0) - S(cm, 0);
        o.dvar[c + N * 1] += S(c, 1) - S(cm, 1);
        o.dvar[c + N * 2] += S(c, 2) - S(cm, 2);
        o.dvar[c + N * 3] += ((S(c, 3) - S(cm, 3)) + (S(c, 4) - S(cm, 4))) + (S(c, 5) - S(cm, 5));
      }
}

// gen3d_func::compute, src/gen3d.cpp:341-369: no boundary conditions, no
// viscous, buoyancy or C-equation terms; calculateFluxesG receives cf.dx as
// its dz argument (:359), and cf.dx = cf.dz = 1 on generalized grids.
void compute_gen3d(Orc &o) {
  // gen_ext (an option the reference does not have): cart3d's ghost fill
  // before the compute, its buoyancy after the pressure gradient and its
  // viscous terms after that (src/cart3d.cpp:221, 268-270, 265-270), the
  // latter in metric form (viscgen3d)
  if (o.c.gen_ext) apply_bcs(o);
  eos3d(o);
  genvel3d(o);
  for (int v = 0; v < o.nv; ++v) {
    flux3d(o, v, o.c.dx);
    advect3d(o, v);
  }
  genpresgrad3d(o);
  if (o.c.gen_ext && o.c.buoyancy) buoyancy3d(o);
  if (o.c.gen_ext && o.c.visc)
    for (int d = 0; d < 3; ++d) viscgen3d(o, d);
}

// gen2d_func::compute, src/gen2d.cpp:100-131
void compute_gen2d(Orc &o) {
  eos2d(o);
  genvel2d(o);
  for (int v = 0; v < o.nv; ++v) {
    flux2d(o, v);
    advect2d(o, v);
  }
  genpresgrad2d(o);
  if (o.c.gen_ext && o.c.buoyancy) buoyancy2d(o);  // gen_ext: cart2d's buoyancy
}

// cart2d_func::compute, src/cart2d.cpp:151-237
void compute2d(Orc &o) {
  eos2d(o);
  vel2d(o);
  for (int v = 0; v < o.nv; ++v) {
    flux2d(o, v);
    advect2d(o, v);
  }
  presgrad2d(o);
  if (o.c.buoyancy) buoyancy2d(o);
  if (o.c.visc) visc2d(o);
  if (o.c.ceq) ceq2d(o);
}

void compute(Orc &o) {
  if (o.gen) {
    if (o.ndim == 3) compute_gen3d(o);
    else compute_gen2d(o);
  } else if (o.ndim == 3) {
    compute3d(o);
  } else {
    compute2d(o);
  }
}

int rk_stages(const Orc &o) { return o.c.rk_order == 3 ? 3 : 2; }

// SSP-RK3 (Shu & Osher 1988) over every cell, as rkAdvance updates every
// cell; tmp1 keeps U^n.  Not in the reference (rk_order = 3 is an option):
//   U1 = U^n + dt R(U^n)
//   U2 = 3/4 U^n + 1/4 (U1 + dt R(U1))
//   U^{n+1} = 1/3 U^n + 2/3 (U2 + dt R(U2))
void rk3_stage(Orc &o, int stage) {
  const double dt = o.c.dt;
  const size_t tot = o.N * o.nvt;
  if (stage == 1) {
#pragma omp parallel for
    for (size_t q = 0; q < tot; ++q) {
      o.tmp1[q] = o.var[q];
      o.var[q] = o.var[q] + dt * o.dvar[q];
    }
  } else {
    const double a = stage == 2 ? 0.75 : 1.0 / 3.0, b = stage == 2 ? 0.25 : 2.0 / 3.0;
#pragma omp parallel for
    for (size_t q = 0; q < tot; ++q) o.var[q] = a * o.tmp1[q] + b * (o.var[q] + dt * o.dvar[q]);
  }
}

// rkAdvance stage updates, src/rk.cpp:42-50 (stage 1) and :63-69 (stage 2)
void rk_stage(Orc &o, int stage) {
  if (rk_stages(o) == 3) {
    rk3_stage(o, stage);
    return;
  }
  const double dt = o.c.dt;
  const size_t tot = o.N * o.nvt;
  if (stage == 1) {
#pragma omp parallel for
    for (size_t q = 0; q < tot; ++q) {
      o.tmp1[q] = o.var[q];
      o.var[q] = o.var[q] + 0.5f * dt * o.dvar[q];
    }
  } else {
#pragma omp parallel for
    for (size_t q = 0; q < tot; ++q) o.var[q] = o.tmp1[q] + dt * o.dvar[q];
  }
}

void pre_step(Orc &o) {
  // src/cart2d.cpp:147-149; gen2d/gen3d: nothing (gen2d with gen_ext as cart2d)
  if (o.ndim == 2 && (!o.gen || o.c.gen_ext)) apply_bcs(o);
}

// secondary variables: calculateRhoPT + computeVelocity + copyExtraVars
// (src/cart3d.cpp:176-185, src/cart2d.cpp:136-145, src/secondary.hpp:25-59)
void secondary(Orc &o) {
  if (o.ndim == 3) { eos3d(o); vel3d(o); }
  else { eos2d(o); vel2d(o); }
  const size_t N = o.N;
  for (size_t c = 0; c < N; ++c) {
    for (int d = 0; d < o.ndim; ++d) o.varx[c + N * d] = o.vel[c + N * d];
    o.varx[c + N * (o.ndim + 0)] = o.p[c];
    o.varx[c + N * (o.ndim + 1)] = o.T[c];
    o.varx[c + N * (o.ndim + 2)] = o.rho[c];
  }
}

// ---------------------------------------------------------------- noise filter
// postStep's noise filter, src/cart3d.cpp:319-374 / src/cart2d.cpp:257-307
// with detectNoise3D/2D and removeNoise3D/2D (src/noise.hpp:26-261).  The
// reference writes the overlapping 3x3(x3) blocks of neighbouring coarse points
// and updates var in place inside parallel loops, so its result depends on the
// loop order.  Defined here without order: a cell is flagged when any coarse
// point whose block covers it flags it (and, with coff > 0, its |C_hat| is
// below coff), and each smoothing sweep reads the values from before the sweep.

// coarse points per direction: 0..noise_half(nc), src/cart3d.cpp:325-341
int noise_half(int nc) { return (nc + 1) % 2 == 0 ? (nc + 1) / 2 : nc / 2; }

// detectNoise3D's stencil, in the reference's summation order (noise.hpp:63-82)
const int NOISE3D_ST[27][4] = {
    {-1, 1, 1, 2}, {0, 1, 1, 2}, {1, 1, 1, 2}, {-1, 0, 1, 2}, {0, 0, 1, 8}, {1, 0, 1, 2},
    {-1, -1, 1, 2}, {0, -1, 1, 2}, {1, -1, 1, 2}, {-1, 1, 0, 2}, {0, 1, 0, 8}, {1, 1, 0, 2},
    {-1, 0, 0, 8}, {0, 0, 0, -88}, {1, 0, 0, 8}, {-1, -1, 0, 2}, {0, -1, 0, 8}, {1, -1, 0, 2},
    {-1, 1, -1, 2}, {0, 1, -1, 2}, {1, 1, -1, 2}, {-1, 0, -1, 2}, {0, 0, -1, 8}, {1, 0, -1, 2},
    {-1, -1, -1, 2}, {0, -1, -1, 2}, {1, -1, -1, 2}};

// Is some coarse point covering cell x along one axis flagged?  Coarse point
// q sits at ng + 2q (0 <= q <= qmax) and covers ng + 2q - 1 .. ng + 2q + 1.
template <typename F>
bool any_cover(int x, int ng, int qmax, F &&f) {
  for (int c = x - 1; c <= x + 1; ++c) {
    const int d = c - ng;
    if (d < 0 || (d & 1) || d / 2 > qmax) continue;
    if (f(d / 2)) return true;
  }
  return false;
}

void noise3d(Orc &o) {
  const int ng = o.ng, M = noise_half(o.nci), Nn = noise_half(o.ncj), L = noise_half(o.nck);
  const double dx = o.cd[1], dy = o.cd[2], dz = o.cd[3];
  double coff = 0.0;
  if (o.c.ceq) {  // maxCvar3D(var, 6) over the cells (src/cart3d.cpp:343-352), one rank
    double mx = 0.0;
    for (int k = ng; k < o.ngk - ng; ++k)
      for (int j = ng; j < o.ngj - ng; ++j)
        for (int i = ng; i < o.ngi - ng; ++i) {
          const double a = std::abs(o.V(o.var, i, j, k, 6));
          if (a > mx) mx = a;
        }
    coff = o.c.n_coff * mx;
  }
  std::vector<int> vars;
  if (o.c.n_mode == 1) vars = {0, 1, 2};
  else vars = {3};
  const int nv = o.nv;
  std::vector<char> flag((size_t)(M + 1) * (Nn + 1) * (L + 1));
  std::vector<double> dn(o.N, 0.0);
  for (int v : vars) {
    for (int kk = 0; kk <= L; ++kk)
      for (int jj = 0; jj <= Nn; ++jj)
        for (int ii = 0; ii <= M; ++ii) {
          const int i = 2 * ii + ng, j = 2 * jj + ng, k = 2 * kk + ng;
          const double a = std::sqrt(165.0 * dx * dy * dz) / 15840.0;
          double sum = 0.0;
          for (int t = 0; t < 27; ++t) {
            const double term = (double)NOISE3D_ST[t][3] * o.V(o.var, i + NOISE3D_ST[t][0], j + NOISE3D_ST[t][1],
                                                             k + NOISE3D_ST[t][2], v);
            sum = t == 0 ? term : sum + term;
          }
          const double c = -a * sum;
          const double cref = o.c.n_dh * std::sqrt(dx * dy * dz / 12.0);
          flag[(size_t)ii + (size_t)(M + 1) * ((size_t)jj + (size_t)(Nn + 1) * kk)] = std::abs(c) >= cref;
        }
    for (int tau = 0; tau < o.c.n_nt; ++tau) {
      for (int k = ng; k < o.ngk - ng; ++k)
        for (int j = ng; j < o.ngj - ng; ++j)
          for (int i = ng; i < o.ngi - ng; ++i) {
            const bool cov = any_cover(i, ng, M, [&](int ii) {
              return any_cover(j, ng, Nn, [&](int jj) {
                return any_cover(k, ng, L, [&](int kk) {
                  return flag[(size_t)ii + (size_t)(M + 1) * ((size_t)jj + (size_t)(Nn + 1) * kk)] != 0;
                });
              });
            });
            const int noise = cov && (coff == 0.0 || std::abs(o.V(o.var, i, j, k, nv + 1)) < coff) ? 1 : 0;
            auto U = [&](int a, int b, int c) { return o.V(o.var, a, b, c, v); };
            const double lap = (U(i - 1, j, k) - 2 * U(i, j, k) + U(i + 1, j, k)) / (dx * dx) +
                               (U(i, j - 1, k) - 2 * U(i, j, k) + U(i, j + 1, k)) / (dy * dy) +
                               (U(i, j, k - 1) - 2 * U(i, j, k) + U(i, j, k + 1)) / (dz * dz);
            dn[o.id(i, j, k)] = o.c.n_eta * (dx * dx + dy * dy + dz * dz) * noise * lap;
          }
      for (int k = ng; k < o.ngk - ng; ++k)
        for (int j = ng; j < o.ngj - ng; ++j)
          for (int i = ng; i < o.ngi - ng; ++i) o.V(o.var, i, j, k, v) += dn[o.id(i, j, k)];
    }
  }
}

// cart2d (src/cart2d.cpp:257-307) and gen2d (src/gen2d.cpp:150-177: coarse
// points 0..M-1 with M = (nc-1)/2 or nc/2, coff = 0, both momentum components)
void noise2d(Orc &o) {
  const int ng = o.ng;
  auto gen_pts = [](int nc) { return (nc - 1) % 2 == 0 ? (nc - 1) / 2 : nc / 2; };
  // M, Nn: the largest coarse index per direction
  const int M = o.gen ? gen_pts(o.nci) - 1 : noise_half(o.nci), Nn = o.gen ? gen_pts(o.ncj) - 1 : noise_half(o.ncj);
  const double dx = o.cd[1], dy = o.cd[2];
  double coff = 0.0;
  if (o.c.ceq && !o.gen) {  // maxCvar2D(var, 1, cd): |C_hat| over the cells (src/cart2d.cpp:274-283)
    double mx = 0.0;
    for (int j = ng; j < o.ngj - ng; ++j)
      for (int i = ng; i < o.ngi - ng; ++i) {
        const double a = std::abs(o.V(o.var, i, j, 0, o.nv + 1));
        if (a > mx) mx = a;
      }
    coff = o.c.n_coff * mx;
  }
  std::vector<int> vars;
  if (o.c.n_mode == 1 || o.gen) vars = {0, 1};
  else vars = {2};
  if (M < 0 || Nn < 0) return;  // no coarse points
  std::vector<char> flag((size_t)(M + 1) * (Nn + 1));
  std::vector<double> dn(o.N, 0.0);
  for (int v : vars) {
    auto U = [&](int a, int b) { return o.V(o.var, a, b, 0, v); };
    for (int jj = 0; jj <= Nn; ++jj)
      for (int ii = 0; ii <= M; ++ii) {
        const int i = 2 * ii + ng, j = 2 * jj + ng;
        const double a = std::sqrt(6.0 * dx * dy);
        const double c = -(a / 192.0) * (U(i - 1, j + 1) + 2 * U(i, j + 1) + U(i + 1, j + 1) + 2 * U(i - 1, j) -
                                         12 * U(i, j) + 2 * U(i + 1, j) + U(i - 1, j - 1) + 2 * U(i, j - 1) +
                                         U(i + 1, j - 1));
        const double cref = o.c.n_dh * a / 16.0;
        flag[(size_t)ii + (size_t)(M + 1) * jj] = std::abs(c) >= cref;
      }
    for (int tau = 0; tau < o.c.n_nt; ++tau) {
      for (int j = ng; j < o.ngj - ng; ++j)
        for (int i = ng; i < o.ngi - ng; ++i) {
          const bool cov = any_cover(i, ng, M, [&](int ii) {
            return any_cover(j, ng, Nn, [&](int jj) { return flag[(size_t)ii + (size_t)(M + 1) * jj] != 0; });
          });
          // no abs here in the reference (noise.hpp:222)
          const int noise = cov && (coff == 0.0 || o.V(o.var, i, j, 0, o.nv + 1) < coff) ? 1 : 0;
          const double lgrad = (U(i, j) - U(i - 1, j)) / dx;
          const double rgrad = (U(i + 1, j) - U(i, j)) / dx;
          const double bgrad = (U(i, j) - U(i, j - 1)) / dy;
          const double tgrad = (U(i, j + 1) - U(i, j)) / dy;
          dn[o.id(i, j, 0)] = o.c.n_eta * (dx * dx + dy * dy) * noise * ((rgrad - lgrad) / dx + (tgrad - bgrad) / dy);
        }
      for (int j = ng; j < o.ngj - ng; ++j)
        for (int i = ng; i < o.ngi - ng; ++i) o.V(o.var, i, j, 0, v) += dn[o.id(i, j, 0)];
    }
  }
}

// rk_func::postStep: secondary variables, then the noise filter (cartesian modules)
void post_step(Orc &o) {
  secondary(o);
  if (o.c.noise && !(o.gen && o.ndim == 3)) {  // gen3d has no noise filter
    if (o.ndim == 3) noise3d(o);
    else noise2d(o);
  }
}

void pre_sim(Orc &o) {
  if (o.gen) {  // gen3d_func::preSim (src/gen3d.cpp:131-304), gen2d_func::preSim (src/gen2d.cpp:187-331)
    if (o.ndim == 3) metrics3d(o);
    else metrics2d(o);
    metric_ghosts(o);
    if (o.c.gen_ext && o.ndim == 2) apply_bcs(o);  // gen_ext: cart2d's preSim fill (src/cart2d.cpp:137)
    return;
  }
  if (o.ndim == 2) apply_bcs(o);  // src/cart2d.cpp:137
  secondary(o);
}

std::vector<double> *array_of(Orc &o, int a, size_t *count) {
  std::vector<double> *v = nullptr;
  switch (a) {
    case FST_VAR: v = &o.var; break;
    case FST_TMP1: v = &o.tmp1; break;
    case FST_DVAR: v = &o.dvar; break;
    case FST_VARX: v = &o.varx; break;
    case FST_P: v = &o.p; break;
    case FST_RHO: v = &o.rho; break;
    case FST_T: v = &o.T; break;
    case FST_VEL: v = &o.vel; break;
    case FST_GRADRHO: v = &o.gradRho; break;
    case FST_GRID: v = &o.grid; break;
    case FST_METRICS: v = &o.metrics; break;
    default: return nullptr;
  }
  if (count) *count = v->size();
  return v;
}

}  // namespace

extern "C" {

void *orc_create(const fst_config *cfg) {
  Orc *o = new Orc();
  o->c = *cfg;
  o->ndim = cfg->ndim;
  o->ns = cfg->ns;
  o->ng = cfg->ng;
  o->nci = cfg->nci;
  o->ncj = cfg->ncj;
  o->nck = cfg->ndim == 3 ? cfg->nck : 1;
  o->ngi = o->nci + 2 * o->ng;
  o->ngj = o->ncj + 2 * o->ng;
  o->ngk = cfg->ndim == 3 ? o->nck + 2 * o->ng : 1;
  o->nv = cfg->ndim == 3 ? 4 + o->ns : 3 + o->ns;
  o->learned = cfg->learned_c != 0 && cfg->ceq;
  o->nvt = o->nv + (cfg->ceq && !o->learned ? 5 : 0);
  o->nvx = o->ndim + 3;
  o->gen = cfg->grid == 1;
  if (o->gen) o->c.dx = o->c.dy = o->c.dz = 1.0;  // src/input.cpp:344-350
  if (cfg->ndim == 2) o->c.dz = o->c.dx;  // src/input.cpp:356
  o->ni = o->nci + 1;
  o->nj = o->ncj + 1;
  o->nk = cfg->ndim == 3 ? o->nck + 1 : 1;
  o->N = (size_t)o->ngi * o->ngj * o->ngk;
  const size_t N = o->N;
  o->cd.assign(6 + 3 * o->ns, 0.0);
  o->cd[0] = o->ns;
  o->cd[1] = o->c.dx;
  o->cd[2] = o->c.dy;
  o->cd[3] = o->c.dz;
  o->cd[4] = o->nv;
  o->cd[5] = o->ng;
  for (int s = 0; s < o->ns; ++s) {
    o->cd[6 + 3 * s] = cfg->gamma[s];
    o->cd[6 + 3 * s + 1] = cfg->R / cfg->M[s];
    o->cd[6 + 3 * s + 2] = cfg->mu[s];
  }
  o->dxmag = o->c.dx * o->c.dx + o->c.dy * o->c.dy + o->c.dz * o->c.dz;
  o->var.assign(N * o->nvt, 0.0);
  o->tmp1.assign(N * o->nvt, 0.0);
  o->dvar.assign(N * o->nvt, 0.0);
  o->varx.assign(N * o->nvx, 0.0);
  o->p.assign(N, 0.0);
  o->rho.assign(N, 0.0);
  o->T.assign(N, 0.0);
  o->vel.assign(N * o->ndim, 0.0);
  o->fluxx.assign(N, 0.0);
  o->fluxy.assign(N, 0.0);
  o->fluxz.assign(o->ndim == 3 ? N : 0, 0.0);
  if (cfg->visc) {
    if (o->ndim == 3) o->stress.assign(N * 6, 0.0);
    else {
      o->stressx.assign(N * 4, 0.0);
      o->stressy.assign(N * 4, 0.0);
      o->qx.assign(N, 0.0);
      o->qy.assign(N, 0.0);
    }
  }
  if (cfg->ceq) {
    if (!o->learned) o->gradRho.assign(N * (o->ndim == 3 ? 5 : 4), 0.0);
    o->mFlux.assign(o->ndim == 3 ? 27 * N : 8 * N, 0.0);
  }
  if (o->gen) {
    o->grid.assign((size_t)o->ni * o->nj * o->nk * 3, 0.0);
    o->metrics.assign(N * o->ndim * o->ndim, 0.0);
  }
  if (o->learned) {
    o->lchat.assign(N, 0.0);
    o->ldir.assign(3 * N, 0.0);
  }
  return o;
}

void orc_destroy(void *h) { delete (Orc *)h; }

long long orc_size(void *h, int array) {
  size_t n = 0;
  if (!array_of(*(Orc *)h, array, &n)) return -1;
  return (long long)n;
}

int orc_set(void *h, int array, const double *src) {
  size_t n;
  auto *v = array_of(*(Orc *)h, array, &n);
  if (!v) return 1;
  std::memcpy(v->data(), src, n * sizeof(double));
  return 0;
}

int orc_get(void *h, int array, double *dst) {
  size_t n;
  auto *v = array_of(*(Orc *)h, array, &n);
  if (!v) return 1;
  std::memcpy(dst, v->data(), n * sizeof(double));
  return 0;
}

void orc_apply_bcs(void *h) { apply_bcs(*(Orc *)h); }
void orc_compute(void *h) { compute(*(Orc *)h); }
void orc_rk_stage(void *h, int stage) { rk_stage(*(Orc *)h, stage); }
void orc_pre_sim(void *h) { pre_sim(*(Orc *)h); }
void orc_pre_step(void *h) { pre_step(*(Orc *)h); }
void orc_post_step(void *h) { post_step(*(Orc *)h); }

// Fiesta::step (src/fiesta.cpp:333-340): preStep, rkAdvance, postStep
void orc_step(void *h, int nsteps) {
  Orc &o = *(Orc *)h;
  for (int t = 0; t < nsteps; ++t) {
    pre_step(o);
    for (int s = 1; s <= rk_stages(o); ++s) {
      compute(o);
      rk_stage(o, s);
    }
  }
}

void orc_maxima(void *h, double *maxS, double *maxCh, double *alpha) {
  Orc &o = *(Orc *)h;
  *maxS = o.maxS;
  *maxCh = o.maxCh;
  *alpha = o.alpha;
}

// Learned form: C_hat [N], direction [3][N] and max C_hat used by compute().
int orc_set_learned(void *h, const double *chat, const double *dir, double maxCh) {
  Orc &o = *(Orc *)h;
  if (!o.learned) return 1;
  std::memcpy(o.lchat.data(), chat, o.N * sizeof(double));
  std::memcpy(o.ldir.data(), dir, 3 * o.N * sizeof(double));
  o.lmaxCh = maxCh;
  return 0;
}

// Primitives (eos3d, vel3d / eos2d, vel2d) and learned-C features of the
// current var (Dir: 3 planes; in 2D the tangent, plane 2 zero).
int orc_features(void *h, float *F, float *M, float *Dir, float *G1) {
  Orc &o = *(Orc *)h;
  if (o.gen) return 1;
  if (o.ndim == 3) {
    eos3d(o);
    vel3d(o);
    features3d(o, F, M, Dir, G1);
  } else {
    eos2d(o);
    vel2d(o);
    features2d(o, F, M, Dir, G1);
  }
  return 0;
}

// blockWriter::dataPack (src/block.cpp:404-454) of variable vn of var, varx
// (cell indices from the first interior cell, offset ng) or the grid nodes
// (offset 0), into out (float or double), i fastest.
int orc_pack(void *h, int array, int vn, const int64_t *start, const int64_t *end, const int64_t *stride,
             int average, int out_f32, void *out, size_t count) {
  Orc &o = *(Orc *)h;
  const std::vector<double> *src = nullptr;
  long long nx = o.ngi, ny = o.ngj, nz = o.ngk;
  int ng = o.ng;
  if (array == FST_VAR) src = &o.var;
  else if (array == FST_VARX) src = &o.varx;
  else if (array == FST_GRID && o.gen) {
    src = &o.grid;
    nx = o.ni; ny = o.nj; nz = o.nk;
    ng = 0;
  }
  if (!src) return 1;
  const size_t plane = (size_t)nx * ny * nz;
  auto S = [&](long long i, long long j, long long k) {
    return (*src)[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k) + plane * vn];
  };
  long long ext[3] = {1, 1, 1};
  for (int a = 0; a < o.ndim; ++a) ext[a] = (end[a] - start[a]) / stride[a] + 1;
  if ((size_t)(ext[0] * ext[1] * ext[2]) != count) return 2;
  auto put = [&](long long idx, double v) {
    if (out_f32) ((float *)out)[idx] = (float)v;
    else ((double *)out)[idx] = v;
  };
  if (o.ndim == 3) {
    for (long long i = start[0]; i < end[0] + 1; i += stride[0])
      for (long long j = start[1]; j < end[1] + 1; j += stride[1])
        for (long long k = start[2]; k < end[2] + 1; k += stride[2]) {
          const long long ii = (i - start[0]) / stride[0], jj = (j - start[1]) / stride[1],
                          kk = (k - start[2]) / stride[2];
          const long long idx = (ext[0] * ext[1]) * kk + ext[0] * jj + ii;
          if (average) {
            double mean = 0.0;
            for (long long ia = 0; ia < stride[0]; ++ia)
              for (long long ja = 0; ja < stride[1]; ++ja)
                for (long long ka = 0; ka < stride[2]; ++ka) mean += S(i + ia + ng, j + ja + ng, k + ka + ng);
            put(idx, mean / (double)(stride[0] * stride[1] * stride[2]));
          } else {
            put(idx, S(i + ng, j + ng, k + ng));
          }
        }
  } else {
    for (long long i = start[0]; i < end[0] + 1; i += stride[0])
      for (long long j = start[1]; j < end[1] + 1; j += stride[1]) {
        const long long ii = (i - start[0]) / stride[0], jj = (j - start[1]) / stride[1];
        const long long idx = ext[0] * jj + ii;
        if (average) {
          double mean = 0.0;
          for (long long ia = 0; ia < stride[0]; ++ia)
            for (long long ja = 0; ja < stride[1]; ++ja) mean += S(i + ia + ng, j + ja + ng, 0);
          put(idx, mean / (double)(stride[0] * stride[1]));
        } else {
          put(idx, S(i + ng, j + ng, 0));
        }
      }
  }
  return 0;
}

double orc_weno3d(double f1, double f2, double f3, double f4, double f5) { return weno3d(f1, f2, f3, f4, f5); }
double orc_weno2d(double f1, double f2, double f3, double f4, double f5) { return weno2d(f1, f2, f3, f4, f5); }

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

}  // extern "C"
