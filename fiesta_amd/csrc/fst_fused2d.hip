// Fused cart2d stage: one kernel per Runge-Kutta stage computes the whole
// right-hand side of cart2d_func::compute (src/cart2d.cpp:151-237) -- EOS,
// WENO5 fluxes and their differences, pressure gradient, buoyancy, the
// viscous terms and the C-equation -- and the stage update of rkAdvance
// (src/rk.cpp:42-69; SSP-RK3 as the option), with every intermediate (the
// primitives, the face fluxes, stresses and artificial-viscosity fluxes) in
// LDS instead of HBM.
//
// Tile: 64 (i) x 8 (j) interior cells per block of 512 threads, one cell per
// thread (wave w owns row w).  The block stages rho, p, T, u, v of its tile
// with a 3-cell halo (70 x 14 cells) in LDS, then for each conserved
// variable stages q = var (+ p for the energy), computes each x and y face of
// the tile once (WENO5, the reference's upwind selection) and differences
// them.  The viscous stresses and the C-equation's face fluxes are computed
// once per face in two passes (x faces, then y faces) with the x parts of
// each cell's sums kept in registers, so every sum has the reference's
// association.  Arithmetic is the reference's term by term; divisions by run
// constants use cdiv_rn (correctly rounded, so equal to IEEE division).
//
// The global maxima the C-equation needs inside the stage (maxS of
// maxWaveSpeed2D, max C_hat of maxCvar2D; src/cart2d.cpp:210-220) are those
// of the stage's input state over the interior: the previous stage's
// epilogue computes them from the values it writes (the ghost fill of
// preStep does not touch the interior), and a separate pass computes them
// for the first stage after the state changed from outside.
//
// Stage states: RK2 reads var (U^n) and writes U* to a second buffer, then
// reads U* and writes U^{n+1} over var pointwise; SSP-RK3 uses U1 and U2 in
// two buffers.  Ghost cells take the stage update with a zero right-hand
// side, as rkAdvance updates every cell.
#include <algorithm>
#include <string>

#include "fst_common.hpp"
#include "fst_fused.hpp"
#include "fst_kernels.hpp"
#include "fst_mlp.hpp"

namespace fst {

namespace {

constexpr int F2X = 64, F2Y = 8;                   // interior cells per tile
constexpr int F2H = 3;                             // halo: the WENO5 reach
constexpr int RW = F2X + 2 * F2H, RH = F2Y + 2 * F2H;
constexpr int RN = RW * RH;                        // staged cells
constexpr int NXF = (F2X + 1) * F2Y;               // x faces: i0-1 .. i0+63
constexpr int NYF = F2X * (F2Y + 1);               // y faces: j0-1 .. j0+7
constexpr int ADV = RN + NXF + NYF;                // one advection buffer set: q, x and y fluxes
constexpr int WORK = 2 * ADV;                      // doubles of the work area
static_assert(3 * NXF + 3 * NYF <= WORK && 4 * NYF <= WORK && 4 * NXF <= WORK, "work area");
constexpr int NT2 = F2X * F2Y;                     // threads per block

enum { UPD_RK2_1 = 0, UPD_ADD = 1, UPD_RK3_1 = 2, UPD_BLEND = 3 };

struct K2 {
  long long N;
  int ngi, ngj, ng, nci, ncj, ns, nv, nvt, nbx, nby;
  double dx, dy, rdx, rdy, dx4, rdx4, dy4, rdy4, dx12, rdx12, dy12, rdy12;
  double cpf[3], cvf[3];    // mixture EOS per species
  double qcp[3], qcv[3];    // maxWaveSpeed2D's per-species factors (stride-2 reads of cd)
  double g, rhoRef;
  int buoy, visc;
  double eps, kap, alpha, dxmag, dd2;  // C-equation constants; dd2 = dx^2 + dy^2
  int upd;
  double dt, hdt, a, b;     // update form: hdt = 0.5f * dt
};

__device__ __forceinline__ double upd_val(const K2 &k, double u, double un, double d) {
  switch (k.upd) {
    case UPD_RK2_1: return u + k.hdt * d;
    case UPD_ADD: return un + k.dt * d;
    case UPD_RK3_1: return u + k.dt * d;
    default: return k.a * un + k.b * (u + k.dt * d);
  }
}

// Cell (i, j) of the block's staged region.
__device__ __forceinline__ int rix(int i0, int j0, int i, int j) { return (i - i0 + F2H) + RW * (j - j0 + F2H); }

// The quantities maxWaveSpeed2D / maxCvar2D take over the interior
// (src/ceq.hpp:23-78) from the conserved values of one cell: the reference's
// stride-2 species constants and |m| kept.
template <int NS>
__device__ __forceinline__ void cell_maxima(const K2 &k, const double *w, double &ms, double &mc, double &mch) {
  double rr = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) rr = rr + w[3 + s];
  double Cp = 0, Cv = 0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double Y = w[3 + s] / rr;
    Cp = Cp + Y * k.cpf[s];
    Cv = Cv + Y * k.cvf[s];
  }
  const double gamma = Cp / Cv;
  const double m0 = w[0], m1 = w[1];
  const double pp = (gamma - 1) * (w[2] - (0.5 / rr) * (m0 * m0 + m1 * m1));
  double qp = 0, qv = 0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    qp = qp + (w[3 + s] / rr) * k.qcp[s];
    qv = qv + (w[3 + s] / rr) * k.qcv[s];
  }
  const double qg = qp / qv;
  const double a = sqrt(qg * pp / rr);
  ms = a + sqrt(m0 * m0 + m1 * m1);
  mc = fabs(w[3 + NS]);
  mch = fabs(w[4 + NS]);
}

// Maxima of a state over the interior (the first stage after the state
// changed from outside).
template <int NS>
__global__ void __launch_bounds__(256) k_max2d(K2 k, const double *__restrict__ U, Scalars *sc) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v3[3] = {0.0, 0.0, 0.0};
  if (t < (long long)k.nci * k.ncj) {
    const int i = k.ng + (int)(t % k.nci), j = k.ng + (int)(t / k.nci);
    const long long c = i + (long long)k.ngi * j;
    double w[NS + 5];
#pragma unroll
    for (int v = 0; v < NS + 5; ++v) w[v] = U[c + k.N * v];
    cell_maxima<NS>(k, w, v3[0], v3[1], v3[2]);
  }
  unsigned long long *const d3[3] = {&sc->maxS, &sc->maxC, &sc->maxCh};
  block_max_atomic<3>(v3, d3);
}

// One stage.  CM: 0 no C-equation, 1 the plain C-equation, 2 the learned
// C-indicator (C_hat and tangent from lchat / ldir, max C_hat from lsc).
template <int NS, int CM>
__global__ void __launch_bounds__(NT2) k_stage2d(K2 k, const double *__restrict__ U, const double *__restrict__ Un,
                                                 double *__restrict__ Out, const Scalars *scIn, Scalars *scOut,
                                                 Scalars *scClr, const float *__restrict__ lchat,
                                                 const float *__restrict__ ldir, const Scalars *lsc) {
  constexpr int NV = 3 + NS;
  constexpr int NVT = NV + (CM == 1 ? 5 : 0);
  __shared__ double sR[RN], sP[RN], sU[RN], sV[RN], sT[RN];
  __shared__ double sW[WORK];
  const int tid = threadIdx.x;
  const long long N = k.N;
  // block -> tile, blocks of one XCD (hardware block b runs on XCD b % 8) on
  // consecutive tiles
  const int nb = k.nbx * k.nby, hb = blockIdx.x;
  int lb = hb;
  const int full = (nb / 8) * 8;
  if (hb < full) lb = (hb % 8) * (full / 8) + hb / 8;
  const int bx = lb % k.nbx, by = lb / k.nbx;
  const int i0 = k.ng + bx * F2X, j0 = k.ng + by * F2Y;
  const int ci = tid & 63, cj = tid >> 6;
  const int i = i0 + ci, j = j0 + cj;
  const int ie = k.ng + k.nci, je = k.ng + k.ncj;  // ends of the interior
  const bool act = i < ie && j < je;
  const long long c = i + (long long)k.ngi * j;
  if (hb == 0 && tid < 4) reinterpret_cast<unsigned long long *>(scClr)[tid] = 0ull;

  // ---- primitives of the region (calculateRhoPT2D, computeVelocity2D); the
  // thread's two staged cells (r = tid, tid + NT2) load everything first
  static_assert(RN <= 2 * NT2, "thread slots");
  {
    double w[2][3 + NS];
    bool ok[2];
    int rr_[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = tid + h * NT2;
      rr_[h] = r;
      const int gi = i0 - F2H + r % RW, gj = j0 - F2H + r / RW;
      ok[h] = r < RN && gi < k.ngi && gj < k.ngj;
      const long long q = gi + (long long)k.ngi * gj;
#pragma unroll
      for (int v = 0; v < 3 + NS; ++v) w[h][v] = ok[h] ? U[q + N * v] : 1.0;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (!ok[h]) continue;
      const int r = rr_[h];
      double rr = 0.0;
#pragma unroll
      for (int s = 0; s < NS; ++s) rr = rr + w[h][3 + s];
      double Cp = 0, Cv = 0;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const double Y = w[h][3 + s] / rr;
        Cp = Cp + Y * k.cpf[s];
        Cv = Cv + Y * k.cvf[s];
      }
      const double gamma = Cp / Cv;
      const double m0 = w[h][0], m1 = w[h][1];
      const double pp = (gamma - 1) * (w[h][2] - (0.5 / rr) * (m0 * m0 + m1 * m1));
      sR[r] = rr;
      sP[r] = pp;
      sT[r] = pp / ((Cp - Cv) * rr);
      sU[r] = m0 / rr;
      sV[r] = m1 / rr;
    }
  }
  double dv[NVT];
#pragma unroll
  for (int v = 0; v < NVT; ++v) dv[v] = 0.0;

  // ---- advection: WENO5 face fluxes and their differences (advect2D)
  double *const sQ = sW, *const sFX = sW + RN, *const sFY = sW + RN + NXF;
  static_assert(RN <= 2 * NT2 && NXF + NYF <= 3 * NT2, "thread slots");
  __syncthreads();  // primitives staged
  // this thread's faces (f = tid + q NT2): the region index of the face's
  // lower cell, the family (0 x, 1 y, -1 none) and the upwind velocity,
  // which does not depend on the variable
  int fbq[3], fty[3];
  double fvq[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int f = tid + q * NT2;
    fty[q] = -1;
    fbq[q] = 0;
    fvq[q] = 0.0;
    if (f < NXF) {
      const int fi = i0 - 1 + f % (F2X + 1), fj = j0 + f / (F2X + 1);
      if (fi < ie && fj < je) {  // else no interior cell of the tile uses it
        const int b0 = rix(i0, j0, fi, fj);
        fty[q] = 0;
        fbq[q] = b0;
        fvq[q] = cdiv_rn(-sU[b0 + 2] + 7.0 * sU[b0 + 1] + 7.0 * sU[b0] - sU[b0 - 1], 12.0, 1.0 / 12.0);
      }
    } else if (f < NXF + NYF) {
      const int g = f - NXF;
      const int fi = i0 + g % F2X, fj = j0 - 1 + g / F2X;
      if (fi < ie && fj < je) {
        const int b0 = rix(i0, j0, fi, fj);
        fty[q] = 1;
        fbq[q] = b0;
        fvq[q] = cdiv_rn(-sV[b0 + 2 * RW] + 7.0 * sV[b0 + RW] + 7.0 * sV[b0] - sV[b0 - RW], 12.0, 1.0 / 12.0);
      }
    }
  }
  // this thread's staged cells r = tid, tid + NT2; the next variable's values
  // are loaded while the current one's faces are computed
  const int ra = tid, rb = tid + NT2;
  const int gia = i0 - F2H + ra % RW, gja = j0 - F2H + ra / RW;
  const int gib = i0 - F2H + rb % RW, gjb = j0 - F2H + rb / RW;
  const bool oka = gia < k.ngi && gja < k.ngj, okb = rb < RN && gib < k.ngi && gjb < k.ngj;
  const long long qia = gia + (long long)k.ngi * gja, qib = gib + (long long)k.ngi * gjb;
  double qa = oka ? U[qia] : 0.0, qb = okb ? U[qib] : 0.0;
  // two buffer sets (q, fluxes) alternate between variables: per variable
  // one barrier separates staging q(v) from its faces, and the faces of v from
  // the differences of v (read after the next barrier, together with the
  // staging of v + 1)
  auto stage_q = [&](int v, double *sq) {
    const double isE = (double)(v == 2);
    if (oka) sq[ra] = qa + isE * sP[ra];
    if (okb) sq[rb] = qb + isE * sP[rb];
  };
  stage_q(0, sW);
#pragma unroll
  for (int v = 0; v <= NV; ++v) {
    __syncthreads();
    double *const sQ = sW + (v & 1) * ADV, *const sFX = sQ + RN, *const sFY = sQ + RN + NXF;
    if (v > 0) {  // differences of variable v - 1
      const double *const pX = sW + ((v - 1) & 1) * ADV + RN, *const pY = pX + NXF;
      const int xr = ci + 1 + (F2X + 1) * cj, xl = ci + (F2X + 1) * cj;
      const int yt = ci + F2X * (cj + 1), yb = ci + F2X * cj;
      dv[v - 1] = -((pX[xr] - pX[xl]) + (pY[yt] - pY[yb]));
    }
    if (v == NV) break;
    if (v + 1 < NV) {
      if (oka) qa = U[qia + N * (v + 1)];
      if (okb) qb = U[qib + N * (v + 1)];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int b0 = fbq[q];
      if (fty[q] == 0) {
        double f1, f2, f3, f4, f5;
        if (fvq[q] < 0.0) {
          f1 = sQ[b0 + 3]; f2 = sQ[b0 + 2]; f3 = sQ[b0 + 1]; f4 = sQ[b0]; f5 = sQ[b0 - 1];
        } else {
          f1 = sQ[b0 - 2]; f2 = sQ[b0 - 1]; f3 = sQ[b0]; f4 = sQ[b0 + 1]; f5 = sQ[b0 + 2];
        }
        sFX[tid + q * NT2] = cdiv_rn(fvq[q] * weno5_f64c(f1, f2, f3, f4, f5), k.dx, k.rdx);
      } else if (fty[q] == 1) {
        double f1, f2, f3, f4, f5;
        if (fvq[q] < 0.0) {
          f1 = sQ[b0 + 3 * RW]; f2 = sQ[b0 + 2 * RW]; f3 = sQ[b0 + RW]; f4 = sQ[b0]; f5 = sQ[b0 - RW];
        } else {
          f1 = sQ[b0 - 2 * RW]; f2 = sQ[b0 - RW]; f3 = sQ[b0]; f4 = sQ[b0 + RW]; f5 = sQ[b0 + 2 * RW];
        }
        sFY[tid + q * NT2 - NXF] = cdiv_rn(fvq[q] * weno5_f64c(f1, f2, f3, f4, f5), k.dy, k.rdy);
      }
    }
    if (v + 1 < NV) stage_q(v + 1, sW + ((v + 1) & 1) * ADV);
  }

  const int rc = rix(i0, j0, i, j);
  // ---- pressure gradient, buoyancy (applyPressureGradient2D, computeBuoyancy)
  {
    const double dxp = cdiv_rn(sP[rc - 2] - 8.0 * sP[rc - 1] + 8.0 * sP[rc + 1] - sP[rc + 2], k.dx12, k.rdx12);
    const double dyp =
        cdiv_rn(sP[rc - 2 * RW] - 8.0 * sP[rc - RW] + 8.0 * sP[rc + RW] - sP[rc + 2 * RW], k.dy12, k.rdy12);
    dv[0] -= dxp;
    dv[1] -= dyp;
    if (k.buoy && act) {
      const double vv = U[c + N] / U[c + 3 * N];
      const double rhop = sR[rc] - k.rhoRef;
      const double eps = 1e-6;
      if (rhop >= eps || rhop <= -eps) {
        const double f = -k.g * rhop;
        dv[1] += f;
        dv[2] += fabs(vv) * f;
      }
    }
  }

  // ---- viscous terms (calculateStressTensor2dv, calculateHeatFlux2dv,
  // applyViscousTerm2dv): x faces, then y faces
  if (k.visc) {
    const double mur = 2.928e-5, mut = 2.928e-5, kk = 0.026;
    double *const SX = sW, *const SY = sW + 3 * NXF;  // x faces: sx00, sx01, qx; y faces: sy11, sy01, qy
    __syncthreads();  // the last variable's fluxes are read
    for (int f = tid; f < NXF + NYF; f += NT2) {
      if (f < NXF) {
        const int fi = i0 - 1 + f % (F2X + 1), fj = j0 + f / (F2X + 1);
        if (fi >= ie || fj >= je) continue;
        const int b0 = rix(i0, j0, fi, fj);
        const double dudx = cdiv_rn(sU[b0 + 1] - sU[b0], k.dx, k.rdx);
        const double dvdx = cdiv_rn(sV[b0 + 1] - sV[b0], k.dx, k.rdx);
        const double dudy =
            cdiv_rn((sU[b0 + RW] + sU[b0 + 1 + RW]) - (sV[b0 - RW] + sU[b0 + 1 - RW]), k.dy4, k.rdy4);
        const double dvdy =
            cdiv_rn((sV[b0 + RW] + sV[b0 + 1 + RW]) - (sV[b0 - RW] + sV[b0 + 1 - RW]), k.dy4, k.rdy4);
        SX[f] = (2.0 / 3.0) * mur * (2.0 * dudx - dvdy);
        SX[NXF + f] = mur * (dudy + dvdx);
        SX[2 * NXF + f] = cdiv_rn(-kk * (sT[b0 + 1] - sT[b0]), k.dx, k.rdx);
      } else {
        const int g = f - NXF;
        const int fi = i0 + g % F2X, fj = j0 - 1 + g / F2X;
        if (fi >= ie || fj >= je) continue;
        const int b0 = rix(i0, j0, fi, fj);
        const double dudx =
            cdiv_rn((sU[b0 + 1] + sU[b0 + 1 + RW]) - (sU[b0 - 1] + sU[b0 - 1 + RW]), k.dx4, k.rdx4);
        const double dvdx =
            cdiv_rn((sV[b0 + 1] + sV[b0 + 1 + RW]) - (sV[b0 - 1] + sV[b0 - 1 + RW]), k.dx4, k.rdx4);
        const double dudy = cdiv_rn(sU[b0 + RW] - sU[b0], k.dy, k.rdy);
        const double dvdy = cdiv_rn(sV[b0 + RW] - sV[b0], k.dy, k.rdy);
        SY[g] = (2.0 / 3.0) * mut * (2.0 * dvdy - dudx);
        SY[NYF + g] = mut * (dudy + dvdx);
        SY[2 * NYF + g] = cdiv_rn(-kk * (sT[b0 + RW] - sT[b0]), k.dy, k.rdy);
      }
    }
    __syncthreads();
    const double ur = (sU[rc + 1] + sU[rc]) / 2.0, ul = (sU[rc - 1] + sU[rc]) / 2.0;
    const double vr = (sV[rc + 1] + sV[rc]) / 2.0, vl = (sV[rc - 1] + sV[rc]) / 2.0;
    const double ut = (sU[rc + RW] + sU[rc]) / 2.0, ub = (sU[rc - RW] + sU[rc]) / 2.0;
    const double vt = (sV[rc + RW] + sV[rc]) / 2.0, vb = (sV[rc - RW] + sV[rc]) / 2.0;
    const int xr = ci + 1 + (F2X + 1) * cj, xl = ci + (F2X + 1) * cj;
    const int yt = ci + F2X * (cj + 1), yb = ci + F2X * cj;
    const double a = cdiv_rn(SX[xr] - SX[xl], k.dx, k.rdx) + cdiv_rn(SY[NYF + yt] - SY[NYF + yb], k.dy, k.rdy);
    const double b = cdiv_rn(SX[NXF + xr] - SX[NXF + xl], k.dx, k.rdx) + cdiv_rn(SY[yt] - SY[yb], k.dy, k.rdy);
    const double c1 = cdiv_rn(ur * SX[xr] - ul * SX[xl], k.dx, k.rdx) +
                      cdiv_rn(vr * SX[NXF + xr] - vl * SX[NXF + xl], k.dx, k.rdx) +
                      cdiv_rn(ut * SY[NYF + yt] - ub * SY[NYF + yb], k.dy, k.rdy) +
                      cdiv_rn(vt * SY[yt] - vb * SY[yb], k.dy, k.rdy);
    const double c2 =
        cdiv_rn(SX[2 * NXF + xr] - SX[2 * NXF + xl], k.dx, k.rdx) + cdiv_rn(SY[2 * NYF + yt] - SY[2 * NYF + yb], k.dy, k.rdy);
    dv[0] = dv[0] + a;
    dv[1] = dv[1] + b;
    dv[2] = dv[2] + c1 - c2;
  }

  // ---- C-equation
  if (CM != 0) {
    const double *const Cv_ = U + N * NV;  // C variables (plain form)
    double alpha;
    if (CM == 1) {
      const double maxS = as_double(scIn->maxS);
      alpha = k.dd2 / (as_double(scIn->maxCh) + 1.0e-6) * k.alpha;
      // sources and relaxation (calculateRhoGrad2D, updateCeq2D)
      __syncthreads();
      for (int r = tid; r < RN; r += NT2) {
        const int gi = i0 - F2H + r % RW, gj = j0 - F2H + r / RW;
        if (gi < k.ngi && gj < k.ngj)
          sW[r] = U[gi + (long long)k.ngi * gj + 2 * N] / sR[r] - 0.5 * sR[r] * (sU[r] * sU[r] + sV[r] * sV[r]);
      }
      __syncthreads();
      if (act) {
        const double *E = sW;
        const double dxr = cdiv_rn(sR[rc - 2] - 8.0 * sR[rc - 1] + 8.0 * sR[rc + 1] - sR[rc + 2], k.dx12, k.rdx12);
        const double dyr = cdiv_rn(sR[rc - 2 * RW] - 8.0 * sR[rc - RW] + 8.0 * sR[rc + RW] - sR[rc + 2 * RW], k.dy12,
                                   k.rdy12);
        const double dxe = cdiv_rn(E[rc - 2] - 8.0 * E[rc - 1] + 8.0 * E[rc + 1] - E[rc + 2], k.dx12, k.rdx12);
        const double dye =
            cdiv_rn(E[rc - 2 * RW] - 8.0 * E[rc - RW] + 8.0 * E[rc + RW] - E[rc + 2 * RW], k.dy12, k.rdy12);
        const double dxu = cdiv_rn(sU[rc - 2] - 8.0 * sU[rc - 1] + 8.0 * sU[rc + 1] - sU[rc + 2], k.dx12, k.rdx12);
        const double dyv = cdiv_rn(sV[rc - 2 * RW] - 8.0 * sV[rc - RW] + 8.0 * sV[rc + RW] - sV[rc + 2 * RW], k.dy12,
                                   k.rdy12);
        const double rgrad = sqrt(dxr * dxr + dyr * dyr);
        const double divu = dxu + dyv;
        const double dnednr = (dxr * dxe + dyr * dye) * (dxr * dxr + dyr * dyr);
        const double i1 = dnednr < 0.0 ? 1.0 : 0.0;
        const double i2 = divu < 0.0 ? 1.0 : 0.0;
        const double gr[4] = {(1 - i1) * i2 * rgrad, i1 * rgrad, -dyr, dxr};
        const double relax = maxS / (k.eps * k.dxmag), diff = k.kap * maxS * k.dxmag;
        const long long sj = k.ngi;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          const double *C = Cv_ + N * n;
          const double cc = C[c];
          const double dx_right = cdiv_rn(C[c + 1] - cc, k.dx, k.rdx);
          const double dx_left = cdiv_rn(cc - C[c - 1], k.dx, k.rdx);
          const double dy_top = cdiv_rn(C[c + sj] - cc, k.dy, k.rdy);
          const double dy_bot = cdiv_rn(cc - C[c - sj], k.dy, k.rdy);
          const double lap = dx_right - dx_left + dy_top - dy_bot;
          dv[NV + n] = relax * (gr[n] - cc) + diff * lap;
        }
      }
    } else {
      alpha = k.dd2 / (as_double(lsc->maxCh) + 1.0e-6) * k.alpha;
    }
    // artificial-viscosity face fluxes (computeCeqFlux2D, computeCeqFaces2D)
    // and their differences (applyCeq2D): x faces, then y faces
    auto face_c = [&](long long q0, long long q1, double &chat, double &ct0, double &ct1) {
      if (CM == 1) {
        chat = (Cv_[q0 + N] + Cv_[q1 + N]) / 2.0;
        ct0 = (Cv_[q0 + 2 * N] + Cv_[q1 + 2 * N]) / 2.0;
        ct1 = (Cv_[q0 + 3 * N] + Cv_[q1 + 3 * N]) / 2.0;
      } else {
        chat = ((double)lchat[q0] + (double)lchat[q1]) / 2.0;
        ct0 = ((double)ldir[q0] + (double)ldir[q1]) / 2.0;
        ct1 = ((double)ldir[q0 + N] + (double)ldir[q1 + N]) / 2.0;
      }
    };
    __syncthreads();
    for (int f = tid; f < NXF; f += NT2) {
      const int fi = i0 - 1 + f % (F2X + 1), fj = j0 + f / (F2X + 1);
      if (fi >= ie || fj >= je) continue;
      const int b0 = rix(i0, j0, fi, fj);
      const long long q0 = fi + (long long)k.ngi * fj;
      double chat, ct0, ct1;
      face_c(q0, q0 + 1, chat, ct0, ct1);
      const double rr = (sR[b0] + sR[b0 + 1]) / 2.0;
      const double ctmag = (ct0 * ct0 + ct1 * ct1) + 1.0e-6;
      const double X00 = alpha * rr * chat * (ct0 * ct0) / ctmag;
      const double X01 = alpha * rr * chat * (ct0 * ct1) / ctmag;
      sW[f] = X00 * cdiv_rn(sU[b0 + 1] - sU[b0], k.dx, k.rdx);
      sW[NXF + f] = X00 * cdiv_rn(sV[b0 + 1] - sV[b0], k.dx, k.rdx);
      sW[2 * NXF + f] =
          X01 * cdiv_rn((sU[b0 + RW] + sU[b0 + 1 + RW]) - (sU[b0 - RW] + sU[b0 + 1 - RW]), k.dy4, k.rdy4);
      sW[3 * NXF + f] =
          X01 * cdiv_rn((sV[b0 + RW] + sV[b0 + 1 + RW]) - (sV[b0 - RW] + sV[b0 + 1 - RW]), k.dy4, k.rdy4);
    }
    __syncthreads();
    const int xr = ci + 1 + (F2X + 1) * cj, xl = ci + (F2X + 1) * cj;
    double d0 = 0, d1 = 0;
    d0 += cdiv_rn(sW[xr] - sW[xl], k.dx, k.rdx);
    d0 += cdiv_rn(sW[2 * NXF + xr] - sW[2 * NXF + xl], k.dx, k.rdx);
    d1 += cdiv_rn(sW[NXF + xr] - sW[NXF + xl], k.dx, k.rdx);
    d1 += cdiv_rn(sW[3 * NXF + xr] - sW[3 * NXF + xl], k.dx, k.rdx);
    __syncthreads();
    for (int f = tid; f < NYF; f += NT2) {
      const int fi = i0 + f % F2X, fj = j0 - 1 + f / F2X;
      if (fi >= ie || fj >= je) continue;
      const int b0 = rix(i0, j0, fi, fj);
      const long long q0 = fi + (long long)k.ngi * fj;
      double chat, ct0, ct1;
      face_c(q0, q0 + k.ngi, chat, ct0, ct1);
      const double rr = (sR[b0] + sR[b0 + RW]) / 2.0;
      const double ctmag = (ct0 * ct0 + ct1 * ct1) + 1.0e-6;
      const double X10 = alpha * rr * chat * (ct1 * ct0) / ctmag;
      const double X11 = alpha * rr * chat * (ct1 * ct1) / ctmag;
      sW[f] = X10 * cdiv_rn((sU[b0 + 1] + sU[b0 + 1 + RW]) - (sU[b0 - 1] + sU[b0 - 1 + RW]), k.dx4, k.rdx4);
      sW[NYF + f] = X10 * cdiv_rn((sV[b0 + 1] + sV[b0 + 1 + RW]) - (sV[b0 - 1] + sV[b0 - 1 + RW]), k.dx4, k.rdx4);
      sW[2 * NYF + f] = X11 * cdiv_rn(sU[b0 + RW] - sU[b0], k.dy, k.rdy);
      sW[3 * NYF + f] = X11 * cdiv_rn(sV[b0 + RW] - sV[b0], k.dy, k.rdy);
    }
    __syncthreads();
    const int yt = ci + F2X * (cj + 1), yb = ci + F2X * cj;
    d0 += cdiv_rn(sW[yt] - sW[yb], k.dy, k.rdy);
    d0 += cdiv_rn(sW[2 * NYF + yt] - sW[2 * NYF + yb], k.dy, k.rdy);
    d1 += cdiv_rn(sW[NYF + yt] - sW[NYF + yb], k.dy, k.rdy);
    d1 += cdiv_rn(sW[3 * NYF + yt] - sW[3 * NYF + yb], k.dy, k.rdy);
    dv[0] += d0;
    dv[1] += d1;
  }

  // ---- stage update and the next stage's maxima
  double v3[3] = {0.0, 0.0, 0.0};
  if (act) {
    double w[NVT];
#pragma unroll
    for (int v = 0; v < NVT; ++v) {
      const long long q = c + N * v;
      const double un = (k.upd == UPD_ADD || k.upd == UPD_BLEND) ? Un[q] : 0.0;
      const double u = k.upd == UPD_ADD ? 0.0 : U[q];
      w[v] = upd_val(k, u, un, dv[v]);
      Out[q] = w[v];
    }
    if (CM == 1) cell_maxima<NS>(k, w, v3[0], v3[1], v3[2]);
  }
  if (CM == 1) {
    unsigned long long *const d3[3] = {&scOut->maxS, &scOut->maxC, &scOut->maxCh};
    block_max_atomic<3>(v3, d3);
  }
  // ghost cells next to the tile: the update with a zero right-hand side
  const int ilo = bx == 0 ? 0 : i0, ihi = bx == k.nbx - 1 ? k.ngi : i0 + F2X;
  const int jlo = by == 0 ? 0 : j0, jhi = by == k.nby - 1 ? k.ngj : j0 + F2Y;
  if (bx == 0 || by == 0 || bx == k.nbx - 1 || by == k.nby - 1) {
    const int wi = ihi - ilo, n = wi * (jhi - jlo);
    for (int r = tid; r < n; r += NT2) {
      const int gi = ilo + r % wi, gj = jlo + r / wi;
      if (gi >= k.ng && gi < k.ng + k.nci && gj >= k.ng && gj < k.ng + k.ncj) continue;
      const long long q0 = gi + (long long)k.ngi * gj;
      for (int v = 0; v < NVT; ++v) {
        const long long q = q0 + N * v;
        const double un = (k.upd == UPD_ADD || k.upd == UPD_BLEND) ? Un[q] : 0.0;
        const double u = k.upd == UPD_ADD ? 0.0 : U[q];
        Out[q] = upd_val(k, u, un, 0.0);
      }
    }
  }
}

__global__ void k_clear2d(Scalars *sc) {
  if (threadIdx.x < 4) reinterpret_cast<unsigned long long *>(sc)[threadIdx.x] = 0ull;
}

typedef void (*Stage2Fn)(dim3, dim3, hipStream_t, const K2 &, const double *, const double *, double *,
                         const Scalars *, Scalars *, Scalars *, const float *, const float *, const Scalars *);
typedef void (*Max2Fn)(unsigned, hipStream_t, const K2 &, const double *, Scalars *);

template <int NS, int CM>
void launch_stage2d(dim3 g, dim3 b, hipStream_t s, const K2 &k, const double *U, const double *Un, double *Out,
                    const Scalars *si, Scalars *so, Scalars *sc, const float *ch, const float *dr, const Scalars *ls) {
  k_stage2d<NS, CM><<<g, b, 0, s>>>(k, U, Un, Out, si, so, sc, ch, dr, ls);
}
template <int NS>
void launch_max2d(unsigned nb, hipStream_t s, const K2 &k, const double *U, Scalars *sc) {
  k_max2d<NS><<<nb, 256, 0, s>>>(k, U, sc);
}

template <int NS>
Stage2Fn pick_cm(int cm) {
  return cm == 0 ? launch_stage2d<NS, 0> : (cm == 1 ? launch_stage2d<NS, 1> : launch_stage2d<NS, 2>);
}

K2 make_k2(const Geo &g, const Phys &ph, const fst_config &cf) {
  K2 k;
  k.N = g.N;
  k.ngi = g.ngi; k.ngj = g.ngj; k.ng = g.ng; k.nci = g.nci; k.ncj = g.ncj;
  k.ns = cf.ns; k.nv = g.nv; k.nvt = g.nvt;
  k.nbx = (g.nci + F2X - 1) / F2X;
  k.nby = (g.ncj + F2Y - 1) / F2Y;
  k.dx = ph.cd[1]; k.dy = ph.cd[2];
  k.rdx = 1.0 / k.dx; k.rdy = 1.0 / k.dy;
  k.dx4 = 4.0 * k.dx; k.rdx4 = 1.0 / k.dx4; k.dy4 = 4.0 * k.dy; k.rdy4 = 1.0 / k.dy4;
  k.dx12 = 12.0 * k.dx; k.rdx12 = 1.0 / k.dx12; k.dy12 = 12.0 * k.dy; k.rdy12 = 1.0 / k.dy12;
  for (int s = 0; s < 3; ++s) {
    k.cpf[s] = s < cf.ns ? ph.cpf[s] : 0.0;
    k.cvf[s] = s < cf.ns ? ph.cvf[s] : 0.0;
    k.qcp[s] = k.qcv[s] = 0.0;
    if (s < cf.ns) {  // maxWaveSpeed2D: gammas = cd(6 + 2s), Rs = cd(6 + 2s + 1)
      const double gammas = ph.cd[6 + 2 * s], Rs = ph.cd[6 + 2 * s + 1];
      k.qcp[s] = gammas * Rs / (gammas - 1);
      k.qcv[s] = Rs / (gammas - 1);
    }
  }
  k.g = ph.g; k.rhoRef = ph.rhoRef;
  k.buoy = cf.buoyancy; k.visc = cf.visc;
  k.eps = ph.eps; k.kap = ph.kap; k.alpha = ph.alpha;
  k.dxmag = sqrt(k.dx * k.dx + k.dy * k.dy);
  k.dd2 = ph.dx * ph.dx + ph.dy * ph.dy;
  k.dt = cf.dt;
  k.hdt = 0.5f * cf.dt;
  k.a = k.b = 0.0;
  k.upd = UPD_RK2_1;
  return k;
}

}  // namespace

bool fused2d_supported(const fst_config &cf) {
  return cf.ndim == 2 && cf.grid == 0 && cf.scheme == FST_WENO5 && cf.precision == FST_FP64 && cf.ns >= 1 &&
         cf.ns <= 3 && cf.nci >= 1 && cf.ncj >= 1;
}

int fused2d_steps(const Geo &g, const Phys &ph, const fst_config &cf, Bufs &b, FusedBufs &fb, LearnedBufs *lb,
                  int nsteps, hipStream_t s, std::string *err) {
  if (!fused2d_supported(cf)) {
    *err = "fused 2D path: configuration not supported";
    return FST_ERR_ARG;
  }
  const int cm = !cf.ceq ? 0 : (cf.learned_c ? 2 : 1);
  if (cm == 2 && (!lb || !lb->have_weights)) {
    *err = "learned C-indicator: no network weights set";
    return FST_ERR_STATE;
  }
  const bool rk3 = rk_stages(cf) == 3;
  const size_t bytes = (size_t)g.N * sizeof(double) * g.nvt;
  if (!fb.stage) {
    if (hipMalloc((void **)&fb.stage, bytes) != hipSuccess) {
      *err = "fused 2D path: device allocation failed";
      return FST_ERR_NOMEM;
    }
    (void)hipMemsetAsync(fb.stage, 0, bytes, s);
  }
  if (rk3 && !fb.stage2) {
    if (hipMalloc((void **)&fb.stage2, bytes) != hipSuccess) {
      *err = "fused 2D path: device allocation failed";
      return FST_ERR_NOMEM;
    }
    (void)hipMemsetAsync(fb.stage2, 0, bytes, s);
  }
  if (!fb.sc) {
    if (hipMalloc((void **)&fb.sc, sizeof(Scalars) * 3) != hipSuccess) {
      *err = "fused 2D path: device allocation failed";
      return FST_ERR_NOMEM;
    }
    (void)hipMemsetAsync(fb.sc, 0, sizeof(Scalars) * 3, s);
  }
  K2 k = make_k2(g, ph, cf);
  Stage2Fn sf = cf.ns == 1 ? pick_cm<1>(cm) : (cf.ns == 2 ? pick_cm<2>(cm) : pick_cm<3>(cm));
  Max2Fn mf = cf.ns == 1 ? launch_max2d<1> : (cf.ns == 2 ? launch_max2d<2> : launch_max2d<3>);
  const dim3 grid((unsigned)(k.nbx * k.nby)), blk(NT2);
  Scalars *const R[3] = {fb.sc, fb.sc + 1, fb.sc + 2};
  if (cm == 1 && !fb.valid) {  // maxima of var (interior)
    fb.cur = 0;
    k_clear2d<<<1, 64, 0, s>>>(R[0]);
    k_clear2d<<<1, 64, 0, s>>>(R[1]);
    const long long nc = (long long)g.nci * g.ncj;
    mf((unsigned)((nc + 255) / 256), s, k, b.var, R[0]);
  }
  fb.valid = true;
  Prof *prof = b.prof && b.prof->on ? b.prof : nullptr;
  if (prof) {
    prof->name = "2D stage: [ghost fill, learned fields: stage 1] + k_stage2d";
    prof->per = 2;
  }
  const int nst = rk3 ? 3 : 2;
  for (int t = 0; t < nsteps; ++t) {
    for (int st = 1; st <= nst; ++st) {
      if (prof) prof->begin(s);
      const double *Uin;
      const double *Un = b.var;
      double *Out;
      if (st == 1) {
        launch_apply_bcs(g, cf, b.var, s);  // cart2d preStep (src/cart2d.cpp:147-149)
        if (cm == 2) {
          const int rc = learned_fields2d(g, cf, ph, b.var, b.rho, *lb, s);
          if (rc) {
            *err = "learned C-indicator: field evaluation failed";
            return rc;
          }
        }
        Uin = b.var;
        Out = fb.stage;
        k.upd = rk3 ? UPD_RK3_1 : UPD_RK2_1;
      } else if (!rk3) {
        Uin = fb.stage;
        Out = b.var;
        k.upd = UPD_ADD;
      } else {
        Uin = st == 2 ? fb.stage : fb.stage2;
        Out = st == 2 ? fb.stage2 : b.var;
        k.upd = UPD_BLEND;
        k.a = st == 2 ? RK3_A2 : RK3_A3;
        k.b = st == 2 ? RK3_B2 : RK3_B3;
      }
      const int cur = fb.cur;
      sf(grid, blk, s, k, Uin, Un, Out, R[cur], R[(cur + 1) % 3], R[(cur + 2) % 3], cm == 2 ? lb->chat : nullptr,
         cm == 2 ? lb->dir : nullptr, cm == 2 ? lb->sc : nullptr);
      if (cm == 1) fb.cur = (cur + 1) % 3;
      if (prof) prof->end(s);
    }
  }
  if (cm == 1) {
    // the maxima the last stage used, where fst_last_maxima reads them
    (void)hipMemcpyAsync(b.sc, R[(fb.cur + 2) % 3], sizeof(Scalars), hipMemcpyDeviceToDevice, s);
  }
  if (hipGetLastError() != hipSuccess) {
    *err = "fused 2D path: launch failed";
    return FST_ERR_HIP;
  }
  return FST_OK;
}

}  // namespace fst
