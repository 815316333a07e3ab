// Learned C-indicator: features, BF16-MFMA MLP inference and training.
//
// The MLP stands in for the relaxation solve of the C-equation
// (updateCeq, src/ceq3d.hpp:189-224): it maps a 5x5x5 stencil of three
// dimensionless features
//   f0 = |grad rho| h / rho,  f1 = |grad p| h / p,  f2 = (div u) h / a
// (2nd-order central differences, a = mixture sound speed, h = min cell size)
// to the smoothed contact indicator C_hat of the plain method (the variable
// var(nv+1) that calculateCeqFaces reads, src/ceq3d.hpp:269), predicted as
// y = C_hat h / rho.  Architecture 375 (padded to 384) -> 64 -> 64 -> 64 -> 1
// with GELU (tanh form) on the hidden layers.  Products run on MFMA
// (v_mfma_f32_16x16x16bf16_1k) with BF16 operands and FP32 accumulation;
// master weights, biases, the output layer and the optimizer are FP32.
//
// MFMA 16x16x16 lane layout (lane l, q = 0..3, v = 0..3):
//   A[i = l%16][k = 4*(l/16)+q], B[k = 4*(l/16)+q][j = l%16],
//   D[i = 4*(l/16)+v][j = l%16].
//
// The direction n of the anisotropic artificial viscosity is the 4th-order
// density gradient of calculateRhoGrad (src/ceq3d.hpp:145-147, 182-184), the
// quantity the relaxed variables var(nv+2..nv+4) follow.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fst_mlp.hpp"

namespace fst {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// LDS row stride (shorts) of W1^T: 8-byte reads spread over the banks
template <int ND>
constexpr int W1S = Net<ND>::IN_PAD + 4;
constexpr int WHS = MLP_H + 4;  // LDS row stride of W2^T, W3^T and the activation tiles

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__host__ __device__ __forceinline__ short f2bf(float x) {
  unsigned u = __builtin_bit_cast(unsigned, x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((unsigned)(unsigned short)b) << 16); }

__device__ __forceinline__ float gelu(float x) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  return 0.5f * x * (1.0f + tanhf(k0 * (x + k1 * x * x * x)));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  const float t = tanhf(k0 * (x + k1 * x * x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * k0 * (1.0f + 3.0f * k1 * x * x);
}

// Orders LDS writes and reads among the lanes of one wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float wave_maxf(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

struct G3 {  // plane geometry
  int ngi, ngj, ngk, ng, nci, ncj, nck;
  long long N, s1, s2;
};

G3 make_g3(const Geo &g) {
  G3 q;
  q.ngi = g.ngi; q.ngj = g.ngj; q.ngk = g.ngk; q.ng = g.ng;
  q.nci = g.nci; q.ncj = g.ncj; q.nck = g.nck;
  q.N = g.N; q.s1 = g.ngi; q.s2 = (long long)g.ngi * g.ngj;
  return q;
}

inline unsigned nblk(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }
// 3D launch grid of the per-cell passes: block (64, 4) over (i, j), one z-slice per k

// ---------------------------------------------------------------- features

// Per cell: F (cells with one neighbour each way), dir and G1 (cells with two
// neighbours each way), M = max |F|.  P planes: rho, p, u, v, w, nrg, Y_s.
// Every value is computed in FP64 in a fixed order and rounded once to FP32.
// Launch: each wave of 64 lanes takes 64 consecutive (i, j) columns of the
// ghosted plane (rows wrap, so no lane idles at the end of a row) and writes
// the middle FEAT_OUT of them; each lane marches its column through FEAT_KCH
// planes, keeping the z stencils (rho and nrg at k-2..k+2, p and w at
// k-1..k+1) in registers, so each plane is read once for z.  x neighbours
// come from the neighbouring lanes (the 2 edge lanes on each side load but do
// not write), y neighbours through L1/L2.  At 262^3 this is 16 % faster than
// 64 x 4 blocks over (i, j), which leave 58 of every 320 lanes idle and load
// every x neighbour.
constexpr int FEAT_KCH = 16;
constexpr int FEAT_WAVES = 4;   // waves per block
constexpr int FEAT_OUT = 60;    // columns written per wave

// Arithmetic of the feature pass: FP64 with IEEE divisions (FP64 primitive
// planes; equal to the oracle's features3d bit for bit), or binary32 with
// hardware reciprocals when the planes are the FP32 forms' (the features
// feed a BF16 network).
template <bool PF>
struct FeatAr {
  using T = double;
  static __device__ __forceinline__ T cdv(T a, double b, double rb) { return cdiv_rn(a, b, rb); }
  static __device__ __forceinline__ T div(T a, T b) { return a / b; }
  static __device__ __forceinline__ T sq(T a) { return sqrt(a); }
};
template <>
struct FeatAr<true> {
  using T = float;
  static __device__ __forceinline__ T cdv(T a, double, double rb) { return a * (float)rb; }
  static __device__ __forceinline__ T div(T a, T b) { return a * __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ T sq(T a) { return __builtin_sqrtf(a); }
};

template <bool PF>
__global__ void __launch_bounds__(64 * FEAT_WAVES) k_features3d(const G3 g, const FeatConst fc,
                                                                const double *__restrict__ P, float *__restrict__ F,
                                                                float *__restrict__ M, float *__restrict__ Dir,
                                                                float *__restrict__ G1, Scalars *sc) {
  const int lane = threadIdx.x & 63;
  // wave w of the grid holds plane columns FEAT_OUT w - 2 .. FEAT_OUT w + 61
  // and writes lanes 2..61; its edge lanes only load
  const long long c = (long long)(blockIdx.x * FEAT_WAVES + (threadIdx.x >> 6)) * FEAT_OUT - 2 + lane;
  const bool outl = lane >= 2 && lane < 62;
  const bool inplane = c >= 0 && c < (long long)g.ngi * g.ngj;
  const int j = inplane ? (int)(c / g.ngi) : 0, i = inplane ? (int)(c - (long long)j * g.ngi) : 0;
  const int k0 = blockIdx.y * FEAT_KCH, k1 = min(k0 + FEAT_KCH, g.ngk);
  const bool col = inplane;
  const long long N = g.N;
  using A = FeatAr<PF>;
  using T = typename A::T;
  auto PR = [&](int q, long long t) { return (T)prim<PF>(P, N, q, t); };  // primitive plane q at cell t
  const long long cij = (long long)i + g.s1 * j;
  auto at = [&](int q, int k) { return (k >= 0 && k < g.ngk) ? PR(q, cij + g.s2 * k) : (T)0; };
  // rho, nrg at k-2..k+2; p, w at k-1..k+1 (zeros on lanes past the plane)
  T rq[5] = {0, 0, 0, 0, 0}, eq[5] = {0, 0, 0, 0, 0}, pq[3] = {0, 0, 0}, wq[3] = {0, 0, 0};
  if (col) {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      rq[q] = at(0, k0 - 2 + q);
      eq[q] = at(5, k0 - 2 + q);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      pq[q] = at(1, k0 - 1 + q);
      wq[q] = at(4, k0 - 1 + q);
    }
  }
  float g1max = 0.f;
  for (int k = k0; k < k1; ++k) {
    // x neighbours from the neighbouring lanes (every lane takes part)
    const T uc = col ? PR(2, cij + g.s2 * k) : (T)0;
    const T xr1 = __shfl_down(rq[2], 1, 64), xl1 = __shfl_up(rq[2], 1, 64);
    const T xr2 = __shfl_down(rq[2], 2, 64), xl2 = __shfl_up(rq[2], 2, 64);
    const T xe1 = __shfl_down(eq[2], 1, 64), xf1 = __shfl_up(eq[2], 1, 64);
    const T xe2 = __shfl_down(eq[2], 2, 64), xf2 = __shfl_up(eq[2], 2, 64);
    const T xpp = __shfl_down(pq[1], 1, 64), xpm = __shfl_up(pq[1], 1, 64);
    const T xup = __shfl_down(uc, 1, 64), xum = __shfl_up(uc, 1, 64);
    float f0 = 0.f, f1 = 0.f, f2 = 0.f, d0 = 0.f, d1 = 0.f, d2 = 0.f, gg = 0.f;
    if (col && outl) {
      const long long t = cij + g.s2 * k;
      const bool in1 = i >= 1 && j >= 1 && k >= 1 && i < g.ngi - 1 && j < g.ngj - 1 && k < g.ngk - 1;
      const bool in2 = i >= 2 && j >= 2 && k >= 2 && i < g.ngi - 2 && j < g.ngj - 2 && k < g.ngk - 2;
      if (in1) {
        const long long s[2] = {1, g.s1};
        T gr2 = 0, gp2 = 0, divu = 0, dr[3] = {0, 0, 0}, de[3] = {0, 0, 0};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const double h2 = fc.h2[a], rh2 = fc.rh2[a];  // divisions by run constants: cdiv_rn == IEEE
          T rp, rm, pp, pm, vp, vm;
          if (a == 0) {
            rp = xr1; rm = xl1; pp = xpp; pm = xpm; vp = xup; vm = xum;
          } else if (a == 1) {
            rp = PR(0, t + s[a]); rm = PR(0, t - s[a]);
            pp = PR(1, t + s[a]); pm = PR(1, t - s[a]);
            vp = PR(2 + a, t + s[a]); vm = PR(2 + a, t - s[a]);
          } else {
            rp = rq[3]; rm = rq[1];
            pp = pq[2]; pm = pq[0];
            vp = wq[2]; vm = wq[0];
          }
          const T r1 = A::cdv(rp - rm, h2, rh2);
          const T p1 = A::cdv(pp - pm, h2, rh2);
          gr2 = gr2 + r1 * r1;
          gp2 = gp2 + p1 * p1;
          divu = divu + A::cdv(vp - vm, h2, rh2);
          if (in2) {
            const double h12 = fc.h12[a], rh12 = fc.rh12[a];
            if (a == 0) {
              dr[a] = A::cdv(xl2 - (T)8.0 * rm + (T)8.0 * rp - xr2, h12, rh12);
              de[a] = A::cdv(xf2 - (T)8.0 * xf1 + (T)8.0 * xe1 - xe2, h12, rh12);
            } else if (a == 1) {
              dr[a] = A::cdv(PR(0, t - 2 * s[a]) - (T)8.0 * rm + (T)8.0 * rp - PR(0, t + 2 * s[a]), h12, rh12);
              de[a] = A::cdv(PR(5, t - 2 * s[a]) - (T)8.0 * PR(5, t - s[a]) + (T)8.0 * PR(5, t + s[a]) -
                                 PR(5, t + 2 * s[a]), h12, rh12);
            } else {
              dr[a] = A::cdv(rq[0] - (T)8.0 * rq[1] + (T)8.0 * rq[3] - rq[4], h12, rh12);
              de[a] = A::cdv(eq[0] - (T)8.0 * eq[1] + (T)8.0 * eq[3] - eq[4], h12, rh12);
            }
          }
        }
        const T rho = rq[2], p = pq[1];
        T Cp = 0, Cv = 0;
        for (int q = 0; q < fc.ns; ++q) {
          const T Y = PR(6 + q, t);
          Cp = Cp + Y * (T)fc.cpf[q];
          Cv = Cv + Y * (T)fc.cvf[q];
        }
        const T hmin = (T)fc.hmin;
        const T a_snd = A::sq(A::div(A::div(Cp, Cv) * p, rho));
        f0 = (float)A::div(A::sq(gr2) * hmin, rho);
        f1 = (float)A::div(A::sq(gp2) * hmin, p);
        f2 = (float)A::div(divu * hmin, a_snd);
        if (in2) {
          d0 = (float)dr[0];
          d1 = (float)dr[1];
          d2 = (float)dr[2];
          const T rg2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
          const T dnednr = (dr[0] * de[0] + dr[1] * de[1] + dr[2] * de[2]) * rg2;
          gg = dnednr < (T)0 ? (float)A::div(A::sq(rg2) * hmin, rho) : 0.f;
          const bool interior = i >= g.ng && j >= g.ng && k >= g.ng && i < g.ng + g.nci && j < g.ng + g.ncj &&
                                k < g.ng + g.nck;
          if (interior) g1max = fmaxf(g1max, gg);
        }
      }
    }
    if (col) {
      const long long t = cij + g.s2 * k;
      if (outl) {
        F[t] = f0;
        F[t + N] = f1;
        F[t + 2 * N] = f2;
        M[t] = fmaxf(fabsf(f0), fmaxf(fabsf(f1), fabsf(f2)));
        Dir[t] = d0;
        Dir[t + N] = d1;
        Dir[t + 2 * N] = d2;
        G1[t] = gg;
      }
      // advance the z stencils
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        rq[q] = rq[q + 1];
        eq[q] = eq[q + 1];
      }
      rq[4] = at(0, k + 3);
      eq[4] = at(5, k + 3);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        pq[q] = pq[q + 1];
        wq[q] = wq[q + 1];
      }
      pq[2] = at(1, k + 2);
      wq[2] = at(4, k + 2);
    }
  }
  // block maximum, then one atomic per block that has a non-zero value
  __shared__ float red[FEAT_WAVES];
  g1max = wave_maxf(g1max);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = g1max;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.f;
    for (int q = 0; q < FEAT_WAVES; ++q) m = fmaxf(m, red[q]);
    if (m > 0.f) atomic_max_nonneg(&sc->maxC, (double)m);
  }
}

constexpr int ACT_KCH = 16;  // planes per block of the activity pass
inline dim3 act3d_grid(const Geo &g) {
  return dim3((unsigned)(((long long)g.ngi * g.ngj + 255) / 256), (g.ngk + ACT_KCH - 1) / ACT_KCH);
}

// Active cells: interior cells whose 3x3x3 neighbourhood has M > tau (every
// stride-th interior cell as well when stride > 0), appended to `list` with
// one atomic per wave; C_hat is cleared on every cell.  256 threads per block
// over 256 consecutive (i, j) columns of the ghosted plane (rows wrap, so no
// lane idles at the end of a row); each thread marches its column through
// ACT_KCH planes, keeping the 3x3 in-plane maxima of planes k-1, k, k+1 in
// registers (9 loads per cell instead of 27; taking the x neighbours' maxima
// from the neighbouring lanes instead, 3 loads per cell, was 14 % slower).
__global__ void __launch_bounds__(256) k_active3d(const G3 g, const float *__restrict__ M, float tau,
                                                  unsigned stride, unsigned *__restrict__ list,
                                                  unsigned *__restrict__ cnt, float *__restrict__ chat) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int j = c / g.ngi, i = c - j * g.ngi;
  const int k0 = blockIdx.y * ACT_KCH, k1 = min(k0 + ACT_KCH, g.ngk);
  const int ng = g.ng;
  const bool col = i < g.ngi && j < g.ngj;
  const bool colin = col && i >= ng && j >= ng && i < ng + g.nci && j < ng + g.ncj;
  const long long cij = (long long)i + g.s1 * j;
  auto m9 = [&](int k) {  // max of M over the 3x3 cells around (i, j) in plane k
    float m = 0.f;
    if (colin && k >= 0 && k < g.ngk) {
      const long long t = cij + g.s2 * k;
      for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di) m = fmaxf(m, M[t + di + dj * g.s1]);
    }
    return m;
  };
  float qm = m9(k0 - 1), qc = m9(k0);
  const int lane = threadIdx.x & 63;
  for (int k = k0; k < k1; ++k) {
    const float qp = m9(k + 1);
    const long long t = cij + g.s2 * k;
    bool act = false;
    if (col) {
      if (chat) chat[t] = 0.f;
      if (colin && k >= ng && k < ng + g.nck) {
        act = fmaxf(qm, fmaxf(qc, qp)) > tau;
        if (stride > 0) {
          const long long lin = (long long)(i - ng) + (long long)g.nci * ((j - ng) + (long long)g.ncj * (k - ng));
          act = act || (lin % stride) == 0;
        }
      }
    }
    const unsigned long long mask = __ballot(act);
    if (mask) {
      const int first = __ffsll((long long)mask) - 1;
      unsigned base = 0;
      if (lane == first) base = atomicAdd(cnt, (unsigned)__popcll(mask));
      base = __shfl(base, first, 64);
      if (act) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = (unsigned)t;
    }
    qm = qc;
    qc = qp;
  }
}

// 2D features (cart2d state, src/cart2d.cpp:151-237), one thread per cell:
// F (cells with one neighbour each way), the tangent and G1 (two
// neighbours each way), M = max |F|, from var.  The tangent (-d rho/dy,
// d rho/dx) and the contact switch follow calculateRhoGrad2D
// (src/ceq.hpp:80-151): 4th-order gradients of rho and of
// nrg = E / rho - rho |u|^2 / 2, G1 = |grad rho| h / rho where
// (grad rho . grad nrg) |grad rho|^2 < 0.  FP64 arithmetic in a fixed order
// (divisions by run constants correctly rounded, cdiv_rn), one rounding to
// FP32 per output.  A block (64 x 4 cells) stages rho, p, u, v and nrg of its
// cells and a 2-cell halo in LDS, computed from var as calculateRhoPT2D and
// computeVelocity2D compute them (src/secondary.hpp:60-107,
// src/velocity.hpp:85-99), and writes rho of its cells to rho_out (the
// network's and the residual check's scale).
constexpr int FW2 = 64 + 4, FH2 = 4 + 4, FN2 = FW2 * FH2;
__global__ void __launch_bounds__(256) k_features2d(const G3 g, const FeatConst fc, const double *__restrict__ var,
                                                    double *__restrict__ rho_out, float *__restrict__ F,
                                                    float *__restrict__ M, float *__restrict__ Dir,
                                                    float *__restrict__ G1, Scalars *sc) {
  __shared__ double sR[FN2], sP[FN2], sU[FN2], sV[FN2], sE[FN2];
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 4;
  const int tid = threadIdx.x + 64 * threadIdx.y;
  const long long N = g.N;
  for (int r = tid; r < FN2; r += 256) {
    const int gi = i0 - 2 + r % FW2, gj = j0 - 2 + r / FW2;
    if (gi >= 0 && gj >= 0 && gi < g.ngi && gj < g.ngj) {
      const long long q = gi + g.s1 * gj;
      double rr = 0.0;
      for (int s = 0; s < fc.ns; ++s) rr = rr + var[q + N * (3 + s)];
      double Cp = 0, Cv = 0;
      for (int s = 0; s < fc.ns; ++s) {
        const double Y = var[q + N * (3 + s)] / rr;
        Cp = Cp + Y * fc.cpf[s];
        Cv = Cv + Y * fc.cvf[s];
      }
      const double gamma = Cp / Cv;
      const double m0 = var[q], m1 = var[q + N];
      const double pp = (gamma - 1) * (var[q + 2 * N] - (0.5 / rr) * (m0 * m0 + m1 * m1));
      const double u = m0 / rr, v = m1 / rr;
      sR[r] = rr;
      sP[r] = pp;
      sU[r] = u;
      sV[r] = v;
      sE[r] = var[q + 2 * N] / rr - 0.5 * rr * (u * u + v * v);
    }
  }
  __syncthreads();
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  float g1max = 0.f;
  if (i < g.ngi && j < g.ngj) {
    const long long t = (long long)i + g.s1 * j;
    const int c = (threadIdx.x + 2) + FW2 * (threadIdx.y + 2);
    const bool in1 = i >= 1 && j >= 1 && i < g.ngi - 1 && j < g.ngj - 1;
    const bool in2 = i >= 2 && j >= 2 && i < g.ngi - 2 && j < g.ngj - 2;
    float f0 = 0.f, f1 = 0.f, f2 = 0.f, d0 = 0.f, d1 = 0.f, gg = 0.f;
    if (in1) {
      const int st[2] = {1, FW2};
      double gr2 = 0.0, gp2 = 0.0, divu = 0.0, dr[2] = {0.0, 0.0}, de[2] = {0.0, 0.0};
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int sp = c + st[a], sm = c - st[a];
        const double h2 = fc.h2[a], rh2 = fc.rh2[a];
        const double *vel = a == 0 ? sU : sV;
        const double r1 = cdiv_rn(sR[sp] - sR[sm], h2, rh2);
        const double p1 = cdiv_rn(sP[sp] - sP[sm], h2, rh2);
        gr2 = gr2 + r1 * r1;
        gp2 = gp2 + p1 * p1;
        divu = divu + cdiv_rn(vel[sp] - vel[sm], h2, rh2);
        if (in2) {
          const int sp2 = c + 2 * st[a], sm2 = c - 2 * st[a];
          const double h12 = fc.h12[a], rh12 = fc.rh12[a];
          dr[a] = cdiv_rn(sR[sm2] - 8.0 * sR[sm] + 8.0 * sR[sp] - sR[sp2], h12, rh12);
          de[a] = cdiv_rn(sE[sm2] - 8.0 * sE[sm] + 8.0 * sE[sp] - sE[sp2], h12, rh12);
        }
      }
      const double rc = sR[c], pc = sP[c];
      double Cp = 0.0, Cv = 0.0;
      for (int q = 0; q < fc.ns; ++q) {
        const double Y = var[t + N * (3 + q)] / rc;
        Cp = Cp + Y * fc.cpf[q];
        Cv = Cv + Y * fc.cvf[q];
      }
      const double hmin = fc.hmin;
      const double a_snd = sqrt((Cp / Cv) * pc / rc);
      f0 = (float)(sqrt(gr2) * hmin / rc);
      f1 = (float)(sqrt(gp2) * hmin / pc);
      f2 = (float)(divu * hmin / a_snd);
      if (in2) {
        d0 = (float)(-dr[1]);
        d1 = (float)dr[0];
        const double rg2 = dr[0] * dr[0] + dr[1] * dr[1];
        const double dnednr = (dr[0] * de[0] + dr[1] * de[1]) * rg2;
        gg = dnednr < 0.0 ? (float)(sqrt(rg2) * hmin / rc) : 0.f;
        const bool interior = i >= g.ng && j >= g.ng && i < g.ng + g.nci && j < g.ng + g.ncj;
        if (interior) g1max = fmaxf(g1max, gg);
      }
    }
    rho_out[t] = sR[c];
    F[t] = f0;
    F[t + N] = f1;
    F[t + 2 * N] = f2;
    M[t] = fmaxf(fabsf(f0), fmaxf(fabsf(f1), fabsf(f2)));
    Dir[t] = d0;
    Dir[t + N] = d1;
    G1[t] = gg;
  }
  __shared__ float red[4];
  g1max = wave_maxf(g1max);
  if (threadIdx.x == 0) red[threadIdx.y] = g1max;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const float m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    if (m > 0.f) atomic_max_nonneg(&sc->maxC, (double)m);
  }
}

constexpr int ACT2_ROWS = 8;  // rows per thread of the 2D activity pass

// Launch: block (64, 4) over (i, j); each thread takes ACT2_ROWS rows of its
// column, keeping the 3-wide row maxima of rows j-1, j, j+1 in registers, and
// the block appends its active cells with one atomic (a prefix sum over the
// block gives each thread its slots).
__global__ void __launch_bounds__(256) k_active2d(const G3 g, const float *__restrict__ A, float tau,
                                                  unsigned stride, unsigned *__restrict__ list,
                                                  unsigned *__restrict__ cnt, float *__restrict__ chat) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int jb = (blockIdx.y * 4 + threadIdx.y) * ACT2_ROWS;
  const int ng = g.ng;
  const bool colin = i >= ng && i < ng + g.nci;
  auto h3 = [&](int j) {  // max of A over (i-1..i+1, j)
    float m = 0.f;
    if (colin && j >= 0 && j < g.ngj) {
      const long long t = (long long)i + g.s1 * j;
      m = fmaxf(A[t - 1], fmaxf(A[t], A[t + 1]));
    }
    return m;
  };
  unsigned bits = 0;  // active rows of this thread
  float hm = h3(jb - 1), hc = h3(jb);
#pragma unroll
  for (int r = 0; r < ACT2_ROWS; ++r) {
    const int j = jb + r;
    const float hp = h3(j + 1);
    if (i < g.ngi && j < g.ngj) {
      const long long t = (long long)i + g.s1 * j;
      if (chat) chat[t] = 0.f;
      if (colin && j >= ng && j < ng + g.ncj) {
        bool act = fmaxf(hm, fmaxf(hc, hp)) > tau;
        if (stride > 0) {
          const long long lin = (long long)(i - ng) + (long long)g.nci * (j - ng);
          act = act || (lin % stride) == 0;
        }
        if (act) bits |= 1u << r;
      }
    }
    hm = hc;
    hc = hp;
  }
  // block prefix sum of the counts
  const unsigned n = (unsigned)__popc(bits);
  const int lane = threadIdx.x, w = threadIdx.y;
  unsigned inc = n;
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  __shared__ unsigned wsum[4], base;
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const unsigned tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    base = tot ? atomicAdd(cnt, tot) : 0u;
  }
  __syncthreads();
  unsigned off = base + inc - n;
  for (int q = 0; q < w; ++q) off += wsum[q];
#pragma unroll
  for (int r = 0; r < ACT2_ROWS; ++r)
    if (bits & (1u << r)) list[off++] = (unsigned)((long long)i + g.s1 * (jb + r));
}

// Ghost layer next to each face (the cells the face interpolation reads):
// periodic image or mirror copy, as the ghost fill treats var(nv+1)
// (src/bc.cpp:33-176).  2D grids (ngk = 1) have x and y faces only.
__global__ void k_chat_ghost(const G3 g, int xper, int yper, int zper, float *__restrict__ chat) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nx = (long long)g.ncj * g.nck, ny = (long long)g.nci * g.nck;
  const long long nz = g.ngk > 1 ? (long long)g.nci * g.ncj : 0;  // 2D: no z faces
  const int ng = g.ng;
  const int k0 = g.ngk > 1 ? ng : 0;  // first interior plane (2D: the only plane)
  long long r = t;
  int face;
  if (r < 2 * nx) face = 0;
  else if ((r -= 2 * nx) < 2 * ny) face = 1;
  else if ((r -= 2 * ny) < 2 * nz) face = 2;
  else return;
  const int side = (int)(r & 1);
  r >>= 1;
  int i, j, k, si, sj, sk;
  if (face == 0) {
    j = ng + (int)(r % g.ncj); k = k0 + (int)(r / g.ncj);
    i = side ? ng + g.nci : ng - 1;
    si = xper ? (side ? ng : ng + g.nci - 1) : (side ? ng + g.nci - 1 : ng);
    sj = j; sk = k;
  } else if (face == 1) {
    i = ng + (int)(r % g.nci); k = k0 + (int)(r / g.nci);
    j = side ? ng + g.ncj : ng - 1;
    sj = yper ? (side ? ng : ng + g.ncj - 1) : (side ? ng + g.ncj - 1 : ng);
    si = i; sk = k;
  } else {
    i = ng + (int)(r % g.nci); j = ng + (int)(r / g.nci);
    k = side ? ng + g.nck : ng - 1;
    sk = zper ? (side ? ng : ng + g.nck - 1) : (side ? ng + g.nck - 1 : ng);
    si = i; sj = j;
  }
  chat[i + j * g.s1 + k * g.s2] = chat[si + sj * g.s1 + sk * g.s2];
}

__global__ void k_clear_learned(Scalars *sc, unsigned *cnt) {
  if (threadIdx.x < 4) {
    reinterpret_cast<unsigned long long *>(sc)[threadIdx.x] = 0ull;
    cnt[threadIdx.x] = 0u;
  }
}

// FP32 master weights -> BF16 inference copy (transposed, W1 padded to IN_PAD).
template <int ND>
__global__ void k_prep_infer(const float *__restrict__ W, short *__restrict__ Wb) {
  using NT = Net<ND>;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < MLP_H * NT::IN_PAD) {
    const int n = t / NT::IN_PAD, kk = t % NT::IN_PAD;
    Wb[NT::INF_W1 + t] = kk < NT::IN ? f2bf(W[NT::OFF_W1 + kk * MLP_H + n]) : (short)0;
  } else if (t < NT::INF_N) {
    const int q = t - NT::INF_W2;
    const int layer = q / (MLP_H * MLP_H), r = q % (MLP_H * MLP_H);
    const int n = r / MLP_H, kk = r % MLP_H;
    const int off = layer == 0 ? NT::OFF_W2 : NT::OFF_W3;
    Wb[t] = f2bf(W[off + kk * MLP_H + n]);
  }
}

// ---------------------------------------------------------------- inference

// Every bf16 array starts at a multiple of 16 bytes (its 8-byte fragment
// reads stay aligned): 64*388*2 (3D), 64*84*2 (2D), 64*68*2 and 16*68*2 bytes
// are multiples of 16.
template <int ND>
struct alignas(16) InfLds {
  short W1[MLP_H * W1S<ND>];  // W1^T [n][k]
  short W2[MLP_H * WHS];      // W2^T
  short W3[MLP_H * WHS];
  short act[4][16 * WHS];  // per wave: activations of 16 rows (bf16)
  float b1[MLP_H], b2[MLP_H], b3[MLP_H], w4[MLP_H];
  long long off[Net<ND>::IN_PAD];  // offset of input kk from its cell in F (channel plane + stencil)
  float b4;
};
static_assert((MLP_H * W1S<3> * 2) % 16 == 0 && (MLP_H * W1S<2> * 2) % 16 == 0 && (MLP_H * WHS * 2) % 16 == 0 &&
                  (16 * WHS * 2) % 16 == 0,
              "LDS fragment alignment");

// Offset of network input kk (channel-major, then the stencil with x
// fastest) from its cell in the feature planes.
template <int ND>
__host__ __device__ __forceinline__ long long input_offset(int kk, long long N, long long s1, long long s2) {
  using NT = Net<ND>;
  const int ch = kk / NT::STEN, r = kk % NT::STEN;
  const int dx = r % 5 - 2, dy = (r / 5) % 5 - 2, dz = ND == 3 ? r / 25 - 2 : 0;
  return N * ch + dx + dy * s1 + dz * s2;
}

// bias + GELU of the D fragments, to bf16 rows in `act` (row stride WHS)
__device__ __forceinline__ void store_act(const f32x4 acc[4], const float *b, short *act, int l) {
  const int i0 = 4 * (l >> 4), j = l & 15;
#pragma unroll
  for (int nb = 0; nb < 4; ++nb)
#pragma unroll
    for (int v = 0; v < 4; ++v) act[(i0 + v) * WHS + nb * 16 + j] = f2bf(gelu(acc[nb][v] + b[nb * 16 + j]));
}

// acc = act (16 x 64, bf16) x W (64 x 64, given as W^T rows of stride WHS)
__device__ __forceinline__ void hidden_mm(const short *act, const short *WT, f32x4 acc[4], int l) {
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < MLP_H / 16; ++kb) {
    const bf16x4 a = *reinterpret_cast<const bf16x4 *>(&act[(l & 15) * WHS + kb * 16 + 4 * (l >> 4)]);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const bf16x4 b = *reinterpret_cast<const bf16x4 *>(&WT[(nb * 16 + (l & 15)) * WHS + kb * 16 + 4 * (l >> 4)]);
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nb], 0, 0, 0);
    }
  }
}

// y of 16 rows through the network; each wave handles tiles of 16 listed
// cells.  Lanes 0..15 write C_hat = max(y, 0) rho / h (NaN for a non-finite
// y, which the residual check then rejects).
// Rho is plane 0 of P (the fused primitive planes in 3D, the rho array in 2D).
template <int ND>
__global__ void __launch_bounds__(256) k_mlp_infer(const G3 g, const float *__restrict__ F,
                                                   const unsigned *__restrict__ list, const unsigned *__restrict__ cnt,
                                                   const short *__restrict__ Wb, const float *__restrict__ W,
                                                   const float *__restrict__ G1, const double *__restrict__ P,
                                                   int pf32, double hmin, float *__restrict__ chat) {
  using NT = Net<ND>;
  constexpr int IP = NT::IN_PAD, S1 = W1S<ND>;
  __shared__ InfLds<ND> L;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  for (int q = tid; q < MLP_H * IP; q += 256) L.W1[(q / IP) * S1 + q % IP] = Wb[NT::INF_W1 + q];
  for (int q = tid; q < MLP_H * MLP_H; q += 256) {
    L.W2[(q / MLP_H) * WHS + q % MLP_H] = Wb[NT::INF_W2 + q];
    L.W3[(q / MLP_H) * WHS + q % MLP_H] = Wb[NT::INF_W3 + q];
  }
  if (tid < MLP_H) {
    L.b1[tid] = W[NT::OFF_B1 + tid];
    L.b2[tid] = W[NT::OFF_B2 + tid];
    L.b3[tid] = W[NT::OFF_B3 + tid];
    L.w4[tid] = W[NT::OFF_W4 + tid];
  }
  if (tid == 0) L.b4 = W[NT::OFF_B4];
  const long long N = g.N, s1 = g.s1, s2 = g.s2;
  for (int kk = tid; kk < IP; kk += 256) L.off[kk] = kk < NT::IN ? input_offset<ND>(kk, N, s1, s2) : 0;
  __syncthreads();
  const unsigned n = cnt[0];
  short *act = L.act[w];
  const unsigned ntiles = (n + 15) / 16;
  for (unsigned tile = blockIdx.x * 4 + w; tile < ntiles; tile += gridDim.x * 4) {
    const unsigned r = tile * 16 + (l & 15);
    // rows past the end repeat the tile's first cell (results discarded)
    const long long cell = (long long)list[r < n ? r : tile * 16];
    f32x4 acc[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    // the lane's inputs are gathered in batches of KBB k-blocks (4 KBB loads
    // in flight together), then fed to the MFMAs
    constexpr int KBB = ND == 3 ? 12 : 5;
    static_assert((IP / 16) % KBB == 0, "k-block batches");
    for (int kb0 = 0; kb0 < IP / 16; kb0 += KBB) {
      float x[KBB][4];
#pragma unroll
      for (int b = 0; b < KBB; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int kk = (kb0 + b) * 16 + 4 * (l >> 4) + q;
          x[b][q] = kk < NT::IN ? F[cell + L.off[kk]] : 0.f;
        }
#pragma unroll
      for (int b = 0; b < KBB; ++b) {
        const int kb = kb0 + b;
        bf16x4 a;
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] = f2bf(x[b][q]);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
          const bf16x4 bw =
              *reinterpret_cast<const bf16x4 *>(&L.W1[(nb * 16 + (l & 15)) * S1 + kb * 16 + 4 * (l >> 4)]);
          acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw, acc[nb], 0, 0, 0);
        }
      }
    }
    store_act(acc, L.b1, act, l);
    wave_sync();
    hidden_mm(act, L.W2, acc, l);
    wave_sync();
    store_act(acc, L.b2, act, l);
    wave_sync();
    hidden_mm(act, L.W3, acc, l);
    wave_sync();
    store_act(acc, L.b3, act, l);
    wave_sync();
    if (l < 16 && tile * 16 + l < n) {
      float y = L.b4;
      for (int q = 0; q < MLP_H; ++q) y += bf2f(act[l * WHS + q]) * L.w4[q];
      const double rho = pf32 ? prim<true>(P, g.N, 0, cell) : prim<false>(P, g.N, 0, cell);
      chat[cell] = isfinite(y) ? (float)((double)fmaxf(y, 0.f) * rho / hmin) : __builtin_nanf("");
    }
    wave_sync();
  }
}

// Residual check of the prediction against the plain method, and its
// fallback.  The relaxation of updateCeq (src/ceq3d.hpp:189-224; in 2D
// updateCeq2D, src/ceq.hpp:153-195, the same form over two directions),
//   d C_hat / dt = maxS / (eps |dx|) (G - C_hat) + kap maxS |dx| lap(C_hat),
//   lap(c) = sum_d (c[-e_d] - 2 c + c[+e_d]) / d_d,   |dx|^2 = sum_d d_d^2,
// is at rest where r = G - C_hat + a lap(C_hat) = 0 with a = kap eps |dx|^2
// (G: the contact source indicator |grad rho| = G1 rho / h).  The plain
// method's own C_hat keeps |r| within 2 % of max G once past its start-up
// transient.  A listed cell whose prediction has |r| h / rho > tol max G1 (or
// is not finite) takes the plain C-equation's value there: the local solution
// of the resting relaxation with the neighbours held,
//   C_hat = (G + a sum_d (c[-e_d] + c[+e_d]) / d_d) / (1 + 2 a sum_d 1 / d_d).
// Reads chat (ghosts filled), writes out[cell] for the listed cells (Jacobi:
// every cell sees the predictions of its neighbours); max C_hat and the
// fallback count go to sc / nfallback.
struct CheckConst {
  double a, rd[3], rden, hmin;
  float tol;
  int nd;  // directions of the Laplacian (2 or 3)
};
__global__ void __launch_bounds__(256) k_ceq_check(const G3 g, const unsigned *__restrict__ list,
                                                   const unsigned *__restrict__ cnt, const float *__restrict__ chat,
                                                   const float *__restrict__ G1, const double *__restrict__ P,
                                                   int pf32, CheckConst cc, float *__restrict__ out, Scalars *sc,
                                                   unsigned *__restrict__ nfallback) {
  const unsigned n = cnt[0];
  const float gmax = (float)as_double(sc->maxC);
  const long long st[3] = {1, g.s1, g.s2};
  const int nd = cc.nd;
  float mloc = 0.f;
  unsigned nfb = 0;
  for (unsigned r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const long long cell = (long long)list[r];
    const double rho = pf32 ? prim<true>(P, g.N, 0, cell) : prim<false>(P, g.N, 0, cell);
    const double G = (double)G1[cell] * rho / cc.hmin;
    const double c = (double)chat[cell];
    double lap = 0.0, nsum = 0.0;
    bool fin = isfinite(c);
    for (int d = 0; d < nd; ++d) {
      double m = (double)chat[cell - st[d]], p = (double)chat[cell + st[d]];
      fin = fin && isfinite(m) && isfinite(p);
      if (!isfinite(m)) m = 0.0;
      if (!isfinite(p)) p = 0.0;
      lap += (m - 2.0 * c + p) * cc.rd[d];
      nsum += (m + p) * cc.rd[d];
    }
    const double res = G - c + cc.a * lap;
    double v = c;
    if (!(fin && fabs(res) * cc.hmin / rho <= (double)cc.tol * (double)gmax)) {
      v = (G + cc.a * nsum) * cc.rden;
      ++nfb;
    }
    const float f = (float)v;
    out[cell] = f;
    mloc = fmaxf(mloc, f);
  }
  mloc = wave_maxf(mloc);
  for (int o = 32; o > 0; o >>= 1) nfb += __shfl_xor(nfb, o, 64);
  if ((threadIdx.x & 63) == 0) {
    if (mloc > 0.f) atomic_max_nonneg(&sc->maxCh, (double)mloc);
    if (nfb) atomicAdd(nfallback, nfb);
  }
}

// chat[cell] = src[cell] for the listed cells
__global__ void k_copy_listed(const unsigned *__restrict__ list, const unsigned *__restrict__ cnt,
                              const float *__restrict__ src, float *__restrict__ chat) {
  const unsigned n = cnt[0];
  for (unsigned r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const unsigned c = list[r];
    chat[c] = src[c];
  }
}

// ---------------------------------------------------------------- training

// Dataset rows: X[r][0..IN_PAD) (bf16, zeros past IN) = the feature stencil
// of cell list[r]; T[r] = var(cvar) h / rho of that cell (rho = P[cell]).
template <int ND>
__global__ void k_gather(const G3 g, const float *__restrict__ F, const unsigned *__restrict__ list, unsigned n,
                         const double *__restrict__ U, int cvar, const double *__restrict__ P, double hmin,
                         short *__restrict__ X, float *__restrict__ T) {
  using NT = Net<ND>;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * NT::IN_PAD) return;
  const unsigned r = (unsigned)(t / NT::IN_PAD);
  const int kk = (int)(t % NT::IN_PAD);
  const long long cell = list[r];
  const float x = kk < NT::IN ? F[cell + input_offset<ND>(kk, g.N, g.s1, g.s2)] : 0.f;
  X[t] = f2bf(x);
  if (kk == 0) T[r] = (float)(U[cell + g.N * cvar] * hmin / P[cell]);
}

// C[m][n] (+)= sum_k A(m,k) B(k,n); bf16 operands at arbitrary strides, one
// wave per 16x16 tile of C, FP32 accumulation on MFMA.  Each element of C is
// written by exactly one lane.
__global__ void k_gemm_bf16(int M, int Nn, int K, const short *__restrict__ A, long long am, long long ak,
                            const short *__restrict__ B, long long bk, long long bn, float *__restrict__ C,
                            long long ldc, int accumulate) {
  const int l = threadIdx.x;
  const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
  const int mi = m0 + (l & 15), nj = n0 + (l & 15);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kk = k0 + 4 * (l >> 4) + q;
      a[q] = (mi < M && kk < K) ? A[mi * am + kk * ak] : (short)0;
      b[q] = (nj < Nn && kk < K) ? B[kk * bk + nj * bn] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int mr = m0 + 4 * (l >> 4) + v;
    if (mr < M && nj < Nn) {
      float *cp = C + mr * ldc + nj;
      *cp = accumulate ? *cp + acc[v] : acc[v];
    }
  }
}

// Z += b (in place), H = bf16(gelu(Z))
__global__ void k_bias_gelu(long long n, float *__restrict__ Z, const float *__restrict__ b, short *__restrict__ H) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float z = Z[t] + b[t % MLP_H];
  Z[t] = z;
  H[t] = f2bf(gelu(z));
}

// y = H3 . w4 + b4 (same order as inference), dY = 2 (y - T) / M,
// loss += sum (y - T)^2 / M.  w4 points at the output weights, b4 = w4[64].
__global__ void k_out_loss(int M, const short *__restrict__ H, const float *__restrict__ w4, const float *__restrict__ T,
                           float *__restrict__ Y, float *__restrict__ dY, float *__restrict__ loss) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (r < M) {
    float y = w4[MLP_H];
    for (int q = 0; q < MLP_H; ++q) y += bf2f(H[(long long)r * MLP_H + q]) * w4[q];
    Y[r] = y;
    const float e = y - T[r];
    dY[r] = 2.0f * e / (float)M;
    e2 = e * e / (float)M;
  }
  for (int o = 32; o > 0; o >>= 1) e2 += __shfl_xor(e2, o, 64);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, e2);
}

// dZ3 = dY w4 gelu'(Z3)
__global__ void k_out_back(int M, const float *__restrict__ Z, const float *__restrict__ w4, const float *__restrict__ dY,
                           float *__restrict__ dZ, short *__restrict__ dZb) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M * MLP_H) return;
  const float d = dY[t / MLP_H] * w4[t % MLP_H] * gelu_grad(Z[t]);
  dZ[t] = d;
  dZb[t] = f2bf(d);
}

// dZ = dH gelu'(Z)
__global__ void k_gelu_back(long long n, const float *__restrict__ dH, const float *__restrict__ Z,
                            float *__restrict__ dZ, short *__restrict__ dZb) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float v = dH[t] * gelu_grad(Z[t]);
  dZ[t] = v;
  dZb[t] = f2bf(v);
}

// g[n] += sum_r A[r][n] * (Bv ? Bv[r] : 1), one block (4 waves) per column n;
// A is FP32 or BF16.
template <bool BF>
__global__ void k_colsum(int M, int Nn, const void *__restrict__ Av, const float *__restrict__ Bv, float *__restrict__ g) {
  const int n = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x; r < M; r += blockDim.x) {
    const long long q = (long long)r * Nn + n;
    const float a = BF ? bf2f(static_cast<const short *>(Av)[q]) : static_cast<const float *>(Av)[q];
    s += Bv ? a * Bv[r] : a;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) g[n] += (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_to_bf16(long long n, const float *__restrict__ a, short *__restrict__ b) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = f2bf(a[t]);
}

__global__ void k_fill(long long n, float *__restrict__ a, float v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// Adam with bias correction, FP32 master weights
__global__ void k_adam(long long n, float *__restrict__ w, const float *__restrict__ gr, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float g = gr[t];
  const float mm = b1 * m[t] + (1.0f - b1) * g;
  const float vv = b2 * v[t] + (1.0f - b2) * g * g;
  m[t] = mm;
  v[t] = vv;
  w[t] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

FeatConst make_feat(const fst_config &cf, const Phys &ph) {
  FeatConst fc;
  fc.d[0] = cf.dx;
  fc.d[1] = cf.dy;
  fc.d[2] = cf.dz;
  for (int a = 0; a < 3; ++a) {
    fc.h2[a] = 2.0 * fc.d[a];
    fc.rh2[a] = 1.0 / fc.h2[a];
    fc.h12[a] = 12.0 * fc.d[a];
    fc.rh12[a] = 1.0 / fc.h12[a];
  }
  fc.hmin = std::min(cf.dx, std::min(cf.dy, cf.dz));
  for (int q = 0; q < FST_MAX_SPECIES; ++q) {
    fc.cpf[q] = ph.cpf[q];
    fc.cvf[q] = ph.cvf[q];
  }
  fc.ns = cf.ns;
  return fc;
}

int features(const Geo &g, const fst_config &cf, const Phys &ph, const double *P, bool pf32, LearnedBufs &L,
             hipStream_t s) {
  const G3 q = make_g3(g);
  k_clear_learned<<<1, 64, 0, s>>>(L.sc, L.cnt);
  const long long nw = ((long long)g.ngi * g.ngj + FEAT_OUT - 1) / FEAT_OUT;  // waves over the plane
  const dim3 fg((unsigned)((nw + FEAT_WAVES - 1) / FEAT_WAVES), (g.ngk + FEAT_KCH - 1) / FEAT_KCH);
  const dim3 fb(64 * FEAT_WAVES);
  if (pf32) k_features3d<true><<<fg, fb, 0, s>>>(q, make_feat(cf, ph), P, L.F, L.M, L.dir, L.G1, L.sc);
  else k_features3d<false><<<fg, fb, 0, s>>>(q, make_feat(cf, ph), P, L.F, L.M, L.dir, L.G1, L.sc);
  return hipGetLastError() == hipSuccess ? FST_OK : FST_ERR_HIP;
}

int features2d(const Geo &g, const fst_config &cf, const Phys &ph, const double *var, double *rho_out,
               LearnedBufs &L, hipStream_t s) {
  const G3 q = make_g3(g);
  k_clear_learned<<<1, 64, 0, s>>>(L.sc, L.cnt);
  const dim3 fg((g.ngi + 63) / 64, (g.ngj + 3) / 4);
  k_features2d<<<fg, dim3(64, 4), 0, s>>>(q, make_feat(cf, ph), var, rho_out, L.F, L.M, L.dir, L.G1, L.sc);
  return hipGetLastError() == hipSuccess ? FST_OK : FST_ERR_HIP;
}

double hmin_of(const fst_config &cf) {
  return cf.ndim == 3 ? std::min(cf.dx, std::min(cf.dy, cf.dz)) : std::min(cf.dx, cf.dy);
}

// C_hat ghost layer, residual check with fallback, ghost layer again; rho is
// plane 0 of P.
int check_fields(const Geo &g, const fst_config &cf, const double *P, bool pf32, LearnedBufs &L, hipStream_t s) {
  const G3 q = make_g3(g);
  long long nf = 2LL * ((long long)g.ncj * g.nck + (long long)g.nci * g.nck);
  if (g.ndim == 3) nf += 2LL * g.nci * g.ncj;
  k_chat_ghost<<<nblk(nf), 256, 0, s>>>(q, cf.xPer, cf.yPer, cf.zPer, L.chat);
  CheckConst cc;
  const double d[3] = {cf.dx, cf.dy, cf.dz};
  cc.nd = g.ndim;
  double dx2 = 0.0, rsum = 0.0;
  for (int a = 0; a < 3; ++a) {
    cc.rd[a] = a < cc.nd ? 1.0 / d[a] : 0.0;
    if (a < cc.nd) {
      dx2 += d[a] * d[a];
      rsum += cc.rd[a];
    }
  }
  cc.a = cf.kap * cf.eps * dx2;
  cc.rden = 1.0 / (1.0 + 2.0 * cc.a * rsum);
  cc.hmin = hmin_of(cf);
  cc.tol = L.tol;
  k_ceq_check<<<1024, 256, 0, s>>>(q, L.list, L.cnt, L.chat, L.G1, P, pf32 ? 1 : 0, cc, L.chat2, L.sc, L.cnt + 1);
  k_copy_listed<<<1024, 256, 0, s>>>(L.list, L.cnt, L.chat2, L.chat);
  k_chat_ghost<<<nblk(nf), 256, 0, s>>>(q, cf.xPer, cf.yPer, cf.zPer, L.chat);
  return hipGetLastError() == hipSuccess ? FST_OK : FST_ERR_HIP;
}

}  // namespace

// ---------------------------------------------------------------- solver side

int learned_alloc(const Geo &g, LearnedBufs &L) {
  if (L.F) return FST_OK;
  L.nd = g.ndim;
  const size_t N = (size_t)g.N;
  const size_t np = (size_t)mlp_nparams(L.nd), ninf = L.nd == 2 ? Net<2>::INF_N : Net<3>::INF_N;
  bool ok = hipMalloc((void **)&L.F, N * 3 * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&L.M, N * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&L.dir, N * 3 * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&L.G1, N * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&L.chat, N * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&L.chat2, N * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&L.list, N * sizeof(unsigned)) == hipSuccess &&
            hipMalloc((void **)&L.cnt, 4 * sizeof(unsigned)) == hipSuccess &&
            hipMalloc((void **)&L.W, np * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&L.Wb, ninf * sizeof(short)) == hipSuccess &&
            hipMalloc((void **)&L.sc, sizeof(Scalars)) == hipSuccess;
  if (!ok) return FST_ERR_NOMEM;
  if (hipMemset(L.chat, 0, N * sizeof(float)) != hipSuccess || hipMemset(L.dir, 0, N * 3 * sizeof(float)) != hipSuccess ||
      hipMemset(L.sc, 0, sizeof(Scalars)) != hipSuccess || hipMemset(L.cnt, 0, 4 * sizeof(unsigned)) != hipSuccess)
    return FST_ERR_HIP;
  return FST_OK;
}

int learned_set_weights(LearnedBufs &L, const float *host, hipStream_t s) {
  if (!L.W) return FST_ERR_STATE;
  if (hipMemcpyAsync(L.W, host, (size_t)mlp_nparams(L.nd) * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess)
    return FST_ERR_HIP;
  if (L.nd == 2) k_prep_infer<2><<<nblk(Net<2>::INF_N), 256, 0, s>>>(L.W, L.Wb);
  else k_prep_infer<3><<<nblk(Net<3>::INF_N), 256, 0, s>>>(L.W, L.Wb);
  if (hipStreamSynchronize(s) != hipSuccess) return FST_ERR_HIP;
  L.have_weights = true;
  return FST_OK;
}

int learned_fields(const Geo &g, const fst_config &cf, const Phys &ph, const double *P, bool pf32, LearnedBufs &L,
                   hipStream_t s) {
  if (!L.have_weights) return FST_ERR_STATE;
  if (L.nd != 3 || g.ndim != 3) return FST_ERR_ARG;
  int rc = features(g, cf, ph, P, pf32, L, s);
  if (rc) return rc;
  const G3 q = make_g3(g);
  k_active3d<<<act3d_grid(g), 256, 0, s>>>(q, L.M, L.tau, 0u, L.list, L.cnt, L.chat);
  k_mlp_infer<3><<<1024, 256, 0, s>>>(q, L.F, L.list, L.cnt, L.Wb, L.W, L.G1, P, pf32 ? 1 : 0, hmin_of(cf), L.chat);
  return check_fields(g, cf, P, pf32, L, s);
}

int learned_fields2d(const Geo &g, const fst_config &cf, const Phys &ph, const double *var, double *rho,
                     LearnedBufs &L, hipStream_t s) {
  if (!L.have_weights) return FST_ERR_STATE;
  if (L.nd != 2 || g.ndim != 2) return FST_ERR_ARG;
  int rc = features2d(g, cf, ph, var, rho, L, s);
  if (rc) return rc;
  const G3 q = make_g3(g);
  const dim3 ag((g.ngi + 63) / 64, (g.ngj + 4 * ACT2_ROWS - 1) / (4 * ACT2_ROWS));
  k_active2d<<<ag, dim3(64, 4), 0, s>>>(q, L.G1, L.tau, 0u, L.list, L.cnt, L.chat);
  k_mlp_infer<2><<<1024, 256, 0, s>>>(q, L.F, L.list, L.cnt, L.Wb, L.W, L.G1, rho, 0, hmin_of(cf), L.chat);
  return check_fields(g, cf, rho, false, L, s);
}

void learned_free(LearnedBufs &L) {
  void *ps[] = {L.F, L.M, L.dir, L.G1, L.chat, L.chat2, L.list, L.cnt, L.W, L.Wb, L.sc};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  L = LearnedBufs();
}

// ---------------------------------------------------------------- trainer

void MlpWork::release() {
  void *ps[] = {Z1, Z2, Z3, dZ, dH, Y, dY, H1, H2, H3, dZb};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  Z1 = Z2 = Z3 = dZ = dH = Y = dY = nullptr;
  H1 = H2 = H3 = dZb = nullptr;
  cap = 0;
}

namespace {

// splitmix64
inline unsigned long long smix(unsigned long long &x) {
  unsigned long long z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline float uni(unsigned long long &x) {
  return (float)((double)(smix(x) >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
}

int ensure_work(Trainer &t, int M) {
  MlpWork &w = t.work;
  if (w.cap >= M) return FST_OK;
  w.release();
  const size_t f = (size_t)M * MLP_H * sizeof(float), h = (size_t)M * MLP_H * sizeof(short);
  bool ok = hipMalloc((void **)&w.Z1, f) == hipSuccess && hipMalloc((void **)&w.Z2, f) == hipSuccess &&
            hipMalloc((void **)&w.Z3, f) == hipSuccess && hipMalloc((void **)&w.dZ, f) == hipSuccess &&
            hipMalloc((void **)&w.dH, f) == hipSuccess && hipMalloc((void **)&w.H1, h) == hipSuccess &&
            hipMalloc((void **)&w.H2, h) == hipSuccess && hipMalloc((void **)&w.H3, h) == hipSuccess &&
            hipMalloc((void **)&w.dZb, h) == hipSuccess && hipMalloc((void **)&w.Y, (size_t)M * 4) == hipSuccess &&
            hipMalloc((void **)&w.dY, (size_t)M * 4) == hipSuccess;
  if (!ok) return FST_ERR_NOMEM;
  w.cap = M;
  return FST_OK;
}

// Forward + backward over M rows X (bf16, stride IN_PAD), targets T.
// Gradients (parameter-vector layout) accumulate into G when non-null; the
// mean squared error accumulates into *loss on the device when non-null.
template <int ND>
int fwd_bwd_nd(Trainer &t, const short *X, const float *T, int M, float *G, float *loss) {
  using NT = Net<ND>;
  int rc = ensure_work(t, M);
  if (rc) return rc;
  MlpWork &w = t.work;
  hipStream_t s = t.s;
  const float *W = t.W;
  k_to_bf16<<<nblk(NT::NPARAM), 256, 0, s>>>(NT::NPARAM, W, t.Wb);
  const short *W1 = t.Wb + NT::OFF_W1, *W2 = t.Wb + NT::OFF_W2, *W3 = t.Wb + NT::OFF_W3;
  const dim3 gm((M + 15) / 16, MLP_H / 16);
  const long long MH = (long long)M * MLP_H;
  k_gemm_bf16<<<gm, 64, 0, s>>>(M, MLP_H, NT::IN, X, NT::IN_PAD, 1, W1, MLP_H, 1, w.Z1, MLP_H, 0);
  k_bias_gelu<<<nblk(MH), 256, 0, s>>>(MH, w.Z1, W + NT::OFF_B1, w.H1);
  k_gemm_bf16<<<gm, 64, 0, s>>>(M, MLP_H, MLP_H, w.H1, MLP_H, 1, W2, MLP_H, 1, w.Z2, MLP_H, 0);
  k_bias_gelu<<<nblk(MH), 256, 0, s>>>(MH, w.Z2, W + NT::OFF_B2, w.H2);
  k_gemm_bf16<<<gm, 64, 0, s>>>(M, MLP_H, MLP_H, w.H2, MLP_H, 1, W3, MLP_H, 1, w.Z3, MLP_H, 0);
  k_bias_gelu<<<nblk(MH), 256, 0, s>>>(MH, w.Z3, W + NT::OFF_B3, w.H3);
  k_out_loss<<<nblk(M), 256, 0, s>>>(M, w.H3, W + NT::OFF_W4, T, w.Y, w.dY, loss);
  if (G) {
    // output layer: dw4 = H3^T dY, db4 = sum dY
    k_colsum<true><<<MLP_H, 256, 0, s>>>(M, MLP_H, w.H3, w.dY, G + NT::OFF_W4);
    k_colsum<false><<<1, 256, 0, s>>>(M, 1, w.dY, nullptr, G + NT::OFF_B4);
    k_out_back<<<nblk(MH), 256, 0, s>>>(M, w.Z3, W + NT::OFF_W4, w.dY, w.dZ, w.dZb);  // dZ3
    k_colsum<false><<<MLP_H, 256, 0, s>>>(M, MLP_H, w.dZ, nullptr, G + NT::OFF_B3);
    // dW3 = H2^T dZ3 ; dH2 = dZ3 W3^T
    k_gemm_bf16<<<dim3(4, 4), 64, 0, s>>>(MLP_H, MLP_H, M, w.H2, 1, MLP_H, w.dZb, MLP_H, 1, G + NT::OFF_W3, MLP_H, 1);
    k_gemm_bf16<<<gm, 64, 0, s>>>(M, MLP_H, MLP_H, w.dZb, MLP_H, 1, W3, 1, MLP_H, w.dH, MLP_H, 0);
    k_gelu_back<<<nblk(MH), 256, 0, s>>>(MH, w.dH, w.Z2, w.dZ, w.dZb);  // dZ2
    k_colsum<false><<<MLP_H, 256, 0, s>>>(M, MLP_H, w.dZ, nullptr, G + NT::OFF_B2);
    k_gemm_bf16<<<dim3(4, 4), 64, 0, s>>>(MLP_H, MLP_H, M, w.H1, 1, MLP_H, w.dZb, MLP_H, 1, G + NT::OFF_W2, MLP_H, 1);
    k_gemm_bf16<<<gm, 64, 0, s>>>(M, MLP_H, MLP_H, w.dZb, MLP_H, 1, W2, 1, MLP_H, w.dH, MLP_H, 0);
    k_gelu_back<<<nblk(MH), 256, 0, s>>>(MH, w.dH, w.Z1, w.dZ, w.dZb);  // dZ1
    k_colsum<false><<<MLP_H, 256, 0, s>>>(M, MLP_H, w.dZ, nullptr, G + NT::OFF_B1);
    // dW1 = X^T dZ1 : (IN x M)(M x 64)
    k_gemm_bf16<<<dim3((NT::IN + 15) / 16, 4), 64, 0, s>>>(NT::IN, MLP_H, M, X, 1, NT::IN_PAD, w.dZb, MLP_H, 1,
                                                          G + NT::OFF_W1, MLP_H, 1);
  }
  return hipGetLastError() == hipSuccess ? FST_OK : FST_ERR_HIP;
}

int fwd_bwd(Trainer &t, const short *X, const float *T, int M, float *G, float *loss) {
  return t.nd == 2 ? fwd_bwd_nd<2>(t, X, T, M, G, loss) : fwd_bwd_nd<3>(t, X, T, M, G, loss);
}

template <int ND>
std::vector<float> initial_weights(unsigned long long seed) {
  using NT = Net<ND>;
  // uniform(-a, a), a = sqrt(6 / fan_in) on the hidden layers, sqrt(3 /
  // fan_in) on the output layer; zero biases
  std::vector<float> w(NT::NPARAM, 0.f);
  unsigned long long x = seed;
  const float a1 = std::sqrt(6.0f / NT::IN), a2 = std::sqrt(6.0f / MLP_H), a4 = std::sqrt(3.0f / MLP_H);
  for (int q = 0; q < NT::IN * MLP_H; ++q) w[NT::OFF_W1 + q] = a1 * uni(x);
  for (int q = 0; q < MLP_H * MLP_H; ++q) w[NT::OFF_W2 + q] = a2 * uni(x);
  for (int q = 0; q < MLP_H * MLP_H; ++q) w[NT::OFF_W3 + q] = a2 * uni(x);
  for (int q = 0; q < MLP_H; ++q) w[NT::OFF_W4 + q] = a4 * uni(x);
  return w;
}

}  // namespace

int trainer_create(Trainer &t, int nd, unsigned long long seed) {
  if (nd != 2 && nd != 3) return FST_ERR_ARG;
  t.nd = nd;
  if (hipStreamCreateWithFlags(&t.s, hipStreamNonBlocking) != hipSuccess) return FST_ERR_HIP;
  const size_t np = (size_t)mlp_nparams(nd), pb = np * sizeof(float);
  bool ok = hipMalloc((void **)&t.W, pb) == hipSuccess && hipMalloc((void **)&t.G, pb) == hipSuccess &&
            hipMalloc((void **)&t.m1, pb) == hipSuccess && hipMalloc((void **)&t.m2, pb) == hipSuccess &&
            hipMalloc((void **)&t.Wb, np * sizeof(short)) == hipSuccess;
  if (!ok) return FST_ERR_NOMEM;
  const std::vector<float> w = nd == 2 ? initial_weights<2>(seed) : initial_weights<3>(seed);
  return trainer_set_weights(t, w.data());
}

void trainer_free(Trainer &t) {
  void *ps[] = {t.W, t.G, t.m1, t.m2, t.Wb, t.loss, t.X, t.T, t.P};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  t.work.release();
  learned_free(t.feat);
  if (t.s) (void)hipStreamDestroy(t.s);
  t.W = t.G = t.m1 = t.m2 = t.loss = t.T = nullptr;
  t.Wb = t.X = nullptr;
  t.P = nullptr;
  t.s = nullptr;
}

int trainer_set_weights(Trainer &t, const float *host) {
  const int np = mlp_nparams(t.nd);
  const size_t pb = (size_t)np * sizeof(float);
  if (hipMemcpyAsync(t.W, host, pb, hipMemcpyHostToDevice, t.s) != hipSuccess) return FST_ERR_HIP;
  k_fill<<<nblk(np), 256, 0, t.s>>>(np, t.m1, 0.f);
  k_fill<<<nblk(np), 256, 0, t.s>>>(np, t.m2, 0.f);
  t.adam_t = 0;
  return hipStreamSynchronize(t.s) == hipSuccess ? FST_OK : FST_ERR_HIP;
}

int trainer_get_weights(Trainer &t, float *host) {
  if (hipMemcpyAsync(host, t.W, (size_t)mlp_nparams(t.nd) * sizeof(float), hipMemcpyDeviceToHost, t.s) != hipSuccess)
    return FST_ERR_HIP;
  return hipStreamSynchronize(t.s) == hipSuccess ? FST_OK : FST_ERR_HIP;
}

namespace {

// Feature buffers of the trainer sized for g.
int feat_bufs(Trainer &t, const Geo &g) {
  if (t.feat_cells == g.N && t.feat.nd == g.ndim) return FST_OK;
  learned_free(t.feat);
  t.feat_cells = 0;
  int rc = learned_alloc(g, t.feat);
  if (rc) return rc;
  t.feat_cells = g.N;
  return FST_OK;
}

// Appends the rows of the listed cells (t.feat.list, count t.feat.cnt[0],
// features t.feat.F) with targets var(nv+1) h / rho (rho = plane 0 of P).
int append_rows(Trainer &t, const Geo &g, const fst_config &cf, const double *U, const double *P,
                long long max_rows, long long *added) {
  hipStream_t s = t.s;
  const int ip = mlp_in_pad(t.nd);
  unsigned n = 0;
  if (hipMemcpyAsync(&n, t.feat.cnt, sizeof(unsigned), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return FST_ERR_HIP;
  const long long nn = std::min<long long>(n, std::max<long long>(0, max_rows));
  if (nn == 0) return FST_OK;
  if (t.nrows + nn > t.cap) {
    const long long cap = std::max(t.nrows + nn, 2 * t.cap + 65536);
    short *X = nullptr;
    float *T = nullptr;
    if (hipMalloc((void **)&X, (size_t)cap * ip * sizeof(short)) != hipSuccess ||
        hipMalloc((void **)&T, (size_t)cap * sizeof(float)) != hipSuccess) {
      if (X) (void)hipFree(X);
      return FST_ERR_NOMEM;
    }
    if (t.nrows) {
      (void)hipMemcpyAsync(X, t.X, (size_t)t.nrows * ip * sizeof(short), hipMemcpyDeviceToDevice, s);
      (void)hipMemcpyAsync(T, t.T, (size_t)t.nrows * sizeof(float), hipMemcpyDeviceToDevice, s);
    }
    if (hipStreamSynchronize(s) != hipSuccess) return FST_ERR_HIP;
    if (t.X) (void)hipFree(t.X);
    if (t.T) (void)hipFree(t.T);
    t.X = X;
    t.T = T;
    t.cap = cap;
  }
  const G3 q = make_g3(g);
  const double hmin = hmin_of(cf);
  if (t.nd == 2)
    k_gather<2><<<nblk(nn * ip), 256, 0, s>>>(q, t.feat.F, t.feat.list, (unsigned)nn, U, g.nv + 1, P, hmin,
                                              t.X + (size_t)t.nrows * ip, t.T + t.nrows);
  else
    k_gather<3><<<nblk(nn * ip), 256, 0, s>>>(q, t.feat.F, t.feat.list, (unsigned)nn, U, g.nv + 1, P, hmin,
                                              t.X + (size_t)t.nrows * ip, t.T + t.nrows);
  if (hipStreamSynchronize(s) != hipSuccess) return FST_ERR_HIP;
  t.nrows += nn;
  *added = nn;
  return FST_OK;
}

}  // namespace

int trainer_harvest(Trainer &t, const Geo &g, const fst_config &cf, const Phys &ph, const double *U,
                    const double *P, unsigned stride, long long max_rows, long long *added) {
  *added = 0;
  if (t.nd != 3 || g.ndim != 3) return FST_ERR_ARG;
  int rc = feat_bufs(t, g);
  if (rc) return rc;
  hipStream_t s = t.s;
  rc = features(g, cf, ph, P, false, t.feat, s);  // plain FP64 contexts: FP64 primitive planes
  if (rc) return rc;
  const G3 q = make_g3(g);
  k_active3d<<<act3d_grid(g), 256, 0, s>>>(q, t.feat.M, t.feat.tau, stride, t.feat.list, t.feat.cnt, nullptr);
  return append_rows(t, g, cf, U, P, max_rows, added);
}

int trainer_harvest2d(Trainer &t, const Geo &g, const fst_config &cf, const Phys &ph, const double *U,
                      double *rho, unsigned stride, long long max_rows, long long *added) {
  *added = 0;
  if (t.nd != 2 || g.ndim != 2) return FST_ERR_ARG;
  int rc = feat_bufs(t, g);
  if (rc) return rc;
  hipStream_t s = t.s;
  rc = features2d(g, cf, ph, U, rho, t.feat, s);
  if (rc) return rc;
  const G3 q = make_g3(g);
  const dim3 ag((g.ngi + 63) / 64, (g.ngj + 4 * ACT2_ROWS - 1) / (4 * ACT2_ROWS));
  k_active2d<<<ag, dim3(64, 4), 0, s>>>(q, t.feat.G1, t.feat.tau, stride, t.feat.list, t.feat.cnt, nullptr);
  return append_rows(t, g, cf, U, rho, max_rows, added);
}

int trainer_train(Trainer &t, int epochs, int batch, float lr, unsigned long long seed, float *loss_hist) {
  if (batch < 16 || t.nrows < batch || epochs < 1) return FST_ERR_ARG;
  if (t.loss_cap < epochs) {
    if (t.loss) (void)hipFree(t.loss);
    t.loss = nullptr;
    t.loss_cap = 0;
    if (hipMalloc((void **)&t.loss, (size_t)epochs * sizeof(float)) != hipSuccess) return FST_ERR_NOMEM;
    t.loss_cap = epochs;
  }
  hipStream_t s = t.s;
  k_fill<<<nblk(epochs), 256, 0, s>>>(epochs, t.loss, 0.f);
  const long long nb = t.nrows / batch;
  std::vector<long long> order(nb);
  unsigned long long x = seed;
  const float b1 = 0.9f, b2 = 0.999f;
  const int np = mlp_nparams(t.nd), ip = mlp_in_pad(t.nd);
  for (int ep = 0; ep < epochs; ++ep) {
    for (long long q = 0; q < nb; ++q) order[q] = q;
    for (long long q = nb - 1; q > 0; --q) std::swap(order[q], order[smix(x) % (unsigned long long)(q + 1)]);
    for (long long q = 0; q < nb; ++q) {
      const long long r0 = order[q] * batch;
      k_fill<<<nblk(np), 256, 0, s>>>(np, t.G, 0.f);
      int rc = fwd_bwd(t, t.X + r0 * ip, t.T + r0, batch, t.G, t.loss + ep);
      if (rc) return rc;
      t.adam_t += 1;
      const float c1 = 1.0f - std::pow(b1, (float)t.adam_t), c2 = 1.0f - std::pow(b2, (float)t.adam_t);
      k_adam<<<nblk(np), 256, 0, s>>>(np, t.W, t.G, t.m1, t.m2, lr, b1, b2, 1e-8f, c1, c2);
    }
  }
  if (loss_hist &&
      hipMemcpyAsync(loss_hist, t.loss, (size_t)epochs * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess)
    return FST_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return FST_ERR_HIP;
  if (loss_hist)
    for (int ep = 0; ep < epochs; ++ep) loss_hist[ep] /= (float)nb;
  return FST_OK;
}

int trainer_forward_backward(Trainer &t, const float *Xh, const float *Th, int M, float *Yh, float *Gh, float *lossh) {
  if (M < 1) return FST_ERR_ARG;
  const int np = mlp_nparams(t.nd), in = mlp_in(t.nd), ip = mlp_in_pad(t.nd);
  std::vector<short> xb((size_t)M * ip, 0);
  for (long long r = 0; r < M; ++r)
    for (int k = 0; k < in; ++k) xb[r * ip + k] = f2bf(Xh[r * in + k]);
  short *X = nullptr;
  float *T = nullptr, *G = nullptr, *L = nullptr;
  hipStream_t s = t.s;
  int rc = FST_OK;
  if (hipMalloc((void **)&X, xb.size() * sizeof(short)) != hipSuccess || hipMalloc((void **)&T, (size_t)M * 4) != hipSuccess ||
      hipMalloc((void **)&G, (size_t)np * 4) != hipSuccess || hipMalloc((void **)&L, 4) != hipSuccess)
    rc = FST_ERR_NOMEM;
  if (!rc) {
    (void)hipMemcpyAsync(X, xb.data(), xb.size() * sizeof(short), hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(T, Th, (size_t)M * 4, hipMemcpyHostToDevice, s);
    k_fill<<<nblk(np), 256, 0, s>>>(np, G, 0.f);
    k_fill<<<1, 64, 0, s>>>(1, L, 0.f);
    rc = fwd_bwd(t, X, T, M, G, L);
    if (!rc) {
      if (Yh) (void)hipMemcpyAsync(Yh, t.work.Y, (size_t)M * 4, hipMemcpyDeviceToHost, s);
      if (Gh) (void)hipMemcpyAsync(Gh, G, (size_t)np * 4, hipMemcpyDeviceToHost, s);
      if (lossh) (void)hipMemcpyAsync(lossh, L, 4, hipMemcpyDeviceToHost, s);
      rc = hipStreamSynchronize(s) == hipSuccess ? FST_OK : FST_ERR_HIP;
    }
  }
  (void)hipStreamSynchronize(s);
  if (X) (void)hipFree(X);
  if (T) (void)hipFree(T);
  if (G) (void)hipFree(G);
  if (L) (void)hipFree(L);
  return rc;
}

int trainer_read_rows(Trainer &t, long long r0, long long n, float *Xh, float *Th) {
  if (r0 < 0 || n < 0 || r0 + n > t.nrows) return FST_ERR_ARG;
  if (n == 0) return FST_OK;
  const int in = mlp_in(t.nd), ip = mlp_in_pad(t.nd);
  std::vector<short> xb((size_t)n * ip);
  if (hipMemcpyAsync(xb.data(), t.X + (size_t)r0 * ip, xb.size() * sizeof(short), hipMemcpyDeviceToHost, t.s) !=
          hipSuccess ||
      (Th && hipMemcpyAsync(Th, t.T + r0, (size_t)n * 4, hipMemcpyDeviceToHost, t.s) != hipSuccess) ||
      hipStreamSynchronize(t.s) != hipSuccess)
    return FST_ERR_HIP;
  if (Xh)
    for (long long r = 0; r < n; ++r)
      for (int k = 0; k < in; ++k) {
        const unsigned u = ((unsigned)(unsigned short)xb[r * ip + k]) << 16;
        std::memcpy(&Xh[r * in + k], &u, 4);
      }
  return FST_OK;
}

}  // namespace fst
