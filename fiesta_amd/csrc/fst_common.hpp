// Shared device-side definitions of the FIESTA RHS/RK library (gfx950).
//
// Layout: every array is a stack of (ngi x ngj x ngk) planes, i fastest
// (LayoutLeft, the reference GPU layout of src/kokkosTypes.hpp:40).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fiesta_hip.h"

namespace fst {

struct Geo {
  int ngi, ngj, ngk, ng;
  int nci, ncj, nck, ndim;
  int nv, nvt, ns, nvx;
  long long N;  // cells including ghosts
  __host__ __device__ __forceinline__ long long id(int i, int j, int k) const {
    return (long long)i + (long long)ngi * ((long long)j + (long long)ngj * (long long)k);
  }
};

// Constants the kernels read.  cd[] is the reference device config array
// [ns, dx, dy, dz, nv, ng, (gamma, R/M, mu) x ns] (src/cart3d.cpp:152-171);
// cpf/cvf are gamma*Rs/(gamma-1) and Rs/(gamma-1), the per-species factors of
// src/secondary.hpp:136-138 evaluated once on the host with the same operations.
struct Phys {
  double cd[6 + 3 * FST_MAX_SPECIES];
  double cpf[FST_MAX_SPECIES], cvf[FST_MAX_SPECIES], mus[FST_MAX_SPECIES];
  double dx, dy, dz;     // cf.dx, cf.dy, cf.dz
  double dxmag2;         // dx*dx+dy*dy+dz*dz (src/cart3d.cpp:173)
  double kap, eps, alpha, g, rhoRef, dt;
};

// Device-side scalars written by the reduction kernels (bit patterns of
// non-negative doubles order like unsigned integers).
struct Scalars {
  unsigned long long maxS, maxC, maxCh, pad;
};

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *a, double v) {
  atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

__host__ __device__ __forceinline__ double as_double(unsigned long long b) {
  return __builtin_bit_cast(double, b);
}

// a/b for a run constant b with r = RN(1/b) precomputed: q0 = a*r,
// e = fma(-b, q0, a), q = fma(e, r, q0) is the correctly rounded quotient
// (Markstein's theorem), so it equals IEEE a/b bit for bit.
__host__ __device__ __forceinline__ double cdiv_rn(double a, double b, double rb) {
  const double q0 = a * rb;
  const double e = fma(-b, q0, a);
  return fma(e, rb, q0);
}

// Primitive planes q = 0 rho, 1 p, 2 u, 3 v, 4 w, 5 nrg, 6.. Y_s of cell t
// (fused path).  FP64 form: doubles, plane q at P + N*q.  FP32 forms: p as
// doubles (the pressure gradient and buoyancy keep their hydrostatic
// cancellation in FP64), then the other planes as floats in the order
// rho, u, v, w, nrg, Y_s.
template <bool PF>
__device__ __forceinline__ double prim(const double *__restrict__ P, long long N, int q, long long t) {
  if (!PF) return P[t + N * q];
  if (q == 1) return P[t];
  return (double)reinterpret_cast<const float *>(P + N)[t + N * (q == 0 ? 0 : q - 1)];
}

// Wave-wide max over 64 lanes.
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Maxima of non-negative values over a 1D block (blockDim.x a multiple of 64,
// at most 1024): one atomic per block and value, none when the block maximum
// is 0 (the targets are cleared to 0 before the pass).  Every thread of the
// block must call it.
template <int NV>
__device__ __forceinline__ void block_max_atomic(double (&v)[NV], unsigned long long *const (&dst)[NV]) {
  __shared__ double red_[NV][16];
  __syncthreads();  // earlier readers of red_ are done
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double m = wave_max(v[q]);
    if ((threadIdx.x & 63) == 0) red_[q][w] = m;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int q = threadIdx.x;
    double m = 0.0;
    for (int u = 0; u < (int)(blockDim.x >> 6); ++u) m = fmax(m, red_[q][u]);
    if (m > 0.0) atomic_max_nonneg(dst[q], m);
  }
}

}  // namespace fst
