// Learned C-indicator: an MLP that stands in for the relaxation solve of the
// C-equation (src/ceq3d.hpp:189-224), its features, inference and training.
#pragma once
#include <string>

#include "fst_common.hpp"

namespace fst {

// Network: 5x5x5 stencil of 3 features (375 inputs, padded to 384 for the
// 16-deep MFMA k-steps) -> 64 -> 64 -> 64 -> 1, GELU on the hidden layers.
constexpr int MLP_NF = 3;
constexpr int MLP_STEN = 125;
constexpr int MLP_IN = MLP_NF * MLP_STEN;
constexpr int MLP_IN_PAD = 384;
constexpr int MLP_H = 64;
// FP32 parameter vector: W1[375][64], b1[64], W2[64][64], b2, W3[64][64], b3, w4[64], b4
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H * MLP_H;
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + MLP_H;
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B4 + 1;

// Constants of the feature pass.
struct FeatConst {
  double d[3];     // dx, dy, dz
  double h2[3], rh2[3], h12[3], rh12[3];  // 2d, 12d and their RN reciprocals
  double hmin;     // min(dx, dy, dz): length scale of the dimensionless features
  double cpf[FST_MAX_SPECIES], cvf[FST_MAX_SPECIES];
  int ns;
};

// Learned-C fields of one solver context.
struct LearnedBufs {
  float *F = nullptr;     // 3 feature planes
  float *M = nullptr;     // per-cell max |feature|
  float *dir = nullptr;   // 3 planes: 4th-order grad rho (the direction n)
  float *G1 = nullptr;    // contact-switch source, dimensionless
  float *chat = nullptr;  // C_hat, dimensional (units of |grad rho|)
  float *chat2 = nullptr; // the residual check's output for the listed cells
  unsigned *list = nullptr;
  unsigned *cnt = nullptr;  // [0] active cells, [1] fallbacks
  float *W = nullptr;       // FP32 parameters
  short *Wb = nullptr;      // BF16 copy used by inference
  Scalars *sc = nullptr;    // maxCh (max C_hat), maxC (max G1 as double)
  float tau = 1.0e-3f;      // activity threshold on the features
  float tol = 0.05f;        // residual-check tolerance (fraction of max G1)
  bool have_weights = false;
};

int learned_alloc(const Geo &g, LearnedBufs &L);
// Features of the primitive planes P (rho, p, u, v, w, nrg, Y_s; ghosts
// filled), then C_hat on the active cells by the MLP, residual check and
// fallback, ghost layer of C_hat; maxCh into L.sc.  Stream-ordered.
int learned_fields(const Geo &g, const fst_config &cf, const Phys &ph, const double *P, bool pf32, LearnedBufs &L,
                   hipStream_t s);
int learned_set_weights(LearnedBufs &L, const float *host, hipStream_t s);
void learned_free(LearnedBufs &L);

// Trainer: dataset of (stencil, target) rows harvested on the device and the
// optimizer state.
struct MlpWork {
  int cap = 0;
  float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *dZ = nullptr, *dH = nullptr, *Y = nullptr, *dY = nullptr;
  short *H1 = nullptr, *H2 = nullptr, *H3 = nullptr, *dZb = nullptr;
  void release();
};

struct Trainer {
  hipStream_t s = nullptr;
  float *W = nullptr, *G = nullptr, *m1 = nullptr, *m2 = nullptr;
  short *Wb = nullptr;
  float *loss = nullptr;  // device loss accumulators
  int loss_cap = 0;
  short *X = nullptr;  // rows x MLP_IN_PAD (bf16)
  float *T = nullptr;  // targets
  long long nrows = 0, cap = 0;
  long long adam_t = 0;
  MlpWork work;
  // harvest scratch
  double *P = nullptr;
  long long Pcells = 0;
  int Pplanes = 0;
  LearnedBufs feat;
  long long feat_cells = 0;
  std::string err;
};

int trainer_create(Trainer &t, unsigned long long seed);
void trainer_free(Trainer &t);
int trainer_set_weights(Trainer &t, const float *host);
int trainer_get_weights(Trainer &t, float *host);
// Rows from the state U of a plain (C-equation) solver with primitives P:
// every active cell plus every stride-th interior cell; targets
// var(nv+1) * hmin / rho.
int trainer_harvest(Trainer &t, const Geo &g, const fst_config &cf, const Phys &ph, const double *U,
                    const double *P, unsigned stride, long long max_rows, long long *added);
int trainer_train(Trainer &t, int epochs, int batch, float lr, unsigned long long seed, float *loss_hist);
int trainer_forward_backward(Trainer &t, const float *X, const float *T, int M, float *Y, float *grad, float *loss);
int trainer_read_rows(Trainer &t, long long r0, long long n, float *X, float *T);

}  // namespace fst
