// Learned C-indicator: an MLP that stands in for the relaxation solve of the
// C-equation (src/ceq3d.hpp:189-224), its features, inference and training.
#pragma once
#include <string>

#include "fst_common.hpp"

namespace fst {

// Network: a 5^ND stencil of 3 features (375 inputs in 3D, 75 in 2D, padded
// to a multiple of 16 for the 16-deep MFMA k-steps) -> 64 -> 64 -> 64 -> 1,
// GELU on the hidden layers.  FP32 parameter vector: W1[IN][64], b1[64],
// W2[64][64], b2, W3[64][64], b3, w4[64], b4.
template <int ND>
struct Net {
  static constexpr int NF = 3;
  static constexpr int STEN = ND == 3 ? 125 : 25;
  static constexpr int IN = NF * STEN;
  static constexpr int IN_PAD = (IN + 15) / 16 * 16;
  static constexpr int H = 64;
  static constexpr int OFF_W1 = 0;
  static constexpr int OFF_B1 = OFF_W1 + IN * H;
  static constexpr int OFF_W2 = OFF_B1 + H;
  static constexpr int OFF_B2 = OFF_W2 + H * H;
  static constexpr int OFF_W3 = OFF_B2 + H;
  static constexpr int OFF_B3 = OFF_W3 + H * H;
  static constexpr int OFF_W4 = OFF_B3 + H;
  static constexpr int OFF_B4 = OFF_W4 + H;
  static constexpr int NPARAM = OFF_B4 + 1;
  // BF16 inference copy: W1^T [64][IN_PAD] (zero columns k >= IN), W2^T, W3^T
  static constexpr int INF_W1 = 0;
  static constexpr int INF_W2 = H * IN_PAD;
  static constexpr int INF_W3 = INF_W2 + H * H;
  static constexpr int INF_N = INF_W3 + H * H;
};
constexpr int MLP_NF = 3;
constexpr int MLP_H = 64;
constexpr int MLP_STEN = Net<3>::STEN;
constexpr int MLP_IN = Net<3>::IN;
constexpr int MLP_IN_PAD = Net<3>::IN_PAD;
constexpr int MLP_NPARAM = Net<3>::NPARAM;
static_assert(Net<3>::IN_PAD == 384 && Net<2>::IN_PAD == 80, "padded input widths");
// parameter count / padded input width / input width of the ND network
inline int mlp_nparams(int nd) { return nd == 2 ? Net<2>::NPARAM : Net<3>::NPARAM; }
inline int mlp_in(int nd) { return nd == 2 ? Net<2>::IN : Net<3>::IN; }
inline int mlp_in_pad(int nd) { return nd == 2 ? Net<2>::IN_PAD : Net<3>::IN_PAD; }

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
  int nd = 3;             // network / grid dimension
  float *F = nullptr;     // 3 feature planes
  float *M = nullptr;     // per-cell max |feature|
  float *dir = nullptr;   // 3D: 4th-order grad rho (the direction n); 2D: the
                          // tangent (-d rho/dy, d rho/dx), 4th order (2 of 3 planes)
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
// 2D (cart2d): features of var (ghosts filled; the pass computes the
// primitives itself and writes rho, which the network and the check read,
// to `rho`), then C_hat as above.
int learned_fields2d(const Geo &g, const fst_config &cf, const Phys &ph, const double *var, double *rho,
                     LearnedBufs &L, hipStream_t s);
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
  int nd = 3;  // network dimension (5^nd stencils)
  hipStream_t s = nullptr;
  float *W = nullptr, *G = nullptr, *m1 = nullptr, *m2 = nullptr;
  short *Wb = nullptr;
  float *loss = nullptr;  // device loss accumulators
  int loss_cap = 0;
  short *X = nullptr;  // rows x IN_PAD (bf16)
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

int trainer_create(Trainer &t, int nd, unsigned long long seed);
void trainer_free(Trainer &t);
int trainer_set_weights(Trainer &t, const float *host);
int trainer_get_weights(Trainer &t, float *host);
// Rows from the state U of a plain (C-equation) solver with primitives P:
// every active cell plus every stride-th interior cell; targets
// var(nv+1) * hmin / rho.  3D: P = the fused primitive planes; 2D: the
// feature pass writes rho into `rho`.
int trainer_harvest(Trainer &t, const Geo &g, const fst_config &cf, const Phys &ph, const double *U,
                    const double *P, unsigned stride, long long max_rows, long long *added);
int trainer_harvest2d(Trainer &t, const Geo &g, const fst_config &cf, const Phys &ph, const double *U,
                      double *rho, unsigned stride, long long max_rows, long long *added);
int trainer_train(Trainer &t, int epochs, int batch, float lr, unsigned long long seed, float *loss_hist);
int trainer_forward_backward(Trainer &t, const float *X, const float *T, int M, float *Y, float *grad, float *loss);
int trainer_read_rows(Trainer &t, long long r0, long long n, float *X, float *T);

}  // namespace fst
