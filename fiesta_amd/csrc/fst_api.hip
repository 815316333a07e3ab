// C-ABI layer (include/fiesta_hip.h): context lifetime, host<->device copies
// and the rk_func method sequence, on one HIP stream per context.
#include <cstdio>
#include <cstring>
#include <string>

#include "fst_common.hpp"
#include "fst_fused.hpp"
#include "fst_kernels.hpp"
#include "fst_mlp.hpp"

using namespace fst;

struct fst_ctx {
  fst_config cf;
  Geo g;
  Phys ph;
  Bufs b;
  FusedBufs fb;
  LearnedBufs lb;         // learned C-indicator fields (cf.learned_c)
  double *lprim = nullptr;  // primitive planes for fst_learned_compute
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  int fused = 1;
  double last_ms = 0.0;
  int device = 0;
  Prof prof;
  double phase_ms[Prof::MAXPER] = {0};  // per-phase sums of the last fst_profile_read
  int nphase = 0;
  void *pack = nullptr;  // device staging of fst_pack
  size_t pack_bytes = 0;
};

namespace {

int fail(fst_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(ctx, FST_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

int check_launch(fst_ctx *ctx, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, FST_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return FST_OK;
}

int derive(const fst_config &c, fst_dims *d, std::string *why) {
  if (c.ndim != 2 && c.ndim != 3) { *why = "ndim must be 2 or 3"; return FST_ERR_ARG; }
  if (c.ns < 1 || c.ns > FST_MAX_SPECIES) { *why = "ns out of range"; return FST_ERR_ARG; }
  if (c.ng < 3) { *why = "ng must be >= 3 (WENO5 reaches 3 cells)"; return FST_ERR_ARG; }
  if (c.nci < 1 || c.ncj < 1 || (c.ndim == 3 && c.nck < 1)) { *why = "empty grid"; return FST_ERR_ARG; }
  if ((c.xPer && c.nci < c.ng) || (c.yPer && c.ncj < c.ng) || (c.ndim == 3 && c.zPer && c.nck < c.ng)) {
    *why = "a periodic direction needs at least ng interior cells";
    return FST_ERR_ARG;
  }
  if (c.grid != 0 && c.grid != 1) {
    *why = "grid: 0 (cartesian) or 1 (generalized); the terrain grid (2) has no module in the reference";
    return FST_ERR_ARG;
  }
  if (c.precision != FST_FP64 && c.precision != FST_FP32_BULK && c.precision != FST_FP64_FAST) {
    *why = "precision: 0 (FP64), 1 (FP32 bulk, FP64 accumulation) or 2 (fast FP64)";
    return FST_ERR_ARG;
  }
  if (c.precision == FST_FP64_FAST && c.learned_c) {
    *why = "the learned C-indicator runs in the FP64 (0) or FP32 bulk (1) forms";
    return FST_ERR_ARG;
  }
  if (c.grid == 1 && (c.learned_c || c.precision == FST_FP64_FAST)) {
    *why = "generalized grids run in FP64 (0) or, on the fused gen3d stage, FP32 bulk (1); no learned C-indicator";
    return FST_ERR_ARG;
  }
  if (c.grid == 1 && c.precision == FST_FP32_BULK) {
    // the fused gen3d stage's conditions (gen_adv_supported): 3D, WENO5,
    // ns <= 3, every array (nvt or the 5 primitive planes) below 2^32 elements
    const double N = (double)(c.nci + 2 * c.ng) * (c.ncj + 2 * c.ng) * (c.nck + 2 * c.ng);
    if (!(c.ndim == 3 && c.scheme == FST_WENO5 && c.ns >= 1 && c.ns <= 3 &&
          N * std::max(4 + c.ns + (c.ceq ? 5 : 0), 5) < 4294967296.0)) {
      *why = "the FP32 form on generalized grids runs on the fused gen3d stage only (3D, WENO5, ns <= 3, every "
             "array below 2^32 elements)";
      return FST_ERR_ARG;
    }
  }
  if (c.gen_ext && c.grid != 1) {
    *why = "gen_ext applies to generalized grids (grid = 1)";
    return FST_ERR_ARG;
  }
  if (c.gen_ext) {
    const int bcs[6] = {c.bcL, c.bcR, c.bcB, c.bcT, c.bcH, c.bcF};
    for (int q = 0; q < 6; ++q)
      if (bcs[q] == 3) {
        *why = "gen_ext: the hydrostatic ghost fill needs the physical cell size, which the index space of a "
               "generalized grid does not give";
        return FST_ERR_ARG;
      }
  }
  if (c.learned_c && !c.ceq) {
    *why = "the learned C-indicator replaces the C-equation's relaxation solve: needs ceq = 1";
    return FST_ERR_ARG;
  }
  if (c.learned_c && c.ndim == 2 && c.precision != FST_FP64) {
    *why = "the 2D learned C-indicator runs on the FP64 cart2d kernel sequence (precision 0)";
    return FST_ERR_ARG;
  }
  if (c.learned_c && c.noise && c.ceq) {
    *why = "the noise filter's C_hat cut-off reads the C-equation variables, which the learned form does not store";
    return FST_ERR_ARG;
  }
  if (c.scheme < 1 || c.scheme > 3 || (c.ndim == 3 && c.scheme != FST_WENO5)) {
    *why = "advection scheme: 3D supports weno5; 2D weno5, centered4, quick";
    return FST_ERR_ARG;
  }
  if (c.noise && c.n_nt < 0) {
    *why = "noise.nt must be >= 0";
    return FST_ERR_ARG;
  }
  if (c.rk_order != 0 && c.rk_order != 2 && c.rk_order != 3) {
    *why = "rk_order: 2 (RK2 midpoint, the reference's rkAdvance; 0 means 2) or 3 (SSP-RK3)";
    return FST_ERR_ARG;
  }
  const int bcs[6] = {c.bcL, c.bcR, c.bcB, c.bcT, c.bcH, c.bcF};
  for (int q = 0; q < 6; ++q)
    if (bcs[q] < 0 || bcs[q] > 3) { *why = "unknown boundary type"; return FST_ERR_ARG; }
  if (c.precision != FST_FP64 && c.grid == 0 && !fused_supported(c)) {
    *why = "the FP32 and fast FP64 forms run on the fused 3D stage kernels only (cartesian, WENO5, ns <= 3, "
           "every array below 2^32 elements)";
    return FST_ERR_ARG;
  }
  d->ngi = c.nci + 2 * c.ng;
  d->ngj = c.ncj + 2 * c.ng;
  d->ngk = c.ndim == 3 ? c.nck + 2 * c.ng : 1;
  d->nv = c.ndim + 1 + c.ns;  // src/input.cpp:354-364
  d->nvt = d->nv + (c.ceq && !c.learned_c ? 5 : 0);  // learned form: no C-equation variables
  d->nvx = c.ndim + 3;
  d->ncells_ghost = (int64_t)d->ngi * d->ngj * d->ngk;
  d->ncells = (int64_t)c.nci * c.ncj * (c.ndim == 3 ? c.nck : 1);
  return FST_OK;
}

// The noise filter runs in the cartesian modules' and gen2d's postStep
// (src/cart3d.cpp:319, src/cart2d.cpp:257, src/gen2d.cpp:150); gen3d has none.
bool noise_on(const fst_config &c) { return c.noise && !(c.grid == 1 && c.ndim == 3); }

// grid nodes (ni*nj*nk): one more than cells per direction, nk = 1 in 2D
// (src/input.cpp:401-406)
size_t grid_nodes(const Geo &g) {
  return (size_t)(g.nci + 1) * (size_t)(g.ncj + 1) * (size_t)(g.ndim == 3 ? g.nck + 1 : 1);
}

double *array_ptr(fst_ctx *ctx, int a, size_t *count) {
  const size_t N = (size_t)ctx->g.N;
  double *p = nullptr;
  size_t n = 0;
  switch (a) {
    case FST_VAR: p = ctx->b.var; n = N * ctx->g.nvt; break;
    case FST_TMP1: p = ctx->b.tmp1; n = N * ctx->g.nvt; break;
    case FST_DVAR: p = ctx->b.dvar; n = N * ctx->g.nvt; break;
    case FST_VARX: p = ctx->b.varx; n = N * ctx->g.nvx; break;
    case FST_P: p = ctx->b.p; n = N; break;
    case FST_RHO: p = ctx->b.rho; n = N; break;
    case FST_T: p = ctx->b.T; n = N; break;
    case FST_VEL: p = ctx->b.vel; n = N * ctx->g.ndim; break;
    case FST_GRADRHO: p = ctx->b.gradRho; n = ctx->b.gradRho ? N * (ctx->g.ndim == 3 ? 5 : 4) : 0; break;
    case FST_GRID: p = ctx->b.grid; n = ctx->b.grid ? grid_nodes(ctx->g) * 3 : 0; break;
    case FST_METRICS: p = ctx->b.metrics; n = ctx->b.metrics ? N * ctx->g.ndim * ctx->g.ndim : 0; break;
    default: return nullptr;
  }
  *count = n;
  return p;
}

void free_all(fst_ctx *ctx) {
  if (ctx->b.prim) {  // rho, p and vel live in this block
    (void)hipFree(ctx->b.prim);
    ctx->b.prim = ctx->b.rho = ctx->b.p = ctx->b.vel = nullptr;
  }
  double **ps[] = {&ctx->b.var, &ctx->b.tmp1, &ctx->b.dvar, &ctx->b.varx, &ctx->b.p, &ctx->b.rho, &ctx->b.T,
                   &ctx->b.vel, &ctx->b.fx, &ctx->b.fy, &ctx->b.fz, &ctx->b.stress, &ctx->b.gradRho,
                   &ctx->b.mflux, &ctx->b.grid, &ctx->b.metrics, &ctx->b.ndn};
  for (auto pp : ps)
    if (*pp) { (void)hipFree(*pp); *pp = nullptr; }
  if (ctx->b.nflag) { (void)hipFree(ctx->b.nflag); ctx->b.nflag = nullptr; }
  if (ctx->b.nmax) { (void)hipFree(ctx->b.nmax); ctx->b.nmax = nullptr; }
  if (ctx->b.mm) { (void)hipFree(ctx->b.mm); ctx->b.mm = nullptr; }
  if (ctx->b.sc) { (void)hipFree(ctx->b.sc); ctx->b.sc = nullptr; }
  fused_free(ctx->fb);
  learned_free(ctx->lb);
  if (ctx->lprim) { (void)hipFree(ctx->lprim); ctx->lprim = nullptr; }
  if (ctx->pack) { (void)hipFree(ctx->pack); ctx->pack = nullptr; ctx->pack_bytes = 0; }
}

// rk_func::compute() of the module the config selects: cart3d applies the
// boundary conditions first (src/cart3d.cpp:221); cart2d applies them in
// preStep; gen2d/gen3d never do (src/gen2d.cpp:100-131, src/gen3d.cpp:341-369).
// The 2D learned form evaluates its fields when `fields` is set (stage 1 of
// a step) and uses the last ones otherwise.
int compute_seq(fst_ctx *ctx, bool fields = false) {
  if (ctx->cf.grid == 1) {
    // gen_ext: the ghost fill cart3d applies before compute (src/cart3d.cpp:221)
    if (ctx->cf.gen_ext && ctx->g.ndim == 3) launch_apply_bcs(ctx->g, ctx->cf, ctx->b.var, ctx->stream);
    launch_compute_gen(ctx->g, ctx->ph, ctx->cf, ctx->b, ctx->stream);
    return FST_OK;
  }
  if (ctx->g.ndim == 3) launch_apply_bcs(ctx->g, ctx->cf, ctx->b.var, ctx->stream);
  if (ctx->cf.learned_c && ctx->g.ndim == 2) {
    const Bufs &b = ctx->b;
    launch_eos_vel(ctx->g, ctx->ph, b, ctx->stream);
    if (fields) {
      const int rc = learned_fields2d(ctx->g, ctx->cf, ctx->ph, b.var, b.rho, ctx->lb, ctx->stream);
      if (rc) return rc;
    }
    Ceq2dFields lf;
    lf.chat = ctx->lb.chat;
    lf.dir = ctx->lb.dir;
    lf.sc = ctx->lb.sc;
    launch_compute_ref(ctx->g, ctx->ph, ctx->cf, b, ctx->stream, &lf, true);
    return FST_OK;
  }
  launch_compute_ref(ctx->g, ctx->ph, ctx->cf, ctx->b, ctx->stream);
  return FST_OK;
}

// gen3d on the fused kernels (fused_gen_steps): FP64 or FP32 bulk, WENO5, ns <= 3
bool gen_fused(const fst_config &c, const Geo &g) {
  return c.grid == 1 && c.ndim == 3 && (c.precision == FST_FP64 || c.precision == FST_FP32_BULK) &&
         gen_adv_supported(c, g);
}

int learned_only(fst_ctx *ctx, const char *what) {
  return fail(ctx, FST_ERR_ARG, std::string(what) + ": the learned C-indicator runs through fst_step only");
}

// The FP32 and fast FP64 forms exist only in the fused stage kernels; the
// per-kernel sequence computes the reference arithmetic whatever the
// precision field says.
int fp32_only(fst_ctx *ctx, const char *what) {
  return fail(ctx, FST_ERR_ARG, std::string(what) + ": the FP32 and fast FP64 forms run through fst_step only");
}

}  // namespace

namespace fst {

int module_check(const fst_config &in, fst_dims *d, std::string *why) { return derive(in, d, why); }

void module_setup(const fst_config &in, const fst_dims &d, fst_config *cf, Geo *g, Phys *ph) {
  *cf = in;
  fst_config &c = *cf;
  if (in.grid == 1) c.dx = c.dy = c.dz = 1.0;  // src/input.cpp:344-350
  if (in.ndim == 2) {  // 2D defaults, src/input.cpp:354-362
    c.dz = c.dx;
    c.nck = 1;
    c.zPer = 0;
  }
  g->ngi = d.ngi; g->ngj = d.ngj; g->ngk = d.ngk; g->ng = c.ng;
  g->nci = c.nci; g->ncj = c.ncj; g->nck = c.ndim == 3 ? c.nck : 1; g->ndim = c.ndim;
  g->nv = d.nv; g->nvt = d.nvt; g->ns = c.ns; g->nvx = d.nvx;
  g->N = d.ncells_ghost;
  std::memset(ph, 0, sizeof(*ph));
  // device config array, src/cart3d.cpp:152-171 / src/cart2d.cpp:92-110
  ph->cd[0] = c.ns; ph->cd[1] = c.dx; ph->cd[2] = c.dy; ph->cd[3] = c.dz; ph->cd[4] = d.nv; ph->cd[5] = c.ng;
  for (int s = 0; s < c.ns; ++s) {
    double gammas = c.gamma[s], Rs = c.R / c.M[s];
    ph->cd[6 + 3 * s] = gammas;
    ph->cd[6 + 3 * s + 1] = Rs;
    ph->cd[6 + 3 * s + 2] = c.mu[s];
    ph->cpf[s] = gammas * Rs / (gammas - 1);
    ph->cvf[s] = Rs / (gammas - 1);
    ph->mus[s] = c.mu[s];
  }
  ph->dx = c.dx; ph->dy = c.dy; ph->dz = c.dz;
  ph->dxmag2 = c.dx * c.dx + c.dy * c.dy + c.dz * c.dz;
  ph->kap = c.kap; ph->eps = c.eps; ph->alpha = c.alpha; ph->g = c.gAccel; ph->rhoRef = c.rhoRef; ph->dt = c.dt;
}

}  // namespace fst

extern "C" {

int fst_abi_version(void) { return FST_ABI_VERSION; }

// Defaults of executeConfiguration, src/input.cpp:130-443.
void fst_config_init(fst_config *c) {
  std::memset(c, 0, sizeof(*c));
  c->ndim = 2;
  c->grid = 0;
  c->ns = 1;
  c->ng = 3;
  c->nci = c->ncj = c->nck = 1;
  c->dx = c->dy = c->dz = 1.0;
  c->R = 8.314462;
  c->scheme = FST_WENO5;
  c->kap = 1.23456;
  c->gAccel = 9.81;
  c->rhoRef = 0.0;
  c->gamma[0] = 1.4;
  c->M[0] = 0.02896;
}

int fst_config_dims(const fst_config *cfg, fst_dims *out) {
  if (!cfg || !out) return FST_ERR_ARG;
  std::string why;
  return derive(*cfg, out, &why);
}

int fst_config_check(const fst_config *cfg, char *why, size_t len) {
  if (!cfg) return FST_ERR_ARG;
  fst_dims d;
  std::string w;
  const int rc = derive(*cfg, &d, &w);
  if (why && len) std::snprintf(why, len, "%s", w.c_str());
  return rc;
}

int fst_create(const fst_config *cfg, fst_ctx **out) {
  if (!cfg || !out) return FST_ERR_ARG;
  *out = nullptr;
  fst_dims d;
  std::string why;
  int rc = derive(*cfg, &d, &why);
  if (rc) {
    fprintf(stderr, "fst_create: %s\n", why.c_str());
    return rc;
  }
  fst_ctx *ctx = new fst_ctx();
  module_setup(*cfg, d, &ctx->cf, &ctx->g, &ctx->ph);
  const fst_config &c = ctx->cf;
  const Geo &g = ctx->g;

  auto alloc = [&](double **p, size_t planes) -> bool {
    if (planes == 0) return true;
    size_t bytes = planes * (size_t)g.N * sizeof(double);
    if (hipMalloc((void **)p, bytes) != hipSuccess) return false;
    return hipMemset(*p, 0, bytes) == hipSuccess;
  };
  bool ok = hipGetDevice(&ctx->device) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipEventCreate(&ctx->ev0) == hipSuccess && hipEventCreate(&ctx->ev1) == hipSuccess;
  ok = ok && alloc(&ctx->b.var, g.nvt) && alloc(&ctx->b.tmp1, g.nvt) && alloc(&ctx->b.dvar, g.nvt);
  if (gen_adv_supported(c, g)) {  // gen3d: the fused advection reads rho, p, vel as one block of planes
    ok = ok && alloc(&ctx->b.prim, 5);
    if (ok) {
      ctx->b.rho = ctx->b.prim;
      ctx->b.p = ctx->b.prim + g.N;
      ctx->b.vel = ctx->b.prim + 2 * g.N;
    }
  } else {
    ok = ok && alloc(&ctx->b.p, 1) && alloc(&ctx->b.rho, 1) && alloc(&ctx->b.vel, g.ndim);
  }
  ok = ok && alloc(&ctx->b.varx, g.nvx) && alloc(&ctx->b.T, 1);
  ok = ok && alloc(&ctx->b.fx, 1) && alloc(&ctx->b.fy, 1);
  ok = ok && alloc(&ctx->b.fz, g.ndim == 3 ? 1 : 0);
  // 3D: 6 stress planes; the gen_ext viscous terms keep the mass fractions there
  ok = ok && alloc(&ctx->b.stress, c.visc ? (g.ndim == 3 ? std::max(6, c.grid == 1 ? c.ns : 0) : 8) : 0);
  const bool ceq_vars = c.ceq && !c.learned_c;
  ok = ok && alloc(&ctx->b.gradRho, ceq_vars ? (g.ndim == 3 ? 5 : 4) : 0);
  // mFlux: the reference sequence's C-equation faces (the 2D learned form too)
  ok = ok && alloc(&ctx->b.mflux, c.ceq && (ceq_vars || g.ndim == 2) ? (g.ndim == 3 ? 27 : 8) : 0);
  if (ok && c.grid == 1) {  // gen3d.cpp:41-43, gen2d.cpp:25-26
    ok = alloc(&ctx->b.metrics, (size_t)g.ndim * g.ndim);
    const size_t gbytes = grid_nodes(g) * 3 * sizeof(double);
    ok = ok && hipMalloc((void **)&ctx->b.grid, gbytes) == hipSuccess && hipMemset(ctx->b.grid, 0, gbytes) == hipSuccess;
  }
  if (ok && noise_on(c)) {  // noise filter buffers (src/cart3d.cpp:91-93, src/cart2d.cpp:64-66)
    ok = alloc(&ctx->b.ndn, 1);
    const size_t nf = (size_t)std::max(noise_flag_count(g, c.grid == 1), 1);
    ok = ok && hipMalloc((void **)&ctx->b.nflag, nf) == hipSuccess &&
         hipMalloc((void **)&ctx->b.nmax, sizeof(unsigned long long)) == hipSuccess &&
         hipMemset(ctx->b.nmax, 0, sizeof(unsigned long long)) == hipSuccess;
  }
  if (ok && c.learned_c) {
    if (g.ndim == 3 && !fused_supported(c)) {
      fprintf(stderr, "fst_create: the 3D learned C-indicator needs the fused 3D path (ns <= 3, FP64 or FP32_BULK)\n");
      (void)fst_destroy(ctx);
      return FST_ERR_ARG;
    }
    ok = learned_alloc(g, ctx->lb) == FST_OK;
  }
  ok = ok && hipMalloc((void **)&ctx->b.mm, 2 * sizeof(unsigned long long) * std::max(g.nvt, g.nvx)) == hipSuccess;
  ok = ok && hipDeviceSynchronize() == hipSuccess;
  ok = ok && hipMalloc((void **)&ctx->b.sc, sizeof(Scalars)) == hipSuccess &&
       hipMemset(ctx->b.sc, 0, sizeof(Scalars)) == hipSuccess;
  if (!ok) {
    fprintf(stderr, "fst_create: device allocation failed\n");
    (void)fst_destroy(ctx);
    return FST_ERR_NOMEM;
  }
  ctx->fused = fused_supported(c) || fused2d_supported(c) || gen_fused(c, g) ? 1 : 0;
  ctx->b.prof = &ctx->prof;
  *out = ctx;
  return FST_OK;
}

int fst_destroy(fst_ctx *ctx) {
  if (!ctx) return FST_ERR_ARG;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_all(ctx);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FST_OK;
}

const char *fst_last_error(const fst_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fst_get_dims(const fst_ctx *ctx, fst_dims *out) {
  if (!ctx || !out) return FST_ERR_ARG;
  std::string why;
  return derive(ctx->cf, out, &why);
}

int fst_upload(fst_ctx *ctx, int array, const double *host, size_t count) {
  if (!ctx || !host) return FST_ERR_ARG;
  size_t n = 0;
  double *p = array_ptr(ctx, array, &n);
  if (!p || n == 0) return fail(ctx, FST_ERR_ARG, "fst_upload: array not available in this configuration");
  if (count != n) return fail(ctx, FST_ERR_ARG, "fst_upload: count does not match the array size");
  HIPCHK(ctx, hipMemcpyAsync(p, host, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (array == FST_VAR) ctx->fb.valid = false;
  ctx->fb.mt32_ok = false;  // the float metrics follow the next preSim or upload
  return FST_OK;
}

int fst_download(fst_ctx *ctx, int array, double *host, size_t count) {
  if (!ctx || !host) return FST_ERR_ARG;
  size_t n = 0;
  double *p = array_ptr(ctx, array, &n);
  if (!p || n == 0) return fail(ctx, FST_ERR_ARG, "fst_download: array not available in this configuration");
  if (count != n) return fail(ctx, FST_ERR_ARG, "fst_download: count does not match the array size");
  HIPCHK(ctx, hipMemcpyAsync(host, p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return FST_OK;
}

int fst_apply_bcs(fst_ctx *ctx) {
  if (!ctx) return FST_ERR_ARG;
  ctx->fb.valid = false;
  launch_apply_bcs(ctx->g, ctx->cf, ctx->b.var, ctx->stream);
  return check_launch(ctx, "fst_apply_bcs");
}

int fst_pre_sim(fst_ctx *ctx) {
  if (!ctx) return FST_ERR_ARG;
  ctx->fb.valid = false;
  ctx->fb.mt32_ok = false;
  if (ctx->cf.grid == 1) {  // gen3d_func::preSim / gen2d_func::preSim: metrics only
    launch_metrics(ctx->g, ctx->b, ctx->stream);
    // gen_ext in 2D: the ghost fill of cart2d's preSim (src/cart2d.cpp:137)
    if (ctx->cf.gen_ext && ctx->g.ndim == 2) launch_apply_bcs(ctx->g, ctx->cf, ctx->b.var, ctx->stream);
    return check_launch(ctx, "fst_pre_sim");
  }
  if (ctx->g.ndim == 2) launch_apply_bcs(ctx->g, ctx->cf, ctx->b.var, ctx->stream);  // src/cart2d.cpp:137
  launch_secondary(ctx->g, ctx->ph, ctx->b, ctx->stream);
  return check_launch(ctx, "fst_pre_sim");
}

int fst_pre_step(fst_ctx *ctx) {
  if (!ctx) return FST_ERR_ARG;
  ctx->fb.valid = false;
  // cart2d: src/cart2d.cpp:147-149 (gen2d with gen_ext likewise); cart3d,
  // gen2d, gen3d: nothing
  if (ctx->g.ndim == 2 && (ctx->cf.grid == 0 || ctx->cf.gen_ext))
    launch_apply_bcs(ctx->g, ctx->cf, ctx->b.var, ctx->stream);
  return check_launch(ctx, "fst_pre_step");
}

int fst_compute(fst_ctx *ctx) {
  if (!ctx) return FST_ERR_ARG;
  if (ctx->cf.learned_c) return learned_only(ctx, "fst_compute");
  if (ctx->cf.precision != FST_FP64) return fp32_only(ctx, "fst_compute");
  ctx->fb.valid = false;
  compute_seq(ctx);
  return check_launch(ctx, "fst_compute");
}

int fst_rk_stage(fst_ctx *ctx, int stage) {
  if (!ctx || stage < 1 || stage > rk_stages(ctx->cf)) return FST_ERR_ARG;
  ctx->fb.valid = false;
  launch_rk_stage(ctx->g, rk_stages(ctx->cf), stage, ctx->cf.dt, ctx->b.var, ctx->b.tmp1, ctx->b.dvar, ctx->stream);
  return check_launch(ctx, "fst_rk_stage");
}

int fst_post_step(fst_ctx *ctx) {
  if (!ctx) return FST_ERR_ARG;
  // generalized grids: rho, p and vel alias the fused gen3d stage's primitive
  // planes, and vel here is the cartesian velocity, not the contravariant one
  if (ctx->cf.grid == 1) ctx->fb.valid = false;
  launch_secondary(ctx->g, ctx->ph, ctx->b, ctx->stream);
  if (noise_on(ctx->cf)) {
    ctx->fb.valid = false;  // var changes
    launch_noise(ctx->g, ctx->ph, ctx->cf, ctx->b, ctx->stream);
  }
  return check_launch(ctx, "fst_post_step");
}

int fst_set_fused(fst_ctx *ctx, int fused) {
  if (!ctx) return FST_ERR_ARG;
  if (fused && !fused_supported(ctx->cf) && !fused2d_supported(ctx->cf) && !gen_fused(ctx->cf, ctx->g))
    return fail(ctx, FST_ERR_ARG,
                "the fused path supports cartesian WENO5 configurations with ns <= 3 (2D: FP64) and every array "
                "below 2^32 elements (3D), and gen3d with WENO5 and ns <= 3");
  if (!fused && ctx->cf.learned_c && ctx->g.ndim == 3) return learned_only(ctx, "fst_set_fused(0)");
  if (!fused && ctx->cf.precision != FST_FP64) return fp32_only(ctx, "fst_set_fused(0)");
  ctx->fused = fused ? 1 : 0;
  return FST_OK;
}

int fst_step(fst_ctx *ctx, int nsteps) {
  if (!ctx || nsteps < 0) return FST_ERR_ARG;
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (ctx->fused) {
    int rc = ctx->g.ndim == 2 ? fused2d_steps(ctx->g, ctx->ph, ctx->cf, ctx->b, ctx->fb, &ctx->lb, nsteps,
                                              ctx->stream, &ctx->err)
             : ctx->cf.grid == 1 ? fused_gen_steps(ctx->g, ctx->ph, ctx->cf, ctx->b, ctx->fb, nsteps, ctx->stream,
                                                   &ctx->err)
                                 : fused_steps(ctx->g, ctx->ph, ctx->cf, ctx->b, ctx->fb, &ctx->lb, nsteps,
                                               ctx->stream, &ctx->err);
    if (rc) return rc;
  } else {
    ctx->fb.valid = false;
    for (int t = 0; t < nsteps; ++t) {
      if (ctx->g.ndim == 2 && (ctx->cf.grid == 0 || ctx->cf.gen_ext))
        launch_apply_bcs(ctx->g, ctx->cf, ctx->b.var, ctx->stream);
      const int ns = rk_stages(ctx->cf);
      for (int stage = 1; stage <= ns; ++stage) {
        const int rc = compute_seq(ctx, stage == 1);
        if (rc)
          return fail(ctx, rc, rc == FST_ERR_STATE ? "fst_step: no network weights set" : "fst_step: launch failed");
        launch_rk_stage(ctx->g, ns, stage, ctx->cf.dt, ctx->b.var, ctx->b.tmp1, ctx->b.dvar, ctx->stream);
      }
    }
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  return check_launch(ctx, "fst_step");
}

int fst_last_step_ms(fst_ctx *ctx, double *ms) {
  if (!ctx || !ms) return FST_ERR_ARG;
  HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
  float f = 0.f;
  HIPCHK(ctx, hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
  *ms = f;
  return FST_OK;
}

int fst_last_maxima(fst_ctx *ctx, double *maxS, double *maxCh) {
  if (!ctx || !maxS || !maxCh) return FST_ERR_ARG;
  Scalars s;
  HIPCHK(ctx, hipMemcpyAsync(&s, ctx->b.sc, sizeof(s), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *maxS = as_double(s.maxS);
  *maxCh = as_double(s.maxCh);
  return FST_OK;
}

int fst_profile_enable(fst_ctx *ctx, int enable) {
  if (!ctx) return FST_ERR_ARG;
  ctx->prof.on = enable != 0;
  ctx->prof.used = 0;
  return FST_OK;
}

int fst_profile_read(fst_ctx *ctx, int64_t *launches, double *total_ms, char *name, size_t name_len) {
  if (!ctx || !launches || !total_ms) return FST_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const size_t per = (size_t)ctx->prof.per;
  double tot = 0.0;
  ctx->nphase = (int)per - 1;
  for (int q = 0; q < Prof::MAXPER; ++q) ctx->phase_ms[q] = 0.0;
  for (size_t q = 0; q + per <= ctx->prof.used; q += per) {
    float f = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&f, ctx->prof.ev[q], ctx->prof.ev[q + per - 1]));
    tot += f;
    for (size_t ph = 0; ph + 1 < per; ++ph) {
      HIPCHK(ctx, hipEventElapsedTime(&f, ctx->prof.ev[q + ph], ctx->prof.ev[q + ph + 1]));
      ctx->phase_ms[ph] += f;
    }
  }
  *launches = (int64_t)(ctx->prof.used / per);
  *total_ms = tot;
  if (name && name_len) {
    std::strncpy(name, ctx->prof.name, name_len - 1);
    name[name_len - 1] = 0;
  }
  ctx->prof.used = 0;
  return FST_OK;
}

int fst_profile_phases(fst_ctx *ctx, double *ms, int n, int *nphase) {
  if (!ctx || !ms || !nphase || n < 0) return FST_ERR_ARG;
  *nphase = ctx->nphase;
  for (int q = 0; q < n && q < ctx->nphase; ++q) ms[q] = ctx->phase_ms[q];
  return FST_OK;
}

int fst_pack(fst_ctx *ctx, int array, int vn, const int64_t start[3], const int64_t end[3], const int64_t stride[3],
             int average, int out_f32, void *out, size_t count) {
  if (!ctx || !start || !end || !stride || !out) return FST_ERR_ARG;
  const Geo &g = ctx->g;
  const int nd = g.ndim;
  // the plane of variable vn, its extents and the offset of index 0
  const double *src = nullptr;
  long long nx = g.ngi, ny = g.ngj, lim[3] = {g.nci, g.ncj, g.nck};
  int nvar = 0, off = g.ng;
  if (array == FST_VAR) { src = ctx->b.var; nvar = g.nvt; }
  else if (array == FST_VARX) { src = ctx->b.varx; nvar = g.nvx; }
  else if (array == FST_GRID && ctx->b.grid) {  // nodes (ni, nj, nk, 3)
    src = ctx->b.grid; nvar = 3; off = 0;
    nx = g.nci + 1; ny = g.ncj + 1;
    lim[0] = g.nci + 1; lim[1] = g.ncj + 1; lim[2] = nd == 3 ? g.nck + 1 : 1;
  }
  if (!src) return fail(ctx, FST_ERR_ARG, "fst_pack: array not available (FST_VAR, FST_VARX, FST_GRID)");
  if (vn < 0 || vn >= nvar) return fail(ctx, FST_ERR_ARG, "fst_pack: variable index out of range");
  const long long plane = array == FST_GRID ? nx * ny * lim[2] : (long long)g.N;
  PackBox bx;
  long long n = 1;
  for (int a = 0; a < 3; ++a) {
    if (a >= nd) { bx.start[a] = 0; bx.stride[a] = 1; bx.ext[a] = 1; continue; }
    if (stride[a] < 1 || start[a] < 0 || end[a] < start[a]) return fail(ctx, FST_ERR_ARG, "fst_pack: bad box");
    bx.start[a] = start[a];
    bx.stride[a] = stride[a];
    bx.ext[a] = (end[a] - start[a]) / stride[a] + 1;
    // every read stays inside the plane: the last point (plus its averaging block)
    const long long last = start[a] + (bx.ext[a] - 1) * stride[a] + (average ? stride[a] - 1 : 0);
    const long long room = array == FST_GRID ? lim[a] : lim[a] + g.ng;  // var/varx may reach into the ghosts
    if (last >= room) return fail(ctx, FST_ERR_ARG, "fst_pack: box leaves the array");
    n *= bx.ext[a];
  }
  if ((size_t)n != count) return fail(ctx, FST_ERR_ARG, "fst_pack: count does not match the packed size");
  const size_t bytes = (size_t)n * (out_f32 ? sizeof(float) : sizeof(double));
  if (bytes > ctx->pack_bytes) {
    if (ctx->pack) (void)hipFree(ctx->pack);
    ctx->pack = nullptr;
    ctx->pack_bytes = 0;
    if (hipMalloc(&ctx->pack, bytes) != hipSuccess) return fail(ctx, FST_ERR_NOMEM, "fst_pack: staging allocation");
    ctx->pack_bytes = bytes;
  }
  launch_pack(src + plane * vn, nx, ny, nd, off, bx, average, out_f32 != 0, ctx->pack, ctx->stream);
  if (int rc = check_launch(ctx, "fst_pack")) return rc;
  HIPCHK(ctx, hipMemcpyAsync(out, ctx->pack, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return FST_OK;
}

int fst_status(fst_ctx *ctx, int array, double *min, double *max, int nvar) {
  if (!ctx || !min || !max) return FST_ERR_ARG;
  const Geo &g = ctx->g;
  const double *a = array == FST_VAR ? ctx->b.var : array == FST_VARX ? ctx->b.varx : nullptr;
  const int na = array == FST_VAR ? g.nvt : g.nvx;
  if (!a) return fail(ctx, FST_ERR_ARG, "fst_status: array must be FST_VAR or FST_VARX");
  if (nvar < 1 || nvar > na) return fail(ctx, FST_ERR_ARG, "fst_status: nvar out of range");
  launch_minmax(g, a, nvar, ctx->b.mm, ctx->stream);
  if (int rc = check_launch(ctx, "fst_status")) return rc;
  unsigned long long k[2 * 64];
  HIPCHK(ctx, hipMemcpyAsync(k, ctx->b.mm, 2 * sizeof(unsigned long long) * nvar, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int v = 0; v < nvar; ++v) {
    max[v] = key_to_double(k[2 * v]);
    min[v] = key_to_double(k[2 * v + 1]);
  }
  return FST_OK;
}

int fst_copy_probe(fst_ctx *ctx, size_t bytes, int reps, double *ms_per_copy) {
  if (!ctx || !ms_per_copy || bytes < 32 || reps < 1) return FST_ERR_ARG;
  bytes &= ~(size_t)31;
  void *a = nullptr, *b = nullptr;
  HIPCHK(ctx, hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return fail(ctx, FST_ERR_NOMEM, "fst_copy_probe: allocation failed");
  }
  (void)hipMemsetAsync(a, 0, bytes, ctx->stream);
  launch_copy8(a, b, bytes, ctx->stream);  // 8-byte accesses: a second counter reference, untimed
  launch_copy(a, b, bytes, ctx->stream);  // warm-up
  (void)hipEventRecord(ctx->ev0, ctx->stream);
  for (int r = 0; r < reps; ++r) launch_copy(a, b, bytes, ctx->stream);
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  hipError_t e = hipEventSynchronize(ctx->ev1);
  float f = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&f, ctx->ev0, ctx->ev1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(ctx, FST_ERR_HIP, std::string("fst_copy_probe: ") + hipGetErrorString(e));
  *ms_per_copy = (double)f / reps;
  return FST_OK;
}

int fst_sync(fst_ctx *ctx) {
  if (!ctx) return FST_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return check_launch(ctx, "fst_sync");
}

// ---------------------------------------------------------------- learned C-indicator

int fst_mlp_nparams(void) { return MLP_NPARAM; }

int fst_mlp_nparams_nd(int ndim) { return ndim == 2 || ndim == 3 ? mlp_nparams(ndim) : -1; }

int fst_set_mlp_weights(fst_ctx *ctx, const float *w, size_t n) {
  if (!ctx || !w) return FST_ERR_ARG;
  if (!ctx->cf.learned_c) return fail(ctx, FST_ERR_ARG, "fst_set_mlp_weights: context has learned_c = 0");
  if (n != (size_t)mlp_nparams(ctx->lb.nd))
    return fail(ctx, FST_ERR_ARG, "fst_set_mlp_weights: wrong parameter count for the context's dimension");
  const int rc = learned_set_weights(ctx->lb, w, ctx->stream);
  return rc ? fail(ctx, rc, "fst_set_mlp_weights: copy failed") : FST_OK;
}

int fst_learned_params(fst_ctx *ctx, float tau, float tol) {
  if (!ctx || !(tau >= 0.f) || !(tol >= 0.f)) return FST_ERR_ARG;
  if (!ctx->cf.learned_c) return fail(ctx, FST_ERR_ARG, "fst_learned_params: context has learned_c = 0");
  ctx->lb.tau = tau;
  ctx->lb.tol = tol;
  return FST_OK;
}

int fst_learned_compute(fst_ctx *ctx) {
  if (!ctx) return FST_ERR_ARG;
  if (!ctx->cf.learned_c) return fail(ctx, FST_ERR_ARG, "fst_learned_compute: context has learned_c = 0");
  int rc = FST_OK;
  launch_apply_bcs(ctx->g, ctx->cf, ctx->b.var, ctx->stream);
  if (ctx->g.ndim == 2) {
    rc = learned_fields2d(ctx->g, ctx->cf, ctx->ph, ctx->b.var, ctx->b.rho, ctx->lb, ctx->stream);
  } else {
    const size_t np = 6 + (size_t)ctx->cf.ns;
    if (!ctx->lprim) HIPCHK(ctx, hipMalloc((void **)&ctx->lprim, np * (size_t)ctx->g.N * sizeof(double)));
    rc = fused_primitives(ctx->g, ctx->ph, ctx->cf, ctx->b.var, ctx->lprim, ctx->stream);
    if (!rc) rc = learned_fields(ctx->g, ctx->cf, ctx->ph, ctx->lprim, false, ctx->lb, ctx->stream);
  }
  if (rc) return fail(ctx, rc, rc == FST_ERR_STATE ? "fst_learned_compute: no network weights set"
                                                   : "fst_learned_compute: launch failed");
  ctx->fb.valid = false;
  return check_launch(ctx, "fst_learned_compute");
}

int fst_learned_stats(fst_ctx *ctx, int64_t *active, int64_t *fallback, double *maxCh) {
  if (!ctx) return FST_ERR_ARG;
  if (!ctx->cf.learned_c) return fail(ctx, FST_ERR_ARG, "fst_learned_stats: context has learned_c = 0");
  unsigned c[4];
  Scalars sc;
  HIPCHK(ctx, hipMemcpyAsync(c, ctx->lb.cnt, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&sc, ctx->lb.sc, sizeof(sc), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (active) *active = c[0];
  if (fallback) *fallback = c[1];
  if (maxCh) *maxCh = as_double(sc.maxCh);
  return FST_OK;
}

int fst_learned_fields(fst_ctx *ctx, float *chat, float *dir, float *feat, float *g1, size_t plane) {
  if (!ctx) return FST_ERR_ARG;
  if (!ctx->cf.learned_c) return fail(ctx, FST_ERR_ARG, "fst_learned_fields: context has learned_c = 0");
  const size_t N = (size_t)ctx->g.N;
  if (plane != N) return fail(ctx, FST_ERR_ARG, "fst_learned_fields: plane size does not match");
  if (chat) HIPCHK(ctx, hipMemcpyAsync(chat, ctx->lb.chat, N * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (dir) HIPCHK(ctx, hipMemcpyAsync(dir, ctx->lb.dir, 3 * N * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (feat) HIPCHK(ctx, hipMemcpyAsync(feat, ctx->lb.F, 3 * N * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (g1) HIPCHK(ctx, hipMemcpyAsync(g1, ctx->lb.G1, N * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return FST_OK;
}

}  // extern "C"

struct fst_mlp {
  Trainer t;
  std::string err;
};

namespace {
int mfail(fst_mlp *m, int code, const std::string &msg) {
  if (m) m->err = msg;
  return code;
}
}  // namespace

extern "C" {

int fst_mlp_create(uint64_t seed, fst_mlp **out) { return fst_mlp_create_nd(3, seed, out); }

int fst_mlp_create_nd(int ndim, uint64_t seed, fst_mlp **out) {
  if (!out || (ndim != 2 && ndim != 3)) return FST_ERR_ARG;
  *out = nullptr;
  fst_mlp *m = new fst_mlp();
  const int rc = trainer_create(m->t, ndim, seed);
  if (rc) {
    trainer_free(m->t);
    delete m;
    return rc;
  }
  *out = m;
  return FST_OK;
}

int fst_mlp_destroy(fst_mlp *m) {
  if (!m) return FST_ERR_ARG;
  if (m->t.s) (void)hipStreamSynchronize(m->t.s);
  trainer_free(m->t);
  delete m;
  return FST_OK;
}

const char *fst_mlp_last_error(const fst_mlp *m) { return m ? m->err.c_str() : "null trainer"; }

int fst_mlp_ndim(const fst_mlp *m) { return m ? m->t.nd : -1; }

int fst_mlp_set_weights(fst_mlp *m, const float *w, size_t n) {
  if (!m || !w) return FST_ERR_ARG;
  if (n != (size_t)mlp_nparams(m->t.nd)) return mfail(m, FST_ERR_ARG, "fst_mlp_set_weights: wrong parameter count");
  const int rc = trainer_set_weights(m->t, w);
  return rc ? mfail(m, rc, "fst_mlp_set_weights: copy failed") : FST_OK;
}

int fst_mlp_get_weights(fst_mlp *m, float *w, size_t n) {
  if (!m || !w) return FST_ERR_ARG;
  if (n != (size_t)mlp_nparams(m->t.nd)) return mfail(m, FST_ERR_ARG, "fst_mlp_get_weights: wrong parameter count");
  const int rc = trainer_get_weights(m->t, w);
  return rc ? mfail(m, rc, "fst_mlp_get_weights: copy failed") : FST_OK;
}

int fst_mlp_harvest(fst_mlp *m, fst_ctx *ctx, uint32_t stride, int64_t max_rows, int64_t *added) {
  if (!m || !ctx) return FST_ERR_ARG;
  const fst_config &c = ctx->cf;
  if (c.ndim != m->t.nd)
    return mfail(m, FST_ERR_ARG, "fst_mlp_harvest: the context's dimension differs from the network's");
  if (c.ndim == 2) {
    if (c.grid != 0 || !c.ceq || c.learned_c)
      return mfail(m, FST_ERR_ARG, "fst_mlp_harvest: needs a 2D cartesian context with the C-equation");
    // cart2d's state as its compute() sees it: ghosts filled in preStep (the
    // feature pass computes the primitives and writes rho into the context's
    // rho plane, which compute() rewrites with the same values)
    launch_apply_bcs(ctx->g, c, ctx->b.var, ctx->stream);
    ctx->fb.valid = false;
    int rc = hipStreamSynchronize(ctx->stream) == hipSuccess ? FST_OK : FST_ERR_HIP;
    long long a = 0;
    if (!rc) rc = trainer_harvest2d(m->t, ctx->g, c, ctx->ph, ctx->b.var, ctx->b.rho, stride, max_rows, &a);
    if (added) *added = a;
    return rc ? mfail(m, rc, "fst_mlp_harvest: failed") : FST_OK;
  }
  if (c.grid != 0 || !c.ceq || c.learned_c || c.ns > 3 || !fused_supported(c))
    return mfail(m, FST_ERR_ARG,
                 "fst_mlp_harvest: needs a 3D cartesian context with the C-equation (ns <= 3, every array "
                 "below 2^32 elements)");
  const long long np = 6 + c.ns;
  if (m->t.Pcells != ctx->g.N || m->t.Pplanes != np) {
    if (m->t.P) (void)hipFree(m->t.P);
    m->t.P = nullptr;
    m->t.Pcells = 0;
    if (hipMalloc((void **)&m->t.P, (size_t)np * ctx->g.N * sizeof(double)) != hipSuccess)
      return mfail(m, FST_ERR_NOMEM, "fst_mlp_harvest: allocation failed");
    m->t.Pcells = ctx->g.N;
    m->t.Pplanes = (int)np;
  }
  // ghost cells of the state as compute() would see them, then primitives
  launch_apply_bcs(ctx->g, c, ctx->b.var, ctx->stream);
  int rc = fused_primitives(ctx->g, ctx->ph, c, ctx->b.var, m->t.P, ctx->stream);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = FST_ERR_HIP;
  long long a = 0;
  if (!rc) rc = trainer_harvest(m->t, ctx->g, c, ctx->ph, ctx->b.var, m->t.P, stride, max_rows, &a);
  if (added) *added = a;
  return rc ? mfail(m, rc, "fst_mlp_harvest: failed") : FST_OK;
}

int fst_mlp_rows(fst_mlp *m, int64_t *rows) {
  if (!m || !rows) return FST_ERR_ARG;
  *rows = m->t.nrows;
  return FST_OK;
}

int fst_mlp_read_rows(fst_mlp *m, int64_t r0, int64_t n, float *X, float *T) {
  if (!m) return FST_ERR_ARG;
  const int rc = trainer_read_rows(m->t, r0, n, X, T);
  return rc ? mfail(m, rc, "fst_mlp_read_rows: bad range or copy failed") : FST_OK;
}

int fst_mlp_train(fst_mlp *m, int epochs, int batch, float lr, uint64_t seed, float *loss_per_epoch) {
  if (!m || !(lr > 0.f)) return FST_ERR_ARG;
  const int rc = trainer_train(m->t, epochs, batch, lr, seed, loss_per_epoch);
  return rc ? mfail(m, rc, "fst_mlp_train: needs epochs >= 1, 16 <= batch <= rows") : FST_OK;
}

int fst_mlp_forward_backward(fst_mlp *m, const float *X, const float *T, int64_t rows, float *Y, float *grad,
                             float *loss) {
  if (!m || !X || !T || rows < 1 || rows > (1 << 24)) return FST_ERR_ARG;
  const int rc = trainer_forward_backward(m->t, X, T, (int)rows, Y, grad, loss);
  return rc ? mfail(m, rc, "fst_mlp_forward_backward: failed") : FST_OK;
}

}  // extern "C"
