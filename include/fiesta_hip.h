/*
 * fiesta_hip.h -- C-ABI of the MI355X-native FIESTA right-hand-side and
 * Runge-Kutta path (gfx950, hand-written HIP kernels).
 *
 * The entry points replace the device side of FIESTA's `rk_func` interface:
 * a host `rk_func` subclass (cart2d / cart3d / gen2d / gen3d) keeps its compute() / preStep()
 * / postStep() / preSim() / postSim() methods and forwards each one to the
 * function named below.  Plain pointers and sizes only; every function
 * returns 0 on success and a nonzero FST_ERR_* code otherwise, with a message
 * available from fst_last_error().  No C++ exception crosses this boundary.
 *
 * Reference interfaces (paths relative to the FIESTA source tree):
 *   class rk_func                      src/rkfunction.hpp:36-69
 *   cart3d_func / cart2d_func          src/cart3d.cpp:44-378, src/cart2d.cpp:30-311
 *   gen3d_func / gen2d_func            src/gen3d.cpp:39-370, src/gen2d.cpp:22-333
 *   rkAdvance(cf, f)                   src/rk.hpp:26, src/rk.cpp:28-72
 *   applyBCs(cf, f)                    src/bc.hpp:33, src/bc.cpp:179-249
 *   struct inputConfig (fields used)   src/input.hpp:59-148
 *
 * Array layout (device and host): Kokkos LayoutLeft of the reference GPU
 * build (src/kokkosTypes.hpp:40): element (i,j,k,v) of an array with
 * extents (ngi,ngj,ngk,nv) sits at  i + ngi*(j + ngj*(k + ngk*v)),
 * i.e. one contiguous plane per variable, i fastest, ghosts included.
 */
#ifndef FIESTA_HIP_H
#define FIESTA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FST_MAX_SPECIES 8
#define FST_ABI_VERSION 7

/* Boundary types, in the order of `enum class BCType` (src/bc.hpp:31). */
enum fst_bc_type {
  FST_BC_OUTFLOW = 0,
  FST_BC_REFLECTIVE = 1,
  FST_BC_NOSLIP = 2,
  FST_BC_HYDROSTATIC = 3
};

/* Advection schemes, numbered as cf.scheme (src/input.cpp:317-323). */
enum fst_scheme { FST_WENO5 = 1, FST_CENTERED4 = 2, FST_QUICK = 3 };

/* Arithmetic forms of the path. */
enum fst_precision {
  FST_FP64 = 0,       /* reference arithmetic, bitwise parity target */
  FST_FP32_BULK = 1,  /* FP32 stencils/fluxes, FP64 EOS and RK accumulation
                         (fused 3D paths: cartesian and gen3d; fst_step only) */
  FST_FP64_FAST = 2   /* FP64 throughout with FMA contraction, reciprocal-based
                         division and one-division WENO5 weights: within
                         relative L2 1e-12 of the reference arithmetic, not
                         bitwise (fused 3D path only) */
};

/* Arrays reachable through fst_upload / fst_download.  Extents in brackets. */
enum fst_array {
  FST_VAR = 0,     /* conserved state var      [ngi,ngj,ngk,nvt]        */
  FST_TMP1 = 1,    /* RK stage copy tmp1       [ngi,ngj,ngk,nvt]        */
  FST_DVAR = 2,    /* right-hand side dvar     [ngi,ngj,ngk,nvt]        */
  FST_VARX = 3,    /* secondary vars varx      [ngi,ngj,ngk,nvx]        */
  FST_P = 4,       /* pressure                 [ngi,ngj,ngk]            */
  FST_RHO = 5,     /* total density            [ngi,ngj,ngk]            */
  FST_T = 6,       /* temperature              [ngi,ngj,ngk]            */
  FST_VEL = 7,     /* velocity                 [ngi,ngj,ngk,ndim]       */
  FST_GRADRHO = 8, /* C-eq source terms        [ngi,ngj,ngk,5 | 4 in 2D] */
  FST_GRID = 9,    /* grid nodes (generalized) [ni,nj,nk,3], ni = nci+1,
                      nj = ncj+1, nk = nck+1 (1 in 2D)                  */
  FST_METRICS = 10 /* inverse Jacobian d xi_m/d x_n [ngi,ngj,ngk,ndim,ndim] */
};

enum fst_err {
  FST_OK = 0,
  FST_ERR_ARG = 1,     /* bad argument / unsupported configuration */
  FST_ERR_HIP = 2,     /* a HIP runtime call failed                */
  FST_ERR_STATE = 3,   /* call out of order (e.g. no state loaded) */
  FST_ERR_NOMEM = 4    /* device allocation failed                 */
};

/*
 * Run configuration: the subset of `struct inputConfig` (src/input.hpp:59-148)
 * that the RHS/RK path reads, with the defaults of executeConfiguration
 * (src/input.cpp:130-443).  fst_config_init() fills those defaults.
 */
typedef struct fst_config {
  int32_t ndim;          /* 2 or 3                                  (grid.ndim) */
  int32_t grid;          /* 0 cartesian, 1 generalized (gen2d/gen3d: FP64,
                            dx = dy = dz = 1, no BCs, no viscous, buoyancy
                            or C-equation terms)                (cf.grid)  */
  int32_t ns;            /* species count, 1..FST_MAX_SPECIES        (cf.ns)    */
  int32_t ng;            /* ghost width, default 3                   (cf.ng)    */
  int32_t nci, ncj, nck; /* interior cells; nck = 1 in 2D                       */
  double dx, dy, dz;     /* cell sizes (1.0 on generalized grids)               */
  double R;              /* universal gas constant, default 8.314462            */
  double gamma[FST_MAX_SPECIES]; /* species ratio of specific heats              */
  double M[FST_MAX_SPECIES];     /* species molar mass                           */
  double mu[FST_MAX_SPECIES];    /* species viscosity                            */
  int32_t visc;          /* viscosity.enabled                                   */
  int32_t ceq;           /* ceq.enabled                                         */
  int32_t buoyancy;      /* buoyancy.enabled                                    */
  int32_t scheme;        /* fst_scheme, default FST_WENO5                       */
  double kap, eps, alpha;/* ceq.kappa (1.23456), ceq.epsilon, ceq.alpha         */
  double gAccel, rhoRef; /* buoyancy.acceleration (9.81), rho_reference (0)     */
  int32_t xPer, yPer, zPer;           /* bc.{x,y,z}periodic                     */
  int32_t bcL, bcR, bcB, bcT, bcH, bcF; /* bc.xmin,xmax,ymin,ymax,zmin,zmax     */
  double dt;             /* time.dt                                             */
  int32_t precision;     /* fst_precision                                       */
  int32_t learned_c;     /* 1: learned C-indicator replaces the C-equation's
                            relaxation solve (ceq = 1; 3D: fused path only;
                            2D: FP64; no noise filter; nvt = nv: no
                            C-equation variables are stored)                 */
  int32_t rk_order;      /* 2 (or 0): the reference's RK2 midpoint (rkAdvance);
                            3: three-stage SSP-RK3 (Shu-Osher), an option the
                            reference does not have                          */
  int32_t noise;         /* noise.enabled: the noise filter of postStep
                            (cartesian modules, src/noise.hpp)               */
  int32_t n_nt;          /* noise.nt, smoothing sweeps per variable (1)      */
  int32_t n_mode;        /* noise.mode: 1 momentum components, else energy   */
  double n_dh;           /* noise.dh, detection threshold                     */
  double n_eta;          /* noise.eta, smoothing coefficient                  */
  double n_coff;         /* noise.coff, C_hat cut-off factor (ceq on)         */
  int32_t gen_ext;       /* generalized grids (grid = 1) only, an option the
                            reference does not have: 1 = the ghost fill of the
                            cartesian modules (index space, the bc.* types)
                            before each compute (3D) / in preStep (2D), and the
                            buoyancy term when buoyancy = 1; 0 = gen2d/gen3d
                            as the reference runs them (neither)             */
  int32_t reserved[3];
} fst_config;

typedef struct fst_ctx fst_ctx;

/* Sizes derived from a config exactly as src/input.cpp:354-392 derives them. */
typedef struct fst_dims {
  int32_t ngi, ngj, ngk; /* extents with ghosts (ngk = 1 in 2D) */
  int32_t nv;            /* ndim + 1 + ns                       */
  int32_t nvt;           /* nv + 5 when the C-equation is on (nv in the learned form) */
  int32_t nvx;           /* secondary variables: ndim + 3       */
  int64_t ncells_ghost;  /* ngi*ngj*ngk                         */
  int64_t ncells;        /* nci*ncj*nck                         */
} fst_dims;

int fst_abi_version(void);
void fst_config_init(fst_config *cfg);
int fst_config_dims(const fst_config *cfg, fst_dims *out);
/* The same validation as fst_config_dims, with the reason for a refusal
 * written to `why` (NUL-terminated, empty on success).  Host only. */
int fst_config_check(const fst_config *cfg, char *why, size_t len);

/* Allocate the device arrays of a module (cart3d_func ctor, src/cart3d.cpp:44-174;
 * cart2d_func ctor, src/cart2d.cpp:30-134).  One HIP stream per context. */
int fst_create(const fst_config *cfg, fst_ctx **out);
int fst_destroy(fst_ctx *ctx);
const char *fst_last_error(const fst_ctx *ctx);
int fst_get_dims(const fst_ctx *ctx, fst_dims *out);

/* Host <-> device copies of a whole array (layout above).  `count` is the
 * number of doubles the host buffer holds; it must equal the array size.
 * Both synchronize the context stream. */
int fst_upload(fst_ctx *ctx, int array, const double *host, size_t count);
int fst_download(fst_ctx *ctx, int array, double *host, size_t count);

/* rk_func::preSim (src/cart3d.cpp:176-185, src/cart2d.cpp:136-145).  On a
 * generalized grid: the metrics of the uploaded FST_GRID and their
 * mirror-copied ghosts (src/gen3d.cpp:131-304, src/gen2d.cpp:187-331). */
int fst_pre_sim(fst_ctx *ctx);
/* rk_func::preStep: cart2d applies the boundary conditions (src/cart2d.cpp:147-149),
 * cart3d, gen2d and gen3d do nothing (src/cart3d.cpp:187-188, src/gen3d.cpp:109). */
int fst_pre_step(fst_ctx *ctx);
/* applyBCs (src/bc.cpp:179-249), no-MPI branch: periodic x,y,z then
 * xmin,xmax,ymin,ymax,zmin,zmax planes. */
int fst_apply_bcs(fst_ctx *ctx);
/* rk_func::compute(): dvar = R(var), reference kernel sequence
 * (src/cart3d.cpp:220-298, src/cart2d.cpp:151-237, src/gen3d.cpp:341-369,
 * src/gen2d.cpp:100-131). */
int fst_compute(fst_ctx *ctx);
/* One Runge-Kutta stage update on the current dvar, over every cell
 * (src/rk.cpp:42-50 for stage 1, src/rk.cpp:63-69 for stage 2).  With
 * rk_order = 3, stages 1..3 of SSP-RK3 (tmp1 = U^n):
 *   1: tmp1 = var; var = var + dt*dvar
 *   2: var = 0.75*tmp1 + 0.25*(var + dt*dvar)
 *   3: var = (1.0/3.0)*tmp1 + (2.0/3.0)*(var + dt*dvar)                  */
int fst_rk_stage(fst_ctx *ctx, int stage);
/* rk_func::postStep: secondary variables varx (src/cart3d.cpp:300-317,
 * src/cart2d.cpp:239-255), then, with noise = 1, the noise filter
 * (src/cart3d.cpp:319-374, src/cart2d.cpp:257-307, src/noise.hpp).  Its
 * detection writes overlapping 3x3(x3) blocks and its smoothing updates var in
 * place inside parallel loops, so the reference's result depends on the loop
 * order; this build defines it order-free: a cell is flagged when any block
 * covering it flags it, and each smoothing sweep reads the values from before
 * the sweep. */
int fst_post_step(fst_ctx *ctx);

/* `nsteps` time steps of preStep + rkAdvance (rk_order stages of compute()
 * + rk_stage) on the fused fast path (one RHS+RK kernel per stage).  Results
 * equal the compute()/rk_stage() sequence.  Fiesta::step
 * (src/fiesta.cpp:333-340) also runs postStep after every step: with the
 * noise filter on, call fst_step(ctx, 1) and fst_post_step(ctx) per step. */
int fst_step(fst_ctx *ctx, int nsteps);

/* Size in MiB of the device scratch array fst_step_static uses on each GPU. */
#define FST_STATIC_SCRATCH_MB 256

/* fst_step without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * `nsteps` time steps of the fused 3D cartesian path (FP64 or FP32 form, RK2
 * or SSP-RK3; not the learned C-indicator) on `var`, a device pointer to the
 * caller's state (LayoutLeft, nvt planes with ghosts, as fst_upload takes
 * it), updated in place; `stream` is a hipStream_t (NULL: the default
 * stream).  The result equals fst_step on a context holding the same var
 * after fst_pre_sim.  The call is asynchronous on `stream`.
 *
 * The stage state, primitive planes, partial right-hand side and stage
 * maxima live in one static device array of FST_STATIC_SCRATCH_MB MiB per
 * GPU, held until the process ends and zeroed at the start of every call.
 * A grid that needs more (about 264 bytes per cell with ghosts at nvt = 11,
 * RK2; 390 at nvt = 12 with SSP-RK3) is refused with FST_ERR_ARG before
 * anything is launched.  Every call on one GPU shares that array: calls
 * that overlap in time on one GPU (other streams, other host threads) race
 * and give wrong results; serialize them.  Errors go to stderr (there is no
 * context to hold a message). */
int fst_step_static(const fst_config *cfg, double *var, int nsteps, void *stream);

/* Per-stage global maxima of the C-equation (max wave speed, max |C_hat|)
 * of the last computed stage (fsMax, src/cart3d.cpp:190-198, 281-283). */
int fst_last_maxima(fst_ctx *ctx, double *maxS, double *maxCh);

/* Block until all work queued on the context stream has finished. */
int fst_sync(fst_ctx *ctx);

/* Which implementation fst_step uses: 1 = fused kernels (default),
 * 0 = the reference kernel sequence through fst_compute/fst_rk_stage. */
int fst_set_fused(fst_ctx *ctx, int fused);

/* Elapsed device milliseconds of the last fst_step call, measured with
 * hipEvents on the context stream (Timer::fiestaTimer, src/timer.hpp:52-119). */
int fst_last_step_ms(fst_ctx *ctx, double *ms);

/* Live timing of the dominant kernel (the per-stage RHS kernel: the fused
 * stage kernel, or the WENO flux kernel on the reference sequence).  While
 * enabled, each launch of it is bracketed by hipEvents on the context stream.
 * fst_profile_read synchronizes and returns the launch count, the summed
 * duration in ms and the kernel's name; it then clears the counters. */
int fst_profile_enable(fst_ctx *ctx, int enable);
int fst_profile_read(fst_ctx *ctx, int64_t *launches, double *total_ms, char *name, size_t name_len);
/* Per-phase sums (ms) of the units read by the last fst_profile_read.  Fused
 * stages have 5 phases: ghost fill + ghost primitives, learned C fields (0 when
 * not used), advection kernel, viscous + flow-variable stage-update kernel,
 * C-equation kernel (0 in the learned form); the reference sequence has 1 (the
 * flux kernel).  Writes min(n, *nphase) values. */
int fst_profile_phases(fst_ctx *ctx, double *ms, int n, int *nphase);

/* blockWriter::dataPack (src/block.cpp:404-454) on the device, for the
 * writers (src/block.cpp:340-352): variable `vn` of FST_VAR, FST_VARX (cell
 * indices relative to the first interior cell) or FST_GRID (node indices),
 * at i = start[a], start[a] + stride[a], ..., end[a] (inclusive) per
 * direction, or, with average = 1, the mean of each stride[0] x stride[1]
 * (x stride[2]) block there; packed i fastest (extent (end - start)/stride
 * + 1 per direction) and copied to `out` as float (out_f32 = 1) or double.
 * `count` is the number of values `out` holds; it must equal the packed size.
 * Only the packed values cross to the host.  2D uses the first two entries. */
int fst_pack(fst_ctx *ctx, int array, int vn, const int64_t start[3], const int64_t end[3],
             const int64_t stride[3], int average, int out_f32, void *out, size_t count);

/* statusCheck's reductions (src/status.cpp:155-231): the maximum and minimum
 * of each of the first `nvar` variables of FST_VAR or FST_VARX over the
 * interior cells, in one pass on the device; only 2*nvar values cross to the
 * host.  As the reference's Kokkos::Max/Min reductions, NaN values are
 * skipped (a variable with no other value reports -DBL_MAX / DBL_MAX); -0
 * counts as smaller than +0. */
int fst_status(fst_ctx *ctx, int array, double *min, double *max, int nvar);

/* Device-to-device copy of `bytes` bytes (rounded down to a multiple of 32)
 * with a streaming HIP kernel (32-byte elements, four in flight per thread),
 * `reps` times, on the context stream; returns the mean ms per copy.  One
 * untimed copy with 8-byte elements runs first.  Used to measure the
 * attainable HBM bandwidth and to scale rocprof byte counters. */
int fst_copy_probe(fst_ctx *ctx, size_t bytes, int reps, double *ms_per_copy);

/* ------------------------------------------------------------------------
 * Learned C-indicator (no counterpart in the reference: it replaces the
 * relaxation solve of the C-equation, updateCeq src/ceq3d.hpp:189-224, and
 * the transport of the five C variables; the artificial-viscosity terms
 * calculateCeqFaces/Grads/applyCeq src/ceq3d.hpp:226-366 stay).
 *
 * Per step, on the stage-1 state: features f0 = |grad rho| h/rho,
 * f1 = |grad p| h/p, f2 = (div u) h/a (h = min cell size) on every cell; an
 * MLP (5x5x5 stencil x 3 features -> 64 -> 64 -> 64 -> 1, GELU, BF16 MFMA,
 * FP32 accumulation) predicts y = C_hat h/rho on the cells whose 3x3x3
 * neighbourhood has a feature above tau (others: C_hat = 0).  The residual
 * check: the prediction must nearly satisfy the C-equation relaxation at
 * rest, |G - C_hat + a lap(C_hat)| h / rho <= tol max G1 (a = kap eps |dx|^2,
 * G = contact-switch source indicator*|grad rho|, lap with the reference's
 * 1/d, src/ceq3d.hpp:189-224); a cell that fails (or is not finite) takes
 * the plain C-equation's value there, the local solution of that resting
 * equation with the neighbouring predictions held.
 * The direction n is the 4th-order density gradient.  Both stages of the
 * step use these fields.
 *
 * 2D (cart2d, src/ceq.hpp:80-293, src/cart2d.cpp:201-236): the same on 5x5
 * stencils (75 inputs); a cell is evaluated when its 3x3 neighbourhood has a
 * contact source G1 (the G above, h / rho scaled) above tau; the network stands in for
 * updateCeq2D's relaxation of var(nv+1) (the contact indicator), the check
 * uses a = kap eps (dx^2 + dy^2) and the two-direction Laplacian, and the
 * tangent (-d rho/dy, d rho/dx) (4th order) takes the place of the relaxed
 * var(nv+2..nv+3) in computeCeqFlux2D.  Fields are evaluated at stage 1 of
 * each fst_step step, after the ghost fill of preStep.
 * ------------------------------------------------------------------------ */

/* Number of FP32 network parameters of the 3D network: W1[375][64], b1[64],
 * W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4 (row-major, input index
 * major; input index = channel*125 + (dz+2)*25 + (dy+2)*5 + (dx+2)). */
int fst_mlp_nparams(void);
/* The same for the ndim network (2: W1[75][64], input index = channel*25 +
 * (dy+2)*5 + (dx+2); 3: as above); -1 for another ndim. */
int fst_mlp_nparams_nd(int ndim);
/* Weights of a learned_c context (copied; required before fst_step); n is
 * fst_mlp_nparams_nd(the context's ndim). */
int fst_set_mlp_weights(fst_ctx *ctx, const float *w, size_t n);
/* Activity threshold tau (default 1e-3) and residual tolerance tol (0.05,
 * a fraction of max G1; a huge value switches the fallback off). */
int fst_learned_params(fst_ctx *ctx, float tau, float tol);
/* Ghost fill, primitives and learned fields of the current var, outside a
 * step (inspection and tests). */
int fst_learned_compute(fst_ctx *ctx);
/* Cells evaluated by the MLP, cells that fell back, and max C_hat of the
 * last field evaluation. */
int fst_learned_stats(fst_ctx *ctx, int64_t *active, int64_t *fallback, double *maxCh);
/* Download the last fields (any pointer may be NULL): C_hat [N], direction
 * [3][N] (2D: the tangent in planes 0-1, plane 2 zero), features [3][N],
 * G [N]; `plane` must equal ncells_ghost. */
int fst_learned_fields(fst_ctx *ctx, float *chat, float *dir, float *feat, float *g, size_t plane);

/* Trainer: a dataset of (stencil, C_hat h/rho) rows harvested on the device
 * from plain C-equation runs, and the network being trained (FP32 master
 * weights, BF16 MFMA forward/backward, Adam).  One HIP stream per trainer. */
typedef struct fst_mlp fst_mlp;
/* A 3D network (5x5x5 stencils). */
int fst_mlp_create(uint64_t seed, fst_mlp **out);
/* A network for ndim = 2 (5x5 stencils) or 3; harvests only from contexts of
 * that ndim. */
int fst_mlp_create_nd(int ndim, uint64_t seed, fst_mlp **out);
int fst_mlp_ndim(const fst_mlp *m);
int fst_mlp_destroy(fst_mlp *m);
const char *fst_mlp_last_error(const fst_mlp *m);
int fst_mlp_set_weights(fst_mlp *m, const float *w, size_t n);
int fst_mlp_get_weights(fst_mlp *m, float *w, size_t n);
/* Rows from the current var of a plain C-equation context (cartesian, the
 * network's ndim; 3D: ns <= 3): every active cell and every `stride`-th
 * interior cell (stride 0: active cells only).  The context's ghost cells
 * are filled first, as its next compute() would see them. */
int fst_mlp_harvest(fst_mlp *m, fst_ctx *plain, uint32_t stride, int64_t max_rows, int64_t *added);
int fst_mlp_rows(fst_mlp *m, int64_t *rows);
/* Rows [r0, r0+n): X [n][375] (2D: [n][75]; BF16 values as FP32), targets [n]. */
int fst_mlp_read_rows(fst_mlp *m, int64_t r0, int64_t n, float *X, float *T);
/* Mini-batch Adam over the dataset (seeded batch order); mean batch loss
 * per epoch into loss_per_epoch (may be NULL). */
int fst_mlp_train(fst_mlp *m, int epochs, int batch, float lr, uint64_t seed, float *loss_per_epoch);
/* One forward + backward pass of the current weights over host rows X
 * [rows][375] (2D: [rows][75]) with targets T: outputs Y [rows], the gradient of the mean
 * squared error [nparams] and the loss (any output may be NULL). */
int fst_mlp_forward_backward(fst_mlp *m, const float *X, const float *T, int64_t rows, float *Y, float *grad,
                             float *loss);

#ifdef __cplusplus
}
#endif
#endif /* FIESTA_HIP_H */
