/*
 * bluebottle.h -- C ABI of the MI355X-native Bluebottle per-timestep hot path.
 *
 * The library libbluebottle_hip.so exports the reference's `cuda_*` host
 * entry points with their exact signatures.  As in the reference, they take
 * (almost) no arguments and communicate through global variables that the
 * caller (the time-stepping driver) defines.  The struct layouts below are
 * byte-identical to the reference's on x86-64 with `real == double`, because
 * the driver memcpy's and fwrite's them raw.
 *
 * Reference interfaces this header replaces:
 *   src/bluebottle.h:42-3391   (real, constants, dom globals, BC, g, gradP,
 *                               cuda_* field entry points)
 *   src/domain.h:39-254        (grid_info, dom_struct, nsubdom)
 *   src/particle.h:41-1149     (cage_struct, part_struct, particle globals,
 *                               cuda_* particle/quadrature entry points)
 *   src/cuda_bicgstab.h:41-150 (Krylov entry points)
 *
 * Everything here is plain C: no HIP or torch types.
 */
#ifndef BLUEBOTTLE_HIP_H
#define BLUEBOTTLE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/bluebottle.h:42-46 -- the reference is built with -DDOUBLE. */
typedef double real;

/* src/bluebottle.h:57-352 */
#define PI 3.14159265358979323846
#define DIV_ST 1e-10
#define FILE_NAME_SIZE 256
#define CHAR_BUF_SIZE 256
#define DOM_BUF 1
#define PERIODIC 0
#define DIRICHLET 1
#define NEUMANN 2
#define PRECURSOR 3
#define HOMOGENEOUS 10
#define WEST 0
#define EAST 1
#define SOUTH 2
#define NORTH 3
#define BOTTOM 4
#define TOP 5
#define MAX_THREADS_1D 128
#define MAX_THREADS_DIM 16
#define MASTER 0
#define QUIESCENT 0
#define SHEAR 1
#define CHANNEL 2
/* src/particle.h:41,53 */
#define NNODES 26
#define MAX_NEIGHBORS 12

/* src/domain.h:39-88 -- 48 ints, 192 bytes. */
typedef struct grid_info {
  int is, ie, in, isb, ieb, inb;
  int js, je, jn, jsb, jeb, jnb;
  int ks, ke, kn, ksb, keb, knb;
  int s1, s1b, s2, s2b, s3, s3b;
  int _is, _ie, _in, _isb, _ieb, _inb;
  int _js, _je, _jn, _jsb, _jeb, _jnb;
  int _ks, _ke, _kn, _ksb, _keb, _knb;
  int _s1, _s1b, _s2, _s2b, _s3, _s3b;
} grid_info;

/* src/domain.h:186-212 -- 912 bytes. */
typedef struct dom_struct {
  grid_info Gcc;
  grid_info Gfx;
  grid_info Gfy;
  grid_info Gfz;
  real xs; real xe; real xl; int xn; real dx;
  real ys; real ye; real yl; int yn; real dy;
  real zs; real ze; real zl; int zn; real dz;
  int E; int W; int N; int S; int T; int B;
} dom_struct;

/* src/bluebottle.h:2061-2071 -- 72 bytes. */
typedef struct g_struct {
  real x; real xm; real xa;
  real y; real ym; real ya;
  real z; real zm; real za;
} g_struct;

/* src/bluebottle.h:2099-2190 -- 720 bytes. */
typedef struct BC {
  int pW; real pWD; int pE; real pED; int pS; real pSD;
  int pN; real pND; int pB; real pBD; int pT; real pTD;
  int uW; real uWDm; real uWD; real uWDa;
  int uE; real uEDm; real uED; real uEDa;
  int uS; real uSDm; real uSD; real uSDa;
  int uN; real uNDm; real uND; real uNDa;
  int uB; real uBDm; real uBD; real uBDa;
  int uT; real uTDm; real uTD; real uTDa;
  int vW; real vWDm; real vWD; real vWDa;
  int vE; real vEDm; real vED; real vEDa;
  int vS; real vSDm; real vSD; real vSDa;
  int vN; real vNDm; real vND; real vNDa;
  int vB; real vBDm; real vBD; real vBDa;
  int vT; real vTDm; real vTD; real vTDa;
  int wW; real wWDm; real wWD; real wWDa;
  int wE; real wEDm; real wED; real wEDa;
  int wS; real wSDm; real wSD; real wSDa;
  int wN; real wNDm; real wND; real wNDa;
  int wB; real wBDm; real wBD; real wBDa;
  int wT; real wTDm; real wTD; real wTDa;
  real dsW; real dsE; real dsS; real dsN; real dsB; real dsT;
} BC;

/* src/bluebottle.h:2312-2322 -- 72 bytes. */
typedef struct gradP_struct {
  real x; real xm; real xa;
  real y; real ym; real ya;
  real z; real zm; real za;
} gradP_struct;

/* src/particle.h:190-210 -- 18 ints. */
typedef struct cage_struct {
  int cx; int cy; int cz;
  int is; int ibs; int ie; int ibe; int in;
  int js; int jbs; int je; int jbe; int jn;
  int ks; int kbs; int ke; int kbe; int kn;
} cage_struct;

/* src/particle.h:244-326 -- 928 bytes. */
typedef struct part_struct {
  real r;
  real x; real y; real z;
  real x0; real y0; real z0;
  real u; real v; real w;
  real u0; real v0; real w0;
  real udot; real vdot; real wdot;
  real udot0; real vdot0; real wdot0;
  real axx; real axy; real axz;
  real ayx; real ayy; real ayz;
  real azx; real azy; real azz;
  real ox; real oy; real oz;
  real ox0; real oy0; real oz0;
  real oxdot; real oydot; real ozdot;
  real oxdot0; real oydot0; real ozdot0;
  real Fx; real Fy; real Fz;
  real Lx; real Ly; real Lz;
  real aFx; real aFy; real aFz;
  real aLx; real aLy; real aLz;
  real kFx; real kFy; real kFz;
  real iFx; real iFy; real iFz;
  real iLx; real iLy; real iLz;
  int nodes[NNODES];
  real rho;
  real E;
  real sigma;
  cage_struct cage;
  int order;
  real rs;
  int ncoeff;
  real spring_k;
  real spring_x; real spring_y; real spring_z;
  real spring_l;
  int translating;
  int rotating;
  int bin;
  real St[MAX_NEIGHBORS];
  int iSt[MAX_NEIGHBORS];
  real e_dry;
  real coeff_fric;
} part_struct;

#if defined(__cplusplus)
static_assert(sizeof(grid_info) == 192, "grid_info layout");
static_assert(sizeof(dom_struct) == 912, "dom_struct layout");
static_assert(sizeof(BC) == 720, "BC layout");
static_assert(sizeof(g_struct) == 72, "g_struct layout");
static_assert(sizeof(gradP_struct) == 72, "gradP_struct layout");
static_assert(sizeof(part_struct) == 928, "part_struct layout");
#else
_Static_assert(sizeof(grid_info) == 192, "grid_info layout");
_Static_assert(sizeof(dom_struct) == 912, "dom_struct layout");
_Static_assert(sizeof(BC) == 720, "BC layout");
_Static_assert(sizeof(g_struct) == 72, "g_struct layout");
_Static_assert(sizeof(gradP_struct) == 72, "gradP_struct layout");
_Static_assert(sizeof(part_struct) == 928, "part_struct layout");
#endif

/* ------------------------------------------------------------------------
 * Global state.  Defined by the driver (src/bluebottle.c:29-187,
 * src/particle.c:29-75, src/domain.c:27); read and written by the library.
 * Host arrays are ghosted, x-fastest: C = i + j*s1b + k*s2b.  The `_x`
 * tables hold one device pointer per subdomain; this library runs one.
 * ---------------------------------------------------------------------- */
extern int nsubdom;
extern int dev_start;
extern int dev_end;
extern dom_struct *dom;
extern dom_struct **_dom;
extern dom_struct Dom;
extern real *p0, *p, *phi;
extern real **_p0, **_p, **_phi;
extern real *u, *u0, *v, *v0, *w, *w0;
extern real **_u, **_u0, **_v, **_v0, **_w, **_w0;
extern real *f_x, *f_y, *f_z;
extern real **_f_x, **_f_y, **_f_z;
extern real *diff0_u, *diff0_v, *diff0_w;
extern real **_diff0_u, **_diff0_v, **_diff0_w;
extern real *conv0_u, *conv0_v, *conv0_w;
extern real **_conv0_u, **_conv0_v, **_conv0_w;
extern real *diff_u, *diff_v, *diff_w;
extern real **_diff_u, **_diff_v, **_diff_w;
extern real *conv_u, *conv_v, *conv_w;
extern real **_conv_u, **_conv_v, **_conv_w;
extern real *u_star, *v_star, *w_star;
extern real **_u_star, **_v_star, **_w_star;
/* precursor inflow planes (src/bluebottle.c:92-121): W/E planes indexed
 * j + k*jnb, S/N planes k + i*knb, B/T planes i + j*inb of the
 * component's grid */
extern real *u_WE, *u_SN_S, *u_SN_N, *u_BT_B, *u_BT_T;
extern real *v_WE_W, *v_WE_E, *v_SN, *v_BT_B, *v_BT_T;
extern real *w_WE_W, *w_WE_E, *w_SN_S, *w_SN_N, *w_BT;
extern real **_u_WE, **_u_SN_S, **_u_SN_N, **_u_BT_B, **_u_BT_T;
extern real **_v_WE_W, **_v_WE_E, **_v_SN, **_v_BT_B, **_v_BT_T;
extern real **_w_WE_W, **_w_WE_E, **_w_SN_S, **_w_SN_N, **_w_BT;
extern real **_rhs_p;
extern real duration, ttime, vel_tDelay, p_tDelay, g_tDelay;
extern real dt, dt0, CFL;
extern int pp_max_iter;
extern real pp_residual;
extern int lamb_max_iter;
extern real lamb_residual, lamb_relax, lamb_cut;
extern int out_plane;
extern int stepnum;
extern g_struct g;
extern real rho_f, mu, nu;
extern BC bc;
extern int init_cond;
extern gradP_struct gradP;
extern long int cpumem, gpumem;
extern real pid_int, pid_back, Kp, Ki, Kd;
/* the precursor domain's linear-forcing amplitude and integral scale
 * (src/bluebottle.h:2342-2360) */
extern real turbA, turbl;

extern int *phase, *phase_shell;
extern int **_phase, **_phase_shell;
extern int *flag_u, *flag_v, *flag_w;
extern int **_flag_u, **_flag_v, **_flag_w;
extern int nparts;
extern real interactionLengthRatio;
extern part_struct *parts;
extern part_struct **_parts;
extern dom_struct binDom;
extern dom_struct *_binDom;
extern int coeff_stride;
extern real *pnm_re, *pnm_im, *phinm_re, *phinm_im, *chinm_re, *chinm_im;
extern real *pnm_re0, *pnm_im0, *phinm_re0, *phinm_im0, *chinm_re0,
  *chinm_im0;
extern real *pnm_re00, *pnm_im00, *phinm_re00, *phinm_im00, *chinm_re00,
  *chinm_im00;
extern real **_pnm_re, **_pnm_im, **_phinm_re, **_phinm_im, **_chinm_re,
  **_chinm_im;
extern real **_pnm_re0, **_pnm_im0, **_phinm_re0, **_phinm_im0, **_chinm_re0,
  **_chinm_im0;
extern real **_pnm_re00, **_pnm_im00, **_phinm_re00, **_phinm_im00,
  **_chinm_re00, **_chinm_im00;

/* Provided by the driver (src/recorder.c:1927-1953): appends
 * `stepnum ttime dt niter resid` to record/<name>. */
void recorder_bicgstab(char *name, int niter, real resid);

/* ------------------------------------------------------------------------
 * Entry points exported by libbluebottle_hip.so.
 * ---------------------------------------------------------------------- */

/* Memory and transfers.  src/bluebottle.h:2373-2450 (cuda_dom_malloc,
 * cuda_dom_push, cuda_dom_pull, cuda_dom_free), implemented at
 * src/cuda_bluebottle.cu:36-1547; src/particle.h:1097-1120
 * (cuda_part_malloc/push/pull/free), src/cuda_particle.cu:28-354. */
void cuda_dom_malloc(void);
void cuda_dom_push(void);
void cuda_dom_pull(void);
void cuda_dom_free(void);
void cuda_part_malloc(void);
void cuda_part_push(void);
void cuda_part_pull(void);
void cuda_part_free(void);

/* Ghost-cell boundary conditions.  src/cuda_bluebottle.cu:1549 (dom_BC),
 * 2302 (dom_BC_star), 3030 (dom_BC_phi), 3226 (dom_BC_p). */
void cuda_dom_BC(void);
void cuda_dom_BC_star(void);
void cuda_dom_BC_phi(void);
void cuda_dom_BC_p(void);

/* Explicit Adams-Bashforth predictor.  src/cuda_bluebottle.cu:2930-3026. */
void cuda_U_star_2(void);

/* Crank-Nicolson Helmholtz solves of the IMPLICIT build (Jacobi-PCG).
 * src/cuda_bicgstab.cu:44-714, src/cuda_bicgstab.h:41-90. */
void cuda_ustar_helmholtz(int rank);
void cuda_vstar_helmholtz(int rank);
void cuda_wstar_helmholtz(int rank);

/* Pressure-Poisson solve.  src/cuda_bicgstab.cu:715-1073. */
void cuda_PP_bicgstab(int rank);

/* Solvability correction of the predicted velocity.
 * src/cuda_bluebottle.cu:4122-4314. */
void cuda_solvability(void);

/* Projection, pressure update, time step, history shift, forcing.
 * src/cuda_bluebottle.cu:3422 (project), 3616 (update_p), 3500 (find_dt),
 * 3580 (store_u), 3674 (compute_forcing). */
void cuda_project(void);
void cuda_update_p(void);
real cuda_find_dt(void);
void cuda_store_u(void);
void cuda_compute_forcing(real *pid_int, real *pid_back, real Kp, real Ki,
  real Kd);

/* Particles.  src/cuda_particle.cu:357 (build_cages), 899 (part_BC),
 * 947 (part_BC_star), 996 (part_BC_p), 1019 (part_p_fill),
 * 1046 (store_coeffs); src/cuda_bluebottle.cu:4317 (move_parts_sub),
 * 4435 (move_parts), 4822 (parts_internal). */
void cuda_build_cages(void);
void cuda_part_BC(void);
void cuda_part_BC_star(void);
void cuda_part_BC_p(int dev);
void cuda_part_p_fill(void);
void cuda_store_coeffs(void);
void cuda_move_parts_sub(void);
void cuda_move_parts(void);
void cuda_parts_internal(void);

/* Lamb-coefficient quadrature.  src/cuda_quadrature.cu:30 (check_nodes),
 * 46 (quad_interp), 64 (Lamb), 316 (lamb_err). */
void cuda_quad_check_nodes(int dev, real *node_t, real *node_p, int nnodes);
void cuda_quad_interp(int dev, real *node_t, real *node_p, int nnodes,
  real *pp, real *ur, real *ut, real *up);
void cuda_Lamb(void);
real cuda_lamb_err(void);

/* ------------------------------------------------------------------------
 * Additions with no counterpart in the reference (solver selection and
 * telemetry the reference only writes to record/solver_expd.rec).
 * ---------------------------------------------------------------------- */

/* Pressure-Poisson solver mode. */
#define BB_SOLVER_FP64  0  /* plain FP64 Jacobi-BiCGSTAB (reference form) */
#define BB_SOLVER_MIXED 1  /* FP32 Jacobi-BiCGSTAB inside FP64 refinement */
/* FP32 flexible BiCGSTAB preconditioned by the learned U-Net inside FP64
 * refinement (Jacobi BiCGSTAB until a network is trained, and from any
 * iterate where |r| falls by less than 10 % over 5 iterations) */
#define BB_SOLVER_LEARNED 2
void bb_set_solver_mode(int mode);
int bb_get_solver_mode(void);

/* Statistics of the most recent cuda_PP_bicgstab call: total Krylov
 * iterations (inner iterations for the mixed mode), outer refinement
 * steps (1 for FP64), and the final relative residual the convergence test
 * used. */
typedef struct bb_solve_stats {
  int iterations;
  int outer_steps;
  real residual;
  int converged;
} bb_solve_stats;
void bb_get_solve_stats(bb_solve_stats *out);

/* Returns a short string naming the HIP device the library runs on. */
const char *bb_device_name(void);

/* Non-convergence of the pressure solve prints the reference's message and
 * exits (on = 1, the default, as src/cuda_bicgstab.cu:1050-1056 does);
 * on = 0 keeps the message and returns. */
void bb_set_exit_on_nonconvergence(int on);

/* Time-integration build the library behaves as.  The reference selects it
 * at compile time (-DIMPLICIT: Crank-Nicolson diffusion through
 * cuda_{u,v,w}star_helmholtz, no diff0 history arrays, find_dt without the
 * viscous limit and capped at 1.1 dt, src/bluebottle.c:60-67,
 * src/cuda_bluebottle.cu:3526-3574).  mode = 1: IMPLICIT, 0: explicit,
 * -1 (default): IMPLICIT exactly when the driver defines no diff0 arrays.
 * Takes effect at the next cuda_dom_malloc. */
void bb_set_implicit(int mode);
/* 1 when the allocated domain runs as the IMPLICIT build. */
int bb_is_implicit(void);



/* Blocks until all work the library queued has finished. */
void bb_synchronize(void);

/* The HIP stream (hipStream_t) all entry points run on, as void *. */
void *bb_stream(void);

/* Krylov convergence is checked on the host every `n` iterations (the
 * device keeps iterating scalars; iterations past convergence are no-ops). */
void bb_set_poll_interval(int n);

/* Mixed mode: inner FP32 relative tolerance, inner and outer iteration
 * caps (values <= 0 leave a parameter unchanged). */
void bb_set_mixed_params(real inner_tol, int inner_max, int outer_max);

/* Mixed and learned modes: the Jacobi BiCGSTAB of the FP32 inner solve in
 * three sweeps per iteration (3, the default: the textbook iteration) or
 * two (2: the update sweep folded into the next iteration's first sweep,
 * its dots taken by linearity in the second; 58 instead of 71 B/cell).
 * Returns the previous setting. */
int bb_set_inner_sweeps(int n);

/* Per-kernel event timing on the library stream.  Kinds index the kernel
 * families (predictor, BCs, Poisson set-up, the four BiCGSTAB sweeps in
 * FP64 and FP32, refinement residual and correction, projection, pressure
 * update, dt, cages, particle BCs, quadrature, particle motion, misc; the
 * learned form's network application, its full-resolution 8->8
 * convolutions, its two products with A and its vector updates), in that
 * order, see BB_PROF_* below. */
enum {
  BB_PROF_USTAR = 0, BB_PROF_BC, BB_PROF_PP_SETUP, BB_PROF_PACK,
  BB_PROF_SPMV_P, BB_PROF_SPMV_S, BB_PROF_UPDATE_XR, BB_PROF_UPDATE_P,
  BB_PROF_FINAL, BB_PROF_RESID64, BB_PROF_CORRECT64, BB_PROF_SPMV_P32,
  BB_PROF_SPMV_S32, BB_PROF_UPDATE_XR32, BB_PROF_UPDATE_P32,
  BB_PROF_PROJECT, BB_PROF_UPDATE_PRESSURE, BB_PROF_FIND_DT, BB_PROF_CAGES,
  BB_PROF_PART_BC, BB_PROF_QUADRATURE, BB_PROF_MOVE_PARTS, BB_PROF_MISC,
  BB_PROF_UNET, BB_PROF_UNET_CONV0, BB_PROF_FSPMV, BB_PROF_FVEC,
  BB_PROF_SWEEP_A32, BB_PROF_SWEEP_B32,
  BB_PROF_NKINDS
};
void bb_prof_enable(int on);
void bb_prof_read(int kind, real *ms, long *count);
void bb_prof_reset(void);

/* A plain streaming device-to-device copy of `bytes`, `reps` times on the
 * library stream; returns GB/s counting every byte read and written.  The
 * reference point for the kernels' HBM rates and for calibrating hardware
 * traffic counters (one read and one write of `bytes` per launch). */
real bb_copy_bandwidth(long bytes, int reps);

/* Device buffers for the entry points that take device pointers
 * (cuda_quad_check_nodes, cuda_quad_interp), allocated on the library's
 * device; the copies go through the library's pinned staging buffer on its
 * stream and return when done. */
void *bb_device_alloc(size_t bytes);
void bb_device_free(void *ptr);
void bb_copy_to_device(void *dev, const void *host, size_t bytes);
void bb_copy_to_host(void *host, const void *dev, size_t bytes);

/* ---------------------------------------------------------------------
 * The learned preconditioner (no counterpart in the reference, whose
 * Poisson solve is diagonally preconditioned: src/cuda_bicgstab.cu:
 * 1016-1018).  A 3-level 3-D U-Net (8/16/32 channels, 3x3x3 convolutions,
 * BF16 weights and activations, FP32 accumulation), trained with Adam on
 * (normalised residual, correction) pairs the mixed form's Jacobi inner
 * solves produce.
 * ------------------------------------------------------------------- */
/* keep the pairs of the next `capacity` inner solves (0: stop and free) */
void bb_learn_collect(int capacity);
int bb_learn_pairs(void);
/* `iters` Adam steps at rate `lr` on pairs drawn with `seed`; returns the
 * mean loss |A (M r - e)|^2 / |r|^2 of the last tenth (-1: nothing to
 * train on, or grid sizes not divisible by 4).  Enables the network. */
real bb_learn_train(int iters, real lr, unsigned long seed);
void bb_learn_enable(int on);
/* the loss of stored pair k under the current network (-1: no such pair) */
real bb_learn_pair_loss(int k);
/* fallback test: every `window` iterations, |r| must fall below `drop`
 * times its value `window` iterations earlier (defaults 5 and 0.9); after
 * a fallback the next `skip` inner solves run Jacobi BiCGSTAB directly,
 * `skip` doubling with each fallback up to 1024 and returning to its start
 * after a retraining or a solve of at least `window` iterations that the
 * network finishes
 * (default 8; values < 0 leave a setting unchanged) */
void bb_learn_set_fallback(int window, real drop, int skip);
void bb_learn_stats(long *applications, long *fallbacks, long *flex_iters,
  long *jacobi_iters, long *skipped, real *train_seconds);
void bb_learn_reset_stats(void);
/* network weights (FP32 masters, layer by layer, each [cout][cin][27]) */
long bb_unet_num_params(void);
/* the factor of the network's output in M(r) = D^-1 r + factor * net(r) on
 * fluid cells (h^2 / 64 for the current domain) */
real bb_unet_output_scale(void);
void bb_unet_init(unsigned long seed);
void bb_unet_set_params(const float *host, long n);
void bb_unet_get_params(float *host, long n);
void bb_unet_get_grads(float *host, long n);
/* out = M(r) for device vectors of Gcc.s3 floats and one operator-flag
 * byte per cell (BB_F* bits, solid 64) */
void bb_unet_apply(const float *r, const unsigned char *flags, float *out);
/* loss of one pair and the weight gradients; lr > 0 also steps Adam */
real bb_unet_loss_grad(const float *r, const float *e,
  const unsigned char *flags, real lr);

/* No counterpart in the reference.  cuda_lamb_err for callers that cannot
 * allocate device memory (code under stream capture, serving loops,
 * harnesses that forbid hipMalloc): the largest relative change of the
 * Lamb coefficients (the value cuda_lamb_err returns) for `np` particles
 * of `stride` coefficients each, with the same copy of the current
 * coefficients into the previous ones.  `cur` and `prev` hold the device
 * arrays in the order p_nm re, im, phi_nm re, im, chi_nm re, im (the
 * reference's _pnm_re, _pnm_re0, ... of src/particle.c:29-75 for the
 * driver's state).  It allocates nothing: its per-particle scratch is a
 * static device array of 32768 + 64 doubles (256.5 KiB, one per device,
 * held until the process ends), so np is limited to 32768 and stride to
 * 21; beyond either it prints why and returns -1.  Runs on the library
 * stream when the library has one on the current device, else on the
 * null stream, and returns when done.  Calls that overlap on one device
 * race on the shared scratch. */
real bb_lamb_err_noalloc(int np, int stride, real cut, real *const cur[6],
  real *const prev[6]);

/* ---------------------------------------------------------------------
 * Precursor inflow and the precursor domain, on one GPU.
 * ------------------------------------------------------------------- */
/* src/cuda_bluebottle.cu:536-1415 (src/bluebottle.h:2411, 2447): copy the
 * inflow planes that bc_configs selects host -> device / device -> host */
void cuda_dom_turb_planes_push(int *bc_configs);
void cuda_dom_turb_planes_pull(int *bc_configs);
/* src/cuda_bluebottle.cu:4556-4820 (src/bluebottle.h:2926): sample the
 * planes at bc_plane_pos out of the current field, plus vel */
void cuda_yank_turb_planes(int *bc_flow_configs, real *bc_plane_pos, real *vel);
/* src/cuda_bluebottle.cu:3817-3943 (src/bluebottle.h:2877): add the
 * linear forcing turbA (k0/k) (u - <u>) to f_x, f_y, f_z, with k the
 * kinetic energy of the cell-centred velocity and k0 = 13.5 turbA^2
 * 0.0361 turbl^2 */
void cuda_compute_turb_forcing(void);
/* src/cuda_bluebottle.cu:3944-4017 (src/bluebottle.h:3341): k, half the
 * mean over the cells of |u|^2 at the cell centres */
real cuda_compute_energy(void);
/* src/cuda_bluebottle.cu:4019-4120 (src/bluebottle.h:3355-3391): the face
 * velocity averaged to the cell centres into a device array of Gcc.s3
 * values (no ghosts); _dom is the device copy of the subdomain, _dom[0] */
void cuda_colocate_Gfx(real *_u, real *_u_co, dom_struct *_dom);
void cuda_colocate_Gfy(real *_v, real *_v_co, dom_struct *_dom);
void cuda_colocate_Gfz(real *_w, real *_w_co, dom_struct *_dom);

/* ---------------------------------------------------------------------
 * Outside this library's path.  Exported so that the reference driver
 * links unchanged; each prints which path it belongs to and exits with
 * EXIT_FAILURE.  The -DTEST checks (src/bluebottle.h:2616, 2628, 2753,
 * 2766, 2780, 2792) and helpers the step does not call (2704, 2717, 2943).
 * ------------------------------------------------------------------- */
void cuda_U_star_test_exp(void);
void cuda_U_star_test_cos(void);
void cuda_project_test(void);
void cuda_BC_test(void);
void cuda_quad_interp_test(void);
void cuda_lamb_test(void);
void cuda_PP_bicgstab_jacobi(void);
void cuda_div_U(void);
void cuda_collisions(void);

#ifdef __cplusplus
}
#endif

#endif /* BLUEBOTTLE_HIP_H */
