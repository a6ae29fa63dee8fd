/*
 * oracle.h -- serial FP64 C restatement of the reference Bluebottle
 * per-timestep path.  TEST INFRASTRUCTURE ONLY: it is the checker that the
 * HIP library is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it; the product path never does.
 *
 * Each function cites the reference kernel/host function it restates.  The
 * oracle keeps its own copy of the simulation state (orc_state), laid out
 * exactly like the reference's host globals (ghosted, x-fastest arrays), so
 * that a test can copy the library's state in, run one entry point on both
 * sides, and compare array by array.
 *
 * Parity status per function: see the comment above each prototype; the
 * pins are exercised in tests/test_oracle_*.py.
 */
#ifndef BB_ORACLE_H
#define BB_ORACLE_H

#include <stddef.h>
#include "../include/bluebottle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state {
  dom_struct dom;             /* the single subdomain (== Dom) */
  BC bc;
  g_struct g;
  gradP_struct gradP;
  real rho_f, mu, nu;
  real dt, dt0, ttime, CFL;
  real pp_residual, lamb_residual, lamb_relax, lamb_cut;
  int pp_max_iter, lamb_max_iter, out_plane, stepnum, init_cond;
  real p_tDelay, g_tDelay, vel_tDelay;
  real Kp, Ki, Kd, pid_int, pid_back;
  /* Gcc arrays */
  real *p0, *p, *phi;
  int *phase, *phase_shell;
  /* Gfx/Gfy/Gfz arrays */
  real *u, *u0, *v, *v0, *w, *w0;
  real *f_x, *f_y, *f_z;
  real *diff0_u, *diff0_v, *diff0_w, *conv0_u, *conv0_v, *conv0_w;
  real *diff_u, *diff_v, *diff_w, *conv_u, *conv_v, *conv_w;
  real *u_star, *v_star, *w_star;
  int *flag_u, *flag_v, *flag_w;
  /* solver vector (no ghosts, Gcc.s3) */
  real *rhs_p;
  /* particles */
  int nparts;
  real interactionLengthRatio;
  part_struct *parts;
  dom_struct binDom;
  int coeff_stride;
  real *coeff[18];            /* pnm_re, pnm_im, phinm_re, phinm_im,
                                 chinm_re, chinm_im, then the same six
                                 for *0 and for *00 */
  /* telemetry of the last Poisson solve */
  int last_iters;
  real last_resid;
  int last_converged;
  /* 1: the reference's IMPLICIT build (Crank-Nicolson Helmholtz solves
   * instead of the explicit predictor, no diff0 history, dt without the
   * viscous limit and capped at 1.1 dt) */
  int implicit;
  /* precursor inflow planes, in the order of orc_plane_names (layouts:
   * W/E j + k*jnb, S/N k + i*knb, B/T i + j*inb of the component's grid) */
  real *plane[15];
  /* the precursor domain's forcing amplitude and integral scale */
  real turbA, turbl;
} orc_state;

/* Names of coeff[] entries, in order. */
extern const char *orc_coeff_names[18];
/* Names of plane[] entries, in order (the reference's globals). */
extern const char *orc_plane_names[15];

/* --- state management (oracle_state.c) --- */
orc_state *orc_create(const dom_struct *Dom, int nparts, int coeff_stride);
void orc_free(orc_state *s);
/* Look up an array by its reference global name; returns its element count
 * in *n and its element type in *is_int (0 = real, 1 = int). */
void *orc_array(orc_state *s, const char *name, size_t *n, int *is_int);
size_t orc_sizeof_state(void);
/* threads of the row-parallel loops (matrix assembly and products; rows are
 * independent, so the count changes no result); 0 keeps the default */
void orc_set_threads(int n);

/* --- field path (oracle_field.c) --- */
/* src/bluebottle_kernel.cu:1504-2003 via src/cuda_bluebottle.cu:2930. */
void orc_U_star_2(orc_state *s);
/* src/cuda_bluebottle.cu:1549-3420 with the BC_* kernels at
 * src/bluebottle_kernel.cu:26-1248 (PRECURSOR faces from s->plane). */
void orc_dom_BC(orc_state *s);
/* src/cuda_bluebottle.cu:4556-4820, src/bluebottle_kernel.cu:3619-3851 */
void orc_yank_turb_planes(orc_state *s, const int *bc_flow_configs,
  const real *pos, const real *vel);
void orc_dom_BC_star(orc_state *s);
void orc_dom_BC_phi(orc_state *s);
void orc_dom_BC_p(orc_state *s);
/* src/cuda_bluebottle.cu:4122-4314, src/bluebottle_kernel.cu:2133-2245. */
void orc_solvability(orc_state *s);
/* src/cuda_bluebottle.cu:3422-3497, src/bluebottle_kernel.cu:1250-1302. */
void orc_project(orc_state *s);
/* src/cuda_bluebottle.cu:3616-3671, src/bluebottle_kernel.cu:1326-1338. */
void orc_update_p(orc_state *s);
/* src/cuda_bluebottle.cu:3500-3577. */
real orc_find_dt(orc_state *s);
/* src/cuda_bluebottle.cu:3580-3613. */
void orc_store_u(orc_state *s);
/* src/cuda_bluebottle.cu:3674-3814 (PID branch included). */
void orc_compute_forcing(orc_state *s);

/* --- the precursor domain (oracle_turb.c) --- */
/* src/bluebottle_kernel.cu:3849-3889: component comp (0 u, 1 v, 2 w) at
 * the cell centres into out[Gcc.s3] */
void orc_colocate(const orc_state *s, int comp, real *out);
/* src/cuda_bluebottle.cu:3944-4017 */
real orc_compute_energy(orc_state *s);
/* src/cuda_bluebottle.cu:3817-3942 */
void orc_compute_turb_forcing(orc_state *s);
/* src/rng.c: the state after rng_init(seed) from `state`, and one double
 * (the state advanced in place) */
unsigned long long orc_rng_seeded(unsigned long long state,
  unsigned long long seed);
double orc_rng_dbl(unsigned long long *state);
/* src/domain.c:2856-3093 (field part), drawing from *rng_state */
void orc_init_turb(orc_state *s, unsigned long long *rng_state);
/* its correction and mean removal alone (src/domain.c:2922-3051) */
void orc_turb_correct(orc_state *s);
/* src/bluebottle.c:945-1016 and 1036-1081: the precursor's set-up after
 * the pushes and one step */
void orc_prepare_turb(orc_state *s);
int orc_step_turb(orc_state *s);

/* --- Poisson solve (oracle_poisson.c) --- */
/* src/bicgstab_kernel.cu:483-562 */
void orc_PP_rhs(orc_state *s);
/* src/particle_kernel.cu:840-940 (non-STEPS branch). */
void orc_part_BC_p(orc_state *s);
/* y = A x with A the 13-diagonal matrix of src/bicgstab_kernel.cu:1654-1856
 * and the offsets of src/cuda_bicgstab.cu:735-747, applied as CUSP's DIA
 * SpMV would (x, y: no-ghost vectors of length Gcc.s3). */
void orc_pp_matvec(const orc_state *s, const real *x, real *y);
/* The 13 DIA coefficient rows for cell C (values[C + d*pitch], d=0..12). */
void orc_pp_coeffs(const orc_state *s, int C, real vals[13]);
/* Jacobi-preconditioned BiCGSTAB with CUSP's monitor semantics
 * (src/cuda_bicgstab.cu:1016-1018; CUSP itself is not vendored, so this is
 * the published right-preconditioned algorithm; its iterates are pinned
 * against SciPy's implementation of the same method, its converged
 * solution by the residual). */
int orc_bicgstab(const orc_state *s, const real *b, real *x, int max_iter,
  real rel_tol, real *resid_out, int *converged);
/* The whole entry point: src/cuda_bicgstab.cu:715-1073. */
void orc_PP_bicgstab(orc_state *s);

/* --- particles (oracle_particle.c) --- */
/* src/cuda_particle.cu:357-896 + src/particle_kernel.cu:25-533. */
void orc_build_cages(orc_state *s);
/* src/cuda_particle.cu:899-992 + src/particle_kernel.cu:535-838. */
void orc_part_BC(orc_state *s);
void orc_part_BC_star(orc_state *s);
/* src/cuda_particle.cu:1019-1042 + src/particle_kernel.cu:942-1018. */
void orc_part_p_fill(orc_state *s);
/* src/cuda_bluebottle.cu:4822-4900 + src/bluebottle_kernel.cu:3916-4022. */
void orc_parts_internal(orc_state *s);
/* src/cuda_particle.cu:1046-1081. */
void orc_store_coeffs(orc_state *s);
/* src/cuda_bluebottle.cu:4317-4554 + src/bluebottle_kernel.cu:2247-3617. */
void orc_move_parts_sub(orc_state *s);
void orc_move_parts(orc_state *s);
/* Lamb analytic velocity at (r, theta, phi) about particle P
 * (src/particle_kernel.cu:1311-1369). */
void orc_lamb_vel(const orc_state *s, int order, real a, real r, real theta,
  real phi, int P, real *Ux, real *Uy, real *Uz);

/* --- quadrature (oracle_quad.c) --- */
/* src/cuda_quadrature.cu:64-313 + src/quadrature_kernel.cu:25-576. */
void orc_Lamb(orc_state *s);
/* src/cuda_quadrature.cu:316-398 + src/quadrature_kernel.cu:578-654 with
 * the cross-particle race removed (per-particle scratch). */
real orc_lamb_err(orc_state *s);
/* cuda_quad_check_nodes and cuda_quad_interp on the node tables below:
 * node flags into parts[].nodes; node values nparts x 26, node fastest. */
void orc_quad_check_nodes(orc_state *s, const real *theta, const real *phi);
void orc_quad_interp(orc_state *s, const real *theta, const real *phi,
  real *pp, real *ur, real *ut, real *up);
/* Node tables of the 26-point Lebedev rule (src/cuda_quadrature.cu:86-157). */
void orc_lebedev_nodes(real theta[26], real phi[26], real weight[26]);
/* Helpers exposed for closed-form tests. */
real orc_Nnm(int n, int m);
real orc_Pnm(int n, int m, real theta);
void orc_xyz2rtp(real x, real y, real z, real *r, real *theta, real *phi);

/* --- Crank-Nicolson velocity solves (oracle_helmholtz.c) --- */
/* cuda_{u,v,w}star_helmholtz, src/cuda_bicgstab.cu:44-714 (comp 0, 1, 2). */
void orc_star_helmholtz(orc_state *s, int comp);
/* The 13 DIA diagonals of the component's operator, values[C + d*n]
 * (src/bicgstab_kernel.cu:643-1652 in the driver's order). */
void orc_helmholtz_coeffs(const orc_state *s, int comp, real *values);
void orc_helmholtz_matvec(const orc_state *s, int comp, const real *values,
  const real *x, real *y);
/* Diagonally preconditioned CG with CUSP's monitor semantics (CUSP is not
 * vendored: the iterates are pinned against SciPy's preconditioned CG, the
 * solution by its residual). */
int orc_pcg(const orc_state *s, int comp, const real *values, const real *b,
  real *x, int max_iter, real rel_tol, real *resid_out, int *converged);

/* --- one full time step (oracle_step.c): src/bluebottle.c:598-711 --- */
/* Returns the number of Lamb sub-iterations taken. */
int orc_step(orc_state *s);
/* The set-up sequence before the loop: src/bluebottle.c:445-545. */
void orc_prepare(orc_state *s);

#ifdef __cplusplus
}
#endif

#endif
