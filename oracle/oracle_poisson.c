/*
 * oracle_poisson.c -- serial FP64 restatement of the pressure-Poisson
 * solve of src/cuda_bicgstab.cu:715-1073.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference assembles a 13-diagonal CUSP DIA matrix (offsets at
 * src/cuda_bicgstab.cu:735-747, values by coeffs_init, coeffs,
 * coeffs_periodic_{W,E,S,N,B,T}, coeffs_particle in
 * src/bicgstab_kernel.cu:1654-1856) and calls
 * cusp::krylov::bicgstab(A, x, b, monitor, diagonal(A)).  CUSP is not
 * vendored (version unpinned), so the solver below is the published
 * right-Jacobi-preconditioned BiCGSTAB (van der Vorst 1992) in the form
 * CUSP implements it, with its monitor semantics: stop when the recursively
 * updated ||r|| <= rel_tol * ||b|| (checked on r at the loop head and on s
 * in mid-iteration) or when the iteration count reaches the limit.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "oracle_internal.h"

/* src/bicgstab_kernel.cu:483-562 */
void orc_PP_rhs(orc_state *s)
{
  const dom_struct *d = &s->dom;
  const grid_info Gc = d->Gcc, Gx = d->Gfx, Gy = d->Gfy, Gz = d->Gfz;
  const real ddx = 1. / d->dx, ddy = 1. / d->dy, ddz = 1. / d->dz;
  for(int k = Gc.ks; k < Gc.ke; k++) for(int j = Gc.js; j < Gc.je; j++)
  for(int i = Gc.is; i < Gc.ie; i++) {
    real r = (s->u_star[OIDX(Gx, i+1, j, k)] - s->u_star[OIDX(Gx, i, j, k)]) * ddx;
    r += (s->v_star[OIDX(Gy, i, j+1, k)] - s->v_star[OIDX(Gy, i, j, k)]) * ddy;
    r += (s->w_star[OIDX(Gz, i, j, k+1)] - s->w_star[OIDX(Gz, i, j, k)]) * ddz;
    r *= s->rho_f / s->dt;
    int C = (i-DOM_BUF) + (j-DOM_BUF)*Gc.s1 + (k-DOM_BUF)*Gc.s2;
    s->rhs_p[C] = r;
  }
}

/* src/particle_kernel.cu:840-940 (non-STEPS branch), launched with p0 as the
 * pressure by src/cuda_particle.cu:996-1015. */
void orc_part_BC_p(orc_state *s)
{
  const dom_struct *d = &s->dom;
  const grid_info Gc = d->Gcc;
  const int stride = s->coeff_stride;
  const real *pnm_re = s->coeff[0], *pnm_im = s->coeff[1];
  const real *phinm_re = s->coeff[2], *phinm_im = s->coeff[3];
  const real mu = s->mu, nu = s->nu, rho_f = s->rho_f;
  for(int tk = Gc.ks; tk < Gc.ke; tk++) for(int tj = Gc.js; tj < Gc.je; tj++)
  for(int i = Gc.is; i < Gc.ie; i++) {
    int CC = OIDX(Gc, i, tj, tk);
    int C = (i-DOM_BUF) + (tj-DOM_BUF)*Gc.s1 + (tk-DOM_BUF)*Gc.s2;
    int P = s->phase[CC];
    real a, X, Y, Z, ox, oy, oz, udot, vdot, wdot;
    int order;
    if(P > -1) {
      const part_struct *pt = &s->parts[P];
      a = pt->r; X = pt->x; Y = pt->y; Z = pt->z; order = pt->order;
      ox = pt->ox; oy = pt->oy; oz = pt->oz;
      udot = pt->udot; vdot = pt->vdot; wdot = pt->wdot;
    } else {
      P = 0;
      a = (d->dx + d->dy + d->dz) / 3.;
      X = (i-0.5) * d->dx + d->xs + a;
      Y = (tj-0.5) * d->dy + d->ys + a;
      Z = (tk-0.5) * d->dz + d->zs + a;
      order = 0;
      ox = oy = oz = 0; udot = vdot = wdot = 0;
    }
    real x = (i-0.5) * d->dx + d->xs - X;
    if(x <= 2.*a-d->xl) x += d->xl;
    if(x >= d->xl-2.*a) x -= d->xl;
    real y = (tj-0.5) * d->dy + d->ys - Y;
    if(y <= 2.*a-d->yl) y += d->yl;
    if(y >= d->yl-2.*a) y -= d->yl;
    real z = (tk-0.5) * d->dz + d->zs - Z;
    if(z <= 2.*a-d->zl) z += d->zl;
    if(z >= d->zl-2.*a) z -= d->zl;
    real r, theta, phi;
    orc_xyz2rtp(x, y, z, &r, &theta, &phi);
    real ar = a / r;
    real ra = r / a;
    real pp_tmp = orc_X(0, theta, phi, pnm_re, pnm_im, P, stride);
    for(int n = 1; n <= order; n++) {
      pp_tmp += (1.-0.5*n*(2.*n-1.)/(n+1.)*pow(ar,2.*n+1.))*pow(ra,n)
        * orc_X(n, theta, phi, pnm_re, pnm_im, P, stride);
      pp_tmp -= n*(2.*n-1.)*(2.*n+1.)/(n+1.)*pow(ar,n+1.)
        * orc_X(n, theta, phi, phinm_re, phinm_im, P, stride);
    }
    pp_tmp *= mu*nu/(a*a);
    real ocrossr2 = (oy*z - oz*y) * (oy*z - oz*y);
    ocrossr2 += (ox*z - oz*x) * (ox*z - oz*x);
    ocrossr2 += (ox*y - oy*x) * (ox*y - oy*x);
    real rhoV = rho_f;
    real accdotr = (-s->gradP.x/rhoV - udot)*x + (-s->gradP.y/rhoV - vdot)*y
      + (-s->gradP.z/rhoV - wdot)*z;
    pp_tmp += 0.5 * rho_f * ocrossr2 + rho_f * accdotr;
    int sh = s->phase_shell[CC];
    s->rhs_p[C] = (real) sh * s->rhs_p[C] + (real) (1 - sh) * (pp_tmp - s->p0[CC]);
    s->rhs_p[C] = (real) (s->phase[CC] < 0 && sh) * s->rhs_p[C];
  }
}

/* The 13 DIA values of row C, built exactly as the assembly kernels do. */
void orc_pp_coeffs(const orc_state *s, int C, real v[13])
{
  const dom_struct *d = &s->dom;
  const grid_info Gc = d->Gcc, Gx = d->Gfx, Gy = d->Gfy, Gz = d->Gfz;
  const real ddx = 1. / (d->dx * d->dx);
  const real ddy = 1. / (d->dy * d->dy);
  const real ddz = 1. / (d->dz * d->dz);
  int i0 = C % Gc.s1, j0 = (C / Gc.s1) % Gc.jn, k0 = C / Gc.s2;
  int i = i0 + DOM_BUF, tj = j0 + DOM_BUF, tk = k0 + DOM_BUF;
  for(int q = 0; q < 13; q++) v[q] = 0.;          /* coeffs_init */
  int fW = abs(s->flag_u[OIDX(Gx, i, tj, tk)]);
  int fE = abs(s->flag_u[OIDX(Gx, i+1, tj, tk)]);
  int fS = abs(s->flag_v[OIDX(Gy, i, tj, tk)]);
  int fN = abs(s->flag_v[OIDX(Gy, i, tj+1, tk)]);
  int fB = abs(s->flag_w[OIDX(Gz, i, tj, tk)]);
  int fT = abs(s->flag_w[OIDX(Gz, i, tj, tk+1)]);
  /* coeffs */
  v[1] += (real) fB * ddz;
  v[3] += (real) fS * ddy;
  v[5] += (real) fW * ddx;
  v[6] -= (real) (fW + fE) * ddx;
  v[6] -= (real) (fS + fN) * ddy;
  v[6] -= (real) (fB + fT) * ddz;
  v[7] += (real) fE * ddx;
  v[9] += (real) fN * ddy;
  v[11] += (real) fT * ddz;
  /* coeffs_periodic_* */
  if(s->bc.pW == PERIODIC && i == Gc.is) {
    v[5] -= (real) fW * ddx; v[8] += (real) fW * ddx; }
  if(s->bc.pE == PERIODIC && i == Gc.ie-1) {
    v[4] += (real) fE * ddx; v[7] -= (real) fE * ddx; }
  if(s->bc.pS == PERIODIC && tj == Gc.js) {
    v[3] -= (real) fS * ddy; v[10] += (real) fS * ddy; }
  if(s->bc.pN == PERIODIC && tj == Gc.je-1) {
    v[2] += (real) fN * ddy; v[9] -= (real) fN * ddy; }
  if(s->bc.pB == PERIODIC && tk == Gc.ks) {
    v[1] -= (real) fB * ddz; v[12] += (real) fB * ddz; }
  if(s->bc.pT == PERIODIC && tk == Gc.ke-1) {
    v[0] += (real) fT * ddz; v[11] -= (real) fT * ddz; }
  /* coeffs_particle */
  if(s->phase[OIDX(Gc, i, tj, tk)] > -1) {
    for(int q = 0; q < 13; q++) v[q] = 0.;
    v[6] = 1.;
  }
}

static void dia_offsets(const grid_info *G, int off[13])
{
  off[0] = -G->s3 + G->s2; off[1] = -G->s2; off[2] = -G->s2 + G->s1;
  off[3] = -G->s1; off[4] = -G->s1 + 1; off[5] = -1; off[6] = 0;
  off[7] = 1; off[8] = G->s1 - 1; off[9] = G->s1; off[10] = G->s2 - G->s1;
  off[11] = G->s2; off[12] = G->s3 - G->s2;
}

void orc_pp_matvec(const orc_state *s, const real *x, real *y)
{
  const grid_info Gc = s->dom.Gcc;
  int off[13];
  dia_offsets(&Gc, off);
  int N = Gc.s3;
  #pragma omp parallel for schedule(static)
  for(int C = 0; C < N; C++) {
    real v[13];
    orc_pp_coeffs(s, C, v);
    real sum = 0.;
    for(int q = 0; q < 13; q++) {
      int col = C + off[q];
      if(col >= 0 && col < N) sum += v[q] * x[col];
    }
    y[C] = sum;
  }
}

/* The assembled matrix of one solve: the 13 DIA values of every row, as
 * the reference stores them (src/cuda_bicgstab.cu:735-747), built once. */
typedef struct { int N; int off[13]; real *vals; } dia_t;

static void dia_build(const orc_state *s, dia_t *A)
{
  const grid_info Gc = s->dom.Gcc;
  A->N = Gc.s3;
  dia_offsets(&Gc, A->off);
  A->vals = malloc((size_t) A->N * 13 * sizeof(real));
  #pragma omp parallel for schedule(static)
  for(int C = 0; C < A->N; C++) orc_pp_coeffs(s, C, A->vals + (size_t) 13 * C);
}

/* y = A x with each row summed over the 13 diagonals in order, exactly as
 * orc_pp_matvec (rows are independent, so the threads do not change a
 * result) */
static void dia_apply(const dia_t *A, const real *x, real *y)
{
  const int N = A->N;
  #pragma omp parallel for schedule(static)
  for(int C = 0; C < N; C++) {
    const real *v = A->vals + (size_t) 13 * C;
    real sum = 0.;
    for(int q = 0; q < 13; q++) {
      int col = C + A->off[q];
      if(col >= 0 && col < N) sum += v[q] * x[col];
    }
    y[C] = sum;
  }
}

/* Dot product in a fixed order independent of the thread count: serial
 * sums over consecutive chunks of DOT_CHUNK entries, then the chunk sums
 * in order. */
#ifndef DOT_CHUNK
#define DOT_CHUNK 4096
#endif
static real dot(const real *a, const real *b, int n)
{
  int nc = (n + DOT_CHUNK - 1) / DOT_CHUNK;
  real *part = malloc((size_t) (nc > 0 ? nc : 1) * sizeof(real));
  #pragma omp parallel for schedule(static)
  for(int c = 0; c < nc; c++) {
    int lo = c * DOT_CHUNK, hi = lo + DOT_CHUNK < n ? lo + DOT_CHUNK : n;
    real t = 0.;
    for(int i = lo; i < hi; i++) t += a[i] * b[i];
    part[c] = t;
  }
  real s = 0.;
  for(int c = 0; c < nc; c++) s += part[c];
  free(part);
  return s;
}

int orc_bicgstab(const orc_state *s, const real *b, real *x, int max_iter,
  real rel_tol, real *resid_out, int *converged)
{
  const int N = s->dom.Gcc.s3;
  real *dinv = malloc(N * sizeof(real));
  real *y = malloc(N * sizeof(real)), *p = malloc(N * sizeof(real));
  real *r = malloc(N * sizeof(real)), *rs = malloc(N * sizeof(real));
  real *sv = malloc(N * sizeof(real)), *Mp = malloc(N * sizeof(real));
  real *AMp = malloc(N * sizeof(real)), *Ms = malloc(N * sizeof(real));
  real *AMs = malloc(N * sizeof(real));
  dia_t A;
  dia_build(s, &A);
  for(int C = 0; C < N; C++) dinv[C] = 1. / A.vals[(size_t) 13 * C + 6];
  real tol = 0. + rel_tol * sqrt(dot(b, b, N));
  int iter = 0;
  real resid;
  dia_apply(&A, x, y);
  #pragma omp parallel for schedule(static)
  for(int i = 0; i < N; i++) r[i] = 1. * b[i] + -1. * y[i];
  memcpy(p, r, N * sizeof(real));
  memcpy(rs, r, N * sizeof(real));
  real rho_old = dot(rs, r, N);
  for(;;) {
    resid = sqrt(dot(r, r, N));
    if(resid <= tol || iter >= max_iter) break;
    #pragma omp parallel for schedule(static)
    for(int i = 0; i < N; i++) Mp[i] = dinv[i] * p[i];
    dia_apply(&A, Mp, AMp);
    real alpha = rho_old / dot(rs, AMp, N);
    #pragma omp parallel for schedule(static)
    for(int i = 0; i < N; i++) sv[i] = 1. * r[i] + -alpha * AMp[i];
    resid = sqrt(dot(sv, sv, N));
    if(resid <= tol || iter >= max_iter) {
      #pragma omp parallel for schedule(static)
      for(int i = 0; i < N; i++) x[i] = 1. * x[i] + alpha * Mp[i];
      break;
    }
    #pragma omp parallel for schedule(static)
    for(int i = 0; i < N; i++) Ms[i] = dinv[i] * sv[i];
    dia_apply(&A, Ms, AMs);
    real omega = dot(AMs, sv, N) / dot(AMs, AMs, N);
    #pragma omp parallel for schedule(static)
    for(int i = 0; i < N; i++) x[i] = 1. * x[i] + alpha * Mp[i] + omega * Ms[i];
    #pragma omp parallel for schedule(static)
    for(int i = 0; i < N; i++) r[i] = 1. * sv[i] + -omega * AMs[i];
    real rho_new = dot(rs, r, N);
    real beta = (rho_new / rho_old) * (alpha / omega);
    rho_old = rho_new;
    #pragma omp parallel for schedule(static)
    for(int i = 0; i < N; i++)
      p[i] = 1. * r[i] + beta * p[i] + -beta * omega * AMp[i];
    iter++;
  }
  *resid_out = resid;
  *converged = resid <= tol;
  free(dinv); free(y); free(p); free(r); free(rs); free(sv); free(Mp);
  free(AMp); free(Ms); free(AMs); free(A.vals);
  return iter;
}

/* src/cuda_bicgstab.cu:715-1073 */
void orc_PP_bicgstab(orc_state *s)
{
  const grid_info Gc = s->dom.Gcc;
  const int N = Gc.s3;
  orc_PP_rhs(s);
  if(s->nparts > 0) orc_part_BC_p(s);
  real *x = malloc(N * sizeof(real));
  for(int k = 0; k < Gc.kn; k++) for(int j = 0; j < Gc.jn; j++)
  for(int i = 0; i < Gc.in; i++)
    x[i + j*Gc.s1 + k*Gc.s2] = s->phi[OIDX(Gc, i+DOM_BUF, j+DOM_BUF, k+DOM_BUF)];
  real norm = sqrt(dot(s->rhs_p, s->rhs_p, N));
  if(norm == 0) norm = 1.;
  for(int i = 0; i < N; i++) s->rhs_p[i] *= 1. / norm;
  for(int i = 0; i < N; i++) x[i] *= 1. / norm;
  real resid;
  int conv;
  int iters = orc_bicgstab(s, s->rhs_p, x, s->pp_max_iter, s->pp_residual,
    &resid, &conv);
  s->last_iters = iters;
  s->last_resid = resid;
  s->last_converged = conv;
  for(int i = 0; i < N; i++) x[i] *= norm;
  real sum = 0.;
  for(int i = 0; i < N; i++) sum += x[i];
  real p_avg = sum / N;
  for(int i = 0; i < N; i++) x[i] = -p_avg * 1. + x[i];
  for(int k = 0; k < Gc.kn; k++) for(int j = 0; j < Gc.jn; j++)
  for(int i = 0; i < Gc.in; i++)
    s->phi[OIDX(Gc, i+DOM_BUF, j+DOM_BUF, k+DOM_BUF)] = x[i + j*Gc.s1 + k*Gc.s2];
  free(x);
}
