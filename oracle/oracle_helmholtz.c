/*
 * oracle_helmholtz.c -- serial FP64 restatement of the Crank-Nicolson
 * velocity solves of the reference's IMPLICIT build:
 * cuda_{u,v,w}star_helmholtz (src/cuda_bicgstab.cu:44-714), their
 * right-hand sides (src/bicgstab_kernel.cu:27-482 via
 * src/cuda_bicgstab.cu:1091-1170), the 13-diagonal operator
 * (src/bicgstab_kernel.cu:643-1652) and CUSP's diagonally preconditioned
 * conjugate gradient.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The operator is built exactly as the reference's DIA assembly does:
 * interior coefficients from the |face flags|, then the W, E, S, N, B, T
 * boundary kernels in that order (periodic: move the coupling to the wrap
 * diagonal; Dirichlet: identity row on the component's own axis, dropped
 * coupling across), then identity rows on faces with flag < 0.  The
 * product applies each diagonal to the flattened index plus its offset
 * whenever that lands inside the vector, as CUSP's DIA SpMV does.
 *
 * The reference's Dirichlet kernels also write into u_star after the
 * right-hand side has been copied out of it; the solution then overwrites
 * those entries, so those writes have no effect and are not restated.
 *
 * CUSP is not vendored: the conjugate gradient is the published
 * preconditioned algorithm with CUSP's monitor semantics (stop when the
 * updated ||r|| <= tol * ||b|| at the loop head, or at the iteration cap);
 * its iterates are pinned against SciPy's preconditioned CG, the converged
 * solution by its residual.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define IDX(G, i, j, k) ((i) + (j) * (G).s1b + (k) * (G).s2b)

/* Right-hand side of the Helmholtz solve for one component
 * (src/bicgstab_kernel.cu:27-482): the explicit predictor with half the
 * diffusion, no Adams-Bashforth on diffusion and no solid mask; conv is
 * stored. */
static void star_rhs(orc_state *s, int comp)
{
  const dom_struct *d = &s->dom;
  const grid_info Gc = d->Gcc, Gx = d->Gfx, Gy = d->Gfy, Gz = d->Gfz;
  const real dt = s->dt, dt0 = s->dt0, nu = s->nu, rho_f = s->rho_f;
  const real ab0 = 0.5 * dt / dt0;
  const real ab = 1. + ab0;
  const real ddx = 1. / d->dx, ddy = 1. / d->dy, ddz = 1. / d->dz;
  const real *u0 = s->u0, *v0 = s->v0, *w0 = s->w0, *p = s->p0;
  if(comp == 0) {
    for(int k = Gx.ks; k < Gx.ke; k++)
    for(int j = Gx.js; j < Gx.je; j++)
    for(int i = Gx.is; i < Gx.ie; i++) {
      int C = IDX(Gx, i, j, k);
      real us = (p[IDX(Gc, i-1, j, k)] - p[IDX(Gc, i, j, k)]) * ddx / rho_f;
      real u011 = u0[IDX(Gx, i-1, j, k)], u111 = u0[C], u211 = u0[IDX(Gx, i+1, j, k)];
      real u101 = u0[IDX(Gx, i, j-1, k)], u121 = u0[IDX(Gx, i, j+1, k)];
      real v011 = v0[IDX(Gy, i-1, j, k)], v111 = v0[IDX(Gy, i, j, k)];
      real v021 = v0[IDX(Gy, i-1, j+1, k)], v121 = v0[IDX(Gy, i, j+1, k)];
      real u110 = u0[IDX(Gx, i, j, k-1)], u112 = u0[IDX(Gx, i, j, k+1)];
      real w011 = w0[IDX(Gz, i-1, j, k)], w111 = w0[IDX(Gz, i, j, k)];
      real w012 = w0[IDX(Gz, i-1, j, k+1)], w112 = w0[IDX(Gz, i, j, k+1)];
      real duudx = (u211 + u111)*(u211 + u111) - (u111 + u011)*(u111 + u011);
      duudx *= 0.25 * ddx;
      real duvdy = (u121 + u111)*(v121 + v021) - (u111 + u101)*(v111 + v011);
      duvdy *= 0.25 * ddy;
      real duwdz = (u112 + u111)*(w112 + w012) - (u111 + u110)*(w111 + w011);
      duwdz *= 0.25 * ddz;
      real c = duudx + duvdy + duwdz;
      if(dt0 > 0) us += (-ab * c + ab0 * s->conv0_u[C]);
      else us += -c;
      real dud1 = (u211 - u111) * ddx, dud0 = (u111 - u011) * ddx;
      real ddudxx = (dud1 - dud0) * ddx;
      dud1 = (u121 - u111) * ddy; dud0 = (u111 - u101) * ddy;
      real ddudyy = (dud1 - dud0) * ddy;
      dud1 = (u112 - u111) * ddz; dud0 = (u111 - u110) * ddz;
      real ddudzz = (dud1 - dud0) * ddz;
      real df = nu * (ddudxx + ddudyy + ddudzz);
      us += 0.5*df;
      us += s->f_x[C];
      us *= dt;
      us += u111;
      s->u_star[C] = us;
      s->conv_u[C] = c;
    }
  } else if(comp == 1) {
    for(int k = Gy.ks; k < Gy.ke; k++)
    for(int j = Gy.js; j < Gy.je; j++)
    for(int i = Gy.is; i < Gy.ie; i++) {
      int C = IDX(Gy, i, j, k);
      real vs = (p[IDX(Gc, i, j-1, k)] - p[IDX(Gc, i, j, k)]) * ddy / rho_f;
      real v101 = v0[IDX(Gy, i, j-1, k)], v111 = v0[C], v121 = v0[IDX(Gy, i, j+1, k)];
      real v110 = v0[IDX(Gy, i, j, k-1)], v112 = v0[IDX(Gy, i, j, k+1)];
      real w101 = w0[IDX(Gz, i, j-1, k)], w111 = w0[IDX(Gz, i, j, k)];
      real w102 = w0[IDX(Gz, i, j-1, k+1)], w112 = w0[IDX(Gz, i, j, k+1)];
      real v011 = v0[IDX(Gy, i-1, j, k)], v211 = v0[IDX(Gy, i+1, j, k)];
      real u101 = u0[IDX(Gx, i, j-1, k)], u111 = u0[IDX(Gx, i, j, k)];
      real u201 = u0[IDX(Gx, i+1, j-1, k)], u211 = u0[IDX(Gx, i+1, j, k)];
      real dvudx = (v211 + v111)*(u211 + u201) - (v111 + v011)*(u111 + u101);
      dvudx *= 0.25 * ddx;
      real dvvdy = (v121 + v111)*(v121 + v111) - (v111 + v101)*(v111 + v101);
      dvvdy *= 0.25 * ddy;
      real dvwdz = (v112 + v111)*(w112 + w102) - (v111 + v110)*(w111 + w101);
      dvwdz *= 0.25 * ddz;
      real c = dvudx + dvvdy + dvwdz;
      if(dt0 > 0) vs += (-ab * c + ab0 * s->conv0_v[C]);
      else vs += -c;
      real dvd1 = (v211 - v111) * ddx, dvd0 = (v111 - v011) * ddx;
      real ddvdxx = (dvd1 - dvd0) * ddx;
      dvd1 = (v121 - v111) * ddy; dvd0 = (v111 - v101) * ddy;
      real ddvdyy = (dvd1 - dvd0) * ddy;
      dvd1 = (v112 - v111) * ddz; dvd0 = (v111 - v110) * ddz;
      real ddvdzz = (dvd1 - dvd0) * ddz;
      real df = nu * (ddvdxx + ddvdyy + ddvdzz);
      vs += 0.5*df;
      vs += s->f_y[C];
      vs *= dt;
      vs += v111;
      s->v_star[C] = vs;
      s->conv_v[C] = c;
    }
  } else {
    for(int k = Gz.ks; k < Gz.ke; k++)
    for(int j = Gz.js; j < Gz.je; j++)
    for(int i = Gz.is; i < Gz.ie; i++) {
      int C = IDX(Gz, i, j, k);
      real ws = (p[IDX(Gc, i, j, k-1)] - p[IDX(Gc, i, j, k)]) * ddz / rho_f;
      real w110 = w0[IDX(Gz, i, j, k-1)], w111 = w0[C], w112 = w0[IDX(Gz, i, j, k+1)];
      real w011 = w0[IDX(Gz, i-1, j, k)], w211 = w0[IDX(Gz, i+1, j, k)];
      real u110 = u0[IDX(Gx, i, j, k-1)], u111 = u0[IDX(Gx, i, j, k)];
      real u210 = u0[IDX(Gx, i+1, j, k-1)], u211 = u0[IDX(Gx, i+1, j, k)];
      real w101 = w0[IDX(Gz, i, j-1, k)], w121 = w0[IDX(Gz, i, j+1, k)];
      real v110 = v0[IDX(Gy, i, j, k-1)], v111 = v0[IDX(Gy, i, j, k)];
      real v120 = v0[IDX(Gy, i, j+1, k-1)], v121 = v0[IDX(Gy, i, j+1, k)];
      real dwudx = (w211 + w111)*(u211 + u210) - (w111 + w011)*(u111 + u110);
      dwudx *= 0.25 * ddx;
      real dwvdy = (w121 + w111)*(v121 + v120) - (w111 + w101)*(v111 + v110);
      dwvdy *= 0.25 * ddy;
      real dwwdz = (w112 + w111)*(w112 + w111) - (w111 + w110)*(w111 + w110);
      dwwdz *= 0.25 * ddz;
      real c = dwudx + dwvdy + dwwdz;
      if(dt0 > 0) ws += (-ab * c + ab0 * s->conv0_w[C]);
      else ws += -c;
      real dwd1 = (w211 - w111) * ddx, dwd0 = (w111 - w011) * ddx;
      real ddwdxx = (dwd1 - dwd0) * ddx;
      dwd1 = (w121 - w111) * ddy; dwd0 = (w111 - w101) * ddy;
      real ddwdyy = (dwd1 - dwd0) * ddy;
      dwd1 = (w112 - w111) * ddz; dwd0 = (w111 - w110) * ddz;
      real ddwdzz = (dwd1 - dwd0) * ddz;
      real df = nu * (ddwdxx + ddwdyy + ddwdzz);
      ws += 0.5*df;
      ws += s->f_z[C];
      ws *= dt;
      ws += w111;
      s->w_star[C] = ws;
      s->conv_w[C] = c;
    }
  }
}

/* Geometry of one component's unknowns (the interior faces of its grid,
 * no ghosts, x fastest) and the thirteen DIA offsets of
 * src/cuda_bicgstab.cu:62-74 / 342-355 / 540-553. */
typedef struct {
  int comp, ni, nj, nk;
  long s1, s2, n;
  long off[13];
} hgeom;

static hgeom h_geom(const orc_state *s, int comp)
{
  const dom_struct *d = &s->dom;
  const grid_info G = comp == 0 ? d->Gfx : comp == 1 ? d->Gfy : d->Gfz;
  hgeom h;
  h.comp = comp;
  h.ni = G.in; h.nj = G.jn; h.nk = G.kn;
  h.s1 = G.s1; h.s2 = G.s2; h.n = G.s3;
  long s1 = h.s1, s2 = h.s2, s3 = h.n;
  /* the wrap partner along the component's own axis skips the duplicated
   * periodic face (2 instead of 1) */
  long wx = comp == 0 ? 2 : 1, wy = comp == 1 ? 2 : 1, wz = comp == 2 ? 2 : 1;
  h.off[0] = -s3 + wz * s2;
  h.off[1] = -s2;
  h.off[2] = -s2 + wy * s1;
  h.off[3] = -s1;
  h.off[4] = -s1 + wx;
  h.off[5] = -1;
  h.off[6] = 0;
  h.off[7] = 1;
  h.off[8] = s1 - wx;
  h.off[9] = s1;
  h.off[10] = s2 - wy * s1;
  h.off[11] = s2;
  h.off[12] = s3 - wz * s2;
  return h;
}

/* BC type of the component at a face (0..5 = W,E,S,N,B,T). */
static int bc_type(const BC *b, int comp, int face)
{
  const int u[6] = {b->uW, b->uE, b->uS, b->uN, b->uB, b->uT};
  const int v[6] = {b->vW, b->vE, b->vS, b->vN, b->vB, b->vT};
  const int w[6] = {b->wW, b->wE, b->wS, b->wN, b->wB, b->wT};
  return comp == 0 ? u[face] : comp == 1 ? v[face] : w[face];
}

/* Assemble the 13 diagonals: values[C + dg * n]. */
void orc_helmholtz_coeffs(const orc_state *s, int comp, real *values)
{
  const dom_struct *d = &s->dom;
  const hgeom h = h_geom(s, comp);
  const grid_info Gx = d->Gfx, Gy = d->Gfy, Gz = d->Gfz;
  const grid_info G = comp == 0 ? Gx : comp == 1 ? Gy : Gz;
  const int *fl_own = comp == 0 ? s->flag_u : comp == 1 ? s->flag_v : s->flag_w;
  const real nu = s->nu, dt = s->dt;
  const real ddx = 1. / (d->dx * d->dx), ddy = 1. / (d->dy * d->dy);
  const real ddz = 1. / (d->dz * d->dz);
  const long n = h.n;
  memset(values, 0, 13 * n * sizeof(real));
#define V(dg) values[C + (long) (dg) * n]
  /* interior rows: *star_coeffs (src/bicgstab_kernel.cu:829-944) */
  for(int k = 0; k < h.nk; k++) for(int j = 0; j < h.nj; j++)
  for(int i = 0; i < h.ni; i++) {
    long C = i + j * h.s1 + k * h.s2;
    int I = i + DOM_BUF, J = j + DOM_BUF, K = k + DOM_BUF;
    int fW = abs(s->flag_u[I + J*Gx.s1b + K*Gx.s2b]);
    int fE = abs(s->flag_u[(I+1) + J*Gx.s1b + K*Gx.s2b]);
    int fS = abs(s->flag_v[I + J*Gy.s1b + K*Gy.s2b]);
    int fN = abs(s->flag_v[I + (J+1)*Gy.s1b + K*Gy.s2b]);
    int fB = abs(s->flag_w[I + J*Gz.s1b + K*Gz.s2b]);
    int fT = abs(s->flag_w[I + J*Gz.s1b + (K+1)*Gz.s2b]);
    V(1)  -= (real)fB*0.5*nu*dt*ddz;
    V(3)  -= (real)fS*0.5*nu*dt*ddy;
    V(5)  -= (real)fW*0.5*nu*dt*ddx;
    V(6)  += 1.;
    V(6)  += (real)fW*0.5*nu*dt*ddx;
    V(6)  += (real)fE*0.5*nu*dt*ddx;
    V(6)  += (real)fS*0.5*nu*dt*ddy;
    V(6)  += (real)fN*0.5*nu*dt*ddy;
    V(6)  += (real)fB*0.5*nu*dt*ddz;
    V(6)  += (real)fT*0.5*nu*dt*ddz;
    V(7)  -= (real)fE*0.5*nu*dt*ddx;
    V(9)  -= (real)fN*0.5*nu*dt*ddy;
    V(11) -= (real)fT*0.5*nu*dt*ddz;
  }
  /* boundary kernels in the driver's order W, E, S, N, B, T
   * (src/cuda_bicgstab.cu:217-264 and the v, w counterparts) */
  const real cc[3] = {0.5*nu*dt*ddx, 0.5*nu*dt*ddy, 0.5*nu*dt*ddz};
  const int lo_in[3] = {5, 3, 1}, hi_in[3] = {7, 9, 11};
  const int lo_wrap[3] = {4, 2, 0}, hi_wrap[3] = {8, 10, 12};
  const int ext[3] = {h.ni, h.nj, h.nk};
  for(int face = 0; face < 6; face++) {
    int a = face / 2, hi = face % 2;
    int t = bc_type(&s->bc, comp, face);
    if(t != PERIODIC && t != DIRICHLET) continue;
    for(int k = 0; k < h.nk; k++) for(int j = 0; j < h.nj; j++)
    for(int i = 0; i < h.ni; i++) {
      int pos = a == 0 ? i : a == 1 ? j : k;
      if(pos != (hi ? ext[a] - 1 : 0)) continue;
      long C = i + j * h.s1 + k * h.s2;
      if(t == PERIODIC) {
        if(!hi) { V(lo_in[a]) += cc[a]; V(hi_wrap[a]) -= cc[a]; }
        else    { V(lo_wrap[a]) -= cc[a]; V(hi_in[a]) += cc[a]; }
      } else if(a == comp) {
        for(int dg = 0; dg < 13; dg++) V(dg) = 0.;
        V(6) = 1.;
      } else {
        V(hi ? hi_in[a] : lo_in[a]) = 0.;
      }
    }
  }
  /* *star_coeffs_particles: identity rows on faces with flag < 0 */
  for(int k = 0; k < h.nk; k++) for(int j = 0; j < h.nj; j++)
  for(int i = 0; i < h.ni; i++) {
    long C = i + j * h.s1 + k * h.s2;
    int CC = (i+DOM_BUF) + (j+DOM_BUF)*G.s1b + (k+DOM_BUF)*G.s2b;
    if(fl_own[CC] < 0) {
      for(int dg = 0; dg < 13; dg++) V(dg) = 0.;
      V(6) = 1.;
    }
  }
#undef V
}

void orc_helmholtz_matvec(const orc_state *s, int comp, const real *values,
  const real *x, real *y)
{
  const hgeom h = h_geom(s, comp);
  for(long C = 0; C < h.n; C++) {
    real sum = 0.;
    for(int dg = 0; dg < 13; dg++) {
      long col = C + h.off[dg];
      if(col >= 0 && col < h.n) sum += values[C + dg * h.n] * x[col];
    }
    y[C] = sum;
  }
}

static real dotv(const real *a, const real *b, long n)
{
  real s = 0.;
  for(long i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* Diagonally preconditioned conjugate gradient with CUSP's monitor
 * semantics; x holds the initial guess.  Returns the iteration count. */
int orc_pcg(const orc_state *s, int comp, const real *values, const real *b,
  real *x, int max_iter, real rel_tol, real *resid_out, int *converged)
{
  const hgeom h = h_geom(s, comp);
  const long n = h.n;
  real *y = malloc(n * sizeof(real)), *z = malloc(n * sizeof(real));
  real *r = malloc(n * sizeof(real)), *p = malloc(n * sizeof(real));
  real *dinv = malloc(n * sizeof(real));
  for(long i = 0; i < n; i++) dinv[i] = 1. / values[i + 6 * n];
  real tol = rel_tol * sqrt(dotv(b, b, n));
  orc_helmholtz_matvec(s, comp, values, x, y);
  for(long i = 0; i < n; i++) r[i] = 1. * b[i] + -1. * y[i];
  for(long i = 0; i < n; i++) z[i] = dinv[i] * r[i];
  memcpy(p, z, n * sizeof(real));
  real rz = dotv(r, z, n);
  int iter = 0;
  real resid = sqrt(dotv(r, r, n));
  while(!(resid <= tol) && iter < max_iter) {
    orc_helmholtz_matvec(s, comp, values, p, y);
    real alpha = rz / dotv(y, p, n);
    for(long i = 0; i < n; i++) x[i] = x[i] + alpha * p[i];
    for(long i = 0; i < n; i++) r[i] = r[i] + -alpha * y[i];
    for(long i = 0; i < n; i++) z[i] = dinv[i] * r[i];
    real rz_old = rz;
    rz = dotv(r, z, n);
    real beta = rz / rz_old;
    for(long i = 0; i < n; i++) p[i] = 1. * z[i] + beta * p[i];
    iter++;
    resid = sqrt(dotv(r, r, n));
  }
  *resid_out = resid;
  *converged = resid <= tol;
  free(y); free(z); free(r); free(p); free(dinv);
  return iter;
}

/* cuda_{u,v,w}star_helmholtz: src/cuda_bicgstab.cu:44-714 */
void orc_star_helmholtz(orc_state *s, int comp)
{
  const dom_struct *d = &s->dom;
  const grid_info G = comp == 0 ? d->Gfx : comp == 1 ? d->Gfy : d->Gfz;
  real *star = comp == 0 ? s->u_star : comp == 1 ? s->v_star : s->w_star;
  star_rhs(s, comp);
  orc_dom_BC_star(s);
  if(s->nparts > 0) orc_part_BC_star(s);
  const hgeom h = h_geom(s, comp);
  const long n = h.n;
  real *b = malloc(n * sizeof(real)), *x = calloc(n, sizeof(real));
  real *values = malloc(13 * n * sizeof(real));
  for(int k = 0; k < h.nk; k++) for(int j = 0; j < h.nj; j++)
  for(int i = 0; i < h.ni; i++)
    b[i + j*h.s1 + k*h.s2] = star[(i+DOM_BUF) + (j+DOM_BUF)*G.s1b + (k+DOM_BUF)*G.s2b];
  orc_helmholtz_coeffs(s, comp, values);
  real norm = sqrt(dotv(b, b, n));
  if(norm == 0) norm = 1.;
  for(long q = 0; q < n; q++) b[q] *= 1. / norm;
  real resid;
  int conv;
  int it = orc_pcg(s, comp, values, b, x, s->pp_max_iter, s->pp_residual, &resid,
    &conv);
  s->last_iters = it;
  s->last_resid = resid;
  s->last_converged = conv;
  for(long q = 0; q < n; q++) x[q] *= norm;
  for(int k = 0; k < h.nk; k++) for(int j = 0; j < h.nj; j++)
  for(int i = 0; i < h.ni; i++)
    star[(i+DOM_BUF) + (j+DOM_BUF)*G.s1b + (k+DOM_BUF)*G.s2b] = x[i + j*h.s1 + k*h.s2];
  free(b); free(x); free(values);
}
