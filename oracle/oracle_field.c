/*
 * oracle_field.c -- serial FP64 restatement of the reference's field
 * kernels: predictor, ghost-cell BCs, solvability, projection, pressure
 * update, time-step size, history shift, forcing.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Operation order inside each expression follows the reference kernels so
 * that the oracle and a faithful device implementation agree to rounding.
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define IDX(G, i, j, k) ((i) + (j) * (G).s1b + (k) * (G).s2b)

/* ---------------------------------------------------------------------
 * Predictor: src/bluebottle_kernel.cu:1504-2003 (u_star_2, v_star_2,
 * w_star_2), launched for all three components by
 * src/cuda_bluebottle.cu:2930-3026 with p0 as the pressure.
 * ------------------------------------------------------------------- */
void orc_U_star_2(orc_state *s)
{
  const dom_struct *d = &s->dom;
  const grid_info Gc = d->Gcc, Gx = d->Gfx, Gy = d->Gfy, Gz = d->Gfz;
  const real dt = s->dt, dt0 = s->dt0, nu = s->nu, rho_f = s->rho_f;
  const real ab0 = 0.5 * dt / dt0;
  const real ab = 1. + ab0;
  const real ddx = 1. / d->dx, ddy = 1. / d->dy, ddz = 1. / d->dz;
  const real *u0 = s->u0, *v0 = s->v0, *w0 = s->w0, *p = s->p0;
  const int *phase = s->phase;

  /* u: src/bluebottle_kernel.cu:1504-1669 */
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

    real dud1 = (u211 - u111) * ddx;
    real dud0 = (u111 - u011) * ddx;
    real ddudxx = (dud1 - dud0) * ddx;
    dud1 = (u121 - u111) * ddy;
    dud0 = (u111 - u101) * ddy;
    real ddudyy = (dud1 - dud0) * ddy;
    dud1 = (u112 - u111) * ddz;
    dud0 = (u111 - u110) * ddz;
    real ddudzz = (dud1 - dud0) * ddz;
    real df = nu * (ddudxx + ddudyy + ddudzz);
    if(dt0 > 0) us += (ab * df - ab0 * s->diff0_u[C]);
    else us += df;
    us += s->f_x[C];
    us *= dt;
    us += u111;
    us *= (phase[IDX(Gc, i-1, j, k)] < 0 && phase[IDX(Gc, i, j, k)] < 0);
    s->u_star[C] = us;
    s->conv_u[C] = c;
    s->diff_u[C] = df;
  }

  /* v: src/bluebottle_kernel.cu:1673-1836 */
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

    real dvd1 = (v211 - v111) * ddx;
    real dvd0 = (v111 - v011) * ddx;
    real ddvdxx = (dvd1 - dvd0) * ddx;
    dvd1 = (v121 - v111) * ddy;
    dvd0 = (v111 - v101) * ddy;
    real ddvdyy = (dvd1 - dvd0) * ddy;
    dvd1 = (v112 - v111) * ddz;
    dvd0 = (v111 - v110) * ddz;
    real ddvdzz = (dvd1 - dvd0) * ddz;
    real df = nu * (ddvdxx + ddvdyy + ddvdzz);
    if(dt0 > 0) vs += (ab * df - ab0 * s->diff0_v[C]);
    else vs += df;
    vs += s->f_y[C];
    vs *= dt;
    vs += v111;
    vs *= (phase[IDX(Gc, i, j-1, k)] < 0 && phase[IDX(Gc, i, j, k)] < 0);
    s->v_star[C] = vs;
    s->conv_v[C] = c;
    s->diff_v[C] = df;
  }

  /* w: src/bluebottle_kernel.cu:1840-2003 */
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

    real dwd1 = (w211 - w111) * ddx;
    real dwd0 = (w111 - w011) * ddx;
    real ddwdxx = (dwd1 - dwd0) * ddx;
    dwd1 = (w121 - w111) * ddy;
    dwd0 = (w111 - w101) * ddy;
    real ddwdyy = (dwd1 - dwd0) * ddy;
    dwd1 = (w112 - w111) * ddz;
    dwd0 = (w111 - w110) * ddz;
    real ddwdzz = (dwd1 - dwd0) * ddz;
    real df = nu * (ddwdxx + ddwdyy + ddwdzz);
    if(dt0 > 0) ws += (ab * df - ab0 * s->diff0_w[C]);
    else ws += df;
    ws += s->f_z[C];
    ws *= dt;
    ws += w111;
    ws *= (phase[IDX(Gc, i, j, k-1)] < 0 && phase[IDX(Gc, i, j, k)] < 0);
    s->w_star[C] = ws;
    s->conv_w[C] = c;
    s->diff_w[C] = df;
  }
}

/* ---------------------------------------------------------------------
 * Ghost-cell boundary conditions: the BC_{p,u,v,w}_{W,E,S,N,B,T}_{P,D,N}
 * kernels of src/bluebottle_kernel.cu:26-1248.  One face of one field is
 * applied over the full ghosted transverse extent; faces are applied in
 * the order W, E, S, N, B, T (src/cuda_bluebottle.cu:1565-2298).
 *   axis: 0 = x, 1 = y, 2 = z;  hi: 0 = W/S/B, 1 = E/N/T;
 *   normal: 1 if the field is staggered along `axis` (u in x, v in y,
 *   w in z), else 0 (cell-centred along the axis, also p).
 * ------------------------------------------------------------------- */
/* PRECURSOR faces (the BC_*_T kernels, src/bluebottle_kernel.cu:228-1247)
 * read inflow planes pl0 (and pl1 for a tangential component), indexed
 * j + k*jnb (W/E), k + i*knb (S/N), i + j*inb (B/T).  PRECURSOR_AT_LO on a
 * high face applies the low face's normal-component kernel, as
 * cuda_dom_BC_star does for the north face of v (src/cuda_bluebottle.cu,
 * its BC_v_S_T call in the north-face switch). */
#define PRECURSOR_AT_LO 103

static void bc_face(real *f, grid_info G, int axis, int hi, int type,
  int normal, real bcv, const real *pl0, const real *pl1)
{
  int s1 = 1, s2, s3;                 /* strides along axis, t1, t2 */
  int n1, n2;                         /* transverse extents incl. ghosts */
  int s, e, sb, eb;                   /* start/end along the axis */
  int str;
  if(axis == 0) {
    str = 1; s2 = G.s1b; s3 = G.s2b; n1 = G.jnb; n2 = G.knb;
    s = G.is; e = G.ie; sb = G.isb; eb = G.ieb;
  } else if(axis == 1) {
    str = G.s1b; s2 = 1; s3 = G.s2b; n1 = G.inb; n2 = G.knb;
    s = G.js; e = G.je; sb = G.jsb; eb = G.jeb;
  } else {
    str = G.s2b; s2 = 1; s3 = G.s1b; n1 = G.inb; n2 = G.jnb;
    s = G.ks; e = G.ke; sb = G.ksb; eb = G.keb;
  }
  (void) s1;
  for(int t2 = 0; t2 < n2; t2++)
  for(int t1 = 0; t1 < n1; t1++) {
    real *F = f + (long) t1 * s2 + (long) t2 * s3;
    const long P = axis == 0 ? t1 + (long) t2 * G.jnb
      : axis == 1 ? t2 + (long) t1 * G.knb : t1 + (long) t2 * G.inb;
#define A(n) F[(long) (n) * str]
    if(type == PRECURSOR) {
      if(normal) {
        if(!hi) { A(sb) = 2. * pl0[P] - A(s+1); A(s) = pl0[P]; }
        else    { A(eb-1) = 2. * pl0[P] - A(e-2); A(e-1) = pl0[P]; }
      } else {
        if(!hi) { A(sb) = pl0[P]; A(s) = pl1[P]; }
        else    { A(e-1) = pl0[P]; A(eb-1) = pl1[P]; }
      }
    } else if(type == PRECURSOR_AT_LO) {
      A(sb) = 2. * pl0[P] - A(s+1); A(s) = pl0[P];
    } else if(type == PERIODIC) {
      if(normal) {
        if(!hi) { A(sb) = A(e-2); A(s) = A(e-1); }
        else    { A(eb-1) = A(s+1); A(e-1) = A(s); }
      } else {
        if(!hi) A(sb) = A(e-1);
        else    A(eb-1) = A(s);
      }
    } else if(type == DIRICHLET) {
      if(normal) {
        if(!hi) { A(sb) = 2. * bcv - A(s+1); A(s) = bcv; }
        else    { A(eb-1) = 2. * bcv - A(e-2); A(e-1) = bcv; }
      } else {
        if(!hi) A(sb) = 8./3. * bcv - 2. * A(s) + 1./3. * A(s+1);
        else    A(eb-1) = 8./3. * bcv - 2. * A(e-1) + 1./3. * A(e-2);
      }
    } else if(type == NEUMANN) {
      if(!hi) A(sb) = A(s);
      else    A(eb-1) = A(e-1);
    }
#undef A
  }
}

/* BC type/value table for one field, faces in W,E,S,N,B,T order. */
static void bc_table(const BC *b, char field, int type[6], real val[6])
{
  switch(field) {
    case 'p':
      type[0] = b->pW; type[1] = b->pE; type[2] = b->pS; type[3] = b->pN;
      type[4] = b->pB; type[5] = b->pT;
      for(int f = 0; f < 6; f++) val[f] = 0.;
      break;
    case 'u':
      type[0] = b->uW; type[1] = b->uE; type[2] = b->uS; type[3] = b->uN;
      type[4] = b->uB; type[5] = b->uT;
      val[0] = b->uWD; val[1] = b->uED; val[2] = b->uSD; val[3] = b->uND;
      val[4] = b->uBD; val[5] = b->uTD;
      break;
    case 'v':
      type[0] = b->vW; type[1] = b->vE; type[2] = b->vS; type[3] = b->vN;
      type[4] = b->vB; type[5] = b->vT;
      val[0] = b->vWD; val[1] = b->vED; val[2] = b->vSD; val[3] = b->vND;
      val[4] = b->vBD; val[5] = b->vTD;
      break;
    default:
      type[0] = b->wW; type[1] = b->wE; type[2] = b->wS; type[3] = b->wN;
      type[4] = b->wB; type[5] = b->wT;
      val[0] = b->wWD; val[1] = b->wED; val[2] = b->wSD; val[3] = b->wND;
      val[4] = b->wBD; val[5] = b->wTD;
      break;
  }
}

/* The planes the reference passes for each PRECURSOR face of a component
 * (plane indices into orc_plane_names; -1 none), from cuda_dom_BC
 * (star = 0) and cuda_dom_BC_star (star = 1), src/cuda_bluebottle.cu:
 * 1549-2298 and 2300-3030.  Kept: cuda_dom_BC passes v_WE_W for both
 * planes of v's west face; cuda_dom_BC_star fills v's north face with the
 * south kernel.  Departure: cuda_dom_BC's north-face call for w writes
 * into u at w's indices (out of u's bounds unless the grids coincide);
 * here it fills w. */
static void precursor_planes(int comp, int star, int pl[6][2])
{
  static const int U[6][2] = {{0, -1}, {0, -1}, {1, 2}, {1, 2}, {3, 4}, {3, 4}};
  static const int V[6][2] = {{5, 5}, {5, 6}, {7, -1}, {7, -1}, {8, 9}, {8, 9}};
  static const int W[6][2] = {{10, 11}, {10, 11}, {12, 13}, {12, 13},
    {14, -1}, {14, -1}};
  const int (*T)[2] = comp == 'u' ? U : comp == 'v' ? V : W;
  for(int f = 0; f < 6; f++) { pl[f][0] = T[f][0]; pl[f][1] = T[f][1]; }
  if(comp == 'v' && star) pl[0][1] = 6;
}

static void bc_field(const orc_state *s, real *f, char field, int star)
{
  const dom_struct *d = &s->dom;
  grid_info G = field == 'p' ? d->Gcc : field == 'u' ? d->Gfx
    : field == 'v' ? d->Gfy : d->Gfz;
  int normal_axis = field == 'u' ? 0 : field == 'v' ? 1 : field == 'w' ? 2 : -1;
  int type[6];
  real val[6];
  bc_table(&s->bc, field, type, val);
  int pl[6][2] = {{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};
  if(field != 'p') precursor_planes(field, star, pl);
  if(field == 'v' && star && type[3] == PRECURSOR) type[3] = PRECURSOR_AT_LO;
  int ext[6] = {d->W, d->E, d->S, d->N, d->B, d->T};
  for(int face = 0; face < 6; face++) {
    if(ext[face] != -1) continue;
    int axis = face / 2, hi = face % 2;
    const real *p0 = pl[face][0] >= 0 ? s->plane[pl[face][0]] : NULL;
    const real *p1 = pl[face][1] >= 0 ? s->plane[pl[face][1]] : NULL;
    bc_face(f, G, axis, hi, type[face], axis == normal_axis, val[face], p0, p1);
  }
}

void orc_dom_BC(orc_state *s)
{
  bc_field(s, s->p, 'p', 0);
  bc_field(s, s->u, 'u', 0);
  bc_field(s, s->v, 'v', 0);
  bc_field(s, s->w, 'w', 0);
}

void orc_dom_BC_star(orc_state *s)
{
  bc_field(s, s->u_star, 'u', 1);
  bc_field(s, s->v_star, 'v', 1);
  bc_field(s, s->w_star, 'w', 1);
}

void orc_dom_BC_phi(orc_state *s) { bc_field(s, s->phi, 'p', 0); }
void orc_dom_BC_p(orc_state *s) { bc_field(s, s->p, 'p', 0); }

/* ---------------------------------------------------------------------
 * Plane extraction of the precursor domain: cuda_yank_turb_planes
 * (src/cuda_bluebottle.cu:4556-4820) with the yank_* kernels
 * (src/bluebottle_kernel.cu:3619-3851).  Along the plane's normal axis the
 * cell index of pos (one periodic wrap), then linear interpolation of the
 * face values: the normal component at pos plus vel; a tangential one
 * half a cell either side of pos.  Kept: yank_w_BT scales by 1/dx;
 * yank_w_SN's high plane is the slope term alone and is also written
 * into w (`plane_n = w[M] = ...`); the north-face w call passes w_SN_N
 * for both planes.  Departure: yank_u_SN bounds k by Gfx.inb (one row
 * past the arrays when inb > knb); here k < min(inb, knb).  Plane
 * positions outside the domain are clamped to the grid (the reference
 * reads out of bounds there).
 * ------------------------------------------------------------------- */
static void yank(real *f, grid_info G, int axis, int normal, real *lo,
  real *hi, real pos, real vel, real h, real xs, real ddh, int write_back,
  int n2_cap)
{
  long str, s1, s2;
  int n1, n2, start, end, nb;
  if(axis == 0) {
    str = 1; s1 = G.s1b; s2 = G.s2b; n1 = G.jnb; n2 = G.knb;
    start = G.is; end = G.ie; nb = G.inb;
  } else if(axis == 1) {
    str = G.s1b; s1 = 1; s2 = G.s2b; n1 = G.inb; n2 = G.knb;
    start = G.js; end = G.je; nb = G.jnb;
  } else {
    str = G.s2b; s1 = 1; s2 = G.s1b; n1 = G.inb; n2 = G.jnb;
    start = G.ks; end = G.ke; nb = G.knb;
  }
  if(n2_cap > 0 && n2_cap < n2) n2 = n2_cap;
  for(int t2 = 0; t2 < n2; t2++)
  for(int t1 = 0; t1 < n1; t1++) {
    double q = floor((pos - xs) * ddh) + DOM_BUF;
    if(q < start) q += nb;
    if(q > end - 1) q -= nb;
    const double lo_i = normal ? 0. : 1.;
    const int i = (int) fmin(fmax(q, lo_i), (double) (nb - 2));
    real *F = f + (long) t1 * s1 + (long) t2 * s2;
    const long P = axis == 0 ? t1 + (long) t2 * G.jnb
      : axis == 1 ? t2 + (long) t1 * G.knb : t1 + (long) t2 * G.inb;
    if(normal) {
      real xx = (i - DOM_BUF) * h + xs;
      real W = F[i * str], E = F[(i + 1) * str];
      real d = (E - W) * ddh;
      lo[P] = W + d * (pos - xx) + vel;
    } else {
      real xw = (i - DOM_BUF - 0.5) * h + xs;
      real xe = (i - DOM_BUF + 0.5) * h + xs;
      real W = F[(i - 1) * str], M = F[i * str], E = F[(i + 1) * str];
      real dw = (M - W) * ddh, de = (E - M) * ddh;
      lo[P] = W + dw * (pos - 0.5 * h - xw);
      if(write_back) { F[i * str] = de * (pos + 0.5 * h - xe); hi[P] = F[i * str]; }
      else hi[P] = M + de * (pos + 0.5 * h - xe);
    }
  }
}

void orc_yank_turb_planes(orc_state *s, const int *c, const real *pos,
  const real *vel)
{
  const dom_struct *d = &s->dom;
  real **P = s->plane;
  const real ddx = 1. / d->dx, ddy = 1. / d->dy, ddz = 1. / d->dz;
  if(c[0] == PRECURSOR)
    yank(s->u, d->Gfx, 0, 1, P[0], NULL, pos[0], vel[0], d->dx, d->xs, ddx, 0, 0);
  else if(c[1] == PRECURSOR)
    yank(s->u, d->Gfx, 0, 1, P[0], NULL, pos[0], vel[1], d->dx, d->xs, ddx, 0, 0);
  if(c[6] == PRECURSOR)
    yank(s->v, d->Gfy, 0, 0, P[5], P[6], pos[0], vel[0], d->dx, d->xs, ddx, 0, 0);
  if(c[7] == PRECURSOR)
    yank(s->v, d->Gfy, 0, 0, P[5], P[6], pos[0], vel[1], d->dx, d->xs, ddx, 0, 0);
  if(c[12] == PRECURSOR)
    yank(s->w, d->Gfz, 0, 0, P[10], P[11], pos[0], vel[0], d->dx, d->xs, ddx, 0, 0);
  if(c[13] == PRECURSOR)
    yank(s->w, d->Gfz, 0, 0, P[10], P[11], pos[0], vel[1], d->dx, d->xs, ddx, 0, 0);
  if(c[2] == PRECURSOR)
    yank(s->u, d->Gfx, 1, 0, P[1], P[2], pos[4], vel[8], d->dy, d->ys, ddy, 0,
      d->Gfx.inb);
  else if(c[3] == PRECURSOR)
    yank(s->u, d->Gfx, 1, 0, P[1], P[2], pos[4], vel[9], d->dy, d->ys, ddy, 0,
      d->Gfx.inb);
  if(c[8] == PRECURSOR)
    yank(s->v, d->Gfy, 1, 1, P[7], NULL, pos[4], vel[8], d->dy, d->ys, ddy, 0, 0);
  if(c[9] == PRECURSOR)
    yank(s->v, d->Gfy, 1, 1, P[7], NULL, pos[4], vel[9], d->dy, d->ys, ddy, 0, 0);
  if(c[14] == PRECURSOR)
    yank(s->w, d->Gfz, 1, 0, P[12], P[13], pos[4], vel[8], d->dy, d->ys, ddy, 1, 0);
  if(c[15] == PRECURSOR)
    yank(s->w, d->Gfz, 1, 0, P[13], P[13], pos[4], vel[9], d->dy, d->ys, ddy, 1, 0);
  if(c[4] == PRECURSOR)
    yank(s->u, d->Gfx, 2, 0, P[3], P[4], pos[8], vel[16], d->dz, d->zs, ddz, 0, 0);
  else if(c[5] == PRECURSOR)
    yank(s->u, d->Gfx, 2, 0, P[3], P[4], pos[8], vel[17], d->dz, d->zs, ddz, 0, 0);
  if(c[10] == PRECURSOR)
    yank(s->v, d->Gfy, 2, 0, P[8], P[9], pos[8], vel[16], d->dz, d->zs, ddz, 0, 0);
  if(c[11] == PRECURSOR)
    yank(s->v, d->Gfy, 2, 0, P[8], P[9], pos[8], vel[17], d->dz, d->zs, ddz, 0, 0);
  if(c[16] == PRECURSOR)
    yank(s->w, d->Gfz, 2, 1, P[14], NULL, pos[8], vel[16], d->dz, d->zs,
      1. / d->dx, 0, 0);
  if(c[17] == PRECURSOR)
    yank(s->w, d->Gfz, 2, 1, P[14], NULL, pos[8], vel[17], d->dz, d->zs,
      1. / d->dx, 0, 0);
}

/* ---------------------------------------------------------------------
 * Solvability: src/cuda_bluebottle.cu:4122-4314 with surf_int_*_copy and
 * plane_eps_* from src/bluebottle_kernel.cu:2133-2245.  The reference sums
 * with thrust::reduce (order unspecified); here the sum runs in array
 * order (first the negated low face, then the high face).
 * ------------------------------------------------------------------- */
void orc_solvability(orc_state *s)
{
  const dom_struct *d = &s->dom;
  const grid_info Gx = d->Gfx, Gy = d->Gfy, Gz = d->Gfz;
  real eps_x = 0., eps_y = 0., eps_z = 0.;
  for(int tk = 0; tk < Gx.kn; tk++) for(int tj = 0; tj < Gx.jn; tj++)
    eps_x += -s->u_star[IDX(Gx, Gx.is, tj+DOM_BUF, tk+DOM_BUF)];
  for(int tk = 0; tk < Gx.kn; tk++) for(int tj = 0; tj < Gx.jn; tj++)
    eps_x += s->u_star[IDX(Gx, Gx.ie-1, tj+DOM_BUF, tk+DOM_BUF)];
  eps_x = eps_x * d->dy * d->dz;
  for(int tk = 0; tk < Gy.kn; tk++) for(int ti = 0; ti < Gy.in; ti++)
    eps_y += -s->v_star[IDX(Gy, ti+DOM_BUF, Gy.js, tk+DOM_BUF)];
  for(int tk = 0; tk < Gy.kn; tk++) for(int ti = 0; ti < Gy.in; ti++)
    eps_y += s->v_star[IDX(Gy, ti+DOM_BUF, Gy.je-1, tk+DOM_BUF)];
  eps_y = eps_y * d->dx * d->dz;
  for(int tj = 0; tj < Gz.jn; tj++) for(int ti = 0; ti < Gz.in; ti++)
    eps_z += -s->w_star[IDX(Gz, ti+DOM_BUF, tj+DOM_BUF, Gz.ks)];
  for(int tj = 0; tj < Gz.jn; tj++) for(int ti = 0; ti < Gz.in; ti++)
    eps_z += s->w_star[IDX(Gz, ti+DOM_BUF, tj+DOM_BUF, Gz.ke-1)];
  eps_z = eps_z * d->dx * d->dy;

  real ex = 0, ey = 0, ez = 0;   /* per-plane increments */
  int plane[6] = {0, 0, 0, 0, 0, 0};
  real add[6] = {0, 0, 0, 0, 0, 0};
  switch(s->out_plane) {
    case WEST: case EAST: {
      eps_x = eps_x / d->yl / d->zl;
      eps_y = eps_y / d->yl / d->zl;
      eps_z = eps_z / d->yl / d->zl;
      int f = s->out_plane == WEST ? 0 : 1;
      plane[f] = 1; add[f] = eps_x + eps_y + eps_z;
      break; }
    case SOUTH: case NORTH: {
      eps_x = eps_x / d->zl / d->xl;
      eps_y = eps_y / d->zl / d->xl;
      eps_z = eps_z / d->zl / d->xl;
      int f = s->out_plane == SOUTH ? 2 : 3;
      plane[f] = 1; add[f] = eps_x + eps_y + eps_z;
      break; }
    case BOTTOM: case TOP: {
      eps_x = eps_x / d->xl / d->yl;
      eps_y = eps_y / d->xl / d->yl;
      eps_z = eps_z / d->xl / d->yl;
      int f = s->out_plane == BOTTOM ? 4 : 5;
      plane[f] = 1; add[f] = eps_x + eps_y + eps_z;
      break; }
    case HOMOGENEOUS:
      ex = 0.5 * eps_x / d->yl / d->zl;
      ey = 0.5 * eps_y / d->zl / d->xl;
      ez = 0.5 * eps_z / d->xl / d->yl;
      for(int f = 0; f < 6; f++) plane[f] = 1;
      add[0] = add[1] = ex; add[2] = add[3] = ey; add[4] = add[5] = ez;
      break;
  }
  /* plane_eps_x_W adds, plane_eps_x_E subtracts, etc. */
  if(plane[0]) for(int tk = 0; tk < Gx.kn; tk++) for(int tj = 0; tj < Gx.jn; tj++) {
    int C = IDX(Gx, Gx.is, tj+DOM_BUF, tk+DOM_BUF); s->u_star[C] = s->u_star[C] + add[0]; }
  if(plane[1]) for(int tk = 0; tk < Gx.kn; tk++) for(int tj = 0; tj < Gx.jn; tj++) {
    int C = IDX(Gx, Gx.ie-1, tj+DOM_BUF, tk+DOM_BUF); s->u_star[C] = s->u_star[C] - add[1]; }
  if(plane[2]) for(int tk = 0; tk < Gy.kn; tk++) for(int ti = 0; ti < Gy.in; ti++) {
    int C = IDX(Gy, ti+DOM_BUF, Gy.js, tk+DOM_BUF); s->v_star[C] = s->v_star[C] + add[2]; }
  if(plane[3]) for(int tk = 0; tk < Gy.kn; tk++) for(int ti = 0; ti < Gy.in; ti++) {
    int C = IDX(Gy, ti+DOM_BUF, Gy.je-1, tk+DOM_BUF); s->v_star[C] = s->v_star[C] - add[3]; }
  if(plane[4]) for(int tj = 0; tj < Gz.jn; tj++) for(int ti = 0; ti < Gz.in; ti++) {
    int C = IDX(Gz, ti+DOM_BUF, tj+DOM_BUF, Gz.ks); s->w_star[C] = s->w_star[C] + add[4]; }
  if(plane[5]) for(int tj = 0; tj < Gz.jn; tj++) for(int ti = 0; ti < Gz.in; ti++) {
    int C = IDX(Gz, ti+DOM_BUF, tj+DOM_BUF, Gz.ke-1); s->w_star[C] = s->w_star[C] - add[5]; }
}

/* ---------------------------------------------------------------------
 * Projection: src/bluebottle_kernel.cu:1250-1302, launched by
 * src/cuda_bluebottle.cu:3422-3497 with ddx = 1/dx and p = phi.
 * ------------------------------------------------------------------- */
void orc_project(orc_state *s)
{
  const dom_struct *d = &s->dom;
  const grid_info Gc = d->Gcc, Gx = d->Gfx, Gy = d->Gfy, Gz = d->Gfz;
  const real ddx = 1. / d->dx, ddy = 1. / d->dy, ddz = 1. / d->dz;
  const real dt = s->dt, rho_f = s->rho_f;
  const real *phi = s->phi;
  for(int k = DOM_BUF; k < Gx.ke; k++) for(int j = DOM_BUF; j < Gx.je; j++)
  for(int i = Gx.is; i < Gx.ie; i++) {
    int C = IDX(Gx, i, j, k);
    real gradPhi = abs(s->flag_u[C]) * ddx
      * (phi[IDX(Gc, i, j, k)] - phi[IDX(Gc, i-1, j, k)]);
    s->u[C] = (s->u_star[C] - dt / rho_f * gradPhi);
  }
  for(int k = DOM_BUF; k < Gy.ke; k++) for(int j = Gy.js; j < Gy.je; j++)
  for(int i = DOM_BUF; i < Gy.ie; i++) {
    int C = IDX(Gy, i, j, k);
    real gradPhi = abs(s->flag_v[C]) * ddy
      * (phi[IDX(Gc, i, j, k)] - phi[IDX(Gc, i, j-1, k)]);
    s->v[C] = (s->v_star[C] - dt / rho_f * gradPhi);
  }
  for(int k = Gz.ks; k < Gz.ke; k++) for(int j = DOM_BUF; j < Gz.je; j++)
  for(int i = DOM_BUF; i < Gz.ie; i++) {
    int C = IDX(Gz, i, j, k);
    real gradPhi = abs(s->flag_w[C]) * ddz
      * (phi[IDX(Gc, i, j, k)] - phi[IDX(Gc, i, j, k-1)]);
    s->w[C] = (s->w_star[C] - dt / rho_f * gradPhi);
  }
}

/* Upper bound (exclusive) of the j or k index that forcing_add_c_const
 * reaches for a launch sized on the interior extent n (16-wide blocks):
 * src/cuda_bluebottle.cu:3631-3669 with src/bluebottle_kernel.cu:2042. */
static int c_const_limit(int n, int nb)
{
  int th = n < MAX_THREADS_DIM ? n : MAX_THREADS_DIM;
  int blocks = (int) ceil((real) n / (real) th);
  int lim = blocks * th + DOM_BUF;
  return lim < nb ? lim : nb;
}

/* ---------------------------------------------------------------------
 * Pressure update: update_p (src/bluebottle_kernel.cu:1326-1338), then the
 * mean over the interior (copy_p_noghost + thrust::reduce, summed here in
 * array order) is subtracted by forcing_add_c_const
 * (src/cuda_bluebottle.cu:3616-3671).
 * ------------------------------------------------------------------- */
void orc_update_p(orc_state *s)
{
  const grid_info Gc = s->dom.Gcc;
  for(int k = Gc.ks; k < Gc.ke; k++) for(int j = Gc.js; j < Gc.je; j++)
  for(int i = Gc.is; i < Gc.ie; i++) {
    int C = IDX(Gc, i, j, k);
    s->p[C] = (s->phase[C] < 0) * (s->p0[C] + s->phi[C]);
  }
  real sum = 0.;
  for(int k = Gc.ks; k < Gc.ke; k++) for(int j = Gc.js; j < Gc.je; j++)
  for(int i = Gc.is; i < Gc.ie; i++) sum += s->p[IDX(Gc, i, j, k)];
  real pmean = sum / Gc.s3;
  real val = -pmean;
  int jl = c_const_limit(Gc.jn, Gc.jnb), kl = c_const_limit(Gc.kn, Gc.knb);
  for(int k = DOM_BUF; k < kl; k++) for(int j = DOM_BUF; j < jl; j++)
  for(int i = Gc.isb; i < Gc.ieb; i++) s->p[IDX(Gc, i, j, k)] += val;
}

/* ---------------------------------------------------------------------
 * Time-step size: src/cuda_bluebottle.cu:3500-3577 (explicit build).  The
 * maxima are of the signed velocities over the whole ghosted arrays,
 * starting from 0.
 * ------------------------------------------------------------------- */
real orc_find_dt(orc_state *s)
{
  const dom_struct *d = &s->dom;
  real umax = 0., vmax = 0., wmax = 0.;
  for(int n = 0; n < d->Gfx.s3b; n++) if(s->u[n] > umax) umax = s->u[n];
  for(int n = 0; n < d->Gfy.s3b; n++) if(s->v[n] > vmax) vmax = s->v[n];
  for(int n = 0; n < d->Gfz.s3b; n++) if(s->w[n] > wmax) wmax = s->w[n];
  real tmp;
  if(!s->implicit) {
    tmp = umax / d->dx + 2.*s->nu/d->dx/d->dx;
    tmp += vmax / d->dy + 2.*s->nu/d->dy/d->dy;
    tmp += wmax / d->dz + 2.*s->nu/d->dz/d->dz;
  } else {
    tmp = umax / d->dx;
    tmp += vmax / d->dy;
    tmp += wmax / d->dz;
  }
  real dts = s->CFL / tmp;
  real mn = FLT_MAX;
  if(dts < mn) mn = dts;
  if(s->implicit && mn > 1.1*s->dt) mn = 1.1*s->dt;
  return mn;
}

/* src/cuda_bluebottle.cu:3580-3613 */
void orc_store_u(orc_state *s)
{
  const dom_struct *d = &s->dom;
  size_t fx = d->Gfx.s3b * sizeof(real), fy = d->Gfy.s3b * sizeof(real);
  size_t fz = d->Gfz.s3b * sizeof(real), cc = d->Gcc.s3b * sizeof(real);
  memcpy(s->conv0_u, s->conv_u, fx);
  memcpy(s->conv0_v, s->conv_v, fy);
  memcpy(s->conv0_w, s->conv_w, fz);
  if(!s->implicit) {
    memcpy(s->diff0_u, s->diff_u, fx);
    memcpy(s->diff0_v, s->diff_v, fy);
    memcpy(s->diff0_w, s->diff_w, fz);
  }
  memcpy(s->p0, s->p, cc);
  memcpy(s->u0, s->u, fx);
  memcpy(s->v0, s->v, fy);
  memcpy(s->w0, s->w, fz);
}

/* Linear ramp of one driving component (src/cuda_bluebottle.cu:3727-3760). */
static real ramp(real delta, real a, real m)
{
  if(a == 0) return m;
  if(fabs(delta * a) > fabs(m)) return m;
  return delta * a;
}

/* src/cuda_bluebottle.cu:3674-3814 */
void orc_compute_forcing(orc_state *s)
{
  const dom_struct *d = &s->dom;
  real delta = s->ttime - s->p_tDelay;
  int pid = (s->Kp > 0 || s->Ki > 0 || s->Kd > 0);
  if(delta >= 0) {
    s->gradP.x = ramp(delta, s->gradP.xa, s->gradP.xm);
    s->gradP.y = ramp(delta, s->gradP.ya, s->gradP.ym);
    if(!pid) s->gradP.z = ramp(delta, s->gradP.za, s->gradP.zm);
  }
  delta = s->ttime - s->g_tDelay;
  if(delta >= 0) {
    s->g.x = ramp(delta, s->g.xa, s->g.xm);
    s->g.y = ramp(delta, s->g.ya, s->g.ym);
    s->g.z = ramp(delta, s->g.za, s->g.zm);
  }
  delta = s->ttime - s->p_tDelay;
  if(delta >= 0 && pid) {
    real acc_z = 0, volp = 0, massp = 0;
    for(int i = 0; i < s->nparts; i++) {
      const part_struct *P = &s->parts[i];
      volp += P->r*P->r*P->r;
      massp += P->rho*P->r*P->r*P->r;
      acc_z += P->wdot;
    }
    volp *= 4./3.*PI;
    massp *= 4./3.*PI;
    real volfrac = volp / (d->xl * d->yl * d->zl);
    real rho_avg = massp/volp*volfrac + s->rho_f*(1.-volfrac);
    acc_z /= s->nparts;
    s->pid_int = s->pid_int + acc_z*s->dt;
    s->gradP.z = s->gradP.z
      + (s->Kp*acc_z + s->Ki*(s->pid_int)/s->ttime
      + (s->Kd)*(acc_z-s->pid_back))*rho_avg;
    s->pid_back = acc_z;
  }
  real fx = 0. + -s->gradP.x / s->rho_f;
  real fy = 0. + -s->gradP.y / s->rho_f;
  real fz = 0. + -s->gradP.z / s->rho_f;
  for(int n = 0; n < d->Gfx.s3b; n++) s->f_x[n] = fx;
  for(int n = 0; n < d->Gfy.s3b; n++) s->f_y[n] = fy;
  for(int n = 0; n < d->Gfz.s3b; n++) s->f_z[n] = fz;
}
