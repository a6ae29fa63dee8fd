/*
 * oracle_particle.c -- serial restatement of the reference's particle
 * path: cage/phase/flag construction (integer, bit-exact), Lamb boundary
 * values on flagged faces, pressure fill inside particles, rigid-body
 * interior velocity, history shift of the Lamb coefficients, and the
 * Newton-Euler update with lubrication/contact/spring/wall forces.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Documented departures from the reference (both also in the HIP library):
 *  - part_BC_*, part_BC_p_fill and internal_* evaluate the Lamb/rigid value
 *    only where the reference's 0/1 blend selects it; elsewhere the
 *    reference adds 0 * (finite value), which leaves the field unchanged.
 *  - The contact-list scans of collision_parts/collision_walls
 *    (src/bluebottle_kernel.cu:2726-2735, 2905-2914, ...) test
 *    iSt[MAX_NEIGHBORS], one past the array; here the index is checked
 *    first, so no byte outside iSt[] is read or written.
 *  - The east-wall contact torque (src/bluebottle_kernel.cu:3114) reads a
 *    tangential force Ftx that is never set on that path; it is taken as 0.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "oracle_internal.h"

/* ---------------------------------------------------------------------
 * Cages: src/cuda_particle.cu:357-896, kernels src/particle_kernel.cu:25-533
 * ------------------------------------------------------------------- */
static void build_phase(orc_state *s, int p, real X, real Y, real Z,
  int is, int ie, int js, int je, int ks, int ke)
{
  const dom_struct *d = &s->dom;
  for(int tk = ks; tk < ke; tk++) for(int tj = js; tj < je; tj++)
  for(int ti = is; ti < ie; ti++) {
    real xx = (ti-0.5)*d->dx - (X-d->xs);
    real yy = (tj-0.5)*d->dy - (Y-d->ys);
    real zz = (tk-0.5)*d->dz - (Z-d->zs);
    real dd = sqrt(xx * xx + yy * yy + zz * zz);
    int C = ti + tj*d->Gcc.s1b + tk*d->Gcc.s2b;
    int cutoff = (floor(dd / (s->parts[p].r))) < 1;
    s->phase[C] = cutoff*p + (1 - cutoff)*s->phase[C];
  }
}

static void periodic_cc(int *a, grid_info G, int axis)
{
  if(axis == 0) {
    for(int tk = 0; tk < G.knb; tk++) for(int tj = 0; tj < G.jnb; tj++) {
      a[G.isb + tj*G.s1b + tk*G.s2b] = a[(G.ie-1) + tj*G.s1b + tk*G.s2b];
      a[(G.ieb-1) + tj*G.s1b + tk*G.s2b] = a[G.is + tj*G.s1b + tk*G.s2b];
    }
  } else if(axis == 1) {
    for(int tk = 0; tk < G.knb; tk++) for(int ti = 0; ti < G.inb; ti++) {
      a[ti + G.jsb*G.s1b + tk*G.s2b] = a[ti + (G.je-1)*G.s1b + tk*G.s2b];
      a[ti + (G.jeb-1)*G.s1b + tk*G.s2b] = a[ti + G.js*G.s1b + tk*G.s2b];
    }
  } else {
    for(int tj = 0; tj < G.jnb; tj++) for(int ti = 0; ti < G.inb; ti++) {
      a[ti + tj*G.s1b + G.ksb*G.s2b] = a[ti + tj*G.s1b + (G.ke-1)*G.s2b];
      a[ti + tj*G.s1b + (G.keb-1)*G.s2b] = a[ti + tj*G.s1b + G.ks*G.s2b];
    }
  }
}

/* periodic ghost copy of a face-centred flag array; `normal` selects the
 * (e-2, s+1) source rows of the staggered direction */
static void periodic_face(int *a, grid_info G, int axis, int normal)
{
  int lo = normal ? 2 : 1, hi = normal ? 1 : 0;
  if(axis == 0) {
    for(int tk = 0; tk < G.knb; tk++) for(int tj = 0; tj < G.jnb; tj++) {
      a[G.isb + tj*G.s1b + tk*G.s2b] = a[(G.ie-lo) + tj*G.s1b + tk*G.s2b];
      a[(G.ieb-1) + tj*G.s1b + tk*G.s2b] = a[(G.is+hi) + tj*G.s1b + tk*G.s2b];
    }
  } else if(axis == 1) {
    for(int tk = 0; tk < G.knb; tk++) for(int ti = 0; ti < G.inb; ti++) {
      a[ti + G.jsb*G.s1b + tk*G.s2b] = a[ti + (G.je-lo)*G.s1b + tk*G.s2b];
      a[ti + (G.jeb-1)*G.s1b + tk*G.s2b] = a[ti + (G.js+hi)*G.s1b + tk*G.s2b];
    }
  } else {
    for(int tj = 0; tj < G.jnb; tj++) for(int ti = 0; ti < G.inb; ti++) {
      a[ti + tj*G.s1b + G.ksb*G.s2b] = a[ti + tj*G.s1b + (G.ke-lo)*G.s2b];
      a[ti + tj*G.s1b + (G.keb-1)*G.s2b] = a[ti + tj*G.s1b + (G.ks+hi)*G.s2b];
    }
  }
}

void orc_build_cages(orc_state *s)
{
  const dom_struct *d = &s->dom;
  const grid_info Gc = d->Gcc, Gx = d->Gfx, Gy = d->Gfy, Gz = d->Gfz;
  const BC *b = &s->bc;
  for(int n = 0; n < Gc.s3b; n++) { s->phase[n] = -1; s->phase_shell[n] = 1; }
  for(int n = 0; n < Gx.s3b; n++) s->flag_u[n] = 1;
  for(int n = 0; n < Gy.s3b; n++) s->flag_v[n] = 1;
  for(int n = 0; n < Gz.s3b; n++) s->flag_w[n] = 1;

  /* host loop, src/cuda_particle.cu:459-531 */
  for(int i = 0; i < s->nparts; i++) {
    part_struct *P = &s->parts[i];
    cage_struct *c = &P->cage;
    c->in = (int)(2.0 * ceil(P->r / d->dx)) + 2;
    c->jn = (int)(2.0 * ceil(P->r / d->dy)) + 2;
    c->kn = (int)(2.0 * ceil(P->r / d->dz)) + 2;
    if(d->xn % 2) c->in = c->in - 1;
    if(d->yn % 2) c->jn = c->jn - 1;
    if(d->zn % 2) c->kn = c->kn - 1;
    c->cx = (int)((P->x - d->xs + 0.5 * d->dx) / d->dx);
    c->cy = (int)((P->y - d->ys + 0.5 * d->dy) / d->dy);
    c->cz = (int)((P->z - d->zs + 0.5 * d->dz) / d->dz);
    c->is = (int)(round((P->x-d->xs)/d->dx) - 0.5 * c->in + DOM_BUF);
    c->ie = c->is + c->in;
    if(c->is < Gc.is) {
      c->is = c->is + Gc.ie; c->ibs = Gc.ie; c->ibe = Gc.is;
    } else if(c->ie > Gc.ie) {
      c->ie = c->ie - Gc.ie; c->ibs = Gc.ie; c->ibe = Gc.is;
    } else {
      c->ibs = c->ie; c->ibe = c->ie;
    }
    c->js = (int)(round((P->y-d->ys)/d->dy) - 0.5 * c->jn + DOM_BUF);
    c->je = c->js + c->jn;
    if(c->js < Gc.js) {
      c->js = c->js + Gc.je; c->jbs = Gc.je; c->jbe = Gc.js;
    } else if(c->je > Gc.je) {
      c->je = c->je - Gc.je; c->jbs = Gc.je; c->jbe = Gc.js;
    } else {
      c->jbs = c->je; c->jbe = c->je;
    }
    c->ks = (int)(round((P->z-d->zs)/d->dz) - 0.5 * c->kn + DOM_BUF);
    c->ke = c->ks + c->kn;
    if(c->ks < Gc.ks) {
      c->ks = c->ks + Gc.ke; c->kbs = Gc.ke; c->kbe = Gc.ks;
    } else if(c->ke > Gc.ke) {
      c->ke = c->ke - Gc.ke; c->kbs = Gc.ke; c->kbe = Gc.ks;
    } else {
      c->kbs = c->ke; c->kbe = c->ke;
    }
  }

  /* build_phase launches, src/cuda_particle.cu:539-667 */
  int xPer = (b->uE == PERIODIC), yPer = (b->vN == PERIODIC);
  int zPer = (b->wT == PERIODIC);
  for(int i = 0; i < s->nparts; i++) {
    const part_struct *P = &s->parts[i];
    const cage_struct *c = &P->cage;
    if(!(c->in > 0 && c->jn > 0 && c->kn > 0)) continue;
    real xlo = P->x + (P->x < d->xs + P->r)*d->xl*xPer;
    real ylo = P->y + (P->y < d->ys + P->r)*d->yl*yPer;
    real zlo = P->z + (P->z < d->zs + P->r)*d->zl*zPer;
    real xhi = P->x - d->xl*(P->x > (d->xe - P->r))*xPer;
    real yhi = P->y - d->yl*(P->y > (d->ye - P->r))*yPer;
    real zhi = P->z - d->zl*(P->z > (d->ze - P->r))*zPer;
    /* the reference writes the + forms with the factor order
     * x + xl*(cond)*per in the split launches; the products are exact
     * (cond, per in {0,1}), so one value per side serves all launches */
    build_phase(s, i, xlo, ylo, zlo, c->is, c->ibs, c->js, c->jbs, c->ks, c->kbs);
    if(c->ibe < c->ie)
      build_phase(s, i, xhi, ylo, zlo, c->ibe, c->ie, c->js, c->jbs, c->ks, c->kbs);
    if(c->jbe < c->je)
      build_phase(s, i, xlo, yhi, zlo, c->is, c->ibs, c->jbe, c->je, c->ks, c->kbs);
    if(c->kbe < c->ke)
      build_phase(s, i, xlo, ylo, zhi, c->is, c->ibs, c->js, c->jbs, c->kbe, c->ke);
    if(c->ibe < c->ie && c->jbe < c->je)
      build_phase(s, i, xhi, yhi, zlo, c->ibe, c->ie, c->jbe, c->je, c->ks, c->kbs);
    if(c->ibe < c->ie && c->kbe < c->ke)
      build_phase(s, i, xhi, ylo, zhi, c->ibe, c->ie, c->js, c->jbs, c->kbe, c->ke);
    if(c->jbe < c->je && c->kbe < c->ke)
      build_phase(s, i, xlo, yhi, zhi, c->is, c->ibs, c->jbe, c->je, c->kbe, c->ke);
    if(c->ibe < c->ie && c->jbe < c->je && c->kbe < c->ke)
      build_phase(s, i, xhi, yhi, zhi, c->ibe, c->ie, c->jbe, c->je, c->kbe, c->ke);
  }

  int perx = (b->uW == PERIODIC && b->uE == PERIODIC);
  int pery = (b->uS == PERIODIC && b->uN == PERIODIC);
  int perz = (b->uB == PERIODIC && b->uT == PERIODIC);
  if(perx) periodic_cc(s->phase, Gc, 0);
  if(pery) periodic_cc(s->phase, Gc, 1);
  if(perz) periodic_cc(s->phase, Gc, 2);

  /* phase_shell_{x,y,z} */
  int *ph = s->phase, *sh = s->phase_shell;
  for(int tk = DOM_BUF; tk < Gc.ke; tk++) for(int tj = DOM_BUF; tj < Gc.je; tj++)
  for(int i = Gc.is; i <= Gc.ie; i++) {
    int W = (i-1) + tj*Gc.s1b + tk*Gc.s2b, E = i + tj*Gc.s1b + tk*Gc.s2b;
    sh[E] = sh[E]*(1 - (ph[W] < 0 && ph[E] > -1));
    sh[W] = sh[W]*(1 - (ph[W] > -1 && ph[E] < 0));
  }
  for(int tk = DOM_BUF; tk < Gc.ke; tk++) for(int ti = DOM_BUF; ti < Gc.ie; ti++)
  for(int j = Gc.js; j <= Gc.je; j++) {
    int S = ti + (j-1)*Gc.s1b + tk*Gc.s2b, N = ti + j*Gc.s1b + tk*Gc.s2b;
    sh[N] = sh[N]*(1 - (ph[S] < 0 && ph[N] > -1));
    sh[S] = sh[S]*(1 - (ph[S] > -1 && ph[N] < 0));
  }
  for(int tj = DOM_BUF; tj < Gc.je; tj++) for(int ti = DOM_BUF; ti < Gc.ie; ti++)
  for(int k = Gc.ks; k <= Gc.ke; k++) {
    int B = ti + tj*Gc.s1b + (k-1)*Gc.s2b, T = ti + tj*Gc.s1b + k*Gc.s2b;
    sh[T] = sh[T]*(1 - (ph[B] < 0 && ph[T] > -1));
    sh[B] = sh[B]*(1 - (ph[B] > -1 && ph[T] < 0));
  }
  if(perx) periodic_cc(sh, Gc, 0);
  if(pery) periodic_cc(sh, Gc, 1);
  if(perz) periodic_cc(sh, Gc, 2);

  /* cage_flag_{u,v,w} */
  for(int tk = DOM_BUF; tk < Gc.ke; tk++) for(int tj = DOM_BUF; tj < Gc.je; tj++)
  for(int i = Gc.is; i <= Gc.ie; i++) {
    int W = (i-1) + tj*Gc.s1b + tk*Gc.s2b, E = i + tj*Gc.s1b + tk*Gc.s2b;
    s->flag_u[i + tj*Gx.s1b + tk*Gx.s2b] = 1 - 2*((ph[W] < 0 && ph[E] > -1)
      || (ph[W] > -1 && ph[E] < 0) || ((sh[W] < 1 && sh[E] < 1)));
  }
  for(int tk = DOM_BUF; tk < Gc.ke; tk++) for(int ti = DOM_BUF; ti < Gc.ie; ti++)
  for(int j = Gc.js; j <= Gc.je; j++) {
    int S = ti + (j-1)*Gc.s1b + tk*Gc.s2b, N = ti + j*Gc.s1b + tk*Gc.s2b;
    s->flag_v[ti + j*Gy.s1b + tk*Gy.s2b] = 1 - 2*((ph[S] < 0 && ph[N] > -1)
      || (ph[S] > -1 && ph[N] < 0) || ((sh[S] < 1 && sh[N] < 1)));
  }
  for(int tj = DOM_BUF; tj < Gc.je; tj++) for(int ti = DOM_BUF; ti < Gc.ie; ti++)
  for(int k = Gc.ks; k <= Gc.ke; k++) {
    int B = ti + tj*Gc.s1b + (k-1)*Gc.s2b, T = ti + tj*Gc.s1b + k*Gc.s2b;
    s->flag_w[ti + tj*Gz.s1b + k*Gz.s2b] = 1 - 2*((ph[B] < 0 && ph[T] > -1)
      || (ph[B] > -1 && ph[T] < 0) || ((sh[B] < 1 && sh[T] < 1)));
  }

  /* flag_external_{u,v,w} */
  if(b->uW != PERIODIC && b->uE != PERIODIC)
    for(int tk = 0; tk < Gx.knb; tk++) for(int tj = 0; tj < Gx.jnb; tj++) {
      s->flag_u[Gx.is + tj*Gx.s1b + tk*Gx.s2b] = 0;
      s->flag_u[Gx.ie-1 + tj*Gx.s1b + tk*Gx.s2b] = 0;
    }
  if(b->vS != PERIODIC && b->vN != PERIODIC)
    for(int tk = 0; tk < Gy.knb; tk++) for(int ti = 0; ti < Gy.inb; ti++) {
      s->flag_v[ti + Gy.js*Gy.s1b + tk*Gy.s2b] = 0;
      s->flag_v[ti + (Gy.je-1)*Gy.s1b + tk*Gy.s2b] = 0;
    }
  if(b->wB != PERIODIC && b->wT != PERIODIC)
    for(int tj = 0; tj < Gz.jnb; tj++) for(int ti = 0; ti < Gz.inb; ti++) {
      s->flag_w[ti + tj*Gz.s1b + Gz.ks*Gz.s2b] = 0;
      s->flag_w[ti + tj*Gz.s1b + (Gz.ke-1)*Gz.s2b] = 0;
    }

  /* periodic flag ghosts; the v/x condition repeats bc.vW as the
   * reference does (src/cuda_particle.cu:845) */
  if(b->uW == PERIODIC && b->uE == PERIODIC) periodic_face(s->flag_u, Gx, 0, 1);
  if(b->uS == PERIODIC && b->uN == PERIODIC) periodic_face(s->flag_u, Gx, 1, 0);
  if(b->uB == PERIODIC && b->uT == PERIODIC) periodic_face(s->flag_u, Gx, 2, 0);
  if(b->vW == PERIODIC && b->vW == PERIODIC) periodic_face(s->flag_v, Gy, 0, 0);
  if(b->vS == PERIODIC && b->vN == PERIODIC) periodic_face(s->flag_v, Gy, 1, 1);
  if(b->vB == PERIODIC && b->vT == PERIODIC) periodic_face(s->flag_v, Gy, 2, 0);
  if(b->wW == PERIODIC && b->wE == PERIODIC) periodic_face(s->flag_w, Gz, 0, 0);
  if(b->wS == PERIODIC && b->wN == PERIODIC) periodic_face(s->flag_w, Gz, 1, 0);
  if(b->wB == PERIODIC && b->wT == PERIODIC) periodic_face(s->flag_w, Gz, 2, 1);
}

/* ---------------------------------------------------------------------
 * Lamb boundary values on flagged faces: part_BC_{u,v,w}
 * (src/particle_kernel.cu:535-838).  comp: 0 = u, 1 = v, 2 = w.
 * ------------------------------------------------------------------- */
static void part_BC_comp(orc_state *s, real *f, const int *flag, int comp)
{
  const dom_struct *d = &s->dom;
  const grid_info Gc = d->Gcc;
  const grid_info G = comp == 0 ? d->Gfx : comp == 1 ? d->Gfy : d->Gfz;
  for(int k = (comp == 2 ? G.ks : G.ksb); k < (comp == 2 ? G.ke : G.keb); k++)
  for(int j = (comp == 1 ? G.js : G.jsb); j < (comp == 1 ? G.je : G.jeb); j++)
  for(int i = (comp == 0 ? G.is : G.isb); i < (comp == 0 ? G.ie : G.ieb); i++) {
    int C = i + j*G.s1b + k*G.s2b;
    int im = i - (comp == 0), jm = j - (comp == 1), km = k - (comp == 2);
    int P0 = s->phase[im + jm*Gc.s1b + km*Gc.s2b];
    int P1 = s->phase[i + j*Gc.s1b + k*Gc.s2b];
    int P, PP;
    if(P0 > -1) { P = P0; PP = P0; }
    else if(P1 > -1) { P = P1; PP = P1; }
    else { P = 0; PP = -1; }
    int check = (flag[C] < 1) && (PP > -1);
    if(!check) continue;
    const part_struct *pt = &s->parts[P];
    real a = pt->r;
    /* face position: staggered coordinate on the normal axis */
    real xf = comp == 0 ? (i-DOM_BUF) * d->dx + d->xs : (i-0.5) * d->dx + d->xs;
    real yf = comp == 1 ? (j-DOM_BUF) * d->dy + d->ys : (j-0.5) * d->dy + d->ys;
    real zf = comp == 2 ? (k-DOM_BUF) * d->dz + d->zs : (k-0.5) * d->dz + d->zs;
    real x = xf - pt->x;
    if(x <= 2.*a-d->xl) x += d->xl;
    if(x >= d->xl-2.*a) x -= d->xl;
    real y = yf - pt->y;
    if(y <= 2.*a-d->yl) y += d->yl;
    if(y >= d->yl-2.*a) y -= d->yl;
    real z = zf - pt->z;
    if(z <= 2.*a-d->zl) z += d->zl;
    if(z >= d->zl-2.*a) z -= d->zl;
    real r, theta, phi;
    orc_xyz2rtp(x, y, z, &r, &theta, &phi);
    real Ux, Uy, Uz;
    orc_lamb_vel(s, pt->order, a, r, theta, phi, P, &Ux, &Uy, &Uz);
    real val;
    if(comp == 0) {
      real ocrossr_x = pt->oy*z - pt->oz*y;
      real odotcrossr_x = pt->oydot*z - pt->ozdot*y;
      Ux += pt->u + ocrossr_x;
      Ux += 0.1/s->nu *(r*r*r*r*r-a*a*a*a*a)/(r*r*r) * odotcrossr_x;
      val = Ux;
    } else if(comp == 1) {
      real ocrossr_y = -(pt->ox*z - pt->oz*x);
      real odotcrossr_y = -(pt->oxdot*z - pt->ozdot*x);
      Uy += pt->v + ocrossr_y;
      Uy += 0.1/s->nu *(r*r*r*r*r-a*a*a*a*a)/(r*r*r) * odotcrossr_y;
      val = Uy;
    } else {
      real ocrossr_z = pt->ox*y - pt->oy*x;
      real odotcrossr_z = pt->oxdot*y - pt->oydot*x;
      Uz += pt->w + ocrossr_z;
      Uz += 0.1/s->nu *(r*r*r*r*r-a*a*a*a*a)/(r*r*r) * odotcrossr_z;
      val = Uz;
    }
    f[C] = val;
  }
}

void orc_part_BC(orc_state *s)
{
  part_BC_comp(s, s->u, s->flag_u, 0);
  part_BC_comp(s, s->v, s->flag_v, 1);
  part_BC_comp(s, s->w, s->flag_w, 2);
}

void orc_part_BC_star(orc_state *s)
{
  part_BC_comp(s, s->u_star, s->flag_u, 0);
  part_BC_comp(s, s->v_star, s->flag_v, 1);
  part_BC_comp(s, s->w_star, s->flag_w, 2);
}

/* src/particle_kernel.cu:942-1018 */
void orc_part_p_fill(orc_state *s)
{
  const dom_struct *d = &s->dom;
  const grid_info Gc = d->Gcc;
  const int stride = s->coeff_stride;
  for(int tk = Gc.ks; tk < Gc.ke; tk++) for(int tj = Gc.js; tj < Gc.je; tj++)
  for(int i = Gc.is; i < Gc.ie; i++) {
    int CC = i + tj*Gc.s1b + tk*Gc.s2b;
    int P = s->phase[CC];
    if(P < 0) continue;
    const part_struct *pt = &s->parts[P];
    real a = pt->r;
    real x = (i-0.5) * d->dx + d->xs - pt->x;
    if(x <= 2.*a-d->xl) x += d->xl;
    if(x >= d->xl-2.*a) x -= d->xl;
    real y = (tj-0.5) * d->dy + d->ys - pt->y;
    if(y <= 2.*a-d->yl) y += d->yl;
    if(y >= d->yl-2.*a) y -= d->yl;
    real z = (tk-0.5) * d->dz + d->zs - pt->z;
    if(z <= 2.*a-d->zl) z += d->zl;
    if(z >= d->zl-2.*a) z -= d->zl;
    real r, theta, phi;
    orc_xyz2rtp(x, y, z, &r, &theta, &phi);
    real pp_tmp = orc_X(0, theta, phi, s->coeff[0], s->coeff[1], P, stride);
    pp_tmp *= s->mu*s->nu/(a*a);
    real ox = pt->ox, oy = pt->oy, oz = pt->oz;
    real ocrossr2 = (oy*z - oz*y) * (oy*z - oz*y);
    ocrossr2 += (ox*z - oz*x) * (ox*z - oz*x);
    ocrossr2 += (ox*y - oy*x) * (ox*y - oy*x);
    real rhoV = s->rho_f;
    real accdotr = (-s->gradP.x/rhoV - pt->udot)*x + (-s->gradP.y/rhoV - pt->vdot)*y
      + (-s->gradP.z/rhoV - pt->wdot)*z;
    pp_tmp += 0.5 * s->rho_f * ocrossr2 + s->rho_f * accdotr;
    s->p[CC] = pp_tmp;
  }
}

/* src/bluebottle_kernel.cu:3916-4022 */
static void internal_comp(orc_state *s, real *f, const int *flag, int comp)
{
  const dom_struct *d = &s->dom;
  const grid_info Gc = d->Gcc;
  const grid_info G = comp == 0 ? d->Gfx : comp == 1 ? d->Gfy : d->Gfz;
  for(int k = (comp == 2 ? G.ks : DOM_BUF); k < G.ke; k++)
  for(int j = (comp == 1 ? G.js : DOM_BUF); j < G.je; j++)
  for(int i = (comp == 0 ? G.is : DOM_BUF); i < G.ie; i++) {
    int C = i + j*G.s1b + k*G.s2b;
    int im = i - (comp == 0), jm = j - (comp == 1), km = k - (comp == 2);
    int p0 = s->phase[im + jm*Gc.s1b + km*Gc.s2b];
    int p1 = s->phase[i + j*Gc.s1b + k*Gc.s2b];
    int fl = flag[C];
    if(!(p0 > -1 && p1 > -1 && fl != -1)) continue;
    const part_struct *pt = &s->parts[p1];
    real xf = comp == 0 ? (i - DOM_BUF) * d->dx + d->xs : (i - 0.5) * d->dx + d->xs;
    real yf = comp == 1 ? (j - DOM_BUF) * d->dy + d->ys : (j - 0.5) * d->dy + d->ys;
    real zf = comp == 2 ? (k - DOM_BUF) * d->dz + d->zs : (k - 0.5) * d->dz + d->zs;
    real rx = xf - pt->x;
    if(rx <= 2.*pt->r-d->xl) rx += d->xl;
    if(rx >= d->xl-2.*pt->r) rx -= d->xl;
    real ry = yf - pt->y;
    if(ry <= 2.*pt->r-d->yl) ry += d->yl;
    if(ry >= d->yl-2.*pt->r) ry -= d->yl;
    real rz = zf - pt->z;
    if(rz <= 2.*pt->r-d->zl) rz += d->zl;
    if(rz >= d->zl-2.*pt->r) rz -= d->zl;
    if(comp == 0) f[C] = (pt->oy*rz - pt->oz*ry) + pt->u;
    else if(comp == 1) f[C] = (pt->oz*rx - pt->ox*rz) + pt->v;
    else f[C] = (pt->ox*ry - pt->oy*rx) + pt->w;
  }
}

void orc_parts_internal(orc_state *s)
{
  if(s->nparts <= 0) return;
  internal_comp(s, s->u, s->flag_u, 0);
  internal_comp(s, s->v, s->flag_v, 1);
  internal_comp(s, s->w, s->flag_w, 2);
}

/* src/cuda_particle.cu:1046-1081: coeff -> coeff00 */
void orc_store_coeffs(orc_state *s)
{
  size_t n = (size_t) s->nparts * s->coeff_stride * sizeof(real);
  for(int c = 0; c < 6; c++) memcpy(s->coeff[12 + c], s->coeff[c], n);
}

/* ---------------------------------------------------------------------
 * Newton-Euler update and interaction forces:
 * src/bluebottle_kernel.cu:2247-3617, driven by
 * src/cuda_bluebottle.cu:4317-4554.
 * ------------------------------------------------------------------- */
static void collision_init(orc_state *s)
{
  for(int j = 0; j < s->nparts; j++) {
    part_struct *P = &s->parts[j];
    P->iFx = 0.; P->iFy = 0.; P->iFz = 0.;
    P->iLx = 0.; P->iLy = 0.; P->iLz = 0.;
  }
}

static void spring_parts(orc_state *s)
{
  for(int i = 0; i < s->nparts; i++) {
    part_struct *P = &s->parts[i];
    if(!(P->spring_k > 0.)) continue;
    real nx = P->x-P->spring_x;
    real ny = P->y-P->spring_y;
    real nz = P->z-P->spring_z;
    real n = sqrt(nx*nx+ny*ny+nz*nz);
    real nhatx = nx / n, nhaty = ny / n, nhatz = nz / n;
    real lx = P->spring_l*nhatx, ly = P->spring_l*nhaty, lz = P->spring_l*nhatz;
    real dx = P->x-P->spring_x-lx;
    real dy = P->y-P->spring_y-ly;
    real dz = P->z-P->spring_z-lz;
    P->kFx = - P->spring_k * dx;
    P->kFy = - P->spring_k * dy;
    P->kFz = - P->spring_k * dz;
  }
}

static void move_a(orc_state *s)
{
  const g_struct g = s->g;
  const gradP_struct gp = s->gradP;
  for(int pp = 0; pp < s->nparts; pp++) {
    part_struct *P = &s->parts[pp];
    real vol = 4./3. * PI * P->r*P->r*P->r;
    real m = vol * P->rho;
    if(P->translating) {
      P->udot = (P->Fx + P->kFx + P->iFx + P->aFx - vol*gp.x) / m
        + (P->rho - s->rho_f) / P->rho * g.x;
      P->vdot = (P->Fy + P->kFy + P->iFy + P->aFy - vol*gp.y) / m
        + (P->rho - s->rho_f) / P->rho * g.y;
      P->wdot = (P->Fz + P->kFz + P->iFz + P->aFz - vol*gp.z) / m
        + (P->rho - s->rho_f) / P->rho * g.z;
      P->u = P->u0 + 0.5*s->dt*(P->udot + P->udot0);
      P->v = P->v0 + 0.5*s->dt*(P->vdot + P->vdot0);
      P->w = P->w0 + 0.5*s->dt*(P->wdot + P->wdot0);
    }
    if(P->rotating) {
      real I = 0.4 * m * P->r*P->r;
      P->oxdot = (P->Lx + P->iLx + P->aLx) / I;
      P->oydot = (P->Ly + P->iLy + P->aLy) / I;
      P->ozdot = (P->Lz + P->iLz + P->aLz) / I;
      P->ox = P->ox0 + 0.5*s->dt*(P->oxdot + P->oxdot0);
      P->oy = P->oy0 + 0.5*s->dt*(P->oydot + P->oydot0);
      P->oz = P->oz0 + 0.5*s->dt*(P->ozdot + P->ozdot0);
    }
  }
}

static void rotate(real qr, real qi, real qj, real qk, real *pi, real *pj,
  real *pk)
{
  real Pr = *pi*qi + *pj*qj + *pk*qk;
  real Pi = *pi*qr - *pj*qk + *pk*qj;
  real Pj = *pi*qk + *pj*qr - *pk*qi;
  real Pk = -*pi*qj + *pj*qi + *pk*qr;
  *pi = qr*Pi + qi*Pr + qj*Pk - qk*Pj;
  *pj = qr*Pj - qi*Pk + qj*Pr + qk*Pi;
  *pk = qr*Pk + qi*Pj - qj*Pi + qk*Pr;
}

static void move_b(orc_state *s)
{
  const dom_struct *d = &s->dom;
  const g_struct g = s->g;
  const gradP_struct gp = s->gradP;
  const real dt = s->dt;
  for(int pp = 0; pp < s->nparts; pp++) {
    part_struct *P = &s->parts[pp];
    real vol = 4./3. * PI * P->r*P->r*P->r;
    real m = vol * P->rho;
    if(P->translating) {
      P->udot = (P->Fx + P->kFx + P->iFx + P->aFx - vol*gp.x) / m
        + (P->rho - s->rho_f) / P->rho * g.x;
      P->vdot = (P->Fy + P->kFy + P->iFy + P->aFy - vol*gp.y) / m
        + (P->rho - s->rho_f) / P->rho * g.y;
      P->wdot = (P->Fz + P->kFz + P->iFz + P->aFz - vol*gp.z) / m
        + (P->rho - s->rho_f) / P->rho * g.z;
      P->u = P->u0 + 0.5*dt*(P->udot + P->udot0);
      P->v = P->v0 + 0.5*dt*(P->vdot + P->vdot0);
      P->w = P->w0 + 0.5*dt*(P->wdot + P->wdot0);
      P->x = P->x0 + 0.5*dt*(P->u + P->u0);
      if(P->x < d->xs) P->x = P->x + d->xl;
      else if(P->x > d->xe) P->x = P->x - d->xl;
      P->y = P->y0 + 0.5*dt*(P->v + P->v0);
      if(P->y < d->ys) P->y = P->y + d->yl;
      else if(P->y > d->ye) P->y = P->y - d->yl;
      P->z = P->z0 + 0.5*dt*(P->w + P->w0);
      if(P->z < d->zs) P->z = P->z + d->zl;
      else if(P->z > d->ze) P->z = P->z - d->zl;
      P->x0 = P->x; P->y0 = P->y; P->z0 = P->z;
      P->u0 = P->u; P->v0 = P->v; P->w0 = P->w;
      P->udot0 = P->udot; P->vdot0 = P->vdot; P->wdot0 = P->wdot;
    }
    if(P->rotating) {
      real I = 0.4 * m * P->r*P->r;
      P->oxdot = (P->Lx + P->iLx + P->aLx) / I;
      P->oydot = (P->Ly + P->iLy + P->aLy) / I;
      P->ozdot = (P->Lz + P->iLz + P->aLz) / I;
      P->ox = P->ox0 + 0.5*dt*(P->oxdot + P->oxdot0);
      P->oy = P->oy0 + 0.5*dt*(P->oydot + P->oydot0);
      P->oz = P->oz0 + 0.5*dt*(P->ozdot + P->ozdot0);
      real mag = 0.5*sqrt(P->ox*P->ox + P->oy*P->oy + P->oz*P->oz);
      mag += 0.5*sqrt(P->ox0*P->ox0 + P->oy0*P->oy0 + P->oz0*P->oz0);
      real X = 0, Y = 0, Z = 0;
      if(mag > 0) {
        X = 0.5 * (P->ox + P->ox0) / mag;
        Y = 0.5 * (P->oy + P->oy0) / mag;
        Z = 0.5 * (P->oz + P->oz0) / mag;
      }
      real theta = mag * dt;
      real qr = cos(0.5*theta);
      real qi = X * sin(0.5*theta);
      real qj = Y * sin(0.5*theta);
      real qk = Z * sin(0.5*theta);
      rotate(qr, qi, qj, qk, &P->axx, &P->axy, &P->axz);
      rotate(qr, qi, qj, qk, &P->ayx, &P->ayy, &P->ayz);
      rotate(qr, qi, qj, qk, &P->azx, &P->azy, &P->azz);
      P->ox0 = P->ox; P->oy0 = P->oy; P->oz0 = P->oz;
      P->oxdot0 = P->oxdot; P->oydot0 = P->oydot; P->ozdot0 = P->ozdot;
    }
  }
}

/* contact-list helpers, bounds checked */
static int st_find(const part_struct *P, int key)
{
  int q = 0;
  while(q < MAX_NEIGHBORS && P->iSt[q] != key) q++;
  return q;
}

static int st_free(const part_struct *P)
{
  int q = 0;
  while(q < MAX_NEIGHBORS && P->iSt[q] != -1) q++;
  return q < MAX_NEIGHBORS ? q : MAX_NEIGHBORS - 1;
}

/* Particle-particle lubrication and contact for particle i against j
 * (src/bluebottle_kernel.cu:2593-2857). */
static void pair_force(orc_state *s, int i, int j, real eps, int ilr)
{
  const dom_struct *d = &s->dom;
  part_struct *Pi = &s->parts[i];
  const part_struct *Pj = &s->parts[j];
  const BC *bc = &s->bc;
  const real mu = s->mu, nu = s->nu, rhof = s->rho_f, dt = s->dt;
  real ai = Pi->r, aj = Pj->r;
  real B = aj / ai;
  real hN = ilr * Pi->r;
  real rx, rx1, rx2, ry, ry1, ry2, rz, rz1, rz2, r;
  real h, ah, lnah;
  real nx, ny, nz, udotn, unx, uny, unz, utx, uty, utz, ut;
  real tx, ty, tz, t, bx, by, bz, b;
  real omegax, omegay, omegaz, omega;
  real ocrossnx, ocrossny, ocrossnz, utcrossnx, utcrossny, utcrossnz, opB;
  real Fnx, Fny, Fnz, Ftx, Fty, Ftz, Lox, Loy, Loz;
  real xi = Pi->x, xj = Pj->x;
  rx = xi - xj; rx1 = xi - (xj + d->xl); rx2 = xi - (xj - d->xl);
  if(rx1*rx1 < rx*rx) rx = rx1;
  if(rx2*rx2 < rx*rx) rx = rx2;
  rx = (bc->uW == PERIODIC) * rx + (bc->uW != PERIODIC) * (xi - xj);
  real yi = Pi->y, yj = Pj->y;
  ry = yi - yj; ry1 = yi - (yj + d->yl); ry2 = yi - (yj - d->yl);
  if(ry1*ry1 < ry*ry) ry = ry1;
  if(ry2*ry2 < ry*ry) ry = ry2;
  ry = (bc->vS == PERIODIC) * ry + (bc->vS != PERIODIC) * (yi - yj);
  real zi = Pi->z, zj = Pj->z;
  rz = zi - zj; rz1 = zi - (zj + d->zl); rz2 = zi - (zj - d->zl);
  if(rz1*rz1 < rz*rz) rz = rz1;
  if(rz2*rz2 < rz*rz) rz = rz2;
  rz = (bc->wB == PERIODIC) * rz + (bc->wB != PERIODIC) * (zi - zj);
  real ux = 0.5*((Pi->u - Pj->u) + (Pi->u0 - Pj->u0));
  real uy = 0.5*((Pi->v - Pj->v) + (Pi->v0 - Pj->v0));
  real uz = 0.5*((Pi->w - Pj->w) + (Pi->w0 - Pj->w0));
  r = sqrt(rx*rx + ry*ry + rz*rz);
  omegax = 0.5*((Pi->ox + Pj->ox) + (Pi->ox0 + Pj->ox0));
  omegay = 0.5*((Pi->oy + Pj->oy) + (Pi->oy0 + Pj->oy0));
  omegaz = 0.5*((Pi->oz + Pj->oz) + (Pi->oz0 + Pj->oz0));
  omega = sqrt(omegax*omegax + omegay*omegay + omegaz*omegaz);
  h = r - ai - aj;
  nx = rx / r; ny = ry / r; nz = rz / r;
  udotn = ux * nx + uy * ny + uz * nz;
  unx = udotn * nx; uny = udotn * ny; unz = udotn * nz;
  utx = ux - unx; uty = uy - uny; utz = uz - unz;
  ut = sqrt(utx*utx + uty*uty + utz*utz);
  if(ut > 0) {
    tx = utx / ut; ty = uty / ut; tz = utz / ut;
    bx = ny*tz - nz*ty; by = -nx*tz + nz*tx; bz = nx*ty - ny*tx;
    b = sqrt(bx*bx + by*by + bz*bz);
    bx = bx / b; by = by / b; bz = bz / b;
  } else if(omega > 0) {
    bx = omegax / omega; by = omegay / omega; bz = omegaz / omega;
    tx = by*nz - bz*ny; ty = -bx*nz + bz*nx; tz = bx*ny - by*nx;
    t = sqrt(tx*tx + ty*ty + tz*tz);
    tx = tx / t; ty = ty / t; tz = tz / t;
  } else {
    tx = 1.; ty = 0.; tz = 0.;
    bx = ny*tz - nz*ty; by = -nx*tz + nz*tx; bz = nx*ty - ny*tx;
    b = sqrt(bx*bx + by*by + bz*bz);
    bx = bx / b; by = by / b; bz = bz / b;
  }
  (void) bx; (void) by; (void) bz; (void) tx; (void) ty; (void) tz;
  opB = 1 + B;
  ocrossnx = omegay*nz - omegaz*ny;
  ocrossny = -omegax*nz + omegaz*nx;
  ocrossnz = omegax*ny - omegay*nx;
  utcrossnx = uty*nz - utz*ny;
  utcrossny = -utx*nz + utz*nx;
  utcrossnz = utx*ny - uty*nx;
  if(h < hN && h > 0) {
    int q = st_find(Pi, j);
    if(q < MAX_NEIGHBORS) { Pi->iSt[q] = -1; Pi->St[q] = 0.; }
    if(h < eps*Pi->r) h = eps*Pi->r;
    ah = ai/h - ai/hN;
    lnah = log(hN/h);
    Fnx = -1. * B*B / (opB*opB) * ah - B*(1.+7.*B+B*B)/(5.*opB*opB*opB)*lnah;
    Fny = Fnx; Fnz = Fnx;
    Fnx *= 6.*PI*mu*ai*unx;
    Fny *= 6.*PI*mu*ai*uny;
    Fnz *= 6.*PI*mu*ai*unz;
    Ftx = -6.*PI*mu*ai*utx*4.*B*(2.+B+2.*B*B)/(15.*opB*opB*opB)*lnah;
    Fty = -6.*PI*mu*ai*uty*4.*B*(2.+B+2.*B*B)/(15.*opB*opB*opB)*lnah;
    Ftz = -6.*PI*mu*ai*utz*4.*B*(2.+B+2.*B*B)/(15.*opB*opB*opB)*lnah;
    Ftx += 8.*PI*mu*ai*ai*ocrossnx*B*(4.+B)/(10.*opB*opB)*lnah;
    Fty += 8.*PI*mu*ai*ai*ocrossny*B*(4.+B)/(10.*opB*opB)*lnah;
    Ftz += 8.*PI*mu*ai*ai*ocrossnz*B*(4.+B)/(10.*opB*opB)*lnah;
    Lox = -8.*PI*mu*ai*ai*utcrossnx*B*(4.+B)/(10.*opB*opB)*lnah;
    Loy = -8.*PI*mu*ai*ai*utcrossny*B*(4.+B)/(10.*opB*opB)*lnah;
    Loz = -8.*PI*mu*ai*ai*utcrossnz*B*(4.+B)/(10.*opB*opB)*lnah;
    Lox += -8.*PI*mu*ai*ai*ai*omegax*2.*B/(5.*opB)*lnah;
    Loy += -8.*PI*mu*ai*ai*ai*omegay*2.*B/(5.*opB)*lnah;
    Loz += -8.*PI*mu*ai*ai*ai*omegaz*2.*B/(5.*opB)*lnah;
  } else {
    ah = 0; lnah = 0;
    Fnx = 0; Fny = 0; Fnz = 0; Ftx = 0; Fty = 0; Ftz = 0;
    Lox = 0; Loy = 0; Loz = 0;
  }
  (void) ah;
  if(h < 0) {
    int q = st_find(Pi, j);
    if(q == MAX_NEIGHBORS) {
      q = st_free(Pi);
      Pi->iSt[q] = j;
      Pi->St[q] = 1./9.*Pi->rho/rhof*2.*Pi->r*fabs(udotn)/nu;
    }
    real Vx = -utx + 0.5*(ai + aj + h)*ocrossnx;
    real Vy = -uty + 0.5*(ai + aj + h)*ocrossny;
    real Vz = -utz + 0.5*(ai + aj + h)*ocrossnz;
    real Hi = 0.5*Pi->E/(1.+Pi->sigma);
    real kt = 8./((1.-Pi->sigma*Pi->sigma)/Hi+(1.-Pi->sigma*Pi->sigma)/Hi)
      /sqrt(1./ai)*sqrt(-h);
    real Vdotn = Vx*nx + Vy*ny + Vz*nz;
    real sx = (Vx - Vdotn * nx) * dt;
    real sy = (Vy - Vdotn * ny) * dt;
    real sz = (Vz - Vdotn * nz) * dt;
    real k = 4./3./((1.-Pi->sigma*Pi->sigma)/Pi->E
      + (1.-Pj->sigma*Pj->sigma)/Pj->E)/sqrt(1./ai + 1./aj);
    real xcx0 = 1.e-4;
    real e = Pi->e_dry + (1.+Pi->e_dry)/Pi->St[q]*log(xcx0);
    if(e < 0) e = 0;
    real alpha = -2.263*pow(e,0.3948)+2.22;
    real eta = alpha*sqrt(4./3.*PI*ai*ai*ai*Pi->rho*k*sqrt(-h));
    Fnx = (sqrt(-h*h*h)*k - eta*udotn)*nx;
    Fny = (sqrt(-h*h*h)*k - eta*udotn)*ny;
    Fnz = (sqrt(-h*h*h)*k - eta*udotn)*nz;
    real coeff_fric = 0.5 * (Pi->coeff_fric + Pj->coeff_fric);
    Ftx = -kt * sx; Fty = -kt * sy; Ftz = -kt * sz;
    real Ftdotn = Ftx*nx + Fty*ny + Ftz*nz;
    Ftx = Ftx - Ftdotn * nx;
    Fty = Fty - Ftdotn * ny;
    Ftz = Ftz - Ftdotn * nz;
    real Ft = sqrt(Ftx*Ftx + Fty*Fty + Ftz*Ftz);
    real Fn = sqrt(Fnx*Fnx + Fny*Fny + Fnz*Fnz);
    if(Ft > coeff_fric * Fn) {
      Ftx = coeff_fric * Fn * Ftx / Ft;
      Fty = coeff_fric * Fn * Fty / Ft;
      Ftz = coeff_fric * Fn * Ftz / Ft;
    }
    Lox = -(ai+0.5*h)*((Fny+Fty)*nz-(Fnz+Ftz)*ny);
    Loy =  (ai+0.5*h)*((Fnx+Ftx)*nz-(Fnz+Ftz)*nx);
    Loz = -(ai+0.5*h)*((Fnx+Ftx)*ny-(Fny+Fty)*nx);
  }
  Pi->iFx += Fnx + Ftx;
  Pi->iFy += Fny + Fty;
  Pi->iFz += Fnz + Ftz;
  Pi->iLx += Lox;
  Pi->iLy += Loy;
  Pi->iLz += Loz;
}

typedef struct { int *partInd, *partBin, *binStart, *binEnd, nBins; } bins_t;

/* bin_fill + stable sort_by_key + bin_start
 * (src/bluebottle_kernel.cu:2412-2498, src/cuda_bluebottle.cu:4358-4381) */
static void make_bins(orc_state *s, bins_t *bn)
{
  const dom_struct *bd = &s->binDom;
  int n = s->nparts;
  bn->nBins = bd->Gcc.s3;
  bn->partInd = malloc(n * sizeof(int));
  bn->partBin = malloc(n * sizeof(int));
  bn->binStart = malloc(bn->nBins * sizeof(int));
  bn->binEnd = malloc(bn->nBins * sizeof(int));
  for(int pp = 0; pp < n; pp++) {
    int ibin = floor((s->parts[pp].x - bd->xs)/bd->dx);
    int jbin = floor((s->parts[pp].y - bd->ys)/bd->dy);
    int kbin = floor((s->parts[pp].z - bd->zs)/bd->dz);
    /* a particle exactly on an upper face would index one past the bins */
    ibin = ibin < 0 ? 0 : (ibin > bd->xn - 1 ? bd->xn - 1 : ibin);
    jbin = jbin < 0 ? 0 : (jbin > bd->yn - 1 ? bd->yn - 1 : jbin);
    kbin = kbin < 0 ? 0 : (kbin > bd->zn - 1 ? bd->zn - 1 : kbin);
    int c = ibin + jbin*bd->Gcc.s1 + kbin*bd->Gcc.s2;
    bn->partInd[pp] = pp;
    bn->partBin[pp] = c;
    s->parts[pp].bin = c;
  }
  /* stable insertion sort by bin */
  for(int a = 1; a < n; a++) {
    int kb = bn->partBin[a], ki = bn->partInd[a];
    int b = a - 1;
    while(b >= 0 && bn->partBin[b] > kb) {
      bn->partBin[b+1] = bn->partBin[b];
      bn->partInd[b+1] = bn->partInd[b];
      b--;
    }
    bn->partBin[b+1] = kb;
    bn->partInd[b+1] = ki;
  }
  for(int c = 0; c < bn->nBins; c++) { bn->binStart[c] = -1; bn->binEnd[c] = -1; }
  for(int index = 0; index < n; index++) {
    int bin = bn->partBin[index];
    if(index == 0 || bin != bn->partBin[index-1]) {
      bn->binStart[bin] = index;
      if(index > 0) bn->binEnd[bn->partBin[index-1]] = index;
    }
    if(index == n - 1) bn->binEnd[bin] = index + 1;
  }
}

static void free_bins(bins_t *bn)
{
  free(bn->partInd); free(bn->partBin); free(bn->binStart); free(bn->binEnd);
}

/* src/bluebottle_kernel.cu:2500-2873 */
static void collision_parts(orc_state *s, const bins_t *bn, real eps)
{
  const dom_struct *bd = &s->binDom;
  const BC *bc = &s->bc;
  int ilr = (int) s->interactionLengthRatio;
  for(int index = 0; index < s->nparts; index++) {
    int i = bn->partInd[index];
    int bin = bn->partBin[index];
    int kbin = bin / bd->Gcc.s2;
    int jbin = (bin - kbin*bd->Gcc.s2) / bd->Gcc.s1;
    int ibin = bin - kbin*bd->Gcc.s2 - jbin*bd->Gcc.s1;
    int fW = bd->Gcc.is - 1, fE = bd->Gcc.ie - 2;
    int fS = bd->Gcc.js - 1, fN = bd->Gcc.je - 2;
    int fB = bd->Gcc.ks - 1, fT = bd->Gcc.ke - 2;
    int xnBin = (bd->xn > 2), ynBin = (bd->yn > 2), znBin = (bd->zn > 2);
    for(int n = -1; n <= 1; n++) {
      int kStride;
      if((n == -1 && kbin == fB && bc->uB != PERIODIC) ||
         (n == 1 && kbin == fT && bc->uT != PERIODIC) ||
         (n == -1 && kbin == fB && bc->uB == PERIODIC && znBin == 0) ||
         (n == 1 && kbin == fT && bc->uT == PERIODIC && znBin == 0)) continue;
      else if(n == -1 && kbin == fB && bc->uB == PERIODIC) kStride = fT*bd->Gcc.s2;
      else if(n == 1 && kbin == fT && bc->uT == PERIODIC) kStride = fB*bd->Gcc.s2;
      else kStride = (kbin + n)*bd->Gcc.s2;
      for(int m = -1; m <= 1; m++) {
        int jStride;
        if((m == -1 && jbin == fS && bc->uS != PERIODIC) ||
           (m == 1 && jbin == fN && bc->uN != PERIODIC) ||
           (m == -1 && jbin == fS && bc->uS == PERIODIC && ynBin == 0) ||
           (m == 1 && jbin == fN && bc->uN == PERIODIC && ynBin == 0)) continue;
        else if(m == -1 && jbin == fS && bc->uS == PERIODIC) jStride = fN*bd->Gcc.s1;
        else if(m == 1 && jbin == fN && bc->uN == PERIODIC) jStride = fS*bd->Gcc.s1;
        else jStride = (jbin + m)*bd->Gcc.s1;
        for(int l = -1; l <= 1; l++) {
          int iStride;
          if((l == -1 && ibin == fW && bc->uW != PERIODIC) ||
             (l == 1 && ibin == fE && bc->uE != PERIODIC) ||
             (l == -1 && ibin == fW && bc->uW == PERIODIC && xnBin == 0) ||
             (l == 1 && ibin == fE && bc->uE == PERIODIC && xnBin == 0)) continue;
          else if(l == -1 && ibin == fW && bc->uW == PERIODIC) iStride = fE;
          else if(l == 1 && ibin == fE && bc->uE == PERIODIC) iStride = fW;
          else iStride = ibin + l;
          int adjBin = iStride + jStride + kStride;
          if(adjBin < 0 || adjBin >= bn->nBins) continue;
          int adjStart = bn->binStart[adjBin];
          int adjEnd = bn->binEnd[adjBin];
          if(adjStart == -1) continue;
          for(int target = adjStart; target < adjEnd; target++) {
            int j = bn->partInd[target];
            if(j != i) pair_force(s, i, j, eps, ilr);
          }
        }
      }
    }
  }
}

/* One wall of collision_walls (src/bluebottle_kernel.cu:2875-3590).
 * axis: 0 x, 1 y, 2 z; hi: 0 low wall (W,S,B), 1 high wall (E,N,T). */
static void wall_force(orc_state *s, int i, int axis, int hi, real eps)
{
  const dom_struct *d = &s->dom;
  const BC *bc = &s->bc;
  part_struct *P = &s->parts[i];
  const real mu = s->mu, nu = s->nu, rhof = s->rho_f, dt = s->dt;
  int ilr = (int) s->interactionLengthRatio;
  real ai = P->r;
  real hN = ilr * P->r;
  int key = -10 - (2*axis + hi);
  real pos = axis == 0 ? P->x : axis == 1 ? P->y : P->z;
  real wall;
  int isTrue;
  /* wall velocity components (u, v, w) and pressure type of this wall */
  real bu, bv, bw;
  switch(2*axis + hi) {
    case 0: wall = d->xs + bc->dsW; isTrue = bc->pW == NEUMANN;
      bu = bc->uWD; bv = bc->vWD; bw = bc->wWD; break;
    case 1: wall = d->xe - bc->dsE; isTrue = bc->pE == NEUMANN;
      bu = bc->uED; bv = bc->vED; bw = bc->wED; break;
    case 2: wall = d->ys + bc->dsS; isTrue = bc->pS == NEUMANN;
      bu = bc->uSD; bv = bc->vSD; bw = bc->wSD; break;
    case 3: wall = d->ye - bc->dsN; isTrue = bc->pN == NEUMANN;
      bu = bc->uND; bv = bc->vND; bw = bc->wND; break;
    case 4: wall = d->zs + bc->dsB; isTrue = bc->pB == NEUMANN;
      bu = bc->uBD; bv = bc->vBD; bw = bc->wBD; break;
    default: wall = d->ze - bc->dsT; isTrue = bc->pT == NEUMANN;
      bu = bc->uTD; bv = bc->vTD; bw = bc->wTD; break;
  }
  real h = fabs(pos - wall) - ai;
  real sgn = hi ? -1. : 1.;   /* sign pattern of the high walls */
  if(h < hN && h > 0) {
    int q = st_find(P, key);
    if(q < MAX_NEIGHBORS) { P->iSt[q] = -1; P->St[q] = 0.; }
    if(h < eps*P->r) h = eps*P->r;
    real ah = ai/h - ai/hN;
    real lnah = log(hN/h);
    real omx = P->ox, omy = P->oy, omz = P->oz;
    real Fnx = 0., Fny = 0., Fnz = 0., Ftx = 0., Fty = 0., Ftz = 0.;
    real Lox = 0., Loy = 0., Loz = 0.;
    if(axis == 0) {
      real Un = P->u - bu, Uty = P->v - bv, Utz = P->w - bw;
      Fnx = -6.*PI*mu*ai*Un*ah;
      Fty = -6.*PI*mu*ai*Uty*8./15.*lnah;
      Ftz = -6.*PI*mu*ai*Utz*8./15.*lnah;
      Ftx += 0.;
      Fty += sgn*8.*PI*mu*ai*ai*omz*1./10.*lnah;
      Ftz += -sgn*8.*PI*mu*ai*ai*omy*1./10.*lnah;
      Loy = -sgn*8.*PI*mu*ai*ai*Utz*1./10.*lnah;
      Loz = sgn*8.*PI*mu*ai*ai*Uty*1./10.*lnah;
      Lox += 0.;
      Loy += -8.*PI*mu*ai*ai*ai*omy*2./5.*lnah;
      Loz += -8.*PI*mu*ai*ai*ai*omz*2./5.*lnah;
    } else if(axis == 1) {
      real Un = P->v - bv, Utx = P->u - bu, Utz = P->w - bw;
      Fny = -6.*PI*mu*ai*Un*ah;
      Ftx = -6.*PI*mu*ai*Utx*8./15.*lnah;
      Ftz = -6.*PI*mu*ai*Utz*8./15.*lnah;
      Ftx += -sgn*8.*PI*mu*ai*ai*omz*1./10.*lnah;
      Fty += 0.;
      Ftz += sgn*8.*PI*mu*ai*ai*omx*1./10.*lnah;
      Lox = sgn*8.*PI*mu*ai*ai*Utz*1./10.*lnah;
      Loz = -sgn*8.*PI*mu*ai*ai*Utx*1./10.*lnah;
      Lox += -8.*PI*mu*ai*ai*ai*omx*2./5.*lnah;
      Loy += 0.;
      Loz += -8.*PI*mu*ai*ai*ai*omz*2./5.*lnah;
    } else {
      real Un = P->w - bw, Utx = P->u - bu, Uty = P->v - bv;
      Fnz = -6.*PI*mu*ai*Un*ah;
      Ftx = -6.*PI*mu*ai*Utx*8./15.*lnah;
      Fty = -6.*PI*mu*ai*Uty*8./15.*lnah;
      Ftx += sgn*8.*PI*mu*ai*ai*omy*1./10.*lnah;
      Fty += -sgn*8.*PI*mu*ai*ai*omx*1./10.*lnah;
      Ftz += 0.;
      Lox = -sgn*8.*PI*mu*ai*ai*Uty*1./10.*lnah;
      Loy = sgn*8.*PI*mu*ai*ai*Utx*1./10.*lnah;
      Lox += -8.*PI*mu*ai*ai*ai*omx*2./5.*lnah;
      Loy += -8.*PI*mu*ai*ai*ai*omy*2./5.*lnah;
      Loz += 0.;
    }
    P->iFx += isTrue * (Fnx + Ftx);
    P->iFy += isTrue * (Fny + Fty);
    P->iFz += isTrue * (Fnz + Ftz);
    P->iLx += isTrue * Lox;
    P->iLy += isTrue * Loy;
    P->iLz += isTrue * Loz;
  }
  if(h < 0) {
    /* the tangential wall velocity of every contact branch is bc.*SD
     * (south-wall values), as written in the reference */
    real Un;
    real Ut1, Ut2;    /* the two tangential components in (x,y,z) order */
    real comp_n = axis == 0 ? P->u - bu : axis == 1 ? P->v - bv : P->w - bw;
    Un = hi ? -comp_n : comp_n;
    if(axis == 0) {
      Ut1 = 0.5*(P->v+P->v0) - bc->vSD; Ut2 = 0.5*(P->w+P->w0) - bc->wSD;
    } else if(axis == 1) {
      Ut1 = 0.5*(P->u+P->u0) - bc->uSD; Ut2 = 0.5*(P->w+P->w0) - bc->wSD;
    } else {
      Ut1 = 0.5*(P->u+P->u0) - bc->uSD; Ut2 = 0.5*(P->v+P->v0) - bc->vSD;
    }
    int q = st_find(P, key);
    if(q == MAX_NEIGHBORS) {
      q = st_free(P);
      P->iSt[q] = key;
      P->St[q] = 1./9.*P->rho/rhof*2.*P->r*fabs(Un)/nu;
    }
    real omx = 0.5*(P->ox+P->ox0);
    real omy = 0.5*(P->oy+P->oy0);
    real omz = 0.5*(P->oz+P->oz0);
    real V1, V2;
    if(axis == 0) {
      V1 = Ut1 - (ai + 0.5*h)*omz; V2 = Ut2 + (ai + 0.5*h)*omx;
    } else if(axis == 1) {
      V1 = Ut1 + (ai + 0.5*h)*omz; V2 = Ut2 - (ai + 0.5*h)*omx;
    } else {
      V1 = Ut1 - (ai + 0.5*h)*omy; V2 = Ut2 + (ai + 0.5*h)*omx;
    }
    if(hi) { V1 = -V1; V2 = -V2; }
    real Hi = 0.5*P->E/(1.+P->sigma);
    real kt = 8./((1.-P->sigma*P->sigma)/Hi+(1.-P->sigma*P->sigma)/Hi)
      /sqrt(1./ai)*sqrt(-h);
    real s1 = V1 * dt, s2 = V2 * dt;
    real k = 4./3./((1.-P->sigma*P->sigma)/P->E
      + (1.-P->sigma*P->sigma)/P->E)/sqrt(1./ai);
    real xcx0 = 1.e-4;
    real e = P->e_dry + (1.+P->e_dry)/P->St[q]*log(xcx0);
    if(e < 0) e = 0;
    real alpha = -2.263*pow(e,0.3948)+2.22;
    real eta = alpha*sqrt(4./3.*PI*ai*ai*ai*P->rho*k*sqrt(-h));
    real coeff_fric = P->coeff_fric;
    real Ft1 = -kt * s1, Ft2 = -kt * s2;
    real Ft = sqrt(Ft1*Ft1 + Ft2*Ft2);
    if(Ft > fabs(coeff_fric * (sqrt(-h*h*h)*k - eta*Un))) {
      Ft1 = coeff_fric * (sqrt(-h*h*h)*k - eta*Un) * Ft1 / Ft;
      Ft2 = coeff_fric * (sqrt(-h*h*h)*k - eta*Un) * Ft2 / Ft;
    }
    real Fn = (sqrt(-h*h*h)*k - eta*Un);
    real arm = (ai+0.5*h);
    if(axis == 0) {
      if(!hi) P->iFx += isTrue * Fn; else P->iFx -= isTrue * Fn;
      P->iFy += isTrue * Ft1;
      P->iFz += isTrue * Ft2;
      if(!hi) { P->iLy += isTrue * arm * Ft2; P->iLz -= isTrue * arm * Ft1; }
      else    { P->iLy -= isTrue * arm * Ft2; P->iLz += isTrue * arm * 0.; }
    } else if(axis == 1) {
      P->iFx += isTrue * Ft1;
      if(!hi) P->iFy += isTrue * Fn; else P->iFy -= isTrue * Fn;
      P->iFz += isTrue * Ft2;
      if(!hi) { P->iLx -= isTrue * arm * Ft2; P->iLz += isTrue * arm * Ft1; }
      else    { P->iLx += isTrue * arm * Ft2; P->iLz -= isTrue * arm * Ft1; }
    } else {
      P->iFx += isTrue * Ft1;
      P->iFy += isTrue * Ft2;
      if(!hi) P->iFz += isTrue * Fn; else P->iFz -= isTrue * Fn;
      if(!hi) { P->iLx += isTrue * arm * Ft2; P->iLy -= isTrue * arm * Ft1; }
      else    { P->iLx -= isTrue * arm * Ft2; P->iLy += isTrue * arm * Ft1; }
    }
  }
}

static void collision_walls(orc_state *s, real eps)
{
  for(int i = 0; i < s->nparts; i++)
    for(int face = 0; face < 6; face++) wall_force(s, i, face / 2, face % 2, eps);
}

static void move_common(orc_state *s, int final)
{
  if(s->nparts <= 0) return;
  real eps = 0.01;
  if(s->nparts == 1) {
    collision_init(s); spring_parts(s); collision_walls(s, eps);
    move_a(s);
    collision_init(s); spring_parts(s); collision_walls(s, eps);
  } else {
    collision_init(s);
    bins_t bn;
    make_bins(s, &bn);
    collision_parts(s, &bn, eps); spring_parts(s); collision_walls(s, eps);
    move_a(s);
    collision_init(s);
    collision_parts(s, &bn, eps); spring_parts(s); collision_walls(s, eps);
    free_bins(&bn);
  }
  if(final) move_b(s);
}

void orc_move_parts_sub(orc_state *s) { move_common(s, 0); }
void orc_move_parts(orc_state *s) { move_common(s, 1); }
