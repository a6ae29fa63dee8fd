/*
 * oracle_quad.c -- serial restatement of the Lamb-coefficient quadrature
 * (src/cuda_quadrature.cu:30-398, src/quadrature_kernel.cu:25-654).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * compute_error's cross-particle race (src/quadrature_kernel.cu:632-649:
 * particle p writes coeffs/errors[p*stride + i] for i < 6*stride, which
 * overlaps particle p+1..p+5's slots) is not reproduced: each particle
 * sorts into its own 6*stride scratch, which is the kernel's evident
 * intent.  The racy result depends on block scheduling and cannot be
 * pinned.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "oracle_internal.h"

/* src/cuda_quadrature.cu:86-157 */
void orc_lebedev_nodes(real theta[26], real phi[26], real weight[26])
{
  real PI14 = 0.25 * PI, PI12 = 0.5 * PI, PI34 = 0.75 * PI;
  real PI54 = 1.25 * PI, PI32 = 1.5 * PI, PI74 = 1.75 * PI;
  real alph1 = 0.955316618124509, alph2 = 2.186276035465284;
  real A1 = 0.598398600683775, A2 = 0.478718880547015, A3 = 0.403919055461543;
  real a1_t[6] = {PI12, PI12, PI12, PI12, 0.+DIV_ST, PI-DIV_ST};
  real a1_p[6] = {0., PI12, PI, PI32, 0., 0.};
  real a2_t[12] = {PI12, PI12, PI12, PI12, PI14, PI14, PI14, PI14,
                   PI34, PI34, PI34, PI34};
  real a2_p[12] = {PI14, PI34, PI54, PI74, 0., PI12, PI, PI32,
                   0., PI12, PI, PI32};
  real a3_t[8] = {alph1, alph1, alph1, alph1, alph2, alph2, alph2, alph2};
  real a3_p[8] = {PI14, PI34, PI54, PI74, PI14, PI34, PI54, PI74};
  for(int i = 0; i < 6; i++) { theta[i] = a1_t[i]; phi[i] = a1_p[i]; weight[i] = A1; }
  for(int i = 0; i < 12; i++) { theta[6+i] = a2_t[i]; phi[6+i] = a2_p[i]; weight[6+i] = A2; }
  for(int i = 0; i < 8; i++) { theta[18+i] = a3_t[i]; phi[18+i] = a3_p[i]; weight[18+i] = A3; }
}

static void rtp2xyz(real r, real theta, real phi, real *x, real *y, real *z)
{
  *x = r * sin(theta) * cos(phi);
  *y = r * sin(theta) * sin(phi);
  *z = r * cos(theta);
}

static void cart2sphere(real u, real v, real w, real theta, real phi,
  real *ur, real *ut, real *up)
{
  real st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
  *ur = st * (u * cp + v * sp) + w * ct;
  *ut = ct * (u * cp + v * sp) - w * st;
  *up = -u * sp + v * cp;
}

/* src/quadrature_kernel.cu:45-93 */
static void check_nodes(orc_state *s, const real *theta, const real *phi)
{
  const dom_struct *d = &s->dom;
  const BC bc = s->bc;
  for(int part = 0; part < s->nparts; part++) for(int node = 0; node < NNODES; node++) {
    part_struct *P = &s->parts[part];
    real xp, yp, zp;
    rtp2xyz(P->rs, theta[node], phi[node], &xp, &yp, &zp);
    real x = xp + P->x, y = yp + P->y, z = zp + P->z;
    P->nodes[node] = -1;
    if(x - d->xs < 0) {
      if(bc.uW == DIRICHLET || bc.vW == DIRICHLET || bc.wW == DIRICHLET)
        if(P->nodes[node] == -1) P->nodes[node] = -10;
    } if(x - d->xe > 0) {
      if(bc.uE == DIRICHLET || bc.vE == DIRICHLET || bc.wE == DIRICHLET)
        if(P->nodes[node] == -1) P->nodes[node] = -11;
    } if(y - d->ys < 0) {
      if(bc.uS == DIRICHLET || bc.vS == DIRICHLET || bc.wS == DIRICHLET)
        if(P->nodes[node] == -1) P->nodes[node] = -12;
    } if(y - d->ye > 0) {
      if(bc.uN == DIRICHLET || bc.vN == DIRICHLET || bc.wN == DIRICHLET)
        if(P->nodes[node] == -1) P->nodes[node] = -13;
    } if(z - d->zs < 0) {
      if(bc.uB == DIRICHLET || bc.vB == DIRICHLET || bc.wB == DIRICHLET)
        if(P->nodes[node] == -1) P->nodes[node] = -14;
    } if(z - d->ze > 0) {
      if(bc.uT == DIRICHLET || bc.vT == DIRICHLET || bc.wT == DIRICHLET)
        if(P->nodes[node] == -1) P->nodes[node] = -15;
    }
  }
}

static int clampi(int v, int lo, int hi) { if(v < lo) v = lo; if(v > hi) v = hi; return v; }

/* src/quadrature_kernel.cu:95-303 */
static void interpolate_nodes(orc_state *s, const real *theta, const real *phi,
  real *pp, real *ur, real *ut, real *up)
{
  const dom_struct *d = &s->dom;
  const BC bc = s->bc;
  const grid_info Gc = d->Gcc, Gx = d->Gfx, Gy = d->Gfy, Gz = d->Gfz;
  const real ddx = 1. / d->dx, ddy = 1. / d->dy, ddz = 1. / d->dz;
  const real nu = s->nu, rho_f = s->rho_f;
  const gradP_struct gradP = s->gradP;
  for(int part = 0; part < s->nparts; part++) for(int node = 0; node < NNODES; node++) {
    const part_struct *P = &s->parts[part];
    real ox = P->ox, oy = P->oy, oz = P->oz;
    real oxdot = P->oxdot, oydot = P->oydot, ozdot = P->ozdot;
    real udot = P->udot, vdot = P->vdot, wdot = P->wdot;
    real xp, yp, zp;
    rtp2xyz(P->rs, theta[node], phi[node], &xp, &yp, &zp);
    real x = xp + P->x, y = yp + P->y, z = zp + P->z;
    if(x < d->xs && bc.uW == PERIODIC) x = x + d->xl;
    else if(x > d->xe && bc.uE == PERIODIC) x = x - d->xl;
    if(y < d->ys && bc.vS == PERIODIC) y = y + d->yl;
    else if(y > d->ye && bc.vN == PERIODIC) y = y - d->yl;
    if(z < d->zs && bc.wB == PERIODIC) z = z + d->zl;
    else if(z > d->ze && bc.wT == PERIODIC) z = z - d->zl;
    int nd = P->nodes[node];
    int o = node + NNODES*part;

    int i = clampi((int) floor((x - d->xs) * ddx) + DOM_BUF, Gc.is, Gc.ie-1);
    int j = clampi((int) floor((y - d->ys) * ddy) + DOM_BUF, Gc.js, Gc.je-1);
    int k = clampi((int) floor((z - d->zs) * ddz) + DOM_BUF, Gc.ks, Gc.ke-1);
    int C = i + j*Gc.s1b + k*Gc.s2b;
    real xx = (i-0.5) * d->dx + d->xs;
    real yy = (j-0.5) * d->dy + d->ys;
    real zz = (k-0.5) * d->dz + d->zs;
    const real *p = s->p;
    real dpdx = 0.5*(p[C+1] - p[C-1]) * ddx;
    real dpdy = 0.5*(p[C+Gc.s1b] - p[C-Gc.s1b]) * ddy;
    real dpdz = 0.5*(p[C+Gc.s2b] - p[C-Gc.s2b]) * ddz;
    pp[o] = p[C] + dpdx*(x-xx) + dpdy*(y-yy) + dpdz*(z-zz);
    real ocrossr2 = (oy*zp - oz*yp) * (oy*zp - oz*yp);
    ocrossr2 += (ox*zp - oz*xp) * (ox*zp - oz*xp);
    ocrossr2 += (ox*yp - oy*xp) * (ox*yp - oy*xp);
    real rhoV = rho_f;
    real accdotr = (-gradP.x/rhoV - udot)*xp + (-gradP.y/rhoV - vdot)*yp
      + (-gradP.z/rhoV - wdot)*zp;
    pp[o] -= 0.5 * rho_f * ocrossr2 + rho_f * accdotr;
    pp[o] = (nd == -1)*pp[o];

    real rs3 = P->rs*P->rs*P->rs;
    real rs5 = rs3*P->rs*P->rs;
    real a5 = P->r*P->r*P->r*P->r*P->r;

    /* u */
    i = clampi((int) round((x - d->xs) * ddx) + DOM_BUF, Gx.is, Gx.ie-1);
    j = clampi((int) floor((y - d->ys) * ddy) + DOM_BUF, Gx.js, Gx.je-1);
    k = clampi((int) floor((z - d->zs) * ddz) + DOM_BUF, Gx.ks, Gx.ke-1);
    xx = (i-DOM_BUF) * d->dx + d->xs;
    yy = (j-0.5) * d->dy + d->ys;
    zz = (k-0.5) * d->dz + d->zs;
    C = i + j*Gx.s1b + k*Gx.s2b;
    const real *u = s->u;
    real dudx = 0.5*(u[C+1] - u[C-1]) * ddx;
    real dudy = 0.5*(u[C+Gx.s1b] - u[C-Gx.s1b]) * ddy;
    real dudz = 0.5*(u[C+Gx.s2b] - u[C-Gx.s2b]) * ddz;
    real uu = u[C] + dudx * (x - xx) + dudy * (y - yy) + dudz * (z - zz);
    real uuwall = (nd == -10)*bc.uWD + (nd == -11)*bc.uED + (nd == -12)*bc.uSD
      + (nd == -13)*bc.uND + (nd == -14)*bc.uBD + (nd == -15)*bc.uTD;
    real ocrossr_x = oy*zp - oz*yp;
    real odotcrossr_x = oydot*zp - ozdot*yp;
    uu -= P->u + ocrossr_x;
    uu -= 0.1/nu *(rs5-a5)/rs3 * odotcrossr_x;
    uuwall -= P->u + ocrossr_x;
    uuwall -= 0.1/nu *(rs5-a5)/rs3 * odotcrossr_x;
    uu = (nd == -1)*uu + (nd < -1)*uuwall;

    /* v */
    i = clampi((int) floor((x - d->xs) * ddx) + DOM_BUF, Gy.is, Gy.ie-1);
    j = clampi((int) round((y - d->ys) * ddy) + DOM_BUF, Gy.js, Gy.je-1);
    k = clampi((int) floor((z - d->zs) * ddz) + DOM_BUF, Gy.ks, Gy.ke-1);
    xx = (i-0.5) * d->dx + d->xs;
    yy = (j-DOM_BUF) * d->dy + d->ys;
    zz = (k-0.5) * d->dz + d->zs;
    C = i + j*Gy.s1b + k*Gy.s2b;
    const real *v = s->v;
    real dvdx = 0.5*(v[C+1] - v[C-1]) * ddx;
    real dvdy = 0.5*(v[C+Gy.s1b] - v[C-Gy.s1b]) * ddy;
    real dvdz = 0.5*(v[C+Gy.s2b] - v[C-Gy.s2b]) * ddz;
    real vv = v[C] + dvdx * (x - xx) + dvdy * (y - yy) + dvdz * (z - zz);
    real vvwall = (nd == -10)*bc.vWD + (nd == -11)*bc.vED + (nd == -12)*bc.vSD
      + (nd == -13)*bc.vND + (nd == -14)*bc.vBD + (nd == -15)*bc.vTD;
    real ocrossr_y = -(ox*zp - oz*xp);
    real odotcrossr_y = -(oxdot*zp - ozdot*xp);
    vv -= P->v + ocrossr_y;
    vv -= 0.1/nu *(rs5-a5)/rs3 * odotcrossr_y;
    vvwall -= P->v + ocrossr_y;
    vvwall -= 0.1/nu *(rs5-a5)/rs3 * odotcrossr_y;
    vv = (nd == -1)*vv + (nd < -1)*vvwall;

    /* w */
    i = clampi((int) floor((x - d->xs) * ddx) + DOM_BUF, Gz.is, Gz.ie-1);
    j = clampi((int) floor((y - d->ys) * ddy) + DOM_BUF, Gz.js, Gz.je-1);
    k = clampi((int) round((z - d->zs) * ddz) + DOM_BUF, Gz.ks, Gz.ke-1);
    xx = (i-0.5) * d->dx + d->xs;
    yy = (j-0.5) * d->dy + d->ys;
    zz = (k-DOM_BUF) * d->dz + d->zs;
    C = i + j*Gz.s1b + k*Gz.s2b;
    const real *w = s->w;
    real dwdx = 0.5*(w[C+1] - w[C-1]) * ddx;
    real dwdy = 0.5*(w[C+Gz.s1b] - w[C-Gz.s1b]) * ddy;
    real dwdz = 0.5*(w[C+Gz.s2b] - w[C-Gz.s2b]) * ddz;
    real ww = w[C] + dwdx * (x - xx) + dwdy * (y - yy) + dwdz * (z - zz);
    real wwwall = (nd == -10)*bc.wWD + (nd == -11)*bc.wED + (nd == -12)*bc.wSD
      + (nd == -13)*bc.wND + (nd == -14)*bc.wBD + (nd == -15)*bc.wTD;
    real ocrossr_z = ox*yp - oy*xp;
    real odotcrossr_z = oxdot*yp - oydot*xp;
    ww -= P->w + ocrossr_z;
    ww -= 0.1/nu *(rs5-a5)/rs3 * odotcrossr_z;
    wwwall -= P->w + ocrossr_z;
    wwwall -= 0.1/nu *(rs5-a5)/rs3 * odotcrossr_z;
    ww = (nd == -1)*ww + (nd < -1)*wwwall;

    cart2sphere(uu, vv, ww, theta[node], phi[node], &ur[o], &ut[o], &up[o]);
  }
}

/* src/quadrature_kernel.cu:377-543 */
static void get_coeffs(orc_state *s, const real *theta, const real *phi,
  const real *pp, const real *ur, const real *ut, const real *up)
{
  static const int nn[15] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4};
  static const int mm[15] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3, 0, 1, 2, 3, 4};
  real A1 = 0.598398600683775, A2 = 0.478718880547015, A3 = 0.403919055461543;
  const int stride = s->coeff_stride;
  const real mu = s->mu, nu = s->nu, relax = s->lamb_relax;
  real *pnm_re = s->coeff[0], *pnm_im = s->coeff[1];
  real *phinm_re = s->coeff[2], *phinm_im = s->coeff[3];
  real *chinm_re = s->coeff[4], *chinm_im = s->coeff[5];
  const real *pnm_re0 = s->coeff[6], *pnm_im0 = s->coeff[7];
  const real *phinm_re0 = s->coeff[8], *phinm_im0 = s->coeff[9];
  const real *chinm_re0 = s->coeff[10], *chinm_im0 = s->coeff[11];
  for(int part = 0; part < s->nparts; part++) {
    const part_struct *P = &s->parts[part];
    real ars = P->r / P->rs;
    real rsa = P->rs / P->r;
    for(int coeff = 0; coeff < P->ncoeff && coeff < 15; coeff++) {
      int n = nn[coeff], m = mm[coeff];
      real iYr[NNODES], iYi[NNODES], iRr[NNODES], iRi[NNODES];
      real iXr[NNODES], iXi[NNODES];
      for(int node = 0; node < NNODES; node++) {
        int o = node + part*NNODES;
        real th = theta[node], ph = phi[node];
        real N_nm = orc_Nnm(n, m);
        real P_nm = orc_Pnm(n, m, th);
        real P_n1m = orc_Pnm(n+1, m, th);
        real dPdt = (n-m+1.)*P_n1m-(n+1.)*cos(th)*P_nm;
        real dPdp = m*P_nm;
        iYr[node] = N_nm*P_nm*pp[o]*cos(m*ph);
        iYi[node] = -N_nm*P_nm*pp[o]*sin(m*ph);
        iRr[node] = N_nm/sin(th)*(dPdt*ut[o]*cos(m*ph) - dPdp*up[o]*sin(m*ph));
        iRi[node] = N_nm/sin(th)*(-dPdt*ut[o]*sin(m*ph) - dPdp*up[o]*cos(m*ph));
        iXr[node] = N_nm/sin(th)*(dPdp*ut[o]*sin(m*ph) + dPdt*up[o]*cos(m*ph));
        iXi[node] = N_nm/sin(th)*(dPdp*ut[o]*cos(m*ph) - dPdt*up[o]*sin(m*ph));
      }
      real *arr[6] = {iYr, iYi, iRr, iRi, iXr, iXi};
      real sum[6];
      for(int a = 0; a < 6; a++) {
        real t = arr[a][0] * A1;
        for(int i = 1; i < 6; i++) t += A1 * arr[a][i];
        for(int i = 6; i < 18; i++) t += A2 * arr[a][i];
        for(int i = 18; i < 26; i++) t += A3 * arr[a][i];
        sum[a] = t;
      }
      int c = stride*part + coeff;
      real r = P->r;
      if(n == 0) {
        pnm_re[c] = pnm_re0[c] + relax*(r*r/mu/nu*sum[0]*pow(ars,n) - pnm_re0[c]);
        pnm_im[c] = pnm_im0[c] + relax*(r*r/mu/nu*sum[1]*pow(ars,n) - pnm_im0[c]);
        phinm_re[c] = 0.; phinm_im[c] = 0.;
        chinm_re[c] = 0.; chinm_im[c] = 0.;
      } else {
        real A = (1.-0.5*n*(2.*n-1.)/(n+1.)*pow(ars,2.*n+1.))*pow(rsa,n);
        real B = n*(2.*n-1.)*(2.*n+1.)/(n+1.)*pow(ars,n+1.);
        real C = 0.25*n*(2.*(n+3.)/(2.*n+3.)
          + (n-2.-n*(2.*n+1.)/(2.*n+3.)*ars*ars)*pow(ars,2.*n+1.))*pow(rsa,n+1.);
        real D = n*(n+1.+0.5*((n-2.)*(2.*n+1.)*rsa*rsa
          - n*(2.*n-1.))*pow(ars,2.*n+1.))*pow(rsa,n-1.);
        pnm_re[c] = (r*r/mu/nu*sum[0]*D + r/nu*sum[2]*B) / (A*D+B*C);
        pnm_im[c] = (r*r/mu/nu*sum[1]*D + r/nu*sum[3]*B) / (A*D+B*C);
        phinm_re[c] = (r/nu*sum[2]*A - r*r/mu/nu*sum[0]*C) / (A*D+B*C);
        phinm_im[c] = (r/nu*sum[3]*A - r*r/mu/nu*sum[1]*C) / (A*D+B*C);
        real E = n*(n+1.)*(pow(ars,2.*n+1.)-1.)*pow(rsa, n);
        chinm_re[c] = r/nu*sum[4] / E;
        chinm_im[c] = r/nu*sum[5] / E;
        pnm_re[c] = pnm_re0[c]*(1.-relax) + relax*pnm_re[c];
        pnm_im[c] = pnm_im0[c]*(1.-relax) + relax*pnm_im[c];
        phinm_re[c] = phinm_re0[c]*(1.-relax) + relax*phinm_re[c];
        phinm_im[c] = phinm_im0[c]*(1.-relax) + relax*phinm_im[c];
        chinm_re[c] = chinm_re0[c]*(1.-relax) + relax*chinm_re[c];
        chinm_im[c] = chinm_im0[c]*(1.-relax) + relax*chinm_im[c];
      }
    }
  }
}

/* src/quadrature_kernel.cu:545-576 */
static void calc_forces(orc_state *s)
{
  const int stride = s->coeff_stride;
  const real rho_f = s->rho_f, mu = s->mu, nu = s->nu;
  const gradP_struct gradP = s->gradP;
  const real *pnm_re = s->coeff[0], *pnm_im = s->coeff[1];
  const real *phinm_re = s->coeff[2], *phinm_im = s->coeff[3];
  const real *chinm_re = s->coeff[4], *chinm_im = s->coeff[5];
  for(int pp = 0; pp < s->nparts; pp++) {
    part_struct *P = &s->parts[pp];
    real vol = 4./3. * PI * P->r*P->r*P->r;
    real N10 = sqrt(3./4./PI);
    real N11 = sqrt(3./8./PI);
    P->Fx = rho_f * vol * (P->udot + gradP.x)
      - PI * mu * nu * 2.*N11 * (pnm_re[stride*pp + 2] + 6.*phinm_re[stride*pp + 2]);
    P->Fy = rho_f * vol * (P->vdot + gradP.y)
      + PI * mu * nu * 2.*N11 * (pnm_im[stride*pp + 2] + 6.*phinm_im[stride*pp + 2]);
    P->Fz = rho_f * vol * (P->wdot + gradP.z)
      + PI * mu * nu * N10 * (pnm_re[stride*pp + 1] + 6.*phinm_re[stride*pp + 1]);
    P->Lx = rho_f * vol * P->r*P->r * P->oxdot
      - 8. * PI * mu * nu * 2.*N11 * P->r * chinm_re[stride*pp + 2];
    P->Ly = rho_f * vol * P->r*P->r * P->oydot
      + 8. * PI * mu * nu * 2.*N11 * P->r * chinm_im[stride*pp + 2];
    P->Lz = rho_f * vol * P->r*P->r * P->ozdot
      + 8. * PI * mu * nu * N10 * P->r * chinm_re[stride*pp + 1];
  }
}

/* cuda_quad_check_nodes (src/cuda_quadrature.cu:30-62,
 * src/quadrature_kernel.cu:25-93): node flags of every particle. */
void orc_quad_check_nodes(orc_state *s, const real *theta, const real *phi)
{
  check_nodes(s, theta, phi);
}

/* cuda_quad_interp (src/cuda_quadrature.cu:30-62,
 * src/quadrature_kernel.cu:95-303): pressure and spherical velocity
 * components at every node, nparts x 26, node fastest. */
void orc_quad_interp(orc_state *s, const real *theta, const real *phi,
  real *pp, real *ur, real *ut, real *up)
{
  interpolate_nodes(s, theta, phi, pp, ur, ut, up);
}

void orc_Lamb(orc_state *s)
{
  if(s->nparts <= 0) return;
  real theta[NNODES], phi[NNODES], wgt[NNODES];
  orc_lebedev_nodes(theta, phi, wgt);
  size_t n = (size_t) s->nparts * NNODES;
  real *pp = calloc(n, sizeof(real)), *ur = calloc(n, sizeof(real));
  real *ut = calloc(n, sizeof(real)), *up = calloc(n, sizeof(real));
  check_nodes(s, theta, phi);
  interpolate_nodes(s, theta, phi, pp, ur, ut, up);
  get_coeffs(s, theta, phi, pp, ur, ut, up);
  calc_forces(s);
  free(pp); free(ur); free(ut); free(up);
}

/* src/quadrature_kernel.cu:578-654 per particle, then the host max and
 * the coeff -> coeff0 copies of src/cuda_quadrature.cu:316-398 */
real orc_lamb_err(orc_state *s)
{
  if(s->nparts <= 0) return 0.;
  const int stride = s->coeff_stride, n6 = 6 * stride;
  real *sc = malloc(n6 * sizeof(real)), *sc0 = malloc(n6 * sizeof(real));
  real *sorted = malloc(n6 * sizeof(real)), *errs = malloc(n6 * sizeof(real));
  real mx = FLT_MIN;
  for(int part = 0; part < s->nparts; part++) {
    for(int c = 0; c < 6; c++) for(int i = 0; i < stride; i++) {
      sc[i + c*stride] = s->coeff[c][part*stride + i];
      sc0[i + c*stride] = s->coeff[6 + c][part*stride + i];
    }
    int loc = 0;
    for(int i = 0; i < n6; i++) {
      real tmp = FLT_MIN;
      for(int j = 0; j < n6; j++) {
        if(sc[j]*sc[j] > tmp) { tmp = sc[j]*sc[j]; loc = j; }
      }
      sorted[i] = sc[loc];
      if(fabs(sc[loc]) > s->lamb_cut*fabs(sorted[0])) {
        real div = fabs(sc[loc]);
        if(div < 1e-16) div = 1e-16;
        errs[i] = fabs((sc[loc] - sc0[loc]) / div);
      } else errs[i] = 0.;
      sc[loc] = 0.;
    }
    real tmp = FLT_MIN;
    for(int i = 0; i < n6; i++) if(errs[i] > tmp) tmp = errs[i];
    if(tmp > mx) mx = tmp;
  }
  free(sc); free(sc0); free(sorted); free(errs);
  size_t bytes = (size_t) s->nparts * stride * sizeof(real);
  for(int c = 0; c < 6; c++) memcpy(s->coeff[6 + c], s->coeff[c], bytes);
  return mx;
}
