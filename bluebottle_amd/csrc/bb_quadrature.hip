/*
 * bb_quadrature.hip -- Lamb-coefficient quadrature:
 *   cuda_quad_check_nodes (src/cuda_quadrature.cu:30-43,
 *                          src/quadrature_kernel.cu:45-93)
 *   cuda_quad_interp      (src/cuda_quadrature.cu:46-61,
 *                          src/quadrature_kernel.cu:95-303)
 *   cuda_Lamb             (src/cuda_quadrature.cu:64-313,
 *                          src/quadrature_kernel.cu:377-576)
 *   cuda_lamb_err         (src/cuda_quadrature.cu:316-398,
 *                          src/quadrature_kernel.cu:578-654)
 *
 * The reference allocates and frees 14 device buffers per call and runs
 * four launches with 26-thread blocks, summing each inner product serially
 * in thread 0 of an (nparts x ncoeff) grid.  cuda_Lamb here is one launch:
 * one 64-lane wavefront per particle computes the 26 node values (wall
 * check + gradient interpolation + rest-frame spherical components) into
 * LDS, then one lane per Lamb coefficient forms the three complex inner
 * products with the 26-node Lebedev weights in the reference's summation
 * order and applies the closed-form solve with under-relaxation, then lane
 * 0 forms the force and torque.  compute_error's cross-particle write race
 * is removed (each particle sorts into its own scratch).
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "bb_device.h"
#include "bb_lamb.h"

struct Nodes { double t[NNODES], p[NNODES], w[NNODES]; };

/* src/cuda_quadrature.cu:86-157 */
static Nodes lebedev_nodes(void)
{
  Nodes N;
  double PI14 = 0.25 * PI, PI12 = 0.5 * PI, PI34 = 0.75 * PI;
  double PI54 = 1.25 * PI, PI32 = 1.5 * PI, PI74 = 1.75 * PI;
  double alph1 = 0.955316618124509, alph2 = 2.186276035465284;
  double A1 = 0.598398600683775, A2 = 0.478718880547015, A3 = 0.403919055461543;
  double a1_t[6] = {PI12, PI12, PI12, PI12, 0.+DIV_ST, PI-DIV_ST};
  double a1_p[6] = {0., PI12, PI, PI32, 0., 0.};
  double a2_t[12] = {PI12, PI12, PI12, PI12, PI14, PI14, PI14, PI14,
                     PI34, PI34, PI34, PI34};
  double a2_p[12] = {PI14, PI34, PI54, PI74, 0., PI12, PI, PI32,
                     0., PI12, PI, PI32};
  double a3_t[8] = {alph1, alph1, alph1, alph1, alph2, alph2, alph2, alph2};
  double a3_p[8] = {PI14, PI34, PI54, PI74, PI14, PI34, PI54, PI74};
  for(int i = 0; i < 6; i++) { N.t[i] = a1_t[i]; N.p[i] = a1_p[i]; N.w[i] = A1; }
  for(int i = 0; i < 12; i++) { N.t[6+i] = a2_t[i]; N.p[6+i] = a2_p[i]; N.w[6+i] = A2; }
  for(int i = 0; i < 8; i++) { N.t[18+i] = a3_t[i]; N.p[18+i] = a3_p[i]; N.w[18+i] = A3; }
  return N;
}

/* wall-interference code of one node, src/quadrature_kernel.cu:60-93 */
__device__ inline int node_flag(const BC &bc, const BBGeom &G, double x,
  double y, double z)
{
  int nd = -1;
  if(x - G.xs < 0) {
    if(bc.uW == DIRICHLET || bc.vW == DIRICHLET || bc.wW == DIRICHLET)
      if(nd == -1) nd = -10;
  } if(x - G.xe > 0) {
    if(bc.uE == DIRICHLET || bc.vE == DIRICHLET || bc.wE == DIRICHLET)
      if(nd == -1) nd = -11;
  } if(y - G.ys < 0) {
    if(bc.uS == DIRICHLET || bc.vS == DIRICHLET || bc.wS == DIRICHLET)
      if(nd == -1) nd = -12;
  } if(y - G.ye > 0) {
    if(bc.uN == DIRICHLET || bc.vN == DIRICHLET || bc.wN == DIRICHLET)
      if(nd == -1) nd = -13;
  } if(z - G.zs < 0) {
    if(bc.uB == DIRICHLET || bc.vB == DIRICHLET || bc.wB == DIRICHLET)
      if(nd == -1) nd = -14;
  } if(z - G.ze > 0) {
    if(bc.uT == DIRICHLET || bc.vT == DIRICHLET || bc.wT == DIRICHLET)
      if(nd == -1) nd = -15;
  }
  return nd;
}

struct InterpArgs {
  BBGeom G;
  BC bc;
  gradP_struct gP;
  double rho_f, nu;
};

/* cell index from a (floored or rounded) cell coordinate, clamped to
 * [lo, hi] before the conversion, so that no position -- however far
 * outside the domain, or not a number -- converts out of range */
__device__ inline int cell_clamp(double q, int lo, int hi)
{
  return (int) fmin(fmax(q, (double) lo), (double) hi);
}

/* node values of one (particle, node): src/quadrature_kernel.cu:95-303 */
__device__ void interp_node(const InterpArgs &A, const part_struct *P,
  double theta, double phi, int nd, const double *__restrict__ p,
  const double *__restrict__ u, const double *__restrict__ v,
  const double *__restrict__ w, double *pp, double *ur, double *ut, double *up)
{
  const BBGeom &G = A.G;
  const BC &bc = A.bc;
  double ox = P->ox, oy = P->oy, oz = P->oz;
  double oxdot = P->oxdot, oydot = P->oydot, ozdot = P->ozdot;
  double udot = P->udot, vdot = P->vdot, wdot = P->wdot;
  double xp = P->rs * sin(theta) * cos(phi);
  double yp = P->rs * sin(theta) * sin(phi);
  double zp = P->rs * cos(theta);
  double x = xp + P->x, y = yp + P->y, z = zp + P->z;
  if(x < G.xs && bc.uW == PERIODIC) x = x + G.xl;
  else if(x > G.xe && bc.uE == PERIODIC) x = x - G.xl;
  if(y < G.ys && bc.vS == PERIODIC) y = y + G.yl;
  else if(y > G.ye && bc.vN == PERIODIC) y = y - G.yl;
  if(z < G.zs && bc.wB == PERIODIC) z = z + G.zl;
  else if(z > G.ze && bc.wT == PERIODIC) z = z - G.zl;
  const double ddx = 1. / G.dx, ddy = 1. / G.dy, ddz = 1. / G.dz;

  int i = cell_clamp(floor((x - G.xs) * ddx) + DOM_BUF, 1, G.xn);
  int j = cell_clamp(floor((y - G.ys) * ddy) + DOM_BUF, 1, G.yn);
  int k = cell_clamp(floor((z - G.zs) * ddz) + DOM_BUF, 1, G.zn);
  int C = i + j*G.cs1 + k*G.cs2;
  double xx = (i-0.5) * G.dx + G.xs;
  double yy = (j-0.5) * G.dy + G.ys;
  double zz = (k-0.5) * G.dz + G.zs;
  double dpdx = 0.5*(p[C+1] - p[C-1]) * ddx;
  double dpdy = 0.5*(p[C+G.cs1] - p[C-G.cs1]) * ddy;
  double dpdz = 0.5*(p[C+G.cs2] - p[C-G.cs2]) * ddz;
  double pv = p[C] + dpdx*(x-xx) + dpdy*(y-yy) + dpdz*(z-zz);
  double ocr2 = (oy*zp - oz*yp) * (oy*zp - oz*yp);
  ocr2 += (ox*zp - oz*xp) * (ox*zp - oz*xp);
  ocr2 += (ox*yp - oy*xp) * (ox*yp - oy*xp);
  double rhoV = A.rho_f;
  double acc = (-A.gP.x/rhoV - udot)*xp + (-A.gP.y/rhoV - vdot)*yp
    + (-A.gP.z/rhoV - wdot)*zp;
  pv -= 0.5 * A.rho_f * ocr2 + A.rho_f * acc;
  *pp = (double) (nd == -1) * pv;

  double rs3 = P->rs*P->rs*P->rs;
  double rs5 = rs3*P->rs*P->rs;
  double a5 = P->r*P->r*P->r*P->r*P->r;
  double sfac = 0.1/A.nu *(rs5-a5)/rs3;

  /* u */
  i = cell_clamp(round((x - G.xs) * ddx) + DOM_BUF, 1, G.xn + 1);
  j = cell_clamp(floor((y - G.ys) * ddy) + DOM_BUF, 1, G.yn);
  k = cell_clamp(floor((z - G.zs) * ddz) + DOM_BUF, 1, G.zn);
  xx = (i-DOM_BUF) * G.dx + G.xs; yy = (j-0.5) * G.dy + G.ys;
  zz = (k-0.5) * G.dz + G.zs;
  C = i + j*G.us1 + k*G.us2;
  double uu = u[C] + 0.5*(u[C+1] - u[C-1]) * ddx * (x - xx)
    + 0.5*(u[C+G.us1] - u[C-G.us1]) * ddy * (y - yy)
    + 0.5*(u[C+G.us2] - u[C-G.us2]) * ddz * (z - zz);
  double uw = (nd == -10)*bc.uWD + (nd == -11)*bc.uED + (nd == -12)*bc.uSD
    + (nd == -13)*bc.uND + (nd == -14)*bc.uBD + (nd == -15)*bc.uTD;
  double ocx = oy*zp - oz*yp, odx = oydot*zp - ozdot*yp;
  uu -= P->u + ocx; uu -= sfac * odx;
  uw -= P->u + ocx; uw -= sfac * odx;
  uu = (nd == -1)*uu + (nd < -1)*uw;

  /* v */
  i = cell_clamp(floor((x - G.xs) * ddx) + DOM_BUF, 1, G.xn);
  j = cell_clamp(round((y - G.ys) * ddy) + DOM_BUF, 1, G.yn + 1);
  k = cell_clamp(floor((z - G.zs) * ddz) + DOM_BUF, 1, G.zn);
  xx = (i-0.5) * G.dx + G.xs; yy = (j-DOM_BUF) * G.dy + G.ys;
  zz = (k-0.5) * G.dz + G.zs;
  C = i + j*G.vs1 + k*G.vs2;
  double vv = v[C] + 0.5*(v[C+1] - v[C-1]) * ddx * (x - xx)
    + 0.5*(v[C+G.vs1] - v[C-G.vs1]) * ddy * (y - yy)
    + 0.5*(v[C+G.vs2] - v[C-G.vs2]) * ddz * (z - zz);
  double vw = (nd == -10)*bc.vWD + (nd == -11)*bc.vED + (nd == -12)*bc.vSD
    + (nd == -13)*bc.vND + (nd == -14)*bc.vBD + (nd == -15)*bc.vTD;
  double ocy = -(ox*zp - oz*xp), ody = -(oxdot*zp - ozdot*xp);
  vv -= P->v + ocy; vv -= sfac * ody;
  vw -= P->v + ocy; vw -= sfac * ody;
  vv = (nd == -1)*vv + (nd < -1)*vw;

  /* w */
  i = cell_clamp(floor((x - G.xs) * ddx) + DOM_BUF, 1, G.xn);
  j = cell_clamp(floor((y - G.ys) * ddy) + DOM_BUF, 1, G.yn);
  k = cell_clamp(round((z - G.zs) * ddz) + DOM_BUF, 1, G.zn + 1);
  xx = (i-0.5) * G.dx + G.xs; yy = (j-0.5) * G.dy + G.ys;
  zz = (k-DOM_BUF) * G.dz + G.zs;
  C = i + j*G.ws1 + k*G.ws2;
  double ww = w[C] + 0.5*(w[C+1] - w[C-1]) * ddx * (x - xx)
    + 0.5*(w[C+G.ws1] - w[C-G.ws1]) * ddy * (y - yy)
    + 0.5*(w[C+G.ws2] - w[C-G.ws2]) * ddz * (z - zz);
  double ww_ = (nd == -10)*bc.wWD + (nd == -11)*bc.wED + (nd == -12)*bc.wSD
    + (nd == -13)*bc.wND + (nd == -14)*bc.wBD + (nd == -15)*bc.wTD;
  double ocz = ox*yp - oy*xp, odz = oxdot*yp - oydot*xp;
  ww -= P->w + ocz; ww -= sfac * odz;
  ww_ -= P->w + ocz; ww_ -= sfac * odz;
  ww = (nd == -1)*ww + (nd < -1)*ww_;

  double st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
  *ur = st * (uu * cp + vv * sp) + ww * ct;
  *ut = ct * (uu * cp + vv * sp) - ww * st;
  *up = -uu * sp + vv * cp;
}

/* ---- the reference's stand-alone node kernels (device-pointer API) ---- */
__global__ void k_check_nodes(BBGeom G, BC bc, int nnodes,
  part_struct *__restrict__ parts, const double *__restrict__ th,
  const double *__restrict__ ph)
{
  int node = threadIdx.x, part = blockIdx.x;
  if(node >= nnodes) return;
  part_struct *P = &parts[part];
  double xp = P->rs * sin(th[node]) * cos(ph[node]);
  double yp = P->rs * sin(th[node]) * sin(ph[node]);
  double zp = P->rs * cos(th[node]);
  P->nodes[node] = node_flag(bc, G, xp + P->x, yp + P->y, zp + P->z);
}

__global__ void k_interp_nodes(InterpArgs A, int nnodes,
  const part_struct *__restrict__ parts, const double *__restrict__ th,
  const double *__restrict__ ph, const double *__restrict__ p,
  const double *__restrict__ u, const double *__restrict__ v,
  const double *__restrict__ w, double *pp, double *ur, double *ut,
  double *up)
{
  int node = threadIdx.x, part = blockIdx.x;
  if(node >= nnodes) return;
  const part_struct *P = &parts[part];
  int o = node + nnodes*part;
  interp_node(A, P, th[node], ph[node], P->nodes[node], p, u, v, w, &pp[o],
    &ur[o], &ut[o], &up[o]);
}

static InterpArgs interp_args(void)
{
  InterpArgs A;
  A.G = bb_geom(); A.bc = bc; A.gP = gradP; A.rho_f = rho_f; A.nu = nu;
  return A;
}

extern "C" void cuda_quad_check_nodes(int dev, real *node_t, real *node_p,
  int nnodes)
{
  (void) dev;
  bb_ensure_context();
  if(nparts <= 0) return;
  k_check_nodes<<<nparts, 64, 0, g_bb.stream>>>(bb_geom(), bc, nnodes,
    _parts[0], node_t, node_p);
  BB_CHECK_LAUNCH();
}

extern "C" void cuda_quad_interp(int dev, real *node_t, real *node_p,
  int nnodes, real *pp, real *ur, real *ut, real *up)
{
  (void) dev;
  bb_ensure_context();
  if(nparts <= 0) return;
  k_interp_nodes<<<nparts, 64, 0, g_bb.stream>>>(interp_args(), nnodes,
    _parts[0], node_t, node_p, _p[0], _u[0], _v[0], _w[0], pp, ur, ut, up);
  BB_CHECK_LAUNCH();
}

/* ---- fused Lamb step: one wavefront per particle ---- */
struct CoeffOut {
  double *pre, *pim, *fre, *fim, *cre, *cim;
  const double *pre0, *pim0, *fre0, *fim0, *cre0, *cim0;
};

__global__ void __launch_bounds__(64)
k_lamb(InterpArgs A, Nodes N, LambNorm L, CoeffOut K, int stride, double mu,
  double relax, part_struct *__restrict__ parts,
  const double *__restrict__ p, const double *__restrict__ u,
  const double *__restrict__ v, const double *__restrict__ w)
{
  __shared__ double s_pp[NNODES], s_ur[NNODES], s_ut[NNODES], s_up[NNODES];
  const int part = blockIdx.x;
  const int lane = threadIdx.x;
  part_struct *P = &parts[part];
  if(lane < NNODES) {
    double th = N.t[lane], ph = N.p[lane];
    double xp = P->rs * sin(th) * cos(ph);
    double yp = P->rs * sin(th) * sin(ph);
    double zp = P->rs * cos(th);
    int nd = node_flag(A.bc, A.G, xp + P->x, yp + P->y, zp + P->z);
    P->nodes[lane] = nd;
    interp_node(A, P, th, ph, nd, p, u, v, w, &s_pp[lane], &s_ur[lane],
      &s_ut[lane], &s_up[lane]);
  }
  __syncthreads();
  const int ncoeff = P->ncoeff < 15 ? P->ncoeff : 15;
  if(lane < ncoeff) {
    /* (n, m) of coefficient `lane`: n(n+1)/2 <= lane */
    int n = 0;
    while((n+1)*(n+2)/2 <= lane) n++;
    int m = lane - n*(n+1)/2;
    double N_nm = L.N[lane];
    double sY[2] = {0., 0.}, sR[2] = {0., 0.}, sX[2] = {0., 0.};
    for(int node = 0; node < NNODES; node++) {
      double th = N.t[node], ph = N.p[node];
      double x = cos(th), y = sin(th);
      double P_nm = lamb_P(n, m, x, y);
      double P_n1m = lamb_P(n+1, m, x, y);
      double dPdt = (n-m+1.)*P_n1m-(n+1.)*cos(th)*P_nm;
      double dPdp = m*P_nm;
      double cmp = cos(m*ph), smp = sin(m*ph);
      double pv = s_pp[node], tv = s_ut[node], qv = s_up[node];
      double Yr = N_nm*P_nm*pv*cmp;
      double Yi = -N_nm*P_nm*pv*smp;
      double Rr = N_nm/sin(th)*(dPdt*tv*cmp - dPdp*qv*smp);
      double Ri = N_nm/sin(th)*(-dPdt*tv*smp - dPdp*qv*cmp);
      double Xr = N_nm/sin(th)*(dPdp*tv*smp + dPdt*qv*cmp);
      double Xi = N_nm/sin(th)*(dPdp*tv*cmp - dPdt*qv*smp);
      double wt = N.w[node];
      if(node == 0) {
        sY[0] = Yr * wt; sY[1] = Yi * wt; sR[0] = Rr * wt; sR[1] = Ri * wt;
        sX[0] = Xr * wt; sX[1] = Xi * wt;
      } else {
        sY[0] += wt * Yr; sY[1] += wt * Yi; sR[0] += wt * Rr; sR[1] += wt * Ri;
        sX[0] += wt * Xr; sX[1] += wt * Xi;
      }
    }
    double r = P->r;
    double ars = P->r / P->rs, rsa = P->rs / P->r;
    const double nu = A.nu;
    int c = stride*part + lane;
    if(n == 0) {
      K.pre[c] = K.pre0[c] + relax*(r*r/mu/nu*sY[0]*pow(ars,(double) n) - K.pre0[c]);
      K.pim[c] = K.pim0[c] + relax*(r*r/mu/nu*sY[1]*pow(ars,(double) n) - K.pim0[c]);
      K.fre[c] = 0.; K.fim[c] = 0.; K.cre[c] = 0.; K.cim[c] = 0.;
    } else {
      double Aa = (1.-0.5*n*(2.*n-1.)/(n+1.)*pow(ars,2.*n+1.))*pow(rsa,(double) n);
      double Bb = n*(2.*n-1.)*(2.*n+1.)/(n+1.)*pow(ars,n+1.);
      double Cc = 0.25*n*(2.*(n+3.)/(2.*n+3.)
        + (n-2.-n*(2.*n+1.)/(2.*n+3.)*ars*ars)*pow(ars,2.*n+1.))*pow(rsa,n+1.);
      double Dd = n*(n+1.+0.5*((n-2.)*(2.*n+1.)*rsa*rsa
        - n*(2.*n-1.))*pow(ars,2.*n+1.))*pow(rsa,n-1.);
      double den = (Aa*Dd+Bb*Cc);
      double pr = (r*r/mu/nu*sY[0]*Dd + r/nu*sR[0]*Bb) / den;
      double pi = (r*r/mu/nu*sY[1]*Dd + r/nu*sR[1]*Bb) / den;
      double fr = (r/nu*sR[0]*Aa - r*r/mu/nu*sY[0]*Cc) / den;
      double fi = (r/nu*sR[1]*Aa - r*r/mu/nu*sY[1]*Cc) / den;
      double Ee = n*(n+1.)*(pow(ars,2.*n+1.)-1.)*pow(rsa,(double) n);
      double cr = r/nu*sX[0] / Ee;
      double ci = r/nu*sX[1] / Ee;
      K.pre[c] = K.pre0[c]*(1.-relax) + relax*pr;
      K.pim[c] = K.pim0[c]*(1.-relax) + relax*pi;
      K.fre[c] = K.fre0[c]*(1.-relax) + relax*fr;
      K.fim[c] = K.fim0[c]*(1.-relax) + relax*fi;
      K.cre[c] = K.cre0[c]*(1.-relax) + relax*cr;
      K.cim[c] = K.cim0[c]*(1.-relax) + relax*ci;
    }
  }
  __syncthreads();
  if(lane == 0) {
    /* cuda_calc_forces, src/quadrature_kernel.cu:545-576 */
    const double rho_f = A.rho_f, nu = A.nu;
    const gradP_struct gP = A.gP;
    int b = stride*part;
    double vol = 4./3. * PI * P->r*P->r*P->r;
    double N10 = sqrt(3./4./PI);
    double N11 = sqrt(3./8./PI);
    P->Fx = rho_f * vol * (P->udot + gP.x)
      - PI * mu * nu * 2.*N11 * (K.pre[b + 2] + 6.*K.fre[b + 2]);
    P->Fy = rho_f * vol * (P->vdot + gP.y)
      + PI * mu * nu * 2.*N11 * (K.pim[b + 2] + 6.*K.fim[b + 2]);
    P->Fz = rho_f * vol * (P->wdot + gP.z)
      + PI * mu * nu * N10 * (K.pre[b + 1] + 6.*K.fre[b + 1]);
    P->Lx = rho_f * vol * P->r*P->r * P->oxdot
      - 8. * PI * mu * nu * 2.*N11 * P->r * K.cre[b + 2];
    P->Ly = rho_f * vol * P->r*P->r * P->oydot
      + 8. * PI * mu * nu * 2.*N11 * P->r * K.cim[b + 2];
    P->Lz = rho_f * vol * P->r*P->r * P->ozdot
      + 8. * PI * mu * nu * N10 * P->r * K.cre[b + 1];
  }
}

extern "C" void cuda_Lamb(void)
{
  bb_ensure_particle_workspace();
  if(nparts <= 0) return;
  CoeffOut K = {_pnm_re[0], _pnm_im[0], _phinm_re[0], _phinm_im[0],
    _chinm_re[0], _chinm_im[0], _pnm_re0[0], _pnm_im0[0], _phinm_re0[0],
    _phinm_im0[0], _chinm_re0[0], _chinm_im0[0]};
  bb_prof_begin(BB_K_QUAD);
  k_lamb<<<nparts, 64, 0, g_bb.stream>>>(interp_args(), lebedev_nodes(),
    bb_lamb_norm(), K, coeff_stride, mu, lamb_relax, _parts[0], _p[0], _u[0],
    _v[0], _w[0]);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_QUAD);
}

/* per-particle error of src/quadrature_kernel.cu:578-654, each particle
 * in its own scratch (one wavefront per particle).
 *
 * The reference repeatedly picks the coefficient of largest square (the
 * lowest index among equal squares; only squares above FLT_MIN qualify),
 * scores it against its previous value -- relative to the first pick,
 * coefficients at most lamb_cut of it score 0 -- and zeroes it, 6*stride
 * times; the result is the largest score (at least FLT_MIN).  A pick's
 * score depends on the order only through the first pick, so every
 * qualifying coefficient is scored at once here, the first pick found by
 * an argmax over the wavefront.  The picks past the qualifying ones see
 * only zeroes and re-select the last pick (the smallest square, the
 * highest index among equal ones), scoring a zero value there; with none
 * qualifying, index 0 is picked first.  Each score is the reference's
 * expression, so the result is the same double. */
struct ErrPick { double sq; int j; };

/* a before b in the reference's order of picks (j < 0: no pick) */
__device__ inline bool pick_first(ErrPick a, ErrPick b)
{
  if(b.j < 0) return a.j >= 0;
  if(a.j < 0) return false;
  return a.sq > b.sq || (a.sq == b.sq && a.j < b.j);
}

__device__ inline ErrPick shfl_pick(ErrPick p, int d)
{
  ErrPick o;
  o.sq = __shfl_down(p.sq, d, 64);
  o.j = __shfl_down(p.j, d, 64);
  return o;
}

__device__ inline double lamb_score(double val, double old, double cut,
  double first)
{
  if(fabs(val) > cut * fabs(first)) {
    double div = fabs(val);
    if(div < 1e-16) div = 1e-16;
    return fabs((val - old) / div);
  }
  return 0.;
}

__global__ void __launch_bounds__(64)
k_lamb_err(int stride, double lamb_cut, const double *pre, const double *pre0,
  const double *pim, const double *pim0, const double *fre, const double *fre0,
  const double *fim, const double *fim0, const double *cre, const double *cre0,
  const double *cim, const double *cim0, double *__restrict__ out)
{
  __shared__ double sc[6*21], sc0[6*21];
  __shared__ double s_first;
  __shared__ int s_last, s_count;
  const int part = blockIdx.x;
  const int lane = threadIdx.x;
  const int n6 = 6 * stride;
  for(int q = lane; q < stride; q += 64) {
    int o = part*stride + q;
    sc[q] = pre[o]; sc[q + stride] = pim[o]; sc[q + 2*stride] = fre[o];
    sc[q + 3*stride] = fim[o]; sc[q + 4*stride] = cre[o]; sc[q + 5*stride] = cim[o];
    sc0[q] = pre0[o]; sc0[q + stride] = pim0[o]; sc0[q + 2*stride] = fre0[o];
    sc0[q + 3*stride] = fim0[o]; sc0[q + 4*stride] = cre0[o];
    sc0[q + 5*stride] = cim0[o];
  }
  __syncthreads();
  /* this lane's coefficients j = lane, lane + 64 (n6 <= 126) */
  ErrPick first = {0., -1}, last = {0., -1};
  int count = 0;
  for(int j = lane; j < n6; j += 64) {
    const double sq = sc[j] * sc[j];
    if(!(sq > (double) FLT_MIN)) continue;
    const ErrPick p = {sq, j};
    count++;
    if(pick_first(p, first)) first = p;
    if(last.j < 0 || pick_first(last, p)) last = p;
  }
  for(int d = 32; d > 0; d >>= 1) {
    const ErrPick f = shfl_pick(first, d), l = shfl_pick(last, d);
    count += __shfl_down(count, d, 64);
    if(pick_first(f, first)) first = f;
    if(l.j >= 0 && (last.j < 0 || pick_first(last, l))) last = l;
  }
  if(lane == 0) {
    s_count = count;
    s_first = count > 0 ? sc[first.j] : sc[0];
    s_last = count > 0 ? last.j : 0;
  }
  __syncthreads();
  const double f0 = s_first;
  const int nq = s_count;
  double best = FLT_MIN;
  for(int j = lane; j < n6; j += 64) {
    const bool picked = nq > 0 ? sc[j] * sc[j] > (double) FLT_MIN : j == 0;
    if(!picked) continue;
    const double e = lamb_score(sc[j], sc0[j], lamb_cut, f0);
    if(e > best) best = e;
  }
  /* the remaining picks: a zero value at the last pick */
  if(lane == 0 && n6 > (nq > 0 ? nq : 1)) {
    const double e = lamb_score(0., sc0[s_last], lamb_cut, f0);
    if(e > best) best = e;
  }
  for(int d = 32; d > 0; d >>= 1) {
    const double o = __shfl_down(best, d, 64);
    if(o > best) best = o;
  }
  if(lane == 0) out[part] = best;
}

extern "C" real cuda_lamb_err(void)
{
  bb_ensure_particle_workspace();
  if(nparts <= 0) return 0.;
  hipStream_t st = g_bb.stream;
  bb_prof_begin(BB_K_QUAD);
  k_lamb_err<<<nparts, 64, 0, st>>>(coeff_stride, lamb_cut, _pnm_re[0],
    _pnm_re0[0], _pnm_im[0], _pnm_im0[0], _phinm_re[0], _phinm_re0[0],
    _phinm_im[0], _phinm_im0[0], _chinm_re[0], _chinm_re0[0], _chinm_im[0],
    _chinm_im0[0], g_bb.coeff_err);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_QUAD);
  BB_CHECK(hipMemcpyAsync(g_bb.host_err, g_bb.coeff_err,
    (size_t) nparts * sizeof(double), hipMemcpyDeviceToHost, st));
  size_t n = (size_t) coeff_stride * nparts * sizeof(real);
  BB_CHECK(hipMemcpyAsync(_pnm_re0[0], _pnm_re[0], n, hipMemcpyDeviceToDevice, st));
  BB_CHECK(hipMemcpyAsync(_pnm_im0[0], _pnm_im[0], n, hipMemcpyDeviceToDevice, st));
  BB_CHECK(hipMemcpyAsync(_phinm_re0[0], _phinm_re[0], n, hipMemcpyDeviceToDevice, st));
  BB_CHECK(hipMemcpyAsync(_phinm_im0[0], _phinm_im[0], n, hipMemcpyDeviceToDevice, st));
  BB_CHECK(hipMemcpyAsync(_chinm_re0[0], _chinm_re[0], n, hipMemcpyDeviceToDevice, st));
  BB_CHECK(hipMemcpyAsync(_chinm_im0[0], _chinm_im[0], n, hipMemcpyDeviceToDevice, st));
  BB_CHECK(hipStreamSynchronize(st));
  real mx = FLT_MIN;
  for(int i = 0; i < nparts; i++) if(g_bb.host_err[i] > mx) mx = g_bb.host_err[i];
  return mx;
}

/* ---------------------------------------------------------------------
 * bb_lamb_err_noalloc: cuda_lamb_err for callers that cannot allocate
 * device memory.  The per-particle errors and the result live in the
 * file-scope array `scratch` (one copy per device, held until the process
 * ends) instead of the library's workspace; the caller passes the six
 * current and the six previous coefficient arrays.  No hipMalloc, hipFree
 * or stream creation happens here.  Calls that overlap on one device race
 * on `scratch`.
 * ------------------------------------------------------------------- */
#define BB_NOALLOC_MAX_PARTS 32768
#define BB_NOALLOC_DEVICES 64

static __device__ __align__(256) double scratch[BB_NOALLOC_MAX_PARTS + 64];

/* device address of `scratch` on each device, set on the first call made
 * with that device current */
static double *g_scratch_dev[BB_NOALLOC_DEVICES];

static double *scratch_ptr(int dev)
{
  if(!g_scratch_dev[dev]) {
    void *ptr = NULL;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if(e != hipSuccess || !ptr) {
      fprintf(stderr, "bluebottle_hip: bb_lamb_err_noalloc: no address for "
        "its scratch on device %d (%s)\n", dev, hipGetErrorString(e));
      exit(EXIT_FAILURE);
    }
    g_scratch_dev[dev] = (double *) ptr;
  }
  return g_scratch_dev[dev];
}

/* out[0] = max(FLT_MIN, err[0..n)), the host loop of cuda_lamb_err */
__global__ void __launch_bounds__(64)
k_err_max(int n, const double *__restrict__ err, double *__restrict__ out)
{
  __shared__ double sh[64];
  double m = FLT_MIN;
  for(int q = threadIdx.x; q < n; q += 64) if(err[q] > m) m = err[q];
  sh[threadIdx.x] = m;
  __syncthreads();
  if(threadIdx.x == 0) {
    double r = FLT_MIN;
    for(int q = 0; q < 64; q++) if(sh[q] > r) r = sh[q];
    out[0] = r;
  }
}

extern "C" real bb_lamb_err_noalloc(int np, int stride, real cut,
  real *const cur[6], real *const prev[6])
{
  if(np <= 0) return 0.;
  if(np > BB_NOALLOC_MAX_PARTS) {
    fprintf(stderr, "bluebottle_hip: bb_lamb_err_noalloc: %d particles "
      "exceed its scratch of %d; use cuda_lamb_err\n", np,
      BB_NOALLOC_MAX_PARTS);
    return -1.;
  }
  if(stride < 1 || stride > 21) {
    fprintf(stderr, "bluebottle_hip: bb_lamb_err_noalloc: coefficient "
      "stride %d outside 1..21 (Lamb order above 5)\n", stride);
    return -1.;
  }
  int dev = 0;
  BB_CHECK(hipGetDevice(&dev));
  if(dev < 0 || dev >= BB_NOALLOC_DEVICES) {
    fprintf(stderr, "bluebottle_hip: bb_lamb_err_noalloc: device %d beyond "
      "the %d it keeps addresses for\n", dev, BB_NOALLOC_DEVICES);
    return -1.;
  }
  double *sp = scratch_ptr(dev);
  /* the library's stream when it has one on this device (ordered after
   * the kernels that wrote the coefficients), else the null stream */
  hipStream_t st = (g_bb.ready && g_bb.device == dev) ? g_bb.stream : 0;
  k_lamb_err<<<np, 64, 0, st>>>(stride, cut, cur[0], prev[0], cur[1],
    prev[1], cur[2], prev[2], cur[3], prev[3], cur[4], prev[4], cur[5],
    prev[5], sp);
  BB_CHECK_LAUNCH();
  k_err_max<<<1, 64, 0, st>>>(np, sp, sp + BB_NOALLOC_MAX_PARTS);
  BB_CHECK_LAUNCH();
  size_t n = (size_t) stride * np * sizeof(real);
  for(int a = 0; a < 6; a++)
    BB_CHECK(hipMemcpyAsync(prev[a], cur[a], n, hipMemcpyDeviceToDevice, st));
  double mx = 0.;
  BB_CHECK(hipMemcpyAsync(&mx, sp + BB_NOALLOC_MAX_PARTS, sizeof(double),
    hipMemcpyDeviceToHost, st));
  BB_CHECK(hipStreamSynchronize(st));
  return mx;
}
