/*
 * bb_field.hip -- field kernels of the per-step path:
 *   cuda_U_star_2        (src/cuda_bluebottle.cu:2930-3026,
 *                         src/bluebottle_kernel.cu:1504-2003)
 *   cuda_dom_BC{,_star,_phi,_p}
 *                        (src/cuda_bluebottle.cu:1549-3420,
 *                         src/bluebottle_kernel.cu:26-1248)
 *   cuda_solvability     (src/cuda_bluebottle.cu:4122-4314)
 *   cuda_project         (src/cuda_bluebottle.cu:3422-3497)
 *   cuda_update_p        (src/cuda_bluebottle.cu:3616-3671)
 *   cuda_find_dt         (src/cuda_bluebottle.cu:3500-3577)
 *   cuda_store_u         (src/cuda_bluebottle.cu:3580-3613)
 *   cuda_compute_forcing (src/cuda_bluebottle.cu:3674-3814)
 *
 * Design: the reference launches one 16x16 thread tile per 2-D column and
 * loops serially over the third axis, with u and v uncoalesced.  Here every
 * kernel maps wavefront lanes to x (the contiguous axis), so each wavefront
 * reads 64 consecutive doubles per array; the three predictor components
 * are fused into one launch; ghost-cell BCs are three launches (one per
 * axis, all fields batched) instead of up to 24; reductions are two-level
 * and deterministic (per-block partials, one finishing block) and never
 * round-trip to the host except where the entry point returns a scalar.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "bb_device.h"

#define TX 64
#define TY 4

/* ---------------------------------------------------------------------
 * Predictor.  Thread (i,j,k) computes the u, v and w faces with those
 * indices when they are interior faces of their grid.  Expressions keep
 * the reference's operation order.
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_u_star(BBGeom G, double rho_f, double nu, double dt0, double dt,
  const double *__restrict__ u0, const double *__restrict__ v0,
  const double *__restrict__ w0, const double *__restrict__ p,
  const double *__restrict__ fx, const double *__restrict__ fy,
  const double *__restrict__ fz,
  const double *__restrict__ diff0_u, const double *__restrict__ diff0_v,
  const double *__restrict__ diff0_w,
  const double *__restrict__ conv0_u, const double *__restrict__ conv0_v,
  const double *__restrict__ conv0_w,
  double *__restrict__ diff_u, double *__restrict__ diff_v,
  double *__restrict__ diff_w,
  double *__restrict__ conv_u, double *__restrict__ conv_v,
  double *__restrict__ conv_w,
  double *__restrict__ us_, double *__restrict__ vs_, double *__restrict__ ws_,
  const int *__restrict__ phase)
{
  const int i = blockIdx.x * TX + threadIdx.x + 1;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  const double ab0 = 0.5 * dt / dt0;
  const double ab = 1. + ab0;
  const double ddx = 1. / G.dx, ddy = 1. / G.dy, ddz = 1. / G.dz;
  const int cs1 = G.cs1, cs2 = G.cs2;
  const int us1 = G.us1, us2 = G.us2, vs1 = G.vs1, vs2 = G.vs2;
  const int ws1 = G.ws1, ws2 = G.ws2;
  const int Cc = i + j*cs1 + k*cs2;

  if(i <= G.xn + 1 && j <= G.yn && k <= G.zn) {
    const int C = i + j*us1 + k*us2;
    double us = (p[Cc-1] - p[Cc]) * ddx / rho_f;
    double u011 = u0[C-1], u111 = u0[C], u211 = u0[C+1];
    double u101 = u0[C-us1], u121 = u0[C+us1];
    double u110 = u0[C-us2], u112 = u0[C+us2];
    const int V = i + j*vs1 + k*vs2;
    double v011 = v0[V-1], v111 = v0[V], v021 = v0[V-1+vs1], v121 = v0[V+vs1];
    const int W = i + j*ws1 + k*ws2;
    double w011 = w0[W-1], w111 = w0[W], w012 = w0[W-1+ws2], w112 = w0[W+ws2];
    double duudx = (u211 + u111)*(u211 + u111) - (u111 + u011)*(u111 + u011);
    duudx *= 0.25 * ddx;
    double duvdy = (u121 + u111)*(v121 + v021) - (u111 + u101)*(v111 + v011);
    duvdy *= 0.25 * ddy;
    double duwdz = (u112 + u111)*(w112 + w012) - (u111 + u110)*(w111 + w011);
    duwdz *= 0.25 * ddz;
    double c = duudx + duvdy + duwdz;
    if(dt0 > 0) us += (-ab * c + ab0 * conv0_u[C]);
    else us += -c;
    double d1 = (u211 - u111) * ddx, d0 = (u111 - u011) * ddx;
    double dxx = (d1 - d0) * ddx;
    d1 = (u121 - u111) * ddy; d0 = (u111 - u101) * ddy;
    double dyy = (d1 - d0) * ddy;
    d1 = (u112 - u111) * ddz; d0 = (u111 - u110) * ddz;
    double dzz = (d1 - d0) * ddz;
    double df = nu * (dxx + dyy + dzz);
    if(dt0 > 0) us += (ab * df - ab0 * diff0_u[C]);
    else us += df;
    us += fx[C];
    us *= dt;
    us += u111;
    us *= (double) (phase[Cc-1] < 0 && phase[Cc] < 0);
    us_[C] = us;
    conv_u[C] = c;
    diff_u[C] = df;
  }
  if(i <= G.xn && j <= G.yn + 1 && k <= G.zn) {
    const int C = i + j*vs1 + k*vs2;
    double vs = (p[Cc-cs1] - p[Cc]) * ddy / rho_f;
    double v101 = v0[C-vs1], v111 = v0[C], v121 = v0[C+vs1];
    double v110 = v0[C-vs2], v112 = v0[C+vs2];
    double v011 = v0[C-1], v211 = v0[C+1];
    const int W = i + j*ws1 + k*ws2;
    double w101 = w0[W-ws1], w111 = w0[W], w102 = w0[W-ws1+ws2], w112 = w0[W+ws2];
    const int U = i + j*us1 + k*us2;
    double u101 = u0[U-us1], u111 = u0[U], u201 = u0[U+1-us1], u211 = u0[U+1];
    double dvudx = (v211 + v111)*(u211 + u201) - (v111 + v011)*(u111 + u101);
    dvudx *= 0.25 * ddx;
    double dvvdy = (v121 + v111)*(v121 + v111) - (v111 + v101)*(v111 + v101);
    dvvdy *= 0.25 * ddy;
    double dvwdz = (v112 + v111)*(w112 + w102) - (v111 + v110)*(w111 + w101);
    dvwdz *= 0.25 * ddz;
    double c = dvudx + dvvdy + dvwdz;
    if(dt0 > 0) vs += (-ab * c + ab0 * conv0_v[C]);
    else vs += -c;
    double d1 = (v211 - v111) * ddx, d0 = (v111 - v011) * ddx;
    double dxx = (d1 - d0) * ddx;
    d1 = (v121 - v111) * ddy; d0 = (v111 - v101) * ddy;
    double dyy = (d1 - d0) * ddy;
    d1 = (v112 - v111) * ddz; d0 = (v111 - v110) * ddz;
    double dzz = (d1 - d0) * ddz;
    double df = nu * (dxx + dyy + dzz);
    if(dt0 > 0) vs += (ab * df - ab0 * diff0_v[C]);
    else vs += df;
    vs += fy[C];
    vs *= dt;
    vs += v111;
    vs *= (double) (phase[Cc-cs1] < 0 && phase[Cc] < 0);
    vs_[C] = vs;
    conv_v[C] = c;
    diff_v[C] = df;
  }
  if(i <= G.xn && j <= G.yn && k <= G.zn + 1) {
    const int C = i + j*ws1 + k*ws2;
    double ws = (p[Cc-cs2] - p[Cc]) * ddz / rho_f;
    double w110 = w0[C-ws2], w111 = w0[C], w112 = w0[C+ws2];
    double w011 = w0[C-1], w211 = w0[C+1];
    double w101 = w0[C-ws1], w121 = w0[C+ws1];
    const int U = i + j*us1 + k*us2;
    double u110 = u0[U-us2], u111 = u0[U], u210 = u0[U+1-us2], u211 = u0[U+1];
    const int V = i + j*vs1 + k*vs2;
    double v110 = v0[V-vs2], v111 = v0[V], v120 = v0[V+vs1-vs2], v121 = v0[V+vs1];
    double dwudx = (w211 + w111)*(u211 + u210) - (w111 + w011)*(u111 + u110);
    dwudx *= 0.25 * ddx;
    double dwvdy = (w121 + w111)*(v121 + v120) - (w111 + w101)*(v111 + v110);
    dwvdy *= 0.25 * ddy;
    double dwwdz = (w112 + w111)*(w112 + w111) - (w111 + w110)*(w111 + w110);
    dwwdz *= 0.25 * ddz;
    double c = dwudx + dwvdy + dwwdz;
    if(dt0 > 0) ws += (-ab * c + ab0 * conv0_w[C]);
    else ws += -c;
    double d1 = (w211 - w111) * ddx, d0 = (w111 - w011) * ddx;
    double dxx = (d1 - d0) * ddx;
    d1 = (w121 - w111) * ddy; d0 = (w111 - w101) * ddy;
    double dyy = (d1 - d0) * ddy;
    d1 = (w112 - w111) * ddz; d0 = (w111 - w110) * ddz;
    double dzz = (d1 - d0) * ddz;
    double df = nu * (dxx + dyy + dzz);
    if(dt0 > 0) ws += (ab * df - ab0 * diff0_w[C]);
    else ws += df;
    ws += fz[C];
    ws *= dt;
    ws += w111;
    ws *= (double) (phase[Cc-cs2] < 0 && phase[Cc] < 0);
    ws_[C] = ws;
    conv_w[C] = c;
    diff_w[C] = df;
  }
}

extern "C" void cuda_U_star_2(void)
{
  bb_ensure_context();
  if(g_bb.implicit) {
    /* the IMPLICIT build has no explicit predictor and no diff0 history */
    fprintf(stderr, "bluebottle_hip: cuda_U_star_2 called in the IMPLICIT "
      "variant; use cuda_{u,v,w}star_helmholtz\n");
    exit(EXIT_FAILURE);
  }
  BBGeom G = bb_geom();
  dim3 blk(TX, TY, 1);
  dim3 grd((G.xn + 1 + TX - 1) / TX, (G.yn + 1 + TY - 1) / TY, G.zn + 1);
  bb_prof_begin(BB_K_USTAR);
  k_u_star<<<grd, blk, 0, g_bb.stream>>>(G, rho_f, nu, dt0, dt,
    _u0[0], _v0[0], _w0[0], _p0[0], _f_x[0], _f_y[0], _f_z[0],
    _diff0_u[0], _diff0_v[0], _diff0_w[0], _conv0_u[0], _conv0_v[0],
    _conv0_w[0], _diff_u[0], _diff_v[0], _diff_w[0], _conv_u[0], _conv_v[0],
    _conv_w[0], _u_star[0], _v_star[0], _w_star[0], _phase[0]);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_USTAR);
}

/* ---------------------------------------------------------------------
 * Ghost-cell boundary conditions.  One launch per axis applies the low
 * then the high face of every field in the batch over the full ghosted
 * transverse extent, matching the reference's W,E,S,N,B,T kernel order.
 * ------------------------------------------------------------------- */
/* A PRECURSOR face reads its values from inflow planes: the normal
 * component from one plane, as a Dirichlet value per position; a
 * tangential component sets the ghost and the first interior value (low
 * face) or the last interior value and the ghost (high face) from two
 * planes.  BC_PRECURSOR_AT_LO on a high face applies the low face's
 * normal-component fill there instead (the reference's cuda_dom_BC_star
 * calls its south kernel for the north face of v). */
#define BC_PRECURSOR_AT_LO 103

struct BCField {
  double *f;
  int s1, s2;          /* strides */
  int n[3];            /* ghosted extents inb, jnb, knb */
  int normal;          /* staggered axis, -1 for cell-centred */
  int type[6];
  double val[6];
  const double *pl[6][2];   /* inflow planes of PRECURSOR faces */
};

struct BCBatch {
  BCField fld[4];
  int nf;
};

__device__ inline void bc_apply(double *F, long str, int lo, int type,
  int normal, double bcv, int s, int e, int sb, int eb,
  const double *const *pl, long P)
{
  /* ghost values are bit-identical to the reference's separately rounded
   * products and sums */
#pragma clang fp contract(off)
  if(type == PRECURSOR) {
    if(normal) {
      const double b = pl[0][P];
      if(lo) { F[sb*str] = 2. * b - F[(s+1)*str]; F[s*str] = b; }
      else   { F[(eb-1)*str] = 2. * b - F[(e-2)*str]; F[(e-1)*str] = b; }
    } else {
      const double b0 = pl[0][P], b1 = pl[1][P];
      if(lo) { F[sb*str] = b0; F[s*str] = b1; }
      else   { F[(e-1)*str] = b0; F[(eb-1)*str] = b1; }
    }
  } else if(type == BC_PRECURSOR_AT_LO) {
    const double b = pl[0][P];
    F[sb*str] = 2. * b - F[(s+1)*str]; F[s*str] = b;
  } else if(type == PERIODIC) {
    if(normal) {
      if(lo) { F[sb*str] = F[(e-2)*str]; F[s*str] = F[(e-1)*str]; }
      else   { F[(eb-1)*str] = F[(s+1)*str]; F[(e-1)*str] = F[s*str]; }
    } else {
      if(lo) F[sb*str] = F[(e-1)*str];
      else   F[(eb-1)*str] = F[s*str];
    }
  } else if(type == DIRICHLET) {
    if(normal) {
      if(lo) { F[sb*str] = 2. * bcv - F[(s+1)*str]; F[s*str] = bcv; }
      else   { F[(eb-1)*str] = 2. * bcv - F[(e-2)*str]; F[(e-1)*str] = bcv; }
    } else {
      if(lo) F[sb*str] = 8./3. * bcv - 2. * F[s*str] + 1./3. * F[(s+1)*str];
      else   F[(eb-1)*str] = 8./3. * bcv - 2. * F[(e-1)*str] + 1./3. * F[(e-2)*str];
    }
  } else if(type == NEUMANN) {
    if(lo) F[sb*str] = F[s*str];
    else   F[(eb-1)*str] = F[(e-1)*str];
  }
}

/* One field's work along one axis, resolved on the host: the kernel reads
 * every value at a fixed kernel-argument offset, with no per-axis
 * selection on the device (a three-way selection of argument fields by a
 * run-time axis index compiled, for gfx950, into loads through scalar
 * registers that the axis-2 path never set). */
struct BCAxisField {
  double *f;
  long s1, s2;         /* strides of the two transverse axes */
  long str;            /* stride along the axis */
  int n1, n2, nax;     /* ghosted extents: transverse, transverse, axis */
  int normal;          /* 1 when the axis is the field's staggered axis */
  int type_lo, type_hi;
  double val_lo, val_hi;
  const double *pl_lo[2], *pl_hi[2];   /* planes of PRECURSOR faces */
  long p1, p2;         /* plane strides of the two transverse axes */
};

struct BCAxisBatch {
  BCAxisField fld[4];
  int nf;
};

__global__ void __launch_bounds__(256) k_bc_axis(BCAxisBatch B)
{
  const int t1 = blockIdx.x * blockDim.x + threadIdx.x;
  const int t2 = blockIdx.y * blockDim.y + threadIdx.y;
#pragma unroll
  for(int q = 0; q < 4; q++) {
    if(q >= B.nf) break;
    const BCAxisField &fd = B.fld[q];
    if(t1 >= fd.n1 || t2 >= fd.n2) continue;
    double *F = fd.f + (long) t1 * fd.s1 + (long) t2 * fd.s2;
    const long P = (long) t1 * fd.p1 + (long) t2 * fd.p2;
    /* along the axis: sb = 0, s = 1, e = nax-1, eb = nax */
    const int s = 1, e = fd.nax - 1, sb = 0, eb = fd.nax;
    bc_apply(F, fd.str, 1, fd.type_lo, fd.normal, fd.val_lo, s, e, sb, eb,
      fd.pl_lo, P);
    bc_apply(F, fd.str, 0, fd.type_hi, fd.normal, fd.val_hi, s, e, sb, eb,
      fd.pl_hi, P);
  }
}

/* the inflow planes the reference's cuda_dom_BC (star = 0) and
 * cuda_dom_BC_star (star = 1) pass for a PRECURSOR face of component
 * `which`; its particulars are kept: cuda_dom_BC gives the west face of v
 * v_WE_W for both planes, cuda_dom_BC_star fills the north face of v with
 * the south kernel.  Departure: cuda_dom_BC's north-face w call writes into
 * u at w's indices (out of u's bounds unless the grids coincide); here it
 * fills w. */
static void precursor_planes(BCField &fd, char which, int star)
{
  for(int q = 0; q < 6; q++) fd.pl[q][0] = fd.pl[q][1] = NULL;
  if(which == 'u') {
    fd.pl[0][0] = fd.pl[1][0] = _u_WE[0];
    fd.pl[2][0] = fd.pl[3][0] = _u_SN_S[0];
    fd.pl[2][1] = fd.pl[3][1] = _u_SN_N[0];
    fd.pl[4][0] = fd.pl[5][0] = _u_BT_B[0];
    fd.pl[4][1] = fd.pl[5][1] = _u_BT_T[0];
  } else if(which == 'v') {
    fd.pl[0][0] = _v_WE_W[0]; fd.pl[0][1] = star ? _v_WE_E[0] : _v_WE_W[0];
    fd.pl[1][0] = _v_WE_W[0]; fd.pl[1][1] = _v_WE_E[0];
    fd.pl[2][0] = fd.pl[3][0] = _v_SN[0];
    fd.pl[4][0] = fd.pl[5][0] = _v_BT_B[0];
    fd.pl[4][1] = fd.pl[5][1] = _v_BT_T[0];
    if(star && fd.type[3] == PRECURSOR) fd.type[3] = BC_PRECURSOR_AT_LO;
  } else if(which == 'w') {
    fd.pl[0][0] = fd.pl[1][0] = _w_WE_W[0];
    fd.pl[0][1] = fd.pl[1][1] = _w_WE_E[0];
    fd.pl[2][0] = fd.pl[3][0] = _w_SN_S[0];
    fd.pl[2][1] = fd.pl[3][1] = _w_SN_N[0];
    fd.pl[4][0] = fd.pl[5][0] = _w_BT[0];
  }
}

static BCField bc_field(double *f, char which, int star = 0)
{
  BCField fd;
  const dom_struct *d = &dom[0];
  const grid_info *G = which == 'p' ? &d->Gcc : which == 'u' ? &d->Gfx
    : which == 'v' ? &d->Gfy : &d->Gfz;
  fd.f = f;
  fd.s1 = G->s1b; fd.s2 = G->s2b;
  fd.n[0] = G->inb; fd.n[1] = G->jnb; fd.n[2] = G->knb;
  fd.normal = which == 'u' ? 0 : which == 'v' ? 1 : which == 'w' ? 2 : -1;
  const BC *b = &bc;
  switch(which) {
    case 'p': {
      int t[6] = {b->pW, b->pE, b->pS, b->pN, b->pB, b->pT};
      for(int q = 0; q < 6; q++) { fd.type[q] = t[q]; fd.val[q] = 0.; }
      break; }
    case 'u': {
      int t[6] = {b->uW, b->uE, b->uS, b->uN, b->uB, b->uT};
      double vv[6] = {b->uWD, b->uED, b->uSD, b->uND, b->uBD, b->uTD};
      for(int q = 0; q < 6; q++) { fd.type[q] = t[q]; fd.val[q] = vv[q]; }
      break; }
    case 'v': {
      int t[6] = {b->vW, b->vE, b->vS, b->vN, b->vB, b->vT};
      double vv[6] = {b->vWD, b->vED, b->vSD, b->vND, b->vBD, b->vTD};
      for(int q = 0; q < 6; q++) { fd.type[q] = t[q]; fd.val[q] = vv[q]; }
      break; }
    default: {
      int t[6] = {b->wW, b->wE, b->wS, b->wN, b->wB, b->wT};
      double vv[6] = {b->wWD, b->wED, b->wSD, b->wND, b->wBD, b->wTD};
      for(int q = 0; q < 6; q++) { fd.type[q] = t[q]; fd.val[q] = vv[q]; }
      break; }
  }
  /* faces of a subdomain with a neighbour are not external; the pressure
   * has no PRECURSOR type */
  int ext[6] = {d->W, d->E, d->S, d->N, d->B, d->T};
  for(int q = 0; q < 6; q++) {
    if(ext[q] != -1) fd.type[q] = -1;
    if(fd.type[q] == PRECURSOR && which == 'p') {
      fprintf(stderr, "bluebottle_hip: a PRECURSOR pressure boundary is not "
        "defined\n");
      exit(EXIT_FAILURE);
    }
  }
  precursor_planes(fd, which, star);
  return fd;
}

/* The per-axis view of field `fd`: transverse extents and strides in
 * (x, y, z) order with the axis removed. */
static BCAxisField bc_axis_field(const BCField &fd, int axis)
{
  BCAxisField a;
  const long st[3] = {1, fd.s1, fd.s2};
  const int t1 = axis == 0 ? 1 : 0;
  const int t2 = axis == 2 ? 1 : 2;
  a.f = fd.f;
  a.n1 = fd.n[t1]; a.n2 = fd.n[t2]; a.nax = fd.n[axis];
  a.s1 = st[t1]; a.s2 = st[t2]; a.str = st[axis];
  a.normal = fd.normal == axis;
  a.type_lo = fd.type[2*axis]; a.type_hi = fd.type[2*axis+1];
  a.val_lo = fd.val[2*axis]; a.val_hi = fd.val[2*axis+1];
  for(int q = 0; q < 2; q++) {
    a.pl_lo[q] = fd.pl[2*axis][q];
    a.pl_hi[q] = fd.pl[2*axis+1][q];
  }
  /* plane layouts: W/E j + k*jnb, S/N k + i*knb, B/T i + j*inb, with
   * (t1, t2) = (j, k), (i, k), (i, j) */
  if(axis == 0) { a.p1 = 1; a.p2 = fd.n[1]; }
  else if(axis == 1) { a.p1 = fd.n[2]; a.p2 = 1; }
  else { a.p1 = 1; a.p2 = fd.n[0]; }
  return a;
}

static void bc_launch(BCBatch &B)
{
  bb_prof_begin(BB_K_BC);
  for(int axis = 0; axis < 3; axis++) {
    BCAxisBatch A;
    memset(&A, 0, sizeof(A));
    A.nf = B.nf;
    int m1 = 0, m2 = 0;
    for(int q = 0; q < B.nf; q++) {
      A.fld[q] = bc_axis_field(B.fld[q], axis);
      if(A.fld[q].n1 > m1) m1 = A.fld[q].n1;
      if(A.fld[q].n2 > m2) m2 = A.fld[q].n2;
    }
    dim3 blk(64, 4);
    dim3 grd((m1 + 63) / 64, (m2 + 3) / 4);
    k_bc_axis<<<grd, blk, 0, g_bb.stream>>>(A);
    BB_CHECK_LAUNCH();
  }
  bb_prof_end(BB_K_BC);
}

extern "C" void cuda_dom_BC(void)
{
  bb_ensure_context();
  BCBatch B;
  B.nf = 4;
  B.fld[0] = bc_field(_p[0], 'p');
  B.fld[1] = bc_field(_u[0], 'u');
  B.fld[2] = bc_field(_v[0], 'v');
  B.fld[3] = bc_field(_w[0], 'w');
  bc_launch(B);
}

extern "C" void cuda_dom_BC_star(void)
{
  bb_ensure_context();
  BCBatch B;
  B.nf = 3;
  B.fld[0] = bc_field(_u_star[0], 'u', 1);
  B.fld[1] = bc_field(_v_star[0], 'v', 1);
  B.fld[2] = bc_field(_w_star[0], 'w', 1);
  bc_launch(B);
}

extern "C" void cuda_dom_BC_phi(void)
{
  bb_ensure_context();
  BCBatch B;
  B.nf = 1;
  B.fld[0] = bc_field(_phi[0], 'p');
  bc_launch(B);
}

extern "C" void cuda_dom_BC_p(void)
{
  bb_ensure_context();
  BCBatch B;
  B.nf = 1;
  B.fld[0] = bc_field(_p[0], 'p');
  bc_launch(B);
}

/* ---------------------------------------------------------------------
 * Solvability: three face-pair flux sums (deterministic two-level
 * reduction), then the outflow-plane correction computed on the device.
 * ------------------------------------------------------------------- */
#define SOLV_BLOCKS 256

/* partial sums of -f(low face) + f(high face) for the three components */
__global__ void __launch_bounds__(256)
k_solv_partial(BBGeom G, const double *__restrict__ us,
  const double *__restrict__ vs, const double *__restrict__ ws,
  double *__restrict__ part)
{
  __shared__ double sh[4];
  const int tid = threadIdx.x;
  double sx = 0., sy = 0., sz = 0.;
  const int nx = G.yn * G.zn, ny = G.xn * G.zn, nz = G.xn * G.yn;
  for(int q = blockIdx.x * 256 + tid; q < 2 * nx; q += gridDim.x * 256) {
    int h = q >= nx, r = h ? q - nx : q;
    int tj = r % G.yn, tk = r / G.yn;
    int C = (h ? G.xn + 1 : 1) + (tj + 1) * G.us1 + (tk + 1) * G.us2;
    sx += h ? us[C] : -us[C];
  }
  for(int q = blockIdx.x * 256 + tid; q < 2 * ny; q += gridDim.x * 256) {
    int h = q >= ny, r = h ? q - ny : q;
    int ti = r % G.xn, tk = r / G.xn;
    int C = (ti + 1) + (h ? G.yn + 1 : 1) * G.vs1 + (tk + 1) * G.vs2;
    sy += h ? vs[C] : -vs[C];
  }
  for(int q = blockIdx.x * 256 + tid; q < 2 * nz; q += gridDim.x * 256) {
    int h = q >= nz, r = h ? q - nz : q;
    int ti = r % G.xn, tj = r / G.xn;
    int C = (ti + 1) + (tj + 1) * G.ws1 + (h ? G.zn + 1 : 1) * G.ws2;
    sz += h ? ws[C] : -ws[C];
  }
  double a = bb_block_sum(sx, sh);
  double b = bb_block_sum(sy, sh);
  double c = bb_block_sum(sz, sh);
  if(tid == 0) {
    part[3*blockIdx.x + 0] = a;
    part[3*blockIdx.x + 1] = b;
    part[3*blockIdx.x + 2] = c;
  }
}

/* one block: the eps values of src/cuda_bluebottle.cu:4196-4307; out[f]
 * is the signed increment of face f (W,E,S,N,B,T; 0 where untouched) */
__global__ void k_solv_final(BBGeom G, int out_plane, int nblk,
  const double *__restrict__ part, double *__restrict__ out)
{
  __shared__ double sh[4];
  double a = 0., b = 0., c = 0.;
  for(int q = threadIdx.x; q < nblk; q += blockDim.x) {
    a += part[3*q]; b += part[3*q+1]; c += part[3*q+2];
  }
  double ex = bb_block_sum(a, sh);
  double ey = bb_block_sum(b, sh);
  double ez = bb_block_sum(c, sh);
  if(threadIdx.x != 0) return;
  ex = ex * G.dy * G.dz;
  ey = ey * G.dx * G.dz;
  ez = ez * G.dx * G.dy;
  for(int f = 0; f < 6; f++) out[f] = 0.;
  switch(out_plane) {
    case WEST: case EAST:
      ex = ex / G.yl / G.zl; ey = ey / G.yl / G.zl; ez = ez / G.yl / G.zl;
      if(out_plane == WEST) out[0] = ex + ey + ez; else out[1] = -(ex + ey + ez);
      break;
    case SOUTH: case NORTH:
      ex = ex / G.zl / G.xl; ey = ey / G.zl / G.xl; ez = ez / G.zl / G.xl;
      if(out_plane == SOUTH) out[2] = ex + ey + ez; else out[3] = -(ex + ey + ez);
      break;
    case BOTTOM: case TOP:
      ex = ex / G.xl / G.yl; ey = ey / G.xl / G.yl; ez = ez / G.xl / G.yl;
      if(out_plane == BOTTOM) out[4] = ex + ey + ez; else out[5] = -(ex + ey + ez);
      break;
    case HOMOGENEOUS:
      ex = 0.5 * ex / G.yl / G.zl;
      ey = 0.5 * ey / G.zl / G.xl;
      ez = 0.5 * ez / G.xl / G.yl;
      out[0] = ex; out[1] = -ex; out[2] = ey; out[3] = -ey; out[4] = ez;
      out[5] = -ez;
      break;
  }
}

/* plane_eps_*: adds out[f] on each outflow face (interior transverse
 * points only, src/bluebottle_kernel.cu:2188-2245) */
__global__ void __launch_bounds__(256)
k_solv_apply(BBGeom G, double *__restrict__ us, double *__restrict__ vs,
  double *__restrict__ ws, const double *__restrict__ out)
{
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int nx = G.yn * G.zn, ny = G.xn * G.zn, nz = G.xn * G.yn;
  if(q < nx) {
    int tj = q % G.yn, tk = q / G.yn;
    int C = (tj + 1) * G.us1 + (tk + 1) * G.us2;
    if(out[0] != 0.) us[1 + C] = us[1 + C] + out[0];
    if(out[1] != 0.) us[G.xn + 1 + C] = us[G.xn + 1 + C] + out[1];
  }
  if(q < ny) {
    int ti = q % G.xn, tk = q / G.xn;
    int C = (ti + 1) + (tk + 1) * G.vs2;
    if(out[2] != 0.) vs[C + G.vs1] = vs[C + G.vs1] + out[2];
    if(out[3] != 0.) vs[C + (G.yn + 1) * G.vs1] = vs[C + (G.yn + 1) * G.vs1] + out[3];
  }
  if(q < nz) {
    int ti = q % G.xn, tj = q / G.xn;
    int C = (ti + 1) + (tj + 1) * G.ws1;
    if(out[4] != 0.) ws[C + G.ws2] = ws[C + G.ws2] + out[4];
    if(out[5] != 0.) ws[C + (G.zn + 1) * G.ws2] = ws[C + (G.zn + 1) * G.ws2] + out[5];
  }
}

extern "C" void cuda_solvability(void)
{
  bb_ensure_workspace();
  BBGeom G = bb_geom();
  double *part = g_bb.ws.partials;
  double *out = g_bb.ws.red;
  bb_prof_begin(BB_K_MISC);
  k_solv_partial<<<SOLV_BLOCKS, 256, 0, g_bb.stream>>>(G, _u_star[0],
    _v_star[0], _w_star[0], part);
  BB_CHECK_LAUNCH();
  k_solv_final<<<1, 256, 0, g_bb.stream>>>(G, out_plane, SOLV_BLOCKS, part, out);
  BB_CHECK_LAUNCH();
  int mx = G.yn * G.zn;
  if(G.xn * G.zn > mx) mx = G.xn * G.zn;
  if(G.xn * G.yn > mx) mx = G.xn * G.yn;
  k_solv_apply<<<(mx + 255) / 256, 256, 0, g_bb.stream>>>(G, _u_star[0],
    _v_star[0], _w_star[0], out);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_MISC);
}

/* ---------------------------------------------------------------------
 * Projection: all three components in one launch.
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_project(BBGeom G, double rho_f, double dt,
  const double *__restrict__ us, const double *__restrict__ vs,
  const double *__restrict__ ws, const double *__restrict__ phi,
  const int *__restrict__ fu, const int *__restrict__ fv,
  const int *__restrict__ fw, double *__restrict__ u,
  double *__restrict__ v, double *__restrict__ w)
{
  const int i = blockIdx.x * TX + threadIdx.x + 1;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  const double ddx = 1. / G.dx, ddy = 1. / G.dy, ddz = 1. / G.dz;
  const double dtr = dt / rho_f;
  const int Cc = i + j*G.cs1 + k*G.cs2;
  if(i <= G.xn + 1 && j <= G.yn && k <= G.zn) {
    int C = i + j*G.us1 + k*G.us2;
    double gp = (double) abs(fu[C]) * ddx * (phi[Cc] - phi[Cc-1]);
    u[C] = (us[C] - dtr * gp);
  }
  if(i <= G.xn && j <= G.yn + 1 && k <= G.zn) {
    int C = i + j*G.vs1 + k*G.vs2;
    double gp = (double) abs(fv[C]) * ddy * (phi[Cc] - phi[Cc-G.cs1]);
    v[C] = (vs[C] - dtr * gp);
  }
  if(i <= G.xn && j <= G.yn && k <= G.zn + 1) {
    int C = i + j*G.ws1 + k*G.ws2;
    double gp = (double) abs(fw[C]) * ddz * (phi[Cc] - phi[Cc-G.cs2]);
    w[C] = (ws[C] - dtr * gp);
  }
}

extern "C" void cuda_project(void)
{
  bb_ensure_context();
  BBGeom G = bb_geom();
  dim3 blk(TX, TY, 1);
  dim3 grd((G.xn + 1 + TX - 1) / TX, (G.yn + 1 + TY - 1) / TY, G.zn + 1);
  bb_prof_begin(BB_K_PROJECT);
  k_project<<<grd, blk, 0, g_bb.stream>>>(G, rho_f, dt, _u_star[0],
    _v_star[0], _w_star[0], _phi[0], _flag_u[0], _flag_v[0], _flag_w[0],
    _u[0], _v[0], _w[0]);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_PROJECT);
}

/* ---------------------------------------------------------------------
 * Pressure update with mean removal.
 * ------------------------------------------------------------------- */
#define RED_BLOCKS 1024

__global__ void __launch_bounds__(256)
k_update_p(BBGeom G, const double *__restrict__ p0,
  const double *__restrict__ phi, const int *__restrict__ phase,
  double *__restrict__ p, double *__restrict__ part)
{
  __shared__ double sh[4];
  const int tid = threadIdx.x;
  const long n = (long) G.xn * G.yn * G.zn;
  double s = 0.;
  for(long q = (long) blockIdx.x * 256 + tid; q < n; q += (long) gridDim.x * 256) {
    int i = (int) (q % G.xn), j = (int) ((q / G.xn) % G.yn), k = (int) (q / ((long) G.xn * G.yn));
    int C = (i + 1) + (j + 1) * G.cs1 + (k + 1) * G.cs2;
    double val = (double) (phase[C] < 0) * (p0[C] + phi[C]);
    p[C] = val;
    s += val;
  }
  double t = bb_block_sum(s, sh);
  if(tid == 0) part[blockIdx.x] = t;
}

__global__ void k_sum_final(int nblk, const double *__restrict__ part,
  double *__restrict__ out)
{
  __shared__ double sh[4];
  double a = 0.;
  for(int q = threadIdx.x; q < nblk; q += blockDim.x) a += part[q];
  double t = bb_block_sum(a, sh);
  if(threadIdx.x == 0) out[0] = t;
}

/* forcing_add_c_const over i in [isb, ieb), j in [1, jl), k in [1, kl) */
__global__ void __launch_bounds__(256)
k_add_mean(BBGeom G, int jl, int kl, const double *__restrict__ sum,
  double *__restrict__ p)
{
  const int i = blockIdx.x * TX + threadIdx.x;
  const int j = blockIdx.y * TY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if(i >= G.xn + 2 || j >= jl || k >= kl) return;
  double n = (double) ((long) G.xn * G.yn * G.zn);
  double val = -(sum[0] / n);
  p[i + j * G.cs1 + k * G.cs2] += val;
}

static int c_const_limit(int n, int nb)
{
  int th = n < MAX_THREADS_DIM ? n : MAX_THREADS_DIM;
  int blocks = (int) ceil((double) n / (double) th);
  int lim = blocks * th + DOM_BUF;
  return lim < nb ? lim : nb;
}

extern "C" void cuda_update_p(void)
{
  bb_ensure_workspace();
  BBGeom G = bb_geom();
  double *part = g_bb.ws.partials;
  double *sum = g_bb.ws.red + 8;
  bb_prof_begin(BB_K_UPDATEP);
  k_update_p<<<RED_BLOCKS, 256, 0, g_bb.stream>>>(G, _p0[0], _phi[0],
    _phase[0], _p[0], part);
  BB_CHECK_LAUNCH();
  k_sum_final<<<1, 256, 0, g_bb.stream>>>(RED_BLOCKS, part, sum);
  BB_CHECK_LAUNCH();
  int jl = c_const_limit(dom[0].Gcc.jn, dom[0].Gcc.jnb);
  int kl = c_const_limit(dom[0].Gcc.kn, dom[0].Gcc.knb);
  dim3 blk(TX, TY, 1);
  dim3 grd((G.xn + 2 + TX - 1) / TX, (jl + TY - 1) / TY, kl);
  k_add_mean<<<grd, blk, 0, g_bb.stream>>>(G, jl, kl, sum, _p[0]);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_UPDATEP);
}

/* ---------------------------------------------------------------------
 * Time-step size: signed maxima of u, v, w over the ghosted arrays,
 * starting from 0 (thrust::reduce(..., 0, maximum) in the reference).
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_max3(const double *__restrict__ a, long na, const double *__restrict__ b,
  long nb, const double *__restrict__ c, long nc, double *__restrict__ part)
{
  __shared__ double sh[3][256];
  const int tid = threadIdx.x;
  double ma = 0., mb = 0., mc = 0.;
  long stride = (long) gridDim.x * 256;
  for(long q = (long) blockIdx.x * 256 + tid; q < na; q += stride) ma = fmax(ma, a[q]);
  for(long q = (long) blockIdx.x * 256 + tid; q < nb; q += stride) mb = fmax(mb, b[q]);
  for(long q = (long) blockIdx.x * 256 + tid; q < nc; q += stride) mc = fmax(mc, c[q]);
  sh[0][tid] = ma; sh[1][tid] = mb; sh[2][tid] = mc;
  __syncthreads();
  for(int off = 128; off > 0; off >>= 1) {
    if(tid < off) {
      sh[0][tid] = fmax(sh[0][tid], sh[0][tid + off]);
      sh[1][tid] = fmax(sh[1][tid], sh[1][tid + off]);
      sh[2][tid] = fmax(sh[2][tid], sh[2][tid + off]);
    }
    __syncthreads();
  }
  if(tid == 0) {
    part[3*blockIdx.x] = sh[0][0];
    part[3*blockIdx.x+1] = sh[1][0];
    part[3*blockIdx.x+2] = sh[2][0];
  }
}

__global__ void k_max3_final(int nblk, const double *__restrict__ part,
  double *__restrict__ out)
{
  if(threadIdx.x != 0) return;
  double a = 0., b = 0., c = 0.;
  for(int q = 0; q < nblk; q++) {
    a = fmax(a, part[3*q]); b = fmax(b, part[3*q+1]); c = fmax(c, part[3*q+2]);
  }
  out[0] = a; out[1] = b; out[2] = c;
}

extern "C" real cuda_find_dt(void)
{
  bb_ensure_workspace();
  BBGeom G = bb_geom();
  double *part = g_bb.ws.partials;
  bb_prof_begin(BB_K_FINDDT);
  k_max3<<<RED_BLOCKS, 256, 0, g_bb.stream>>>(_u[0], G.ucount, _v[0],
    G.vcount, _w[0], G.wcount, part);
  BB_CHECK_LAUNCH();
  k_max3_final<<<1, 64, 0, g_bb.stream>>>(RED_BLOCKS, part, g_bb.ws.red + 16);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_FINDDT);
  BB_CHECK(hipMemcpyAsync(g_bb.ws.red_host, g_bb.ws.red + 16, 3 * sizeof(double),
    hipMemcpyDeviceToHost, g_bb.stream));
  BB_CHECK(hipStreamSynchronize(g_bb.stream));
  double u_max = g_bb.ws.red_host[0], v_max = g_bb.ws.red_host[1];
  double w_max = g_bb.ws.red_host[2];
  const dom_struct *d = &dom[0];
  real tmp;
  if(!g_bb.implicit) {
    tmp = u_max / d->dx + 2.*nu/d->dx/d->dx;
    tmp += v_max / d->dy + 2.*nu/d->dy/d->dy;
    tmp += w_max / d->dz + 2.*nu/d->dz/d->dz;
  } else {
    /* Crank-Nicolson diffusion: advective limit only, growth capped */
    tmp = u_max / d->dx;
    tmp += v_max / d->dy;
    tmp += w_max / d->dz;
  }
  real dts = CFL / tmp;
  real mn = FLT_MAX;
  if(dts < mn) mn = dts;
  if(g_bb.implicit && mn > 1.1*dt) mn = 1.1*dt;
  return mn;
}

/* ---------------------------------------------------------------------
 * History shift: device-to-device copies on the library stream.
 * ------------------------------------------------------------------- */
extern "C" void cuda_store_u(void)
{
  bb_ensure_context();
  const dom_struct *d = &dom[0];
  size_t fx = d->Gfx.s3b * sizeof(real), fy = d->Gfy.s3b * sizeof(real);
  size_t fz = d->Gfz.s3b * sizeof(real), cc = d->Gcc.s3b * sizeof(real);
  hipStream_t s = g_bb.stream;
  BB_CHECK(hipMemcpyAsync(_conv0_u[0], _conv_u[0], fx, hipMemcpyDeviceToDevice, s));
  BB_CHECK(hipMemcpyAsync(_conv0_v[0], _conv_v[0], fy, hipMemcpyDeviceToDevice, s));
  BB_CHECK(hipMemcpyAsync(_conv0_w[0], _conv_w[0], fz, hipMemcpyDeviceToDevice, s));
  if(!g_bb.implicit) {
    BB_CHECK(hipMemcpyAsync(_diff0_u[0], _diff_u[0], fx, hipMemcpyDeviceToDevice, s));
    BB_CHECK(hipMemcpyAsync(_diff0_v[0], _diff_v[0], fy, hipMemcpyDeviceToDevice, s));
    BB_CHECK(hipMemcpyAsync(_diff0_w[0], _diff_w[0], fz, hipMemcpyDeviceToDevice, s));
  }
  BB_CHECK(hipMemcpyAsync(_p0[0], _p[0], cc, hipMemcpyDeviceToDevice, s));
  BB_CHECK(hipMemcpyAsync(_u0[0], _u[0], fx, hipMemcpyDeviceToDevice, s));
  BB_CHECK(hipMemcpyAsync(_v0[0], _v[0], fy, hipMemcpyDeviceToDevice, s));
  BB_CHECK(hipMemcpyAsync(_w0[0], _w[0], fz, hipMemcpyDeviceToDevice, s));
}

/* ---------------------------------------------------------------------
 * Forcing: host-side ramps (and PID), then a fill of the three force
 * arrays with the spatially constant -gradP/rho_f.
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_fill3(double *__restrict__ a, long na, double va, double *__restrict__ b,
  long nb, double vb, double *__restrict__ c, long nc, double vc)
{
  long stride = (long) gridDim.x * 256;
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < na; q += stride) a[q] = va;
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < nb; q += stride) b[q] = vb;
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < nc; q += stride) c[q] = vc;
}

static real ramp(real delta, real a, real m)
{
  if(a == 0) return m;
  if(fabs(delta * a) > fabs(m)) return m;
  return delta * a;
}

extern "C" void cuda_compute_forcing(real *pid_int_, real *pid_back_,
  real Kp_, real Ki_, real Kd_)
{
  bb_ensure_context();
  real delta = ttime - p_tDelay;
  int pid = (Kp_ > 0 || Ki_ > 0 || Kd_ > 0);
  if(delta >= 0) {
    gradP.x = ramp(delta, gradP.xa, gradP.xm);
    gradP.y = ramp(delta, gradP.ya, gradP.ym);
    if(!pid) gradP.z = ramp(delta, gradP.za, gradP.zm);
  }
  delta = ttime - g_tDelay;
  if(delta >= 0) {
    g.x = ramp(delta, g.xa, g.xm);
    g.y = ramp(delta, g.ya, g.ym);
    g.z = ramp(delta, g.za, g.zm);
  }
  delta = ttime - p_tDelay;
  if(delta >= 0 && pid) {
    cuda_part_pull();
    real acc_z = 0, volp = 0, massp = 0;
    for(int i = 0; i < nparts; i++) {
      volp += parts[i].r*parts[i].r*parts[i].r;
      massp += parts[i].rho*parts[i].r*parts[i].r*parts[i].r;
      acc_z += parts[i].wdot;
    }
    volp *= 4./3.*PI;
    massp *= 4./3.*PI;
    real volfrac = volp / (Dom.xl * Dom.yl * Dom.zl);
    real rho_avg = massp/volp*volfrac + rho_f*(1.-volfrac);
    acc_z /= nparts;
    *pid_int_ = *pid_int_ + acc_z*dt;
    gradP.z = gradP.z
      + (Kp_*acc_z + Ki_*(*pid_int_)/ttime + (Kd_)*(acc_z-*pid_back_))*rho_avg;
    *pid_back_ = acc_z;
  }
  BBGeom G = bb_geom();
  bb_prof_begin(BB_K_MISC);
  k_fill3<<<RED_BLOCKS, 256, 0, g_bb.stream>>>(_f_x[0], G.ucount,
    0. + -gradP.x / rho_f, _f_y[0], G.vcount, 0. + -gradP.y / rho_f,
    _f_z[0], G.wcount, 0. + -gradP.z / rho_f);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_MISC);
}
