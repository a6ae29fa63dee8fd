/*
 * bb_helmholtz.hip -- Crank-Nicolson velocity solves of the reference's
 * IMPLICIT build: cuda_ustar_helmholtz, cuda_vstar_helmholtz,
 * cuda_wstar_helmholtz (src/cuda_bicgstab.cu:44-714).
 *
 * For one velocity component:
 *   rhs  = u0 + dt [ -dp0/dx / rho_f - (AB2 or Euler) conv + nu/2 lap(u0) + f ]
 *          (src/bicgstab_kernel.cu:27-482; conv is stored, no solid mask),
 *   then the domain and particle BCs of u*, then
 *   (I - nu dt/2 L) u* = rhs   by diagonally preconditioned CG on the
 *   interior faces, normalised by ||rhs||, from a zero start, to
 *   pp_residual (src/cuda_bicgstab.cu:60-307).
 *
 * The reference assembles a 13-diagonal CUSP matrix every call; here the
 * operator is matrix-free.  A set-up pass packs each row's six |face flag|
 * bits and its solid-face bit into one byte; the product kernel rebuilds
 * the row's thirteen coefficients from that byte by replaying the
 * reference's assembly (interior coefficients, then the W, E, S, N, B, T
 * boundary kernels in that order, then identity rows on faces inside
 * particles) and applies every non-zero coefficient at the flattened index
 * plus the diagonal's offset when that lands inside the vector, as the DIA
 * product does -- including the reference's wrap couplings and its
 * couplings across non-periodic, non-Dirichlet faces.
 *
 * CG per iteration, three sweeps (dots reduced deterministically, scalars
 * on the device, host polls every few iterations):
 *   H1  p = z + beta p (on the fly at the stencil points), y = A p, <y, p>
 *   H2  x += alpha p, r -= alpha y, z = M r, <r, z>, <r, r>
 * with alpha, beta and the convergence test in one-block finishing kernels.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "bb_device.h"

#define HX 64
#define HY 4
#define HZ 16

/* ---------------------------------------------------------------------
 * Right-hand side, one component (reference operation order)
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_star_rhs(BBGeom G, int comp, double rho_f, double nu, double dt0, double dt,
  const double *__restrict__ u0, const double *__restrict__ v0,
  const double *__restrict__ w0, const double *__restrict__ p,
  const double *__restrict__ f, const double *__restrict__ conv0,
  double *__restrict__ conv, double *__restrict__ star)
{
  const int i = blockIdx.x * HX + threadIdx.x + 1;
  const int j = blockIdx.y * HY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  const double ab0 = 0.5 * dt / dt0;
  const double ab = 1. + ab0;
  const double ddx = 1. / G.dx, ddy = 1. / G.dy, ddz = 1. / G.dz;
  const int cs1 = G.cs1, cs2 = G.cs2;
  const int us1 = G.us1, us2 = G.us2, vs1 = G.vs1, vs2 = G.vs2;
  const int ws1 = G.ws1, ws2 = G.ws2;
  const int Cc = i + j*cs1 + k*cs2;
  if(comp == 0) {
    if(!(i <= G.xn + 1 && j <= G.yn && k <= G.zn)) return;
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
    if(dt0 > 0) us += (-ab * c + ab0 * conv0[C]);
    else us += -c;
    double d1 = (u211 - u111) * ddx, d0 = (u111 - u011) * ddx;
    double dxx = (d1 - d0) * ddx;
    d1 = (u121 - u111) * ddy; d0 = (u111 - u101) * ddy;
    double dyy = (d1 - d0) * ddy;
    d1 = (u112 - u111) * ddz; d0 = (u111 - u110) * ddz;
    double dzz = (d1 - d0) * ddz;
    double df = nu * (dxx + dyy + dzz);
    us += 0.5*df;
    us += f[C];
    us *= dt;
    us += u111;
    star[C] = us;
    conv[C] = c;
  } else if(comp == 1) {
    if(!(i <= G.xn && j <= G.yn + 1 && k <= G.zn)) return;
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
    if(dt0 > 0) vs += (-ab * c + ab0 * conv0[C]);
    else vs += -c;
    double d1 = (v211 - v111) * ddx, d0 = (v111 - v011) * ddx;
    double dxx = (d1 - d0) * ddx;
    d1 = (v121 - v111) * ddy; d0 = (v111 - v101) * ddy;
    double dyy = (d1 - d0) * ddy;
    d1 = (v112 - v111) * ddz; d0 = (v111 - v110) * ddz;
    double dzz = (d1 - d0) * ddz;
    double df = nu * (dxx + dyy + dzz);
    vs += 0.5*df;
    vs += f[C];
    vs *= dt;
    vs += v111;
    star[C] = vs;
    conv[C] = c;
  } else {
    if(!(i <= G.xn && j <= G.yn && k <= G.zn + 1)) return;
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
    if(dt0 > 0) ws += (-ab * c + ab0 * conv0[C]);
    else ws += -c;
    double d1 = (w211 - w111) * ddx, d0 = (w111 - w011) * ddx;
    double dxx = (d1 - d0) * ddx;
    d1 = (w121 - w111) * ddy; d0 = (w111 - w101) * ddy;
    double dyy = (d1 - d0) * ddy;
    d1 = (w112 - w111) * ddz; d0 = (w111 - w110) * ddz;
    double dzz = (d1 - d0) * ddz;
    double df = nu * (dxx + dyy + dzz);
    ws += 0.5*df;
    ws += f[C];
    ws *= dt;
    ws += w111;
    star[C] = ws;
    conv[C] = c;
  }
}

/* ---------------------------------------------------------------------
 * Operator geometry of one component's unknowns: its interior faces
 * (ni x nj x nk, x fastest, no ghosts) and the thirteen DIA offsets of
 * src/cuda_bicgstab.cu:62-74, 342-355, 540-553 (the wrap partner along the
 * component's own axis skips the duplicated periodic face).
 * ------------------------------------------------------------------- */
struct HGeom {
  int comp, ni, nj, nk;
  int s1, s2, n;
  int off[13];
  int tx, ty, tz, nblk;
  int bcl[3], bch[3];      /* BC type at the low / high face of each axis */
  double c[3];             /* 0.5 nu dt / h^2, the reference's product order */
  int us1, us2, vs1, vs2, ws1, ws2;   /* ghosted strides of the face grids */
  int gs1, gs2;            /* ghosted strides of this component's grid */
};

static HGeom h_geom(int comp)
{
  const dom_struct *d = &dom[0];
  const grid_info *Gr = comp == 0 ? &d->Gfx : comp == 1 ? &d->Gfy : &d->Gfz;
  HGeom H;
  H.comp = comp;
  H.ni = Gr->in; H.nj = Gr->jn; H.nk = Gr->kn;
  H.s1 = Gr->s1; H.s2 = Gr->s2; H.n = Gr->s3;
  const int s1 = H.s1, s2 = H.s2, s3 = H.n;
  const int wx = comp == 0 ? 2 : 1, wy = comp == 1 ? 2 : 1, wz = comp == 2 ? 2 : 1;
  int off[13] = {-s3 + wz*s2, -s2, -s2 + wy*s1, -s1, -s1 + wx, -1, 0, 1,
    s1 - wx, s1, s2 - wy*s1, s2, s3 - wz*s2};
  for(int q = 0; q < 13; q++) H.off[q] = off[q];
  H.tx = (H.ni + HX - 1) / HX;
  H.ty = (H.nj + HY - 1) / HY;
  H.tz = (H.nk + HZ - 1) / HZ;
  H.nblk = H.tx * H.ty * H.tz;
  const BC *b = &bc;
  const int tu[6] = {b->uW, b->uE, b->uS, b->uN, b->uB, b->uT};
  const int tv[6] = {b->vW, b->vE, b->vS, b->vN, b->vB, b->vT};
  const int tw[6] = {b->wW, b->wE, b->wS, b->wN, b->wB, b->wT};
  const int *t = comp == 0 ? tu : comp == 1 ? tv : tw;
  for(int a = 0; a < 3; a++) { H.bcl[a] = t[2*a]; H.bch[a] = t[2*a+1]; }
  const double ddx = 1. / (d->dx * d->dx), ddy = 1. / (d->dy * d->dy);
  const double ddz = 1. / (d->dz * d->dz);
  H.c[0] = 0.5*nu*dt*ddx;
  H.c[1] = 0.5*nu*dt*ddy;
  H.c[2] = 0.5*nu*dt*ddz;
  H.us1 = d->Gfx.s1b; H.us2 = d->Gfx.s2b;
  H.vs1 = d->Gfy.s1b; H.vs2 = d->Gfy.s2b;
  H.ws1 = d->Gfz.s1b; H.ws2 = d->Gfz.s2b;
  H.gs1 = Gr->s1b; H.gs2 = Gr->s2b;
  return H;
}

/* Row flags: bits 0..5 = |flag| of the W, E, S, N, B, T faces the
 * reference's coefficient kernel reads, bit 6 = this face inside a
 * particle (flag < 0: identity row). */
enum { HF_W = 1, HF_E = 2, HF_S = 4, HF_N = 8, HF_B = 16, HF_T = 32, HF_ID = 64 };

__device__ inline void h_identity(double v[13])
{
  for(int q = 0; q < 13; q++) v[q] = 0.;
  v[6] = 1.;
}

/* The thirteen coefficients of row (i, j, k), replaying the reference's
 * assembly order (src/bicgstab_kernel.cu:643-1652). */
__device__ inline void h_row(const HGeom &H, uint8_t code, int i, int j, int k,
  double v[13])
{
  if(code & HF_ID) { h_identity(v); return; }
  const double c0 = H.c[0], c1 = H.c[1], c2 = H.c[2];
  for(int q = 0; q < 13; q++) v[q] = 0.;
  v[1] = (code & HF_B) ? -c2 : 0.;
  v[3] = (code & HF_S) ? -c1 : 0.;
  v[5] = (code & HF_W) ? -c0 : 0.;
  double dg = 1.;
  if(code & HF_W) dg += c0;
  if(code & HF_E) dg += c0;
  if(code & HF_S) dg += c1;
  if(code & HF_N) dg += c1;
  if(code & HF_B) dg += c2;
  if(code & HF_T) dg += c2;
  v[6] = dg;
  v[7] = (code & HF_E) ? -c0 : 0.;
  v[9] = (code & HF_N) ? -c1 : 0.;
  v[11] = (code & HF_T) ? -c2 : 0.;
  /* W, E, S, N, B, T boundary kernels */
  const int pos[3] = {i, j, k};
  const int ext[3] = {H.ni, H.nj, H.nk};
  const int lo_in[3] = {5, 3, 1}, hi_in[3] = {7, 9, 11};
  const int lo_wrap[3] = {4, 2, 0}, hi_wrap[3] = {8, 10, 12};
#pragma unroll
  for(int a = 0; a < 3; a++) {
    const double ca = H.c[a];
    if(pos[a] == 0) {
      if(H.bcl[a] == PERIODIC) { v[lo_in[a]] += ca; v[hi_wrap[a]] -= ca; }
      else if(H.bcl[a] == DIRICHLET) {
        if(a == H.comp) h_identity(v); else v[lo_in[a]] = 0.;
      }
    }
    if(pos[a] == ext[a] - 1) {
      if(H.bch[a] == PERIODIC) { v[lo_wrap[a]] -= ca; v[hi_in[a]] += ca; }
      else if(H.bch[a] == DIRICHLET) {
        if(a == H.comp) h_identity(v); else v[hi_in[a]] = 0.;
      }
    }
  }
}

struct HTile { int i, j, k0, k1; };

__device__ inline HTile h_tile(const HGeom &H, int lb)
{
  HTile t;
  int tx = lb % H.tx, rest = lb / H.tx;
  int ty = rest % H.ty, tz = rest / H.ty;
  t.i = tx * HX + threadIdx.x;
  t.j = ty * HY + threadIdx.y;
  t.k0 = tz * HZ;
  t.k1 = min(t.k0 + HZ, H.nk);
  return t;
}

/* XCD-contiguous tile ranges (blocks are dealt round-robin to 8 XCDs) */
__device__ inline int h_block(int nblk)
{
  int b = blockIdx.x;
  if(nblk % 8 == 0) return (b % 8) * (nblk / 8) + b / 8;
  return b;
}

/* Set-up: row flags, b = interior of u*, partial ||b||^2 */
__global__ void __launch_bounds__(256)
k_h_setup(HGeom H, const int *__restrict__ fu, const int *__restrict__ fv,
  const int *__restrict__ fw, const double *__restrict__ star,
  uint8_t *__restrict__ code, double *__restrict__ b,
  double *__restrict__ part)
{
  __shared__ double sh[4];
  const int lb = h_block(H.nblk);
  HTile T = h_tile(H, lb);
  const int *own = H.comp == 0 ? fu : H.comp == 1 ? fv : fw;
  double s2 = 0.;
  if(T.i < H.ni && T.j < H.nj) {
    for(int k = T.k0; k < T.k1; k++) {
      const int I = T.i + DOM_BUF, J = T.j + DOM_BUF, K = k + DOM_BUF;
      uint8_t c = 0;
      c |= fu[I + J*H.us1 + K*H.us2] ? HF_W : 0;
      c |= fu[(I+1) + J*H.us1 + K*H.us2] ? HF_E : 0;
      c |= fv[I + J*H.vs1 + K*H.vs2] ? HF_S : 0;
      c |= fv[I + (J+1)*H.vs1 + K*H.vs2] ? HF_N : 0;
      c |= fw[I + J*H.ws1 + K*H.ws2] ? HF_B : 0;
      c |= fw[I + J*H.ws1 + (K+1)*H.ws2] ? HF_T : 0;
      c |= own[I + J*H.gs1 + K*H.gs2] < 0 ? HF_ID : 0;
      const int C = T.i + T.j*H.s1 + k*H.s2;
      code[C] = c;
      double bv = star[I + J*H.gs1 + K*H.gs2];
      b[C] = bv;
      s2 += bv * bv;
    }
  }
  double t = bb_block_sum(s2, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) part[lb] = t;
}

/* CG scalars on the device */
struct CGScalars {
  double rz, alpha, beta, tol, resid, bnorm;
  int iter, max_iter, done, converged;
};

__device__ inline void h_sum(const double *part, int nblk, int nq, double *out)
{
  __shared__ double sh[4];
  for(int q = 0; q < nq; q++) {
    double a = 0.;
    for(int b = threadIdx.x; b < nblk; b += blockDim.x) a += part[b*nq + q];
    out[q] = bb_block_sum(a, sh);
  }
}

__global__ void k_h_norm(int nblk, const double *part, CGScalars *sc)
{
  double o[1];
  h_sum(part, nblk, 1, o);
  if(threadIdx.x == 0) {
    double nrm = sqrt(o[0]);
    if(nrm == 0) nrm = 1.;
    sc->bnorm = nrm;
  }
}

/* b /= ||b||, x = 0, r = b, z = M r, p = z; partials <r,z>, <r,r>, <b,b>.
 * The first product of the loop forms p = z + 0 p, so p needs no zeroing
 * of stale contents beyond this copy. */
__global__ void __launch_bounds__(256)
k_h_init(HGeom H, const CGScalars *sc, const uint8_t *__restrict__ code,
  double *__restrict__ b, double *__restrict__ x, double *__restrict__ r,
  double *__restrict__ z, double *__restrict__ p, double *__restrict__ part)
{
  __shared__ double sh[4];
  const int lb = h_block(H.nblk);
  HTile T = h_tile(H, lb);
  const double inv = 1. / sc->bnorm;
  double rz = 0., rr = 0., bb = 0.;
  if(T.i < H.ni && T.j < H.nj) {
    for(int k = T.k0; k < T.k1; k++) {
      const int C = T.i + T.j*H.s1 + k*H.s2;
      double v[13];
      h_row(H, code[C], T.i, T.j, k, v);
      double bv = b[C] * inv;
      b[C] = bv;
      x[C] = 0.;
      double rv = 1. * bv + -1. * 0.;
      double zv = (1. / v[6]) * rv;
      r[C] = rv; z[C] = zv; p[C] = zv;
      rz += rv * zv;
      rr += rv * rv;
      bb += bv * bv;
    }
  }
  double a0 = bb_block_sum(rz, sh);
  double a1 = bb_block_sum(rr, sh);
  double a2 = bb_block_sum(bb, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) {
    part[3*lb] = a0; part[3*lb+1] = a1; part[3*lb+2] = a2;
  }
}

__global__ void k_h_f0(int nblk, const double *part, CGScalars *sc,
  double rel_tol, int max_iter)
{
  double o[3];
  h_sum(part, nblk, 3, o);
  if(threadIdx.x == 0) {
    sc->rz = o[0];
    sc->resid = sqrt(o[1]);
    sc->tol = rel_tol * sqrt(o[2]);
    sc->beta = 0.;
    sc->alpha = 0.;
    sc->iter = 0;
    sc->max_iter = max_iter;
    sc->converged = sc->resid <= sc->tol;
    sc->done = sc->converged || max_iter <= 0;
  }
}

/* H1: p = z + beta p at the stencil points, y = A p, partial <y, p> */
__global__ void __launch_bounds__(256)
k_h_spmv(HGeom H, const CGScalars *sc, const uint8_t *__restrict__ code,
  const double *__restrict__ z, double *__restrict__ p,
  double *__restrict__ pn, double *__restrict__ y, double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[4];
  const int lb = h_block(H.nblk);
  HTile T = h_tile(H, lb);
  const double beta = sc->beta;
  double yp = 0.;
  if(T.i < H.ni && T.j < H.nj) {
    for(int k = T.k0; k < T.k1; k++) {
      const int C = T.i + T.j*H.s1 + k*H.s2;
      double v[13];
      h_row(H, code[C], T.i, T.j, k, v);
      double sum = 0.;
#pragma unroll
      for(int q = 0; q < 13; q++) {
        const int col = C + H.off[q];
        if(v[q] != 0. && col >= 0 && col < H.n)
          sum += v[q] * (1. * z[col] + beta * p[col]);
      }
      const double pc = 1. * z[C] + beta * p[C];
      pn[C] = pc;
      y[C] = sum;
      yp += sum * pc;
    }
  }
  double t = bb_block_sum(yp, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) part[lb] = t;
}

__global__ void k_h_f1(int nblk, const double *part, CGScalars *sc)
{
  if(sc->done) return;
  double o[1];
  h_sum(part, nblk, 1, o);
  if(threadIdx.x == 0) sc->alpha = sc->rz / o[0];
}

/* H2: x += alpha p, r -= alpha y, z = M r, partials <r,z>, <r,r> */
__global__ void __launch_bounds__(256)
k_h_update(HGeom H, const CGScalars *sc, const uint8_t *__restrict__ code,
  const double *__restrict__ p, const double *__restrict__ y,
  double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
  double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[4];
  const int lb = h_block(H.nblk);
  HTile T = h_tile(H, lb);
  const double alpha = sc->alpha;
  double rz = 0., rr = 0.;
  if(T.i < H.ni && T.j < H.nj) {
    for(int k = T.k0; k < T.k1; k++) {
      const int C = T.i + T.j*H.s1 + k*H.s2;
      double v[13];
      h_row(H, code[C], T.i, T.j, k, v);
      x[C] = x[C] + alpha * p[C];
      const double rv = r[C] + -alpha * y[C];
      r[C] = rv;
      const double zv = (1. / v[6]) * rv;
      z[C] = zv;
      rz += rv * zv;
      rr += rv * rv;
    }
  }
  double a0 = bb_block_sum(rz, sh);
  double a1 = bb_block_sum(rr, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) { part[2*lb] = a0; part[2*lb+1] = a1; }
}

__global__ void k_h_f2(int nblk, const double *part, CGScalars *sc)
{
  if(sc->done) return;
  double o[2];
  h_sum(part, nblk, 2, o);
  if(threadIdx.x == 0) {
    double rz_old = sc->rz;
    sc->rz = o[0];
    sc->beta = o[0] / rz_old;
    sc->iter += 1;
    sc->resid = sqrt(o[1]);
    sc->converged = sc->resid <= sc->tol;
    sc->done = sc->converged || sc->iter >= sc->max_iter;
  }
}

/* u* interior = x * ||b|| */
__global__ void __launch_bounds__(256)
k_h_copy_back(HGeom H, const CGScalars *sc, const double *__restrict__ x,
  double *__restrict__ star)
{
  HTile T = h_tile(H, blockIdx.x);
  if(T.i >= H.ni || T.j >= H.nj) return;
  const double nrm = sc->bnorm;
  for(int k = T.k0; k < T.k1; k++) {
    const int C = T.i + T.j*H.s1 + k*H.s2;
    star[(T.i + DOM_BUF) + (T.j + DOM_BUF)*H.gs1 + (k + DOM_BUF)*H.gs2] = x[C] * nrm;
  }
}

/* ---------------------------------------------------------------------
 * Host driver
 * ------------------------------------------------------------------- */
struct HWork {
  int n;
  double *b, *x, *r, *z, *p, *pn, *y, *part;
  uint8_t *code;
  CGScalars *sc, *sc_host;
};
static HWork g_h;

static void h_ensure(int n, int nblk)
{
  if(g_h.n >= n && g_h.b) return;
  if(g_h.b) {
    BB_CHECK(hipStreamSynchronize(g_bb.stream));
    void *ptrs[] = {g_h.b, g_h.x, g_h.r, g_h.z, g_h.p, g_h.pn, g_h.y, g_h.part,
      g_h.code, g_h.sc};
    for(size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); q++)
      BB_CHECK(hipFree(ptrs[q]));
  }
  size_t bytes = (size_t) n * sizeof(double);
  BB_CHECK(hipMalloc(&g_h.b, bytes));
  BB_CHECK(hipMalloc(&g_h.x, bytes));
  BB_CHECK(hipMalloc(&g_h.r, bytes));
  BB_CHECK(hipMalloc(&g_h.z, bytes));
  BB_CHECK(hipMalloc(&g_h.p, bytes));
  BB_CHECK(hipMalloc(&g_h.pn, bytes));
  BB_CHECK(hipMalloc(&g_h.y, bytes));
  BB_CHECK(hipMalloc(&g_h.code, (size_t) n));
  int slots = nblk > 4096 ? nblk : 4096;
  BB_CHECK(hipMalloc(&g_h.part, (size_t) slots * 3 * sizeof(double)));
  BB_CHECK(hipMalloc(&g_h.sc, sizeof(CGScalars)));
  if(!g_h.sc_host)
    BB_CHECK(hipHostMalloc((void **) &g_h.sc_host, sizeof(CGScalars),
      hipHostMallocDefault));
  g_h.n = n;
}

/* frees the solver workspace (called from cuda_dom_free) */
void bb_free_helmholtz(void)
{
  if(!g_h.b) return;
  void *ptrs[] = {g_h.b, g_h.x, g_h.r, g_h.z, g_h.p, g_h.pn, g_h.y, g_h.part,
    g_h.code, g_h.sc};
  for(size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); q++) BB_CHECK(hipFree(ptrs[q]));
  if(g_h.sc_host) BB_CHECK(hipHostFree(g_h.sc_host));
  memset(&g_h, 0, sizeof(g_h));
}

static int g_h_poll = 4;

static void star_helmholtz(int comp, int rank)
{
  bb_ensure_context();
  hipStream_t st = g_bb.stream;
  BBGeom G = bb_geom();
  real **star_p = comp == 0 ? _u_star : comp == 1 ? _v_star : _w_star;
  real *star = star_p[0];
  real *f = comp == 0 ? _f_x[0] : comp == 1 ? _f_y[0] : _f_z[0];
  real *conv0 = comp == 0 ? _conv0_u[0] : comp == 1 ? _conv0_v[0] : _conv0_w[0];
  real *conv = comp == 0 ? _conv_u[0] : comp == 1 ? _conv_v[0] : _conv_w[0];

  /* right-hand side, then the domain and particle BCs of u* */
  dim3 blk(HX, HY, 1);
  dim3 grd((G.xn + 1 + HX - 1) / HX, (G.yn + 1 + HY - 1) / HY, G.zn + 1);
  bb_prof_begin(BB_K_USTAR);
  k_star_rhs<<<grd, blk, 0, st>>>(G, comp, rho_f, nu, dt0, dt, _u0[0], _v0[0],
    _w0[0], _p0[0], f, conv0, conv, star);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_USTAR);
  cuda_dom_BC_star();
  if(nparts > 0) cuda_part_BC_star();

  HGeom H = h_geom(comp);
  h_ensure(H.n, H.nblk);
  dim3 hb(HX, HY);
  k_h_setup<<<H.nblk, hb, 0, st>>>(H, _flag_u[0], _flag_v[0], _flag_w[0], star,
    g_h.code, g_h.b, g_h.part);
  BB_CHECK_LAUNCH();
  k_h_norm<<<1, 256, 0, st>>>(H.nblk, g_h.part, g_h.sc);
  BB_CHECK_LAUNCH();
  k_h_init<<<H.nblk, hb, 0, st>>>(H, g_h.sc, g_h.code, g_h.b, g_h.x, g_h.r,
    g_h.z, g_h.p, g_h.part);
  BB_CHECK_LAUNCH();
  k_h_f0<<<1, 256, 0, st>>>(H.nblk, g_h.part, g_h.sc, pp_residual, pp_max_iter);
  BB_CHECK_LAUNCH();
  /* p and pn alternate: H1 reads p (old) and writes pn (new) */
  double *pa = g_h.p, *pb = g_h.pn;
  for(int it = 0; it < pp_max_iter + 1; it++) {
    bb_prof_begin(BB_K_MISC);
    k_h_spmv<<<H.nblk, hb, 0, st>>>(H, g_h.sc, g_h.code, g_h.z, pa, pb, g_h.y,
      g_h.part);
    BB_CHECK_LAUNCH();
    k_h_f1<<<1, 256, 0, st>>>(H.nblk, g_h.part, g_h.sc);
    BB_CHECK_LAUNCH();
    k_h_update<<<H.nblk, hb, 0, st>>>(H, g_h.sc, g_h.code, pb, g_h.y, g_h.x,
      g_h.r, g_h.z, g_h.part);
    BB_CHECK_LAUNCH();
    k_h_f2<<<1, 256, 0, st>>>(H.nblk, g_h.part, g_h.sc);
    BB_CHECK_LAUNCH();
    bb_prof_end(BB_K_MISC);
    double *tmp = pa; pa = pb; pb = tmp;
    if((it + 1) % g_h_poll == 0 || it == pp_max_iter) {
      BB_CHECK(hipMemcpyAsync(g_h.sc_host, g_h.sc, sizeof(CGScalars),
        hipMemcpyDeviceToHost, st));
      BB_CHECK(hipStreamSynchronize(st));
      if(g_h.sc_host->done) break;
    }
  }
  BB_CHECK(hipMemcpyAsync(g_h.sc_host, g_h.sc, sizeof(CGScalars),
    hipMemcpyDeviceToHost, st));
  BB_CHECK(hipStreamSynchronize(st));

  g_bb.stats.iterations = g_h.sc_host->iter;
  g_bb.stats.outer_steps = 1;
  g_bb.stats.residual = g_h.sc_host->resid;
  g_bb.stats.converged = g_h.sc_host->converged;
  char nam[FILE_NAME_SIZE] = "solver_helmholtz_expd.rec";
  char namp[FILE_NAME_SIZE] = "solver_helmholtz_prec.rec";
  recorder_bicgstab(rank == 0 ? nam : namp, g_h.sc_host->iter,
    g_h.sc_host->resid);
  if(!g_h.sc_host->converged) {
    const char *which = comp == 0 ? "u_star" : comp == 1 ? "v_star" : "w_star";
    printf("The %s Helmholtz equation did not converge.              \n", which);
    fflush(stdout);
    if(g_bb.exit_on_fail) exit(EXIT_FAILURE);
  }
  k_h_copy_back<<<H.nblk, hb, 0, st>>>(H, g_h.sc, g_h.x, star);
  BB_CHECK_LAUNCH();
}

extern "C" void cuda_ustar_helmholtz(int rank) { star_helmholtz(0, rank); }
extern "C" void cuda_vstar_helmholtz(int rank) { star_helmholtz(1, rank); }
extern "C" void cuda_wstar_helmholtz(int rank) { star_helmholtz(2, rank); }
