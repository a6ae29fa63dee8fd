/*
 * bb_poisson.hip -- the pressure-Poisson solve, cuda_PP_bicgstab
 * (src/cuda_bicgstab.cu:715-1073).
 *
 * The reference assembles a 13-diagonal CUSP DIA matrix every call
 * (13 x 8 B = 104 B/cell, src/bicgstab_kernel.cu:1654-1856) and runs
 * cusp::krylov::bicgstab with a diagonal preconditioner, paying a
 * host-synchronous reduction per dot product.  Here the operator is
 * matrix-free: one byte per cell packs the six |face flag| bits and the
 * solid bit, and each off-diagonal is that bit times 1/h^2 -- exactly the
 * values the DIA assembly stores, including the periodic wrap diagonals
 * (offsets -s3+s2, -s2+s1, -s1+1, s1-1, s2-s1, s3-s2 of
 * src/cuda_bicgstab.cu:735-747 become index wraps) and the identity rows
 * of solid cells.  Jacobi scaling uses a 128-entry table of 1/diag
 * computed on the host with the reference's diagonal arithmetic.
 *
 * Iteration structure (per BiCGSTAB step, CUSP semantics):
 *   K1  v = A (M p),              partial <r*, v>
 *   F1  alpha = rho / <r*, v>
 *   K2  s = r - alpha v (on the fly), t = A (M s),  partials <t,s>,<t,t>,<s,s>
 *   F2  ||s|| test, omega = <t,s>/<t,t>
 *   K3  x += alpha M p + omega M s,  r = s - omega t,  partials <r*,r>,<r,r>
 *   F3  beta, rho, ||r|| test, iteration count
 *   K4  p = r + beta (p - omega v)
 * Dots are reduced deterministically (per-block partials, then one
 * finishing block); the scalars stay on the device and the host polls the
 * `done` flag every few iterations.
 *
 * Mixed mode (bb_set_solver_mode(BB_SOLVER_MIXED)): FP64 iterative
 * refinement around the same BiCGSTAB run in FP32 (FP32 vectors, FP64
 * accumulation of FP32 products in every dot).  Convergence is judged on
 * the true FP64 residual ||b - A x|| <= pp_residual * ||b||.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "bb_device.h"
#include "bb_unet.h"

#define PX 64
#define PY 4
#define PZ 16

/* The two SpMV sweeps march in z over SX x SY columns (one 512-thread
 * block per column tile and z chunk of SZ planes). */
#define SX 64
#define SY 8
#define SZ 32

struct PPGeom {
  int nx, ny, nz;
  int tx, ty, tz;          /* tiles of PX x PY x PZ */
  int perx, pery, perz;
  double ddx, ddy, ddz;    /* 1/h^2 */
  int nblk;
  int sx, sy, sz;          /* SpMV tiles of SX x SY x SZ */
  int snblk;
};

struct DinvTable { double d[128]; double g[128]; };

static DinvTable make_dinv(double ddx, double ddy, double ddz)
{
  DinvTable t;
  for(int f = 0; f < 128; f++) {
    if(f & BB_SOLID) { t.d[f] = 1.; t.g[f] = 1.; continue; }
    int fx = ((f & BB_FW) ? 1 : 0) + ((f & BB_FE) ? 1 : 0);
    int fy = ((f & BB_FS) ? 1 : 0) + ((f & BB_FN) ? 1 : 0);
    int fz = ((f & BB_FB) ? 1 : 0) + ((f & BB_FT) ? 1 : 0);
    /* values[6] of src/bicgstab_kernel.cu:1710-1712 */
    double dg = 0.;
    dg -= (double) fx * ddx;
    dg -= (double) fy * ddy;
    dg -= (double) fz * ddz;
    t.g[f] = dg;
    t.d[f] = 1. / dg;
  }
  return t;
}

static PPGeom pp_geom(void)
{
  PPGeom g;
  const dom_struct *d = &dom[0];
  g.nx = d->xn; g.ny = d->yn; g.nz = d->zn;
  g.tx = (g.nx + PX - 1) / PX;
  g.ty = (g.ny + PY - 1) / PY;
  g.tz = (g.nz + PZ - 1) / PZ;
  g.perx = bc.pW == PERIODIC;
  g.pery = bc.pS == PERIODIC;
  g.perz = bc.pB == PERIODIC;
  g.ddx = 1. / (d->dx * d->dx);
  g.ddy = 1. / (d->dy * d->dy);
  g.ddz = 1. / (d->dz * d->dz);
  g.nblk = g.tx * g.ty * g.tz;
  g.sx = (g.nx + SX - 1) / SX;
  g.sy = (g.ny + SY - 1) / SY;
  g.sz = (g.nz + SZ - 1) / SZ;
  g.snblk = g.sx * g.sy * g.sz;
  return g;
}

/* Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD
 * works on a contiguous range of tiles (its own L2 then holds the y and z
 * halos of neighbouring tiles). */
__device__ inline int xcd_block(int nblk)
{
  int b = blockIdx.x;
  if(nblk % 8 == 0) return (b % 8) * (nblk / 8) + b / 8;
  return b;
}

struct TileIdx { int i, j, k0, k1; };

__device__ inline TileIdx tile_of(const PPGeom &G, int lb)
{
  TileIdx t;
  int tx = lb % G.tx;
  int rest = lb / G.tx;
  int ty = rest % G.ty;
  int tz = rest / G.ty;
  t.i = tx * PX + threadIdx.x;
  t.j = ty * PY + threadIdx.y;
  t.k0 = tz * PZ;
  t.k1 = min(t.k0 + PZ, G.nz);
  return t;
}

/* Neighbour offsets of cell (i,j,k) honouring the periodic wraps of the
 * DIA matrix; non-periodic boundary neighbours carry a zero coefficient
 * (their face flag is 0), so any in-range index serves. */
struct Nbr { int w, e, s, n, b, t; };

__device__ inline Nbr neighbours(const PPGeom &G, int i, int j, int k)
{
  const int s1 = G.nx, s2 = G.nx * G.ny;
  Nbr o;
  o.w = (i > 0) ? -1 : (G.perx ? G.nx - 1 : 0);
  o.e = (i < G.nx - 1) ? 1 : (G.perx ? -(G.nx - 1) : 0);
  o.s = (j > 0) ? -s1 : (G.pery ? (G.ny - 1) * s1 : 0);
  o.n = (j < G.ny - 1) ? s1 : (G.pery ? -(G.ny - 1) * s1 : 0);
  o.b = (k > 0) ? -s2 : (G.perz ? (G.nz - 1) * s2 : 0);
  o.t = (k < G.nz - 1) ? s2 : (G.perz ? -(G.nz - 1) * s2 : 0);
  return o;
}

/* (A y)_C for the row of cell C given y at the centre and six neighbours,
 * in the DIA summation order (diagonals 0..12). */
__device__ inline double apply_row(uint8_t f, const PPGeom &G, double yc,
  double yw, double ye, double ys, double yn, double yb, double yt,
  double dg)
{
  if(f & BB_SOLID) return yc;
  double a = 0.;
  a += (f & BB_FB) ? G.ddz * yb : 0.;
  a += (f & BB_FS) ? G.ddy * ys : 0.;
  a += (f & BB_FW) ? G.ddx * yw : 0.;
  a += dg * yc;
  a += (f & BB_FE) ? G.ddx * ye : 0.;
  a += (f & BB_FN) ? G.ddy * yn : 0.;
  a += (f & BB_FT) ? G.ddz * yt : 0.;
  return a;
}

/* diag entry from the flag byte (the diagonal table sits right after the
 * 128 reciprocals in LDS) */
__device__ inline double diag_of(const double *dinv, uint8_t f)
{
  return dinv[128 + (f & 127)];
}

/* ---------------------------------------------------------------------
 * RHS + flag packing + warm start (one pass over the fields).
 *   rhs = rho_f/dt * div(u*)                (PP_rhs)
 *   rhs *= (phase < 0 && phase_shell)       (part_BC_p's final mask; its
 *                                            Lamb term is multiplied by 0)
 *   flags, x = phi(interior), partial sum of rhs^2
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_pp_setup(PPGeom G, BBGeom B, double rho_dt, int mask,
  const double *__restrict__ us, const double *__restrict__ vs,
  const double *__restrict__ ws, const double *__restrict__ phi,
  const int *__restrict__ phase, const int *__restrict__ shell,
  const int *__restrict__ fu, const int *__restrict__ fv,
  const int *__restrict__ fw, double *__restrict__ rhs,
  double *__restrict__ x, uint8_t *__restrict__ flags,
  double *__restrict__ part)
{
  __shared__ double sh[4];
  const int lb = xcd_block(G.nblk);
  TileIdx T = tile_of(G, lb);
  const double ddx = 1. / B.dx, ddy = 1. / B.dy, ddz = 1. / B.dz;
  double s2 = 0.;
  if(T.i < G.nx && T.j < G.ny) {
    for(int k = T.k0; k < T.k1; k++) {
      const int i = T.i + 1, j = T.j + 1, kk = k + 1;
      const int U = i + j*B.us1 + kk*B.us2;
      const int V = i + j*B.vs1 + kk*B.vs2;
      const int W = i + j*B.ws1 + kk*B.ws2;
      const int Cc = i + j*B.cs1 + kk*B.cs2;
      double r = (us[U+1] - us[U]) * ddx;
      r += (vs[V+B.vs1] - vs[V]) * ddy;
      r += (ws[W+B.ws2] - ws[W]) * ddz;
      r *= rho_dt;
      int ph = phase[Cc];
      if(mask) r = (double) (ph < 0 && shell[Cc]) * r;
      uint8_t f = 0;
      f |= fu[U] ? BB_FW : 0;
      f |= fu[U+1] ? BB_FE : 0;
      f |= fv[V] ? BB_FS : 0;
      f |= fv[V+B.vs1] ? BB_FN : 0;
      f |= fw[W] ? BB_FB : 0;
      f |= fw[W+B.ws2] ? BB_FT : 0;
      f |= (ph > -1) ? BB_SOLID : 0;
      const int C = T.i + T.j*G.nx + k*G.nx*G.ny;
      rhs[C] = r;
      flags[C] = f;
      x[C] = phi[Cc];
      s2 += r * r;
    }
  }
  double t = bb_block_sum(s2, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) part[lb] = t;
}

/* ---------------------------------------------------------------------
 * Finishing kernels (one block of up to 1024 threads): sum the per-block
 * partials (NQ quantities per block, interleaved) in a fixed order and
 * advance the scalar state.  Each thread walks its blocks b = tid,
 * tid + blockDim, ... in order; the loads of four of them are issued
 * before their adds, so the walk costs one memory latency per four
 * blocks instead of one per block.
 * ------------------------------------------------------------------- */
template<int NQ>
__device__ inline void sum_partials(const double *part, int nblk, double *out)
{
  __shared__ double sh[NQ][16];
  const int tid = threadIdx.x, nth = blockDim.x;
  double a[NQ];
#pragma unroll
  for(int q = 0; q < NQ; q++) a[q] = 0.;
  int b = tid;
  for(; b + 3 * nth < nblk; b += 4 * nth) {
    double v[4][NQ];
#pragma unroll
    for(int u = 0; u < 4; u++)
#pragma unroll
      for(int q = 0; q < NQ; q++) v[u][q] = part[(b + u * nth) * NQ + q];
#pragma unroll
    for(int u = 0; u < 4; u++)
#pragma unroll
      for(int q = 0; q < NQ; q++) a[q] += v[u][q];
  }
  for(; b < nblk; b += nth)
#pragma unroll
    for(int q = 0; q < NQ; q++) a[q] += part[b * NQ + q];
#pragma unroll
  for(int q = 0; q < NQ; q++) {
    double w = bb_wave_sum(a[q]);
    if((tid & 63) == 0) sh[q][tid >> 6] = w;
  }
  __syncthreads();
  if(tid == 0) {
#pragma unroll
    for(int q = 0; q < NQ; q++) {
      double t = 0.;
      for(int w = 0; w < (nth + 63) >> 6; w++) t += sh[q][w];
      out[q] = t;
    }
  }
}

/* threads of the per-iteration finishing kernels */
#define FIN_THREADS 1024

/* norm of the raw rhs; 1 when it is 0 (src/cuda_bicgstab.cu:996-1000) */
__global__ void k_f_norm(int nblk, const double *part, KrylovScalars *sc)
{
  double o[1];
  sum_partials<1>(part, nblk, o);
  if(threadIdx.x == 0) {
    double nrm = sqrt(o[0]);
    if(nrm == 0) nrm = 1.;
    sc->bnorm = nrm;
  }
}

/* r0 = b - A x0, tolerance, first convergence test */
__global__ void k_f_init(int nblk, const double *part, KrylovScalars *sc,
  double rel_tol, int max_iter)
{
  double o[2];
  sum_partials<2>(part, nblk, o);
  if(threadIdx.x == 0) {
    sc->tol = 0. + rel_tol * sqrt(o[0]);
    sc->rho = o[1];
    sc->resid = sqrt(o[1]);
    sc->iter = 0;
    sc->max_iter = max_iter;
    sc->s_conv = 0;
    sc->alpha = sc->omega = sc->beta = 0.;
    sc->done = (sc->resid <= sc->tol || 0 >= max_iter);
    sc->converged = sc->resid <= sc->tol;
  }
}

/* The scalar steps of one iteration, from the summed partials: alpha
 * after <r*, v>; omega and the s-convergence test after <t,s>, <t,t>,
 * <s,s>; beta, rho and the convergence test after <r*, r>, <r, r>. */
__device__ inline void fin_alpha(const double *o, KrylovScalars *sc)
{
  sc->alpha = sc->rho / o[0];
}

__device__ inline void fin_omega(const double *o, KrylovScalars *sc)
{
  double rs = sqrt(o[2]);
  if(rs <= sc->tol) {   /* monitor.finished(s) */
    sc->s_conv = 1;
    sc->resid = rs;
    sc->converged = 1;
  }
  sc->omega = o[0] / o[1];
}

__device__ inline void fin_beta(const double *o, KrylovScalars *sc)
{
  if(sc->s_conv) { sc->done = 1; return; }
  double rho_new = o[0];
  sc->beta = (rho_new / sc->rho) * (sc->alpha / sc->omega);
  sc->rho = rho_new;
  sc->iter += 1;
  sc->resid = sqrt(o[1]);
  sc->converged = sc->resid <= sc->tol;
  sc->done = sc->converged || sc->iter >= sc->max_iter;
}

__global__ void k_f1(int nblk, const double *part, KrylovScalars *sc)
{
  if(sc->done) return;
  double o[1];
  sum_partials<1>(part, nblk, o);
  if(threadIdx.x == 0) fin_alpha(o, sc);
}

__global__ void k_f2(int nblk, const double *part, KrylovScalars *sc)
{
  if(sc->done) return;
  double o[3];
  sum_partials<3>(part, nblk, o);
  if(threadIdx.x == 0) fin_omega(o, sc);
}

__global__ void k_f3(int nblk, const double *part, KrylovScalars *sc)
{
  if(sc->done) return;
  double o[2];
  sum_partials<2>(part, nblk, o);
  if(threadIdx.x == 0) fin_beta(o, sc);
}

/* ---------------------------------------------------------------------
 * Sweeps, templated on the vector type (double for the plain form, float
 * for the inner solve of the mixed form; dots accumulate in double).
 * ------------------------------------------------------------------- */
template<typename T>
__device__ inline double mval(const T *__restrict__ y,
  const uint8_t *__restrict__ fl, const double *dinv, int idx)
{
  return dinv[fl[idx] & 127] * (double) y[idx];
}

/* r = b/nrm - A (x/nrm); x and b are scaled in place; r* = p = r; v = 0
 * (and t = 0 when tz is given) */
template<typename T>
__global__ void __launch_bounds__(256)
k_init(PPGeom G, DinvTable tab, const KrylovScalars *sc,
  T *__restrict__ b, T *__restrict__ x, const uint8_t *__restrict__ fl,
  T *__restrict__ r, T *__restrict__ rs, T *__restrict__ p,
  T *__restrict__ v, double *__restrict__ part, int scale_inputs,
  T *__restrict__ tz)
{
  __shared__ double sh[4];
  __shared__ double dinv[256];
  const int tid = threadIdx.x + threadIdx.y * PX;
  dinv[tid] = tid < 128 ? tab.d[tid] : tab.g[tid - 128];
  __syncthreads();
  const int lb = xcd_block(G.nblk);
  TileIdx Ti = tile_of(G, lb);
  const double inv = scale_inputs ? 1. / sc->bnorm : 1.;
  double bb = 0., rr = 0.;
  if(Ti.i < G.nx && Ti.j < G.ny) {
    for(int k = Ti.k0; k < Ti.k1; k++) {
      const int C = Ti.i + Ti.j*G.nx + k*G.nx*G.ny;
      Nbr o = neighbours(G, Ti.i, Ti.j, k);
      uint8_t f = fl[C];
      double xc = (double) x[C] * inv;
      double Ax = apply_row(f, G, xc,
        (double) x[C+o.w] * inv, (double) x[C+o.e] * inv,
        (double) x[C+o.s] * inv, (double) x[C+o.n] * inv,
        (double) x[C+o.b] * inv, (double) x[C+o.t] * inv,
        diag_of(dinv, f));
      double bc_ = (double) b[C] * inv;
      T rv = (T) (1. * bc_ + -1. * Ax);
      r[C] = rv; rs[C] = rv; p[C] = rv; v[C] = (T) 0;
      if(tz) tz[C] = (T) 0;
      bb += bc_ * bc_;
      rr += (double) rv * (double) rv;
    }
  }
  double t0 = bb_block_sum(bb, sh);
  double t1 = bb_block_sum(rr, sh);
  if(tid == 0) { part[2*lb] = t0; part[2*lb+1] = t1; }
}

/* in-place scaling of x and b by 1/nrm (after k_init has used them) */
template<typename T>
__global__ void __launch_bounds__(256)
k_scale2(size_t n, const KrylovScalars *sc, T *__restrict__ x,
  T *__restrict__ b)
{
  const double inv = 1. / sc->bnorm;
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for(size_t q = (size_t) blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    x[q] = (T) ((double) x[q] * inv);
    b[q] = (T) ((double) b[q] * inv);
  }
}

/* ---------------------------------------------------------------------
 * The SpMV sweeps K1 and K2 march in z.  A 512-thread block owns an
 * SX x SY column tile and SZ consecutive planes.  The operand of A (M p,
 * or M s with s = r - alpha v) is formed once per cell: the block stages
 * plane k with its one-cell x/y halo in LDS, while each thread keeps its
 * column's planes k-1, k, k+1 in registers.  All global loads of an
 * iteration are for planes k+1 (halo) and k+2 (column), so their latency
 * overlaps a whole iteration.  A row is summed in the DIA order of
 * apply_row, from the same values the reference's matrix product sees.
 * ------------------------------------------------------------------- */

/* The operand of A at a cell: its value m (M p for K1, M s for K2), the
 * unscaled s (K2's dots) and the cell's operator flags f.  Sweeps load the
 * raw inputs of a cell one iteration before they form its value, so that
 * no load is waited on in the iteration that issues it. */
struct CellVal {
  double m, s;        /* operand value m = M y, and the unscaled y */
  uint8_t f;
  double p0, sv, rn;  /* two-sweep form only: old p, s, new r */
};

template<typename T>
struct OperandP {
  struct Raw { T p; uint8_t f; };
  const T *__restrict__ p;
  __device__ inline Raw load(const uint8_t *__restrict__ fl, long C) const
  {
    Raw w;
    w.p = p[C];
    w.f = fl[C];
    return w;
  }
  __device__ inline CellVal val(const Raw &w, const double *dinv) const
  {
    CellVal c = {};
    c.f = w.f;
    c.m = dinv[w.f & 127] * (double) w.p;
    c.s = 0.;
    return c;
  }
};

/* K1's operand with K4 folded in: p = r + beta p - beta omega v rounded to
 * T (the reference's update of p), then M p.  s carries the new p. */
template<typename T>
struct OperandPNew {
  struct Raw { T r, p, v; uint8_t f; };
  const T *__restrict__ r;
  const T *__restrict__ p;
  const T *__restrict__ v;
  double beta, bo;
  __device__ inline Raw load(const uint8_t *__restrict__ fl, long C) const
  {
    Raw w;
    w.r = r[C];
    w.p = p[C];
    w.v = v[C];
    w.f = fl[C];
    return w;
  }
  __device__ inline CellVal val(const Raw &w, const double *dinv) const
  {
    CellVal c = {};
    c.f = w.f;
    c.s = (double) (T) (1. * (double) w.r + beta * (double) w.p +
      bo * (double) w.v);
    c.m = dinv[w.f & 127] * c.s;
    return c;
  }
};

/* s = r - alpha v rounded to T, as the reference's vector operation */
template<typename T>
struct OperandS {
  struct Raw { T r, v; uint8_t f; };
  const T *__restrict__ r;
  const T *__restrict__ v;
  double alpha;
  __device__ inline Raw load(const uint8_t *__restrict__ fl, long C) const
  {
    Raw w;
    w.r = r[C];
    w.v = v[C];
    w.f = fl[C];
    return w;
  }
  __device__ inline CellVal val(const Raw &w, const double *dinv) const
  {
    CellVal c = {};
    c.f = w.f;
    c.s = (double) (T) (1. * (double) w.r + -alpha * (double) w.v);
    c.m = dinv[w.f & 127] * c.s;
    return c;
  }
};

/* Tile of a sweep block: interior cell (i, j) of this thread, planes
 * [k0, k1) of the block. */
struct SweepTile { int i, j, k0, k1, tx, ty; bool act; };

__device__ inline SweepTile sweep_tile(const PPGeom &G, int lb)
{
  SweepTile t;
  const int bx = lb % G.sx;
  const int rest = lb / G.sx;
  const int by = rest % G.sy;
  const int bz = rest / G.sy;
  t.tx = threadIdx.x;
  t.ty = threadIdx.y;
  t.i = bx * SX + t.tx;
  t.j = by * SY + t.ty;
  t.k0 = bz * SZ;
  t.k1 = min(t.k0 + SZ, G.nz);
  t.act = t.i < G.nx && t.j < G.ny;
  return t;
}

/* Plane k of the column of cell (i, j): index of its neighbour above
 * (honours the periodic wrap; any in-range cell where the coefficient is
 * zero, as in neighbours()). */
__device__ inline int plane_above(const PPGeom &G, int k)
{
  return (k < G.nz - 1) ? k + 1 : (G.perz ? 0 : k);
}

__device__ inline int plane_below(const PPGeom &G, int k)
{
  return (k > 0) ? k - 1 : (G.perz ? G.nz - 1 : k);
}

/* In-plane neighbour offsets of (i, j) (periodic wraps; 0 where the
 * coefficient is zero). */
struct PlaneNbr { int w, e, s, n; };

__device__ inline PlaneNbr plane_nbr(const PPGeom &G, int i, int j)
{
  const int s1 = G.nx;
  PlaneNbr o;
  o.w = (i > 0) ? -1 : (G.perx ? G.nx - 1 : 0);
  o.e = (i < G.nx - 1) ? 1 : (G.perx ? -(G.nx - 1) : 0);
  o.s = (j > 0) ? -s1 : (G.pery ? (G.ny - 1) * s1 : 0);
  o.n = (j < G.ny - 1) ? s1 : (G.pery ? -(G.ny - 1) * s1 : 0);
  return o;
}

/* The one halo cell of the staged plane this thread stages, if any: the
 * first SX threads take the row south of the tile, the next SX the row
 * north of its last row, then SY threads the column west of it and SY the
 * column east of its last column (144 of the 512 threads).  (row, col) is
 * its place in the staged plane, off its in-plane cell index; threads
 * without a halo cell get their own column's in-plane index, so that every
 * thread issues the same loads. */
struct HaloCell { bool v; int row, col; long off; };

__device__ inline HaloCell halo_cell(const PPGeom &G, const SweepTile &t,
  long own)
{
  HaloCell h = {false, 0, 0, own};
  const int tid = t.tx + t.ty * SX;
  const int bx0 = t.i - t.tx, by0 = t.j - t.ty;
  const int nc = min(SX, G.nx - bx0), nr = min(SY, G.ny - by0);
  int i = 0, j = 0;
  if(tid < SX) {
    if(tid >= nc) return h;
    i = bx0 + tid; j = by0;
    h.off = i + (long) j * G.nx + plane_nbr(G, i, j).s;
    h.row = 0; h.col = tid + 1;
  } else if(tid < 2 * SX) {
    const int x = tid - SX;
    if(x >= nc) return h;
    i = bx0 + x; j = by0 + nr - 1;
    h.off = i + (long) j * G.nx + plane_nbr(G, i, j).n;
    h.row = nr + 1; h.col = x + 1;
  } else if(tid < 2 * SX + SY) {
    const int y = tid - 2 * SX;
    if(y >= nr) return h;
    i = bx0; j = by0 + y;
    h.off = i + (long) j * G.nx + plane_nbr(G, i, j).w;
    h.row = y + 1; h.col = 0;
  } else if(tid < 2 * SX + 2 * SY) {
    const int y = tid - 2 * SX - SY;
    if(y >= nr) return h;
    i = bx0 + nc - 1; j = by0 + y;
    h.off = i + (long) j * G.nx + plane_nbr(G, i, j).e;
    h.row = y + 1; h.col = nc + 1;
  } else {
    return h;
  }
  h.v = true;
  return h;
}

/* One z-marching sweep: for every interior cell of the tile, y = A (op)
 * at that cell is handed to `out(C, y, s)`.  Iteration k issues the loads
 * of the column plane above k+1 and of plane k+1's halo cell, forms the
 * values loaded by iteration k-1, stages plane k in one of two LDS buffers
 * and applies the row.  Every thread issues the same loads in the same
 * order (threads outside the grid or without a halo cell load in-range
 * cells they do not use), so that the compiler waits only for the loads of
 * the previous iteration. */
template<typename Op, typename Out>
__device__ inline void sweep_z(const PPGeom &G, const SweepTile &t,
  const uint8_t *__restrict__ fl, const double *dinv,
  double (*pl)[SY + 2][SX + 2], const Op &op, Out &out)
{
  typedef typename Op::Raw Raw;
  if(t.k0 >= t.k1) return;   /* uniform over the block */
  const long s2 = (long) G.nx * G.ny;
  const long col = min(t.i, G.nx - 1) + (long) min(t.j, G.ny - 1) * G.nx;
  const HaloCell h = halo_cell(G, t, col);
  Raw raw_c = op.load(fl, col + t.k0 * s2);
  Raw raw_b = op.load(fl, col + (long) plane_below(G, t.k0) * s2);
  Raw raw_a = op.load(fl, col + (long) plane_above(G, t.k0) * s2);
  Raw hraw = op.load(fl, h.off + t.k0 * s2);
  CellVal cur = op.val(raw_c, dinv);
  double m_prev = op.val(raw_b, dinv).m;
  for(int k = t.k0; k < t.k1; k++) {
    /* on the last plane these reload plane k: in range, unused */
    const int kn = min(k + 1, t.k1 - 1);
    const Raw raw_nn = op.load(fl, col + (long) plane_above(G, kn) * s2);
    const Raw hraw_n = op.load(fl, h.off + (long) kn * s2);
    const CellVal above = op.val(raw_a, dinv);
    const double hm = op.val(hraw, dinv).m;
    double (*P)[SX + 2] = pl[k & 1];
    if(t.act) P[t.ty + 1][t.tx + 1] = cur.m;
    if(h.v) P[h.row][h.col] = hm;
    __syncthreads();
    if(t.act) {
      const double y = apply_row(cur.f, G, cur.m,
        P[t.ty + 1][t.tx], P[t.ty + 1][t.tx + 2],
        P[t.ty][t.tx + 1], P[t.ty + 2][t.tx + 1],
        m_prev, above.m, diag_of(dinv, cur.f));
      out(col + (long) k * s2, y, cur);
    }
    m_prev = cur.m; cur = above;
    raw_a = raw_nn; hraw = hraw_n;
  }
}

__device__ inline void load_dinv(const DinvTable &tab, double *dinv)
{
  const int tid = threadIdx.x + threadIdx.y * SX;
  if(tid < 256) dinv[tid] = tid < 128 ? tab.d[tid] : tab.g[tid - 128];
  __syncthreads();
}

/* Block sum over the 512 threads of a sweep block; valid in thread 0. */
__device__ inline double sweep_block_sum(double v, double *sh)
{
  return bb_block_sum(v, sh);
}

template<typename T>
struct OutP {
  const T *__restrict__ rs;
  T *__restrict__ v;
  double acc;
  __device__ inline void operator()(long C, double y, const CellVal &) {
    const T vt = (T) y;
    v[C] = vt;
    acc += (double) rs[C] * (double) vt;
  }
};

template<typename T>
struct OutS {
  T *__restrict__ t;
  double ts, tt, ss;
  __device__ inline void operator()(long C, double y, const CellVal &c) {
    const T tv = (T) y;
    const double s = c.s;
    t[C] = tv;
    ts += (double) tv * s;
    tt += (double) tv * (double) tv;
    ss += s * s;
  }
};

template<typename T>
struct OutPNew {
  const T *__restrict__ rs;
  T *__restrict__ p;
  T *__restrict__ v;
  double acc;
  __device__ inline void operator()(long C, double y, const CellVal &c) {
    const T vt = (T) y;
    p[C] = (T) c.s;
    v[C] = vt;
    acc += (double) rs[C] * (double) vt;
  }
};

/* K4 + K1: p' = r + beta (p - omega v), v' = A M p', partial <r*, v'>.
 * p and v are read from one buffer pair and written to the other: a block
 * reads its halo cells' old p and v, which neighbouring blocks replace. */
template<typename T>
__global__ void __launch_bounds__(SX * SY)
k_spmv_p(PPGeom G, DinvTable tab, const KrylovScalars *sc,
  const T *__restrict__ r, const T *__restrict__ p_old,
  const T *__restrict__ v_old, const T *__restrict__ rs,
  const uint8_t *__restrict__ fl, T *__restrict__ p_new,
  T *__restrict__ v_new, double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[SX * SY / 64];
  __shared__ double dinv[256];
  __shared__ double pl[2][SY + 2][SX + 2];
  load_dinv(tab, dinv);
  const int lb = xcd_block(G.snblk);
  const SweepTile t = sweep_tile(G, lb);
  OperandPNew<T> op{r, p_old, v_old, sc->beta, -sc->beta * sc->omega};
  OutPNew<T> out{rs, p_new, v_new, 0.};
  sweep_z(G, t, fl, dinv, pl, op, out);
  const double a = sweep_block_sum(out.acc, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) part[lb] = a;
}

/* K2: s = r - alpha v (formed once per cell), t = A M s,
 * partials <t,s>, <t,t>, <s,s> */
template<typename T>
__global__ void __launch_bounds__(SX * SY)
k_spmv_s(PPGeom G, DinvTable tab, const KrylovScalars *sc,
  const T *__restrict__ r, const T *__restrict__ v,
  const uint8_t *__restrict__ fl, T *__restrict__ t_out,
  double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[SX * SY / 64];
  __shared__ double dinv[256];
  __shared__ double pl[2][SY + 2][SX + 2];
  load_dinv(tab, dinv);
  const int lb = xcd_block(G.snblk);
  const SweepTile t = sweep_tile(G, lb);
  OperandS<T> op{r, v, sc->alpha};
  OutS<T> out{t_out, 0., 0., 0.};
  sweep_z(G, t, fl, dinv, pl, op, out);
  const double a = sweep_block_sum(out.ts, sh);
  const double b = sweep_block_sum(out.tt, sh);
  const double c = sweep_block_sum(out.ss, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) {
    part[3*lb] = a; part[3*lb+1] = b; part[3*lb+2] = c;
  }
}

/* K3: x += alpha M p + omega M s; r = s - omega t; partials <r*,r>, <r,r>.
 * After an s-convergence only x += alpha M p is applied. */
template<typename T>
__global__ void __launch_bounds__(256)
k_update_xr(PPGeom G, DinvTable tab, const KrylovScalars *sc,
  T *__restrict__ x, const T *__restrict__ p, T *__restrict__ r,
  const T *__restrict__ v, const T *__restrict__ t, const T *__restrict__ rs,
  const uint8_t *__restrict__ fl, double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[4];
  __shared__ double dinv[256];
  const int tid = threadIdx.x + threadIdx.y * PX;
  dinv[tid] = tid < 128 ? tab.d[tid] : tab.g[tid - 128];
  __syncthreads();
  const int lb = xcd_block(G.nblk);
  TileIdx Ti = tile_of(G, lb);
  const double alpha = sc->alpha, omega = sc->omega;
  const int sconv = sc->s_conv;
  double a0 = 0., a1 = 0.;
  if(Ti.i < G.nx && Ti.j < G.ny) {
    for(int k = Ti.k0; k < Ti.k1; k++) {
      const int C = Ti.i + Ti.j*G.nx + k*G.nx*G.ny;
      double di = dinv[fl[C] & 127];
      double Mp = di * (double) p[C];
      if(sconv) {
        x[C] = (T) (1. * (double) x[C] + alpha * Mp);
        continue;
      }
      T s = (T) (1. * (double) r[C] + -alpha * (double) v[C]);
      double Ms = di * (double) s;
      x[C] = (T) (1. * (double) x[C] + alpha * Mp + omega * Ms);
      T rn = (T) (1. * (double) s + -omega * (double) t[C]);
      r[C] = rn;
      a0 += (double) rs[C] * (double) rn;
      a1 += (double) rn * (double) rn;
    }
  }
  double b0 = bb_block_sum(a0, sh);
  double b1 = bb_block_sum(a1, sh);
  if(tid == 0) { part[2*lb] = b0; part[2*lb+1] = b1; }
}

/* ---------------------------------------------------------------------
 * Two-sweep BiCGSTAB (the FP32 inner solve of the mixed form).
 *
 * The update sweep K3 exists only to produce r' = s - omega t and its dots
 * <r*, r'> and <r', r'>, which beta and the convergence test need before
 * the next K1.  Both follow from dots the K2 sweep can take, by linearity:
 *   <r*, r'> = <r*, s> - omega <r*, t>
 *   <r', r'> = <s, s> - 2 omega <t, s> + omega^2 <t, t>
 * (each evaluated afresh from this iteration's vectors, not a recurrence).
 * So K2 takes five dots and the finishing block computes omega, beta and
 * the residual at once; the next iteration's first sweep forms
 * s = r - alpha v, r' = s - omega t and x += alpha M p + omega M s on the
 * fly for its own cells, then p' = r' + beta (p - omega v) and v' = A M p'
 * as K1 does.  Every vector is rounded to T exactly where the three-sweep
 * form rounds it; r, p and v alternate between two buffer sets, because a
 * block reads its halo cells' old values while their owners write new
 * ones.  Per iteration: two sweeps and 58 B/cell at FP32 (three sweeps and
 * 71 B/cell before).  After the last iteration one pass applies the
 * pending x update.
 * ------------------------------------------------------------------- */
template<typename T>
struct OperandA {
  struct Raw { T r, p, v, t; uint8_t f; };
  const T *__restrict__ r;
  const T *__restrict__ p;
  const T *__restrict__ v;
  const T *__restrict__ t;
  double alpha, omega, beta, bo;
  __device__ inline Raw load(const uint8_t *__restrict__ fl, long C) const
  {
    Raw w;
    w.r = r[C];
    w.p = p[C];
    w.v = v[C];
    w.t = t[C];
    w.f = fl[C];
    return w;
  }
  __device__ inline CellVal val(const Raw &w, const double *dinv) const
  {
    CellVal c = {};
    c.f = w.f;
    const T sv = (T) (1. * (double) w.r + -alpha * (double) w.v);
    const T rn = (T) (1. * (double) sv + -omega * (double) w.t);
    c.s = (double) (T) (1. * (double) rn + beta * (double) w.p +
      bo * (double) w.v);
    c.m = dinv[w.f & 127] * c.s;
    c.p0 = (double) w.p;
    c.sv = (double) sv;
    c.rn = (double) rn;
    return c;
  }
};

template<typename T>
struct OutA {
  const T *__restrict__ rs;
  T *__restrict__ x;
  T *__restrict__ r_new;
  T *__restrict__ p_new;
  T *__restrict__ v_new;
  const double *dinv;
  double alpha, omega;
  double acc;
  __device__ inline void operator()(long C, double y, const CellVal &c) {
    const double di = dinv[c.f & 127];
    x[C] = (T) (1. * (double) x[C] + alpha * (di * c.p0) + omega * (di * c.sv));
    r_new[C] = (T) c.rn;
    p_new[C] = (T) c.s;
    const T vt = (T) y;
    v_new[C] = vt;
    acc += (double) rs[C] * (double) vt;
  }
};

template<typename T>
struct OutB {
  const T *__restrict__ rs;
  T *__restrict__ t;
  double ts, tt, ss, rss, rst;
  __device__ inline void operator()(long C, double y, const CellVal &c) {
    const T tv = (T) y;
    const double s = c.s, rv = (double) rs[C];
    t[C] = tv;
    ts += (double) tv * s;
    tt += (double) tv * (double) tv;
    ss += s * s;
    rss += rv * s;
    rst += rv * (double) tv;
  }
};

/* omega, the s-convergence test, then beta and the convergence test from
 * the five dots <t,s>, <t,t>, <s,s>, <r*,s>, <r*,t>; `set` is the buffer
 * set this iteration's vectors are in */
__device__ inline void fin_two_sweep(const double *o, KrylovScalars *sc,
  int set)
{
  sc->cur = set;
  const double rs_ = sqrt(o[2]);
  if(rs_ <= sc->tol) {   /* monitor.finished(s) */
    sc->s_conv = 1;
    sc->resid = rs_;
    sc->converged = 1;
    sc->done = 1;
    return;
  }
  const double omega = o[0] / o[1];
  sc->omega = omega;
  const double rho_new = o[3] - omega * o[4];
  const double rr = o[2] - 2. * omega * o[0] + omega * omega * o[1];
  sc->beta = (rho_new / sc->rho) * (sc->alpha / omega);
  sc->rho = rho_new;
  sc->iter += 1;
  sc->resid = sqrt(rr > 0. ? rr : 0.);
  sc->converged = sc->resid <= sc->tol;
  sc->done = sc->converged || sc->iter >= sc->max_iter;
}

/* sweep A: pending update of the previous iteration, then K1 */
template<typename T>
__global__ void __launch_bounds__(SX * SY)
k_sweep_a(PPGeom G, DinvTable tab, const KrylovScalars *sc,
  const T *__restrict__ r, const T *__restrict__ p, const T *__restrict__ v,
  const T *__restrict__ t, const T *__restrict__ rs,
  const uint8_t *__restrict__ fl, T *__restrict__ x, T *__restrict__ r_new,
  T *__restrict__ p_new, T *__restrict__ v_new, double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[SX * SY / 64];
  __shared__ double dinv[256];
  __shared__ double pl[2][SY + 2][SX + 2];
  load_dinv(tab, dinv);
  const int lb = xcd_block(G.snblk);
  const SweepTile tt = sweep_tile(G, lb);
  const double alpha = sc->alpha, omega = sc->omega, beta = sc->beta;
  OperandA<T> op{r, p, v, t, alpha, omega, beta, -beta * omega};
  OutA<T> out{rs, x, r_new, p_new, v_new, dinv, alpha, omega, 0.};
  sweep_z(G, tt, fl, dinv, pl, op, out);
  const double a = sweep_block_sum(out.acc, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) part[lb] = a;
}

/* sweep B: s = r - alpha v, t = A M s and the five dots */
template<typename T>
__global__ void __launch_bounds__(SX * SY)
k_sweep_b(PPGeom G, DinvTable tab, const KrylovScalars *sc,
  const T *__restrict__ r, const T *__restrict__ v,
  const T *__restrict__ rs, const uint8_t *__restrict__ fl,
  T *__restrict__ t_out, double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[SX * SY / 64];
  __shared__ double dinv[256];
  __shared__ double pl[2][SY + 2][SX + 2];
  load_dinv(tab, dinv);
  const int lb = xcd_block(G.snblk);
  const SweepTile t = sweep_tile(G, lb);
  OperandS<T> op{r, v, sc->alpha};
  OutB<T> out{rs, t_out, 0., 0., 0., 0., 0.};
  sweep_z(G, t, fl, dinv, pl, op, out);
  const double q0 = sweep_block_sum(out.ts, sh);
  const double q1 = sweep_block_sum(out.tt, sh);
  const double q2 = sweep_block_sum(out.ss, sh);
  const double q3 = sweep_block_sum(out.rss, sh);
  const double q4 = sweep_block_sum(out.rst, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) {
    part[5*lb] = q0; part[5*lb+1] = q1; part[5*lb+2] = q2;
    part[5*lb+3] = q3; part[5*lb+4] = q4;
  }
}

/* the finishing kernel of sweep B */
__global__ void k_f_two(int nblk, const double *part, KrylovScalars *sc,
  int set)
{
  if(sc->done) return;
  double o[5];
  sum_partials<5>(part, nblk, o);
  if(threadIdx.x == 0) fin_two_sweep(o, sc, set);
}

/* the update the last iteration left pending: x += alpha M p (+ omega M s
 * with s = r - alpha v, unless s converged), on buffer set sc->cur */
template<typename T>
__global__ void __launch_bounds__(256)
k_final_x(size_t n, DinvTable tab, const KrylovScalars *sc,
  const uint8_t *__restrict__ fl, T *__restrict__ x,
  const T *r0, const T *p0, const T *v0, const T *r1, const T *p1,
  const T *v1)
{
  __shared__ double dinv[256];
  if(threadIdx.x < 256)
    dinv[threadIdx.x] = threadIdx.x < 128 ? tab.d[threadIdx.x]
                                          : tab.g[threadIdx.x - 128];
  __syncthreads();
  if(sc->iter == 0 && !sc->s_conv) return;   /* no iteration ran */
  const int set = sc->cur;
  const T *r = set ? r1 : r0, *p = set ? p1 : p0, *v = set ? v1 : v0;
  const double alpha = sc->alpha, omega = sc->omega;
  const int sconv = sc->s_conv;
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for(size_t q = (size_t) blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const double di = dinv[fl[q] & 127];
    const double Mp = di * (double) p[q];
    if(sconv) {
      x[q] = (T) (1. * (double) x[q] + alpha * Mp);
      continue;
    }
    const T sv = (T) (1. * (double) r[q] + -alpha * (double) v[q]);
    x[q] = (T) (1. * (double) x[q] + alpha * Mp + omega * (di * (double) sv));
  }
}

/* ---------------------------------------------------------------------
 * Flexible BiCGSTAB (right preconditioner M that may change between
 * applications: the learned network).  Per iteration
 *   p^ = M p;  v = A p^, <r*,v>;  alpha
 *   s = r - alpha v, <s,s>         (s-convergence test)
 *   s^ = M s;  t = A s^, <t,s>, <t,t>;  omega
 *   x += alpha p^ + omega s^;  r = s - omega t, <r*,r>, <r,r>;  beta
 *   p = r + beta (p - omega v)
 * The products with A use the same z-marching sweep with the stored
 * preconditioned vector as the operand.
 * ------------------------------------------------------------------- */
template<typename T>
struct OperandPlain {
  struct Raw { T p; uint8_t f; };
  const T *__restrict__ p;
  __device__ inline Raw load(const uint8_t *__restrict__ fl, long C) const
  {
    Raw w;
    w.p = p[C];
    w.f = fl[C];
    return w;
  }
  __device__ inline CellVal val(const Raw &w, const double *) const
  {
    CellVal c = {};
    c.f = w.f;
    c.m = (double) w.p;
    c.s = 0.;
    return c;
  }
};

template<typename T>
struct OutTS {
  T *__restrict__ t;
  const T *__restrict__ s;
  double ts, tt;
  __device__ inline void operator()(long C, double y, const CellVal &) {
    const T tv = (T) y;
    t[C] = tv;
    const double sv = (double) s[C];
    ts += (double) tv * sv;
    tt += (double) tv * (double) tv;
  }
};

/* v = A p^, partial <r*, v> */
template<typename T>
__global__ void __launch_bounds__(SX * SY)
k_fspmv_v(PPGeom G, DinvTable tab, const KrylovScalars *sc,
  const T *__restrict__ ph, const T *__restrict__ rs,
  const uint8_t *__restrict__ fl, T *__restrict__ v,
  double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[SX * SY / 64];
  __shared__ double dinv[256];
  __shared__ double pl[2][SY + 2][SX + 2];
  load_dinv(tab, dinv);
  const int lb = xcd_block(G.snblk);
  const SweepTile t = sweep_tile(G, lb);
  OperandPlain<T> op{ph};
  OutP<T> out{rs, v, 0.};
  sweep_z(G, t, fl, dinv, pl, op, out);
  const double a = sweep_block_sum(out.acc, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) part[lb] = a;
}

/* t = A s^, partials <t, s>, <t, t> */
template<typename T>
__global__ void __launch_bounds__(SX * SY)
k_fspmv_t(PPGeom G, DinvTable tab, const KrylovScalars *sc,
  const T *__restrict__ shat, const T *__restrict__ s,
  const uint8_t *__restrict__ fl, T *__restrict__ t_out,
  double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[SX * SY / 64];
  __shared__ double dinv[256];
  __shared__ double pl[2][SY + 2][SX + 2];
  load_dinv(tab, dinv);
  const int lb = xcd_block(G.snblk);
  const SweepTile t = sweep_tile(G, lb);
  OperandPlain<T> op{shat};
  OutTS<T> out{t_out, s, 0., 0.};
  sweep_z(G, t, fl, dinv, pl, op, out);
  const double a = sweep_block_sum(out.ts, sh);
  const double b = sweep_block_sum(out.tt, sh);
  if(threadIdx.x == 0 && threadIdx.y == 0) { part[2*lb] = a; part[2*lb+1] = b; }
}

/* s = r - alpha v, partial <s, s> */
template<typename T>
__global__ void __launch_bounds__(256)
k_form_s(PPGeom G, const KrylovScalars *sc, const T *__restrict__ r,
  const T *__restrict__ v, T *__restrict__ s, double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[4];
  const int lb = xcd_block(G.nblk);
  TileIdx Ti = tile_of(G, lb);
  const double alpha = sc->alpha;
  double ss = 0.;
  if(Ti.i < G.nx && Ti.j < G.ny) {
    for(int k = Ti.k0; k < Ti.k1; k++) {
      const int C = Ti.i + Ti.j*G.nx + k*G.nx*G.ny;
      const T sv = (T) (1. * (double) r[C] + -alpha * (double) v[C]);
      s[C] = sv;
      ss += (double) sv * (double) sv;
    }
  }
  const double b = bb_block_sum(ss, sh);
  if(threadIdx.x + threadIdx.y * PX == 0) part[lb] = b;
}

__global__ void k_ff_s(int nblk, const double *part, KrylovScalars *sc)
{
  if(sc->done) return;
  double o[1];
  sum_partials<1>(part, nblk, o);
  if(threadIdx.x == 0) {
    double rs = sqrt(o[0]);
    if(rs <= sc->tol) {
      sc->s_conv = 1;
      sc->resid = rs;
      sc->converged = 1;
    }
  }
}

__global__ void k_ff_omega(int nblk, const double *part, KrylovScalars *sc)
{
  if(sc->done || sc->s_conv) return;
  double o[2];
  sum_partials<2>(part, nblk, o);
  if(threadIdx.x == 0) sc->omega = o[0] / o[1];
}

/* x += alpha p^ + omega s^, r = s - omega t, partials <r*,r>, <r,r>;
 * after an s-convergence only x += alpha p^ */
template<typename T>
__global__ void __launch_bounds__(256)
k_fupdate_xr(PPGeom G, const KrylovScalars *sc, T *__restrict__ x,
  const T *__restrict__ ph, const T *__restrict__ shat, T *__restrict__ r,
  const T *__restrict__ s, const T *__restrict__ t, const T *__restrict__ rs,
  double *__restrict__ part)
{
  if(sc->done) return;
  __shared__ double sh[4];
  const int lb = xcd_block(G.nblk);
  TileIdx Ti = tile_of(G, lb);
  const double alpha = sc->alpha, omega = sc->omega;
  const int sconv = sc->s_conv;
  double a0 = 0., a1 = 0.;
  if(Ti.i < G.nx && Ti.j < G.ny) {
    for(int k = Ti.k0; k < Ti.k1; k++) {
      const int C = Ti.i + Ti.j*G.nx + k*G.nx*G.ny;
      if(sconv) {
        x[C] = (T) (1. * (double) x[C] + alpha * (double) ph[C]);
        continue;
      }
      x[C] = (T) (1. * (double) x[C] + alpha * (double) ph[C]
        + omega * (double) shat[C]);
      const T rn = (T) (1. * (double) s[C] + -omega * (double) t[C]);
      r[C] = rn;
      a0 += (double) rs[C] * (double) rn;
      a1 += (double) rn * (double) rn;
    }
  }
  const double b0 = bb_block_sum(a0, sh);
  const double b1 = bb_block_sum(a1, sh);
  if(threadIdx.x + threadIdx.y * PX == 0) { part[2*lb] = b0; part[2*lb+1] = b1; }
}

/* p = r + beta (p - omega v) */
template<typename T>
__global__ void __launch_bounds__(256)
k_fupdate_p(size_t n, const KrylovScalars *sc, T *__restrict__ p,
  const T *__restrict__ r, const T *__restrict__ v)
{
  if(sc->done) return;
  const double beta = sc->beta, bo = -sc->beta * sc->omega;
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for(size_t q = (size_t) blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    p[q] = (T) (1. * (double) r[q] + beta * (double) p[q] + bo * (double) v[q]);
}

/* ---------------------------------------------------------------------
 * Mean removal and copy back: x *= nrm; x -= mean(x); phi(interior) = x
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_unscale_sum(size_t n, const KrylovScalars *sc, double *__restrict__ x,
  double *__restrict__ part)
{
  __shared__ double sh[4];
  const double nrm = sc->bnorm;
  double s = 0.;
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for(size_t q = (size_t) blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    double val = x[q] * nrm;
    x[q] = val;
    s += val;
  }
  double t = bb_block_sum(s, sh);
  if(threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_f_mean(int nblk, const double *part, double *out, double n)
{
  double o[1];
  sum_partials<1>(part, nblk, o);
  if(threadIdx.x == 0) out[0] = o[0] / n;
}

__global__ void __launch_bounds__(256)
k_copy_back(PPGeom G, BBGeom B, const double *__restrict__ mean,
  const double *__restrict__ x, double *__restrict__ phi)
{
  const int lb = blockIdx.x;
  TileIdx Ti = tile_of(G, lb);
  if(Ti.i >= G.nx || Ti.j >= G.ny) return;
  const double pa = mean[0];
  for(int k = Ti.k0; k < Ti.k1; k++) {
    const int C = Ti.i + Ti.j*G.nx + k*G.nx*G.ny;
    phi[(Ti.i + 1) + (Ti.j + 1)*B.cs1 + (k + 1)*B.cs2] = -pa * 1. + x[C];
  }
}

/* ---------------------------------------------------------------------
 * Mixed-precision helpers.
 *   FP64 residual: r = b - A x  (x, b FP64), writes r/|r| into FP32 and
 *   partial |r|^2; FP64 correction x += scale * e (e FP32).
 * ------------------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_resid64(PPGeom G, DinvTable tab, const double *__restrict__ b,
  const double *__restrict__ x, const uint8_t *__restrict__ fl,
  double *__restrict__ r, double *__restrict__ part)
{
  __shared__ double sh[4];
  __shared__ double dinv[256];
  const int tid = threadIdx.x + threadIdx.y * PX;
  dinv[tid] = tid < 128 ? tab.d[tid] : tab.g[tid - 128];
  __syncthreads();
  const int lb = xcd_block(G.nblk);
  TileIdx Ti = tile_of(G, lb);
  double rr = 0.;
  if(Ti.i < G.nx && Ti.j < G.ny) {
    for(int k = Ti.k0; k < Ti.k1; k++) {
      const int C = Ti.i + Ti.j*G.nx + k*G.nx*G.ny;
      Nbr o = neighbours(G, Ti.i, Ti.j, k);
      uint8_t f = fl[C];
      double Ax = apply_row(f, G, x[C], x[C+o.w], x[C+o.e], x[C+o.s],
        x[C+o.n], x[C+o.b], x[C+o.t], diag_of(dinv, f));
      double rv = b[C] - Ax;
      r[C] = rv;
      rr += rv * rv;
    }
  }
  double t = bb_block_sum(rr, sh);
  if(tid == 0) part[lb] = t;
}

/* FP32 copies for the inner solve: rhs32 = r / |r|, x32 = 0 */
__global__ void __launch_bounds__(256)
k_to32(size_t n, const double *__restrict__ scal, const double *__restrict__ r,
  float *__restrict__ rf, float *__restrict__ xf)
{
  const double inv = 1. / scal[0];
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for(size_t q = (size_t) blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    rf[q] = (float) (r[q] * inv);
    xf[q] = 0.f;
  }
}

__global__ void __launch_bounds__(256)
k_correct64(size_t n, const double *__restrict__ scal,
  const float *__restrict__ e, double *__restrict__ x)
{
  const double s = scal[0];
  size_t stride = (size_t) gridDim.x * blockDim.x;
  for(size_t q = (size_t) blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    x[q] += s * (double) e[q];
}

__global__ void k_f_rnorm(int nblk, const double *part, double *out)
{
  double o[1];
  sum_partials<1>(part, nblk, o);
  if(threadIdx.x == 0) out[0] = sqrt(o[0]);
}

/* ---------------------------------------------------------------------
 * Host drivers
 * ------------------------------------------------------------------- */
static int elementwise_blocks(size_t n)
{
  size_t b = (n + 255) / 256;
  if(b > 2048) b = 2048;
  return (int) (b ? b : 1);
}

/* One BiCGSTAB solve of A x = b on vectors of type T, with x already
 * holding the initial guess and b the right-hand side (both already
 * normalised when `scale_inputs` is 0).  Returns the iteration count. */
template<typename T>
static void bicgstab_run(const PPGeom &G, const DinvTable &tab, T *b, T *x,
  T *r, T *rs, T *p, T *v, T *p2, T *v2, T *t, double rel_tol, int max_iter,
  int scale_inputs, int poll, int kinds[4])
{
  BBWorkspace *ws = &g_bb.ws;
  hipStream_t st = g_bb.stream;
  dim3 blk(PX, PY);
  dim3 sblk(SX, SY);
  size_t n = ws->n;
  int eb = elementwise_blocks(n);
  if(G.snblk > ws->nblk) {
    fprintf(stderr, "bluebottle_hip: %d sweep tiles exceed the %d reduction "
      "slots\n", G.snblk, ws->nblk);
    exit(EXIT_FAILURE);
  }
  k_init<T><<<G.nblk, blk, 0, st>>>(G, tab, ws->sc, b, x, ws->flags, r, rs,
    p, v, ws->partials, scale_inputs, (T *) nullptr);
  BB_CHECK_LAUNCH();
  if(scale_inputs) {
    k_scale2<T><<<eb, 256, 0, st>>>(n, ws->sc, x, b);
    BB_CHECK_LAUNCH();
  }
  k_f_init<<<1, FIN_THREADS, 0, st>>>(G.nblk, ws->partials, ws->sc, rel_tol, max_iter);
  BB_CHECK_LAUNCH();
  /* p and v alternate between two buffers; k_init left p = r, v = 0 and
   * beta = omega = 0, so the first sweep forms p = r */
  T *pa = p, *pb = p2, *va = v, *vb = v2;
  for(int it = 0; it < max_iter + 1; it++) {
    bb_prof_begin(kinds[0]);
    k_spmv_p<T><<<G.snblk, sblk, 0, st>>>(G, tab, ws->sc, r, pa, va, rs,
      ws->flags, pb, vb, ws->partials);
    BB_CHECK_LAUNCH();
    bb_prof_end(kinds[0]);
    k_f1<<<1, FIN_THREADS, 0, st>>>(G.snblk, ws->partials, ws->sc);
    BB_CHECK_LAUNCH();
    bb_prof_begin(kinds[1]);
    k_spmv_s<T><<<G.snblk, sblk, 0, st>>>(G, tab, ws->sc, r, vb, ws->flags,
      t, ws->partials);
    BB_CHECK_LAUNCH();
    bb_prof_end(kinds[1]);
    k_f2<<<1, FIN_THREADS, 0, st>>>(G.snblk, ws->partials, ws->sc);
    BB_CHECK_LAUNCH();
    bb_prof_begin(kinds[2]);
    k_update_xr<T><<<G.nblk, blk, 0, st>>>(G, tab, ws->sc, x, pb, r, vb, t,
      rs, ws->flags, ws->partials);
    BB_CHECK_LAUNCH();
    bb_prof_end(kinds[2]);
    k_f3<<<1, FIN_THREADS, 0, st>>>(G.nblk, ws->partials, ws->sc);
    BB_CHECK_LAUNCH();
    T *tp = pa; pa = pb; pb = tp;
    T *tv = va; va = vb; vb = tv;
    if((it + 1) % poll == 0 || it == max_iter) {
      BB_CHECK(hipMemcpyAsync(ws->sc_host, ws->sc, sizeof(KrylovScalars),
        hipMemcpyDeviceToHost, st));
      BB_CHECK(hipStreamSynchronize(st));
      if(ws->sc_host->done) break;
    }
  }
  BB_CHECK(hipMemcpyAsync(ws->sc_host, ws->sc, sizeof(KrylovScalars),
    hipMemcpyDeviceToHost, st));
  BB_CHECK(hipStreamSynchronize(st));
}

/* Two-sweep BiCGSTAB of A x = b on vectors of type T (b and x already
 * normalised, x holding the initial guess); r, p and v alternate between
 * the sets (r0, p0, v0) and (r1, p1, v1).  Leaves the final scalars in
 * ws->sc_host. */
template<typename T>
static void bicgstab2_run(const PPGeom &G, const DinvTable &tab, T *b, T *x,
  T *r0, T *r1, T *rs, T *p0, T *p1, T *v0, T *v1, T *t, double rel_tol,
  int max_iter, int poll, const int kinds[2])
{
  BBWorkspace *ws = &g_bb.ws;
  hipStream_t st = g_bb.stream;
  dim3 blk(PX, PY);
  dim3 sblk(SX, SY);
  size_t n = ws->n;
  if(G.snblk > ws->nblk || 5 * G.snblk > 4 * ws->nblk) {
    fprintf(stderr, "bluebottle_hip: %d sweep tiles exceed the %d reduction "
      "slots\n", G.snblk, ws->nblk);
    exit(EXIT_FAILURE);
  }
  k_init<T><<<G.nblk, blk, 0, st>>>(G, tab, ws->sc, b, x, ws->flags, r0, rs,
    p0, v0, ws->partials, 0, t);
  BB_CHECK_LAUNCH();
  k_f_init<<<1, FIN_THREADS, 0, st>>>(G.nblk, ws->partials, ws->sc, rel_tol,
    max_iter);
  BB_CHECK_LAUNCH();
  T *r[2] = {r0, r1}, *p[2] = {p0, p1}, *v[2] = {v0, v1};
  for(int it = 0; it < max_iter + 1; it++) {
    const int a = it & 1, c = a ^ 1;
    bb_prof_begin(kinds[0]);
    k_sweep_a<T><<<G.snblk, sblk, 0, st>>>(G, tab, ws->sc, r[a], p[a], v[a],
      t, rs, ws->flags, x, r[c], p[c], v[c], ws->partials);
    BB_CHECK_LAUNCH();
    bb_prof_end(kinds[0]);
    k_f1<<<1, FIN_THREADS, 0, st>>>(G.snblk, ws->partials, ws->sc);
    BB_CHECK_LAUNCH();
    bb_prof_begin(kinds[1]);
    k_sweep_b<T><<<G.snblk, sblk, 0, st>>>(G, tab, ws->sc, r[c], v[c], rs,
      ws->flags, t, ws->partials);
    BB_CHECK_LAUNCH();
    bb_prof_end(kinds[1]);
    k_f_two<<<1, FIN_THREADS, 0, st>>>(G.snblk, ws->partials, ws->sc, c);
    BB_CHECK_LAUNCH();
    if((it + 1) % poll == 0 || it == max_iter) {
      BB_CHECK(hipMemcpyAsync(ws->sc_host, ws->sc, sizeof(KrylovScalars),
        hipMemcpyDeviceToHost, st));
      BB_CHECK(hipStreamSynchronize(st));
      if(ws->sc_host->done) break;
    }
  }
  k_final_x<T><<<elementwise_blocks(n), 256, 0, st>>>(n, tab, ws->sc,
    ws->flags, x, r0, p0, v0, r1, p1, v1);
  BB_CHECK_LAUNCH();
  BB_CHECK(hipMemcpyAsync(ws->sc_host, ws->sc, sizeof(KrylovScalars),
    hipMemcpyDeviceToHost, st));
  BB_CHECK(hipStreamSynchronize(st));
}

static int g_poll = 8;
static int g_inner_sweeps = 3;

extern "C" int bb_set_inner_sweeps(int n)
{
  const int old = g_inner_sweeps;
  if(n == 2 || n == 3) g_inner_sweeps = n;
  return old;
}

/* The FP32 Jacobi BiCGSTAB of the mixed form's inner solves (and of the
 * learned form's fallback): A x = b from x, b and x normalised. */
static void inner_jacobi(const PPGeom &G, const DinvTable &tab, float *b,
  float *x, double rel_tol, int max_iter)
{
  BBWorkspace *ws = &g_bb.ws;
  if(g_inner_sweeps == 2) {
    const int kinds[2] = {BB_K_SWEEP_A32, BB_K_SWEEP_B32};
    bicgstab2_run<float>(G, tab, b, x, ws->rf, ws->rf2, ws->rsf, ws->pf,
      ws->pf2, ws->vf, ws->vf2, ws->tf, rel_tol, max_iter, g_poll, kinds);
  } else {
    int kinds[4] = {BB_K_SPMV_P32, BB_K_SPMV_S32, BB_K_UPDATE_XR32,
      BB_K_UPDATE_P32};
    bicgstab_run<float>(G, tab, b, x, ws->rf, ws->rsf, ws->pf, ws->vf,
      ws->pf2, ws->vf2, ws->tf, rel_tol, max_iter, 0, g_poll, kinds);
  }
}

extern "C" void bb_set_poll_interval(int n) { g_poll = n > 0 ? n : 1; }

static void solve_fp64(const PPGeom &G, const DinvTable &tab)
{
  BBWorkspace *ws = &g_bb.ws;
  int kinds[4] = {BB_K_SPMV_P, BB_K_SPMV_S, BB_K_UPDATE_XR, BB_K_UPDATE_P};
  bicgstab_run<double>(G, tab, _rhs_p[0], ws->x, ws->r, ws->rs, ws->p, ws->v,
    ws->p2, ws->v2, ws->t, pp_residual, pp_max_iter, 1, g_poll, kinds);
  KrylovScalars *h = ws->sc_host;
  g_bb.stats.iterations = h->iter;
  g_bb.stats.outer_steps = 1;
  g_bb.stats.residual = h->resid;
  g_bb.stats.converged = h->converged;
}

/* ---------------------------------------------------------------------
 * Learned preconditioner: the flexible inner solve, the store of
 * (residual, correction) pairs produced by the Jacobi inner solves, and
 * the training loop (network in bb_unet.hip).
 * ------------------------------------------------------------------- */
struct LearnState {
  int active;              /* network trained and in use */
  int stall_window;        /* fallback test every this many iterations */
  double stall_drop;       /* ... requires |r| to fall below this ratio */
  /* pair store */
  int cap, count, next;
  float *pr, *pe;          /* cap x n each */
  uint8_t *pf;             /* cap x n flags */
  size_t n;
  /* after a fallback the next `skip` inner solves use Jacobi directly,
   * then the network is tried again; `skip` starts at skip_after_stall and
   * doubles with every fallback (up to 1024), and returns to
   * skip_after_stall after a solve of at least one stall window that the
   * network finishes, or a retraining */
  int skip_after_stall, skip_left, skip_next;
  /* statistics */
  long applications, fallbacks, flex_iters, jacobi_iters, skipped;
  double train_seconds;
};

static LearnState g_learn = {0, 5, 0.9, 0, 0, 0, NULL, NULL, NULL, 0,
  8, 0, 8, 0, 0, 0, 0, 0, 0.};

static void learn_free_pairs(void)
{
  if(g_learn.pr) BB_CHECK(hipFree(g_learn.pr));
  if(g_learn.pe) BB_CHECK(hipFree(g_learn.pe));
  if(g_learn.pf) BB_CHECK(hipFree(g_learn.pf));
  g_learn.pr = g_learn.pe = NULL;
  g_learn.pf = NULL;
  g_learn.cap = g_learn.count = g_learn.next = 0;
}

/* keep (b, x) of the Jacobi inner solve just finished: the normalised
 * residual and its correction, with the operator flags of this solve */
static void collect_pair(BBWorkspace *ws)
{
  if(g_learn.cap <= 0 || g_learn.n != ws->n) return;
  /* only corrections solved to 1e-4 relative (|b| = 1): the outer loop
   * lets an inner solve stop early when the outer test needs less */
  if(!(ws->sc_host->resid <= 1e-4)) return;
  hipStream_t st = g_bb.stream;
  size_t n = ws->n;
  int slot = g_learn.next;
  BB_CHECK(hipMemcpyAsync(g_learn.pr + slot * n, ws->bf, n * sizeof(float),
    hipMemcpyDeviceToDevice, st));
  BB_CHECK(hipMemcpyAsync(g_learn.pe + slot * n, ws->xf, n * sizeof(float),
    hipMemcpyDeviceToDevice, st));
  BB_CHECK(hipMemcpyAsync(g_learn.pf + slot * n, ws->flags, n,
    hipMemcpyDeviceToDevice, st));
  g_learn.next = (slot + 1) % g_learn.cap;
  if(g_learn.count < g_learn.cap) g_learn.count++;
}

static void apply_net(const float *in, float *out)
{
  bb_prof_begin(BB_K_UNET);
  bb_unet_apply_dev(in, g_bb.ws.flags, out, g_bb.stream);
  bb_prof_end(BB_K_UNET);
  g_learn.applications++;
}

/* Flexible BiCGSTAB on the FP32 inner system A e = b (e from 0), judged
 * by the same relative tolerance; falls back to Jacobi BiCGSTAB from the
 * current iterate when |r| drops by less than 10 % over 5 iterations. */
static void flexible_run(const PPGeom &G, const DinvTable &tab, double rel_tol,
  int max_iter)
{
  BBWorkspace *ws = &g_bb.ws;
  hipStream_t st = g_bb.stream;
  dim3 blk(PX, PY), sblk(SX, SY);
  size_t n = ws->n;
  int eb = elementwise_blocks(n);
  float *b = ws->bf, *x = ws->xf, *r = ws->rf, *rs = ws->rsf, *p = ws->pf;
  float *v = ws->vf, *sv = ws->pf2, *t = ws->tf, *ph = ws->phf, *sh = ws->shf;
  if(G.snblk > ws->nblk || G.nblk > ws->nblk) {
    fprintf(stderr, "bluebottle_hip: %d sweep tiles exceed the %d reduction "
      "slots\n", G.snblk, ws->nblk);
    exit(EXIT_FAILURE);
  }
  k_init<float><<<G.nblk, blk, 0, st>>>(G, tab, ws->sc, b, x, ws->flags, r,
    rs, p, v, ws->partials, 0, (float *) nullptr);
  BB_CHECK_LAUNCH();
  k_f_init<<<1, FIN_THREADS, 0, st>>>(G.nblk, ws->partials, ws->sc, rel_tol,
    max_iter);
  BB_CHECK_LAUNCH();
  BB_CHECK(hipMemcpyAsync(ws->sc_host, ws->sc, sizeof(KrylovScalars),
    hipMemcpyDeviceToHost, st));
  BB_CHECK(hipStreamSynchronize(st));
  if(ws->sc_host->done) return;
  double hist[64];
  int it;
  for(it = 0; it < max_iter; it++) {
    apply_net(p, ph);
    bb_prof_begin(BB_K_FSPMV);
    k_fspmv_v<float><<<G.snblk, sblk, 0, st>>>(G, tab, ws->sc, ph, rs,
      ws->flags, v, ws->partials);
    BB_CHECK_LAUNCH();
    bb_prof_end(BB_K_FSPMV);
    k_f1<<<1, FIN_THREADS, 0, st>>>(G.snblk, ws->partials, ws->sc);
    BB_CHECK_LAUNCH();
    bb_prof_begin(BB_K_FVEC);
    k_form_s<float><<<G.nblk, blk, 0, st>>>(G, ws->sc, r, v, sv, ws->partials);
    BB_CHECK_LAUNCH();
    bb_prof_end(BB_K_FVEC);
    k_ff_s<<<1, FIN_THREADS, 0, st>>>(G.nblk, ws->partials, ws->sc);
    BB_CHECK_LAUNCH();
    apply_net(sv, sh);
    bb_prof_begin(BB_K_FSPMV);
    k_fspmv_t<float><<<G.snblk, sblk, 0, st>>>(G, tab, ws->sc, sh, sv,
      ws->flags, t, ws->partials);
    BB_CHECK_LAUNCH();
    bb_prof_end(BB_K_FSPMV);
    k_ff_omega<<<1, FIN_THREADS, 0, st>>>(G.snblk, ws->partials, ws->sc);
    BB_CHECK_LAUNCH();
    bb_prof_begin(BB_K_FVEC);
    k_fupdate_xr<float><<<G.nblk, blk, 0, st>>>(G, ws->sc, x, ph, sh, r, sv,
      t, rs, ws->partials);
    BB_CHECK_LAUNCH();
    bb_prof_end(BB_K_FVEC);
    k_f3<<<1, FIN_THREADS, 0, st>>>(G.nblk, ws->partials, ws->sc);
    BB_CHECK_LAUNCH();
    bb_prof_begin(BB_K_FVEC);
    k_fupdate_p<float><<<eb, 256, 0, st>>>(n, ws->sc, p, r, v);
    BB_CHECK_LAUNCH();
    bb_prof_end(BB_K_FVEC);
    BB_CHECK(hipMemcpyAsync(ws->sc_host, ws->sc, sizeof(KrylovScalars),
      hipMemcpyDeviceToHost, st));
    BB_CHECK(hipStreamSynchronize(st));
    hist[it % 64] = ws->sc_host->resid;
    if(ws->sc_host->done) break;
    const int w = g_learn.stall_window;
    if(it + 1 >= w && (it + 1) % w == 0) {
      double before = it + 1 == w ? ws->sc_host->tol / rel_tol
                                  : hist[(it - w) % 64];
      if(ws->sc_host->resid > g_learn.stall_drop * before) {
        /* stalled: Jacobi BiCGSTAB from the current iterate */
        int done_it = ws->sc_host->iter;
        g_learn.flex_iters += done_it;
        g_learn.fallbacks++;
        inner_jacobi(G, tab, b, x, rel_tol, max_iter - done_it);
        g_learn.jacobi_iters += ws->sc_host->iter;
        ws->sc_host->iter += done_it;
        g_learn.skip_left = g_learn.skip_next;
        if(g_learn.skip_next < 1024) g_learn.skip_next *= 2;
        return;
      }
    }
  }
  g_learn.flex_iters += ws->sc_host->iter;
  /* a solve the network carried past one stall window resets the backoff;
   * shorter ones say nothing about it */
  if(ws->sc_host->iter >= g_learn.stall_window)
    g_learn.skip_next = g_learn.skip_after_stall;
}

/* ---- C ABI of the learned form (no reference counterpart) ---- */
extern "C" void bb_learn_collect(int capacity)
{
  bb_ensure_workspace();
  learn_free_pairs();
  if(capacity <= 0) return;
  size_t n = g_bb.ws.n;
  g_learn.n = n;
  BB_CHECK(hipMalloc(&g_learn.pr, (size_t) capacity * n * sizeof(float)));
  BB_CHECK(hipMalloc(&g_learn.pe, (size_t) capacity * n * sizeof(float)));
  BB_CHECK(hipMalloc(&g_learn.pf, (size_t) capacity * n));
  g_learn.cap = capacity;
}

extern "C" int bb_learn_pairs(void) { return g_learn.count; }

/* Train the network on the stored pairs: `iters` Adam steps at learning
 * rate `lr`, pairs drawn from a splitmix64 stream seeded with `seed`.
 * Returns the mean loss of the last tenth of the steps (-1 without pairs
 * or a suitable grid).  Afterwards the learned form uses the network. */
extern "C" real bb_learn_train(int iters, real lr, unsigned long seed)
{
  bb_ensure_workspace();
  if(g_learn.count <= 0 || !bb_unet_ready()) return -1.;
  hipStream_t st = g_bb.stream;
  size_t n = g_learn.n;
  uint64_t sd = seed;
  double tail = 0.;
  int ntail = 0;
  hipEvent_t e0, e1;
  BB_CHECK(hipEventCreate(&e0));
  BB_CHECK(hipEventCreate(&e1));
  BB_CHECK(hipEventRecord(e0, st));
  for(int i = 0; i < iters; i++) {
    uint64_t z = (sd += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    int k = (int) (z % (uint64_t) g_learn.count);
    double loss = bb_unet_train_pair(g_learn.pr + k * n, g_learn.pe + k * n,
      g_learn.pf + k * n, (float) lr, 1, st);
    if(i >= iters - (iters + 9) / 10) { tail += loss; ntail++; }
  }
  BB_CHECK(hipEventRecord(e1, st));
  BB_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  BB_CHECK(hipEventElapsedTime(&ms, e0, e1));
  BB_CHECK(hipEventDestroy(e0));
  BB_CHECK(hipEventDestroy(e1));
  g_learn.train_seconds += 1e-3 * ms;
  g_learn.active = 1;
  g_learn.skip_left = 0;
  g_learn.skip_next = g_learn.skip_after_stall;
  return ntail ? tail / ntail : -1.;
}

/* loss of stored pair k under the current network (no training step) */
extern "C" real bb_learn_pair_loss(int k)
{
  if(k < 0 || k >= g_learn.count || !bb_unet_ready()) return -1.;
  size_t n = g_learn.n;
  return bb_unet_train_pair(g_learn.pr + k * n, g_learn.pe + k * n,
    g_learn.pf + k * n, 0.f, 0, g_bb.stream);
}

/* |A e - r| / |r| of stored pair k (fluid rows) */
extern "C" real bb_learn_pair_residual(int k)
{
  if(k < 0 || k >= g_learn.count || !bb_unet_ready()) return -1.;
  size_t n = g_learn.n;
  return bb_unet_pair_residual(g_learn.pr + k * n, g_learn.pe + k * n,
    g_learn.pf + k * n, g_bb.stream);
}

/* use (1) or ignore (0) the network in the learned form */
extern "C" void bb_learn_enable(int on)
{
  g_learn.active = on && bb_unet_ready();
}

extern "C" void bb_learn_set_fallback(int window, real drop, int skip)
{
  if(window > 0 && window <= 60) g_learn.stall_window = window;
  if(drop > 0) g_learn.stall_drop = drop;
  if(skip >= 0) g_learn.skip_after_stall = skip;
  g_learn.skip_left = 0;
  g_learn.skip_next = g_learn.skip_after_stall;
}

extern "C" void bb_learn_stats(long *applications, long *fallbacks,
  long *flex_iters, long *jacobi_iters, long *skipped, real *train_seconds)
{
  if(skipped) *skipped = g_learn.skipped;
  if(applications) *applications = g_learn.applications;
  if(fallbacks) *fallbacks = g_learn.fallbacks;
  if(flex_iters) *flex_iters = g_learn.flex_iters;
  if(jacobi_iters) *jacobi_iters = g_learn.jacobi_iters;
  if(train_seconds) *train_seconds = g_learn.train_seconds;
}

extern "C" void bb_learn_reset_stats(void)
{
  g_learn.applications = g_learn.fallbacks = g_learn.skipped = 0;
  g_learn.flex_iters = g_learn.jacobi_iters = 0;
}

void bb_learn_free(void)
{
  learn_free_pairs();
  g_learn.active = 0;
  bb_unet_free();
}

/* FP64 refinement around FP32 BiCGSTAB.  Inner solves run to a relative
 * reduction of `inner_tol` (or `inner_max` iterations); each outer step
 * recomputes the true residual in FP64. */
static double g_inner_tol = 1e-5;
static int g_inner_max = 2000;
static int g_outer_max = 50;

extern "C" void bb_set_mixed_params(double inner_tol, int inner_max, int outer_max)
{
  if(inner_tol > 0) g_inner_tol = inner_tol;
  if(inner_max > 0) g_inner_max = inner_max;
  if(outer_max > 0) g_outer_max = outer_max;
}

static void solve_mixed(const PPGeom &G, const DinvTable &tab)
{
  BBWorkspace *ws = &g_bb.ws;
  hipStream_t st = g_bb.stream;
  size_t n = ws->n;
  int eb = elementwise_blocks(n);
  dim3 blk(PX, PY);
  double *b = _rhs_p[0];
  /* normalise b and x like the plain form (b := rhs/|rhs|) */
  k_scale2<double><<<eb, 256, 0, st>>>(n, ws->sc, ws->x, b);
  BB_CHECK_LAUNCH();
  /* |b| (close to 1) for the tolerance */
  double *rn = ws->red + 24;   /* device: |r| of the current outer step */
  double bnorm = 0.;
  int total_inner = 0, outer = 0, converged = 0;
  double resid = 0.;
  for(outer = 0; outer <= g_outer_max; outer++) {
    bb_prof_begin(BB_K_RESID64);
    k_resid64<<<G.nblk, blk, 0, st>>>(G, tab, b, ws->x, ws->flags, ws->r,
      ws->partials);
    BB_CHECK_LAUNCH();
    bb_prof_end(BB_K_RESID64);
    k_f_rnorm<<<1, 256, 0, st>>>(G.nblk, ws->partials, rn);
    BB_CHECK_LAUNCH();
    if(outer == 0) {
      /* |b|: reuse the residual kernel on x = 0 would cost a pass; the
       * normalised rhs has |b| = 1 up to rounding, computed exactly here */
      k_resid64<<<G.nblk, blk, 0, st>>>(G, tab, b, ws->t /* zero */, ws->flags,
        ws->v, ws->partials);
      BB_CHECK_LAUNCH();
      k_f_rnorm<<<1, 256, 0, st>>>(G.nblk, ws->partials, rn + 1);
      BB_CHECK_LAUNCH();
    }
    BB_CHECK(hipMemcpyAsync(ws->red_host, rn, 2 * sizeof(double),
      hipMemcpyDeviceToHost, st));
    BB_CHECK(hipStreamSynchronize(st));
    if(outer == 0) bnorm = ws->red_host[1];
    resid = ws->red_host[0];
    double tol = pp_residual * bnorm;
    if(resid <= tol) { converged = 1; break; }
    if(outer == g_outer_max || total_inner >= pp_max_iter) break;
    /* inner: A e = r/|r| in FP32 from e = 0 */
    k_to32<<<eb, 256, 0, st>>>(n, rn, ws->r, ws->bf, ws->xf);
    BB_CHECK_LAUNCH();
    double inner_tol = g_inner_tol;
    /* do not solve far below what the outer test needs */
    double need = 0.5 * tol / resid;
    if(need > inner_tol) inner_tol = need;
    int imax = g_inner_max;
    if(imax > pp_max_iter - total_inner) imax = pp_max_iter - total_inner;
    if(g_bb.solver_mode == BB_SOLVER_LEARNED && g_learn.active &&
       g_learn.skip_left <= 0) {
      flexible_run(G, tab, inner_tol, imax);
    } else {
      if(g_bb.solver_mode == BB_SOLVER_LEARNED && g_learn.active) {
        g_learn.skip_left--;
        g_learn.skipped++;
      }
      inner_jacobi(G, tab, ws->bf, ws->xf, inner_tol, imax);
      collect_pair(ws);
    }
    total_inner += ws->sc_host->iter + (ws->sc_host->s_conv ? 1 : 0);
    bb_prof_begin(BB_K_CORRECT64);
    k_correct64<<<eb, 256, 0, st>>>(n, rn, ws->xf, ws->x);
    BB_CHECK_LAUNCH();
    bb_prof_end(BB_K_CORRECT64);
  }
  g_bb.stats.iterations = total_inner;
  g_bb.stats.outer_steps = outer;
  g_bb.stats.residual = bnorm > 0 ? resid / bnorm : resid;
  g_bb.stats.converged = converged;
  ws->sc_host->iter = total_inner;
  ws->sc_host->resid = g_bb.stats.residual;
  ws->sc_host->converged = converged;
}

extern "C" void cuda_PP_bicgstab(int rank)
{
  bb_ensure_workspace();
  BBWorkspace *ws = &g_bb.ws;
  hipStream_t st = g_bb.stream;
  PPGeom G = pp_geom();
  BBGeom B = bb_geom();
  DinvTable tab = make_dinv(G.ddx, G.ddy, G.ddz);
  dim3 blk(PX, PY);

  bb_prof_begin(BB_K_RHS);
  k_pp_setup<<<G.nblk, blk, 0, st>>>(G, B, rho_f / dt, nparts > 0,
    _u_star[0], _v_star[0], _w_star[0], _phi[0], _phase[0], _phase_shell[0],
    _flag_u[0], _flag_v[0], _flag_w[0], _rhs_p[0], ws->x, ws->flags,
    ws->partials);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_RHS);
  k_f_norm<<<1, 256, 0, st>>>(G.nblk, ws->partials, ws->sc);
  BB_CHECK_LAUNCH();

  if(g_bb.solver_mode == BB_SOLVER_MIXED ||
     g_bb.solver_mode == BB_SOLVER_LEARNED) {
    /* a zero FP64 vector for the |b| pass of the refinement loop */
    BB_CHECK(hipMemsetAsync(ws->t, 0, ws->n * sizeof(double), st));
    solve_mixed(G, tab);
  } else {
    solve_fp64(G, tab);
  }

  char nam[FILE_NAME_SIZE] = "solver_expd.rec";
  char namp[FILE_NAME_SIZE] = "solver_prec.rec";
  recorder_bicgstab(rank == 0 ? nam : namp, g_bb.stats.iterations,
    g_bb.stats.residual);
  if(!g_bb.stats.converged) {
    printf("The pressure-Poisson equation did not converge.              \n");
    fflush(stdout);
    if(g_bb.exit_on_fail) exit(EXIT_FAILURE);
  }

  /* unnormalise, subtract the mean over all cells, copy into phi */
  int eb = elementwise_blocks(ws->n);
  k_unscale_sum<<<eb, 256, 0, st>>>(ws->n, ws->sc, ws->x, ws->partials);
  BB_CHECK_LAUNCH();
  k_f_mean<<<1, 256, 0, st>>>(eb, ws->partials, ws->red + 32, (double) ws->n);
  BB_CHECK_LAUNCH();
  k_copy_back<<<G.nblk, blk, 0, st>>>(G, B, ws->red + 32, ws->x, _phi[0]);
  BB_CHECK_LAUNCH();
}
