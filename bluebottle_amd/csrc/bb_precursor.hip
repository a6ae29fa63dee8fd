/*
 * bb_precursor.hip -- the one-GPU part of the precursor (turbulent inflow)
 * path: the 15 inflow planes and their device copies
 * (src/cuda_bluebottle.cu:99-129,239-284,1480-1544), the host <-> device
 * plane copies cuda_dom_turb_planes_push / _pull (src/cuda_bluebottle.cu:
 * 536-1415) and the plane extraction of the precursor domain,
 * cuda_yank_turb_planes (src/cuda_bluebottle.cu:4556-4820,
 * src/bluebottle_kernel.cu:3619-3851).  The ghost fill from the planes is
 * in bb_field.hip (PRECURSOR faces of cuda_dom_BC / cuda_dom_BC_star).
 * The transfer of planes between the precursor and the main domain (MPI
 * in the reference) is the driver's.
 *
 * Plane layouts (the transverse ghosted extents of the component's grid):
 *   W/E planes  j + k * jnb      S/N planes  k + i * knb
 *   B/T planes  i + j * inb
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "bb_device.h"

/* every plane: its driver globals, grid (1 Gfx, 2 Gfy, 3 Gfz), the axis it
 * is normal to, and the pair of config entries (face pair of the
 * component) whose PRECURSOR setting selects it for push and pull */
struct PlaneDesc { real ***dev; real **host; int grid, axis, cfg; };

static const PlaneDesc *planes(int *count)
{
  static const PlaneDesc P[] = {
    {&_u_WE, &u_WE, 1, 0, 0}, {&_u_SN_S, &u_SN_S, 1, 1, 2},
    {&_u_SN_N, &u_SN_N, 1, 1, 2}, {&_u_BT_B, &u_BT_B, 1, 2, 4},
    {&_u_BT_T, &u_BT_T, 1, 2, 4},
    {&_v_WE_W, &v_WE_W, 2, 0, 6}, {&_v_WE_E, &v_WE_E, 2, 0, 6},
    {&_v_SN, &v_SN, 2, 1, 8}, {&_v_BT_B, &v_BT_B, 2, 2, 10},
    {&_v_BT_T, &v_BT_T, 2, 2, 10},
    {&_w_WE_W, &w_WE_W, 3, 0, 12}, {&_w_WE_E, &w_WE_E, 3, 0, 12},
    {&_w_SN_S, &w_SN_S, 3, 1, 14}, {&_w_SN_N, &w_SN_N, 3, 1, 14},
    {&_w_BT, &w_BT, 3, 2, 16},
  };
  *count = (int) (sizeof(P) / sizeof(P[0]));
  return P;
}

static size_t plane_count(const PlaneDesc &p)
{
  const dom_struct *d = &dom[0];
  const grid_info *G = p.grid == 1 ? &d->Gfx : p.grid == 2 ? &d->Gfy : &d->Gfz;
  if(p.axis == 0) return (size_t) G->jnb * G->knb;
  if(p.axis == 1) return (size_t) G->inb * G->knb;
  return (size_t) G->inb * G->jnb;
}

/* called by cuda_dom_malloc: the reference allocates every plane whatever
 * the boundary types */
void bb_alloc_planes(void)
{
  int n;
  const PlaneDesc *P = planes(&n);
  for(int q = 0; q < n; q++) {
    const size_t bytes = plane_count(P[q]) * sizeof(real);
    *P[q].dev = (real **) malloc(nsubdom * sizeof(real *));
    BB_CHECK(hipMalloc((void **) &(*P[q].dev)[0], bytes));
    BB_CHECK(hipMemsetAsync((*P[q].dev)[0], 0, bytes, g_bb.stream));
  }
}

void bb_free_planes(void)
{
  int n;
  const PlaneDesc *P = planes(&n);
  for(int q = 0; q < n; q++) {
    if(*P[q].dev == NULL) continue;
    BB_CHECK(hipFree((*P[q].dev)[0]));
    free(*P[q].dev);
    *P[q].dev = NULL;
  }
}

/* bc_configs: the 18 velocity BC types, u W,E,S,N,B,T then v, then w */
static void copy_planes(const int *bc_configs, int to_device)
{
  bb_ensure_context();
  int n;
  const PlaneDesc *P = planes(&n);
  for(int q = 0; q < n; q++) {
    const int c = P[q].cfg;
    if(bc_configs[c] != PRECURSOR && bc_configs[c + 1] != PRECURSOR) continue;
    if(*P[q].host == NULL || *P[q].dev == NULL) continue;
    const size_t bytes = plane_count(P[q]) * sizeof(real);
    if(to_device) bb_h2d((*P[q].dev)[0], *P[q].host, bytes);
    else bb_d2h(*P[q].host, (*P[q].dev)[0], bytes);
  }
}

extern "C" void cuda_dom_turb_planes_push(int *bc_configs)
{
  copy_planes(bc_configs, 1);
}

extern "C" void cuda_dom_turb_planes_pull(int *bc_configs)
{
  copy_planes(bc_configs, 0);
}

/* ---------------------------------------------------------------------
 * Plane extraction.  One thread per plane entry.  Along the plane's
 * normal axis, the cell holding `pos` (periodic wrap of the index as in
 * the reference) and the face values either side give, by linear
 * interpolation: for the normal component the value at pos plus `vel`;
 * for a tangential component the values half a cell before and after
 * pos, into the low and high planes.  The reference's particulars are
 * kept: the B/T normal component scales by 1/dx, and the S/N w plane
 * stores the slope term alone and writes it back into w (the assignment
 * `plane_n = w[M] = ...`).  The reference's S/N u extraction bounds its k
 * index by inb instead of knb: where inb < knb the rows past inb are left
 * as they were, as there; where inb > knb (the usual case, xn + 3 against
 * zn + 2) the reference reads and writes a row past its arrays, which
 * here is not done (the departure).
 * ------------------------------------------------------------------- */
struct YankArgs {
  double *f;
  double *pl_lo, *pl_hi;     /* pl_hi unused for the normal component */
  long str, s1, s2;          /* strides: along the normal, transverse */
  int n1, n2;                /* plane extents (transverse, ghosted) */
  long p1, p2;               /* plane strides of the two transverse axes */
  int start, end, nb;        /* the grid's s, e, nb along the normal */
  int normal;                /* the component is the normal one */
  int write_back;            /* the S/N w quirk */
  double pos, vel, h, xs, ddh;
};

__global__ void __launch_bounds__(256) k_yank(YankArgs A)
{
  /* separately rounded products and sums, as the reference's */
#pragma clang fp contract(off)
  const int t1 = blockIdx.x * blockDim.x + threadIdx.x;
  const int t2 = blockIdx.y * blockDim.y + threadIdx.y;
  if(t1 >= A.n1 || t2 >= A.n2) return;
  double q = floor((A.pos - A.xs) * A.ddh) + DOM_BUF;
  if(q < A.start) q += A.nb;
  if(q > A.end - 1) q -= A.nb;
  /* a plane position inside the domain needs no clamp; one outside (which
   * the reference reads out of bounds for) is held to the grid */
  const int i = (int) fmin(fmax(q, A.normal ? 0. : 1.), (double) (A.nb - 2));
  double *F = A.f + (long) t1 * A.s1 + (long) t2 * A.s2;
  const long P = (long) t1 * A.p1 + (long) t2 * A.p2;
  if(A.normal) {
    const double xx = (i - DOM_BUF) * A.h + A.xs;
    const double W = F[i * A.str], E = F[(i + 1) * A.str];
    const double d = (E - W) * A.ddh;
    A.pl_lo[P] = W + d * (A.pos - xx) + A.vel;
  } else {
    const double xw = (i - DOM_BUF - 0.5) * A.h + A.xs;
    const double xe = (i - DOM_BUF + 0.5) * A.h + A.xs;
    const double W = F[(i - 1) * A.str], M = F[i * A.str];
    const double E = F[(i + 1) * A.str];
    const double dw = (M - W) * A.ddh, de = (E - M) * A.ddh;
    A.pl_lo[P] = W + dw * (A.pos - 0.5 * A.h - xw);
    if(A.write_back) {
      const double v = de * (A.pos + 0.5 * A.h - xe);
      F[i * A.str] = v;
      A.pl_hi[P] = v;
    } else {
      A.pl_hi[P] = M + de * (A.pos + 0.5 * A.h - xe);
    }
  }
}

static void yank(double *f, const grid_info &G, int axis, int normal,
  double *lo, double *hi, double pos, double vel, double h, double xs,
  double ddh, int write_back, int n2_cap = 0)
{
  YankArgs A;
  A.f = f; A.pl_lo = lo; A.pl_hi = hi;
  A.normal = normal; A.write_back = write_back;
  A.pos = pos; A.vel = vel; A.h = h; A.xs = xs; A.ddh = ddh;
  if(axis == 0) {
    A.str = 1; A.s1 = G.s1b; A.s2 = G.s2b; A.n1 = G.jnb; A.n2 = G.knb;
    A.p1 = 1; A.p2 = G.jnb;
    A.start = G.is; A.end = G.ie; A.nb = G.inb;
  } else if(axis == 1) {
    A.str = G.s1b; A.s1 = 1; A.s2 = G.s2b; A.n1 = G.inb; A.n2 = G.knb;
    A.p1 = G.knb; A.p2 = 1;
    A.start = G.js; A.end = G.je; A.nb = G.jnb;
  } else {
    A.str = G.s2b; A.s1 = 1; A.s2 = G.s1b; A.n1 = G.inb; A.n2 = G.jnb;
    A.p1 = 1; A.p2 = G.inb;
    A.start = G.ks; A.end = G.ke; A.nb = G.knb;
  }
  if(n2_cap > 0 && n2_cap < A.n2) A.n2 = n2_cap;
  dim3 blk(64, 4);
  dim3 grd((A.n1 + 63) / 64, (A.n2 + 3) / 4);
  k_yank<<<grd, blk, 0, g_bb.stream>>>(A);
  BB_CHECK_LAUNCH();
}

/* bc_flow_configs: the 18 velocity BC types of the main domain; pos: the
 * plane positions (x of the W/E planes in pos[0], y of S/N in pos[4], z of
 * B/T in pos[8]); vel: the velocities added to the normal components
 * (vel[0] or vel[1] for W/E, vel[8] or vel[9] for S/N, vel[16] or vel[17]
 * for B/T, by which face of the pair is PRECURSOR) */
extern "C" void cuda_yank_turb_planes(int *bc_flow_configs, real *pos,
  real *vel)
{
  bb_ensure_context();
  const int *c = bc_flow_configs;
  const dom_struct *d = &dom[0];
  const double ddx = 1. / d->dx, ddy = 1. / d->dy, ddz = 1. / d->dz;
  /* W/E planes */
  if(c[0] == PRECURSOR)
    yank(_u[0], d->Gfx, 0, 1, _u_WE[0], NULL, pos[0], vel[0], d->dx, d->xs,
      ddx, 0);
  else if(c[1] == PRECURSOR)
    yank(_u[0], d->Gfx, 0, 1, _u_WE[0], NULL, pos[0], vel[1], d->dx, d->xs,
      ddx, 0);
  if(c[6] == PRECURSOR)
    yank(_v[0], d->Gfy, 0, 0, _v_WE_W[0], _v_WE_E[0], pos[0], vel[0], d->dx,
      d->xs, ddx, 0);
  if(c[7] == PRECURSOR)
    yank(_v[0], d->Gfy, 0, 0, _v_WE_W[0], _v_WE_E[0], pos[0], vel[1], d->dx,
      d->xs, ddx, 0);
  if(c[12] == PRECURSOR)
    yank(_w[0], d->Gfz, 0, 0, _w_WE_W[0], _w_WE_E[0], pos[0], vel[0], d->dx,
      d->xs, ddx, 0);
  if(c[13] == PRECURSOR)
    yank(_w[0], d->Gfz, 0, 0, _w_WE_W[0], _w_WE_E[0], pos[0], vel[1], d->dx,
      d->xs, ddx, 0);
  /* S/N planes */
  if(c[2] == PRECURSOR)
    yank(_u[0], d->Gfx, 1, 0, _u_SN_S[0], _u_SN_N[0], pos[4], vel[8], d->dy,
      d->ys, ddy, 0, d->Gfx.inb);
  else if(c[3] == PRECURSOR)
    yank(_u[0], d->Gfx, 1, 0, _u_SN_S[0], _u_SN_N[0], pos[4], vel[9], d->dy,
      d->ys, ddy, 0, d->Gfx.inb);
  if(c[8] == PRECURSOR)
    yank(_v[0], d->Gfy, 1, 1, _v_SN[0], NULL, pos[4], vel[8], d->dy, d->ys,
      ddy, 0);
  if(c[9] == PRECURSOR)
    yank(_v[0], d->Gfy, 1, 1, _v_SN[0], NULL, pos[4], vel[9], d->dy, d->ys,
      ddy, 0);
  /* the reference's N-face call passes _w_SN_N as both planes */
  if(c[14] == PRECURSOR)
    yank(_w[0], d->Gfz, 1, 0, _w_SN_S[0], _w_SN_N[0], pos[4], vel[8], d->dy,
      d->ys, ddy, 1);
  if(c[15] == PRECURSOR)
    yank(_w[0], d->Gfz, 1, 0, _w_SN_N[0], _w_SN_N[0], pos[4], vel[9], d->dy,
      d->ys, ddy, 1);
  /* B/T planes */
  if(c[4] == PRECURSOR)
    yank(_u[0], d->Gfx, 2, 0, _u_BT_B[0], _u_BT_T[0], pos[8], vel[16], d->dz,
      d->zs, ddz, 0);
  else if(c[5] == PRECURSOR)
    yank(_u[0], d->Gfx, 2, 0, _u_BT_B[0], _u_BT_T[0], pos[8], vel[17], d->dz,
      d->zs, ddz, 0);
  if(c[10] == PRECURSOR)
    yank(_v[0], d->Gfy, 2, 0, _v_BT_B[0], _v_BT_T[0], pos[8], vel[16], d->dz,
      d->zs, ddz, 0);
  if(c[11] == PRECURSOR)
    yank(_v[0], d->Gfy, 2, 0, _v_BT_B[0], _v_BT_T[0], pos[8], vel[17], d->dz,
      d->zs, ddz, 0);
  /* the reference scales the B/T normal component by 1/dx */
  if(c[16] == PRECURSOR)
    yank(_w[0], d->Gfz, 2, 1, _w_BT[0], NULL, pos[8], vel[16], d->dz, d->zs,
      1. / d->dx, 0);
  if(c[17] == PRECURSOR)
    yank(_w[0], d->Gfz, 2, 1, _w_BT[0], NULL, pos[8], vel[17], d->dz, d->zs,
      1. / d->dx, 0);
}
