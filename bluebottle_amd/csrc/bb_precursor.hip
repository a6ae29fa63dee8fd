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

/* ---------------------------------------------------------------------
 * The precursor domain's linear forcing (src/cuda_bluebottle.cu:3817-4120,
 * src/bluebottle_kernel.cu:2054-2131,3849-3906).
 *
 * The reference colocates u, v and w into three cell arrays (twice: once
 * for the energy, once for the means), multiplies them into a fourth,
 * reduces each with Thrust on the host's clock, then adds the field and
 * the constant in six launches.  Here one pass over the cells forms the
 * three cell-centred values in registers and leaves per-block partials of
 * sum(uc), sum(vc), sum(wc) and sum(uc^2 + vc^2 + wc^2); one block sums
 * them in a fixed order and computes k, the means, the gain turbA k0 / k
 * and the three constants on the device; one launch over the three face
 * arrays adds both terms.  Nothing returns to the host unless the entry
 * point returns k.  Each term is rounded where the reference rounds it
 * (no fused multiply-add), so the result differs from a serial sum only
 * by the order of the sums.
 * ------------------------------------------------------------------- */
#define TURB_BLOCKS 512

/* slots of g_bb.ws.red used here */
#define TURB_OUT 40   /* k, umean, vmean, wmean, gain, cu, cv, cw */

__global__ void __launch_bounds__(256)
k_turb_sums(BBGeom G, const double *__restrict__ u,
  const double *__restrict__ v, const double *__restrict__ w,
  double *__restrict__ part)
{
#pragma clang fp contract(off)
  __shared__ double sh[4];
  const long n = (long) G.xn * G.yn * G.zn;
  double su = 0., sv = 0., sw = 0., se = 0.;
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < n;
      q += (long) gridDim.x * 256) {
    const int i = (int) (q % G.xn) + 1;
    const int j = (int) ((q / G.xn) % G.yn) + 1;
    const int k = (int) (q / ((long) G.xn * G.yn)) + 1;
    const int cx = i + j * G.us1 + k * G.us2;
    const int cy = i + j * G.vs1 + k * G.vs2;
    const int cz = i + j * G.ws1 + k * G.ws2;
    const double uc = 0.5 * (u[cx] + u[cx + 1]);
    const double vc = 0.5 * (v[cy] + v[cy + G.vs1]);
    const double wc = 0.5 * (w[cz] + w[cz + G.ws2]);
    su += uc;
    sv += vc;
    sw += wc;
    se += uc * uc + vc * vc + wc * wc;
  }
  const double a = bb_block_sum(su, sh);
  const double b = bb_block_sum(sv, sh);
  const double c = bb_block_sum(sw, sh);
  const double e = bb_block_sum(se, sh);
  if(threadIdx.x == 0) {
    part[4 * blockIdx.x + 0] = a;
    part[4 * blockIdx.x + 1] = b;
    part[4 * blockIdx.x + 2] = c;
    part[4 * blockIdx.x + 3] = e;
  }
}

/* one block: out[] = k, the three means, the gain and the three constants,
 * each in the reference's operation order */
__global__ void __launch_bounds__(256)
k_turb_final(int nblk, double ncell, double turbA_, double turbl_,
  const double *__restrict__ part, double *__restrict__ out)
{
#pragma clang fp contract(off)
  __shared__ double sh[4];
  double a = 0., b = 0., c = 0., e = 0.;
  for(int q = threadIdx.x; q < nblk; q += 256) {
    a += part[4 * q];
    b += part[4 * q + 1];
    c += part[4 * q + 2];
    e += part[4 * q + 3];
  }
  a = bb_block_sum(a, sh);
  b = bb_block_sum(b, sh);
  c = bb_block_sum(c, sh);
  e = bb_block_sum(e, sh);
  if(threadIdx.x != 0) return;
  const double k = e / ncell * 0.5;
  const double um = a / ncell, vm = b / ncell, wm = c / ncell;
  const double k0 = 13.5 * turbA_ * turbA_ * 0.0361 * turbl_ * turbl_;
  out[0] = k;
  out[1] = um;
  out[2] = vm;
  out[3] = wm;
  out[4] = turbA_ * k0 / k;
  out[5] = -turbA_ * k0 / k * um;
  out[6] = -turbA_ * k0 / k * vm;
  out[7] = -turbA_ * k0 / k * wm;
}

/* f += gain * vel, then f += constant, over every value of the three
 * ghosted face arrays (one flat index over u's, then v's, then w's) */
__global__ void __launch_bounds__(256)
k_turb_apply(int nu_, int nv_, int nw_, const double *__restrict__ u,
  const double *__restrict__ v, const double *__restrict__ w,
  double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz,
  const double *__restrict__ out)
{
#pragma clang fp contract(off)
  const long n = (long) nu_ + nv_ + nw_;
  const double gain = out[4];
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < n;
      q += (long) gridDim.x * 256) {
    const double *x;
    double *f;
    double c;
    long r = q;
    if(r < nu_) { x = u; f = fx; c = out[5]; }
    else if((r -= nu_) < nv_) { x = v; f = fy; c = out[6]; }
    else { r -= nv_; x = w; f = fz; c = out[7]; }
    double t = f[r] + gain * x[r];
    f[r] = t + c;
  }
}

/* the sums and the constants, left in g_bb.ws.red + TURB_OUT */
static void turb_reduce(void)
{
  bb_ensure_workspace();
  BBGeom G = bb_geom();
  k_turb_sums<<<TURB_BLOCKS, 256, 0, g_bb.stream>>>(G, _u[0], _v[0], _w[0],
    g_bb.ws.partials);
  BB_CHECK_LAUNCH();
  const double ncell = (double) G.xn * G.yn * G.zn;
  k_turb_final<<<1, 256, 0, g_bb.stream>>>(TURB_BLOCKS, ncell, turbA, turbl,
    g_bb.ws.partials, g_bb.ws.red + TURB_OUT);
  BB_CHECK_LAUNCH();
}

extern "C" real cuda_compute_energy(void)
{
  bb_prof_begin(BB_K_MISC);
  turb_reduce();
  bb_prof_end(BB_K_MISC);
  BB_CHECK(hipMemcpyAsync(g_bb.ws.red_host, g_bb.ws.red + TURB_OUT,
    sizeof(double), hipMemcpyDeviceToHost, g_bb.stream));
  BB_CHECK(hipStreamSynchronize(g_bb.stream));
  return g_bb.ws.red_host[0];
}

extern "C" void cuda_compute_turb_forcing(void)
{
  bb_prof_begin(BB_K_MISC);
  turb_reduce();
  BBGeom G = bb_geom();
  const long n = (long) G.ucount + G.vcount + G.wcount;
  long nb = (n + 255) / 256;
  if(nb > 4096) nb = 4096;
  k_turb_apply<<<(int) nb, 256, 0, g_bb.stream>>>(G.ucount, G.vcount,
    G.wcount, _u[0], _v[0], _w[0], _f_x[0], _f_y[0], _f_z[0],
    g_bb.ws.red + TURB_OUT);
  BB_CHECK_LAUNCH();
  bb_prof_end(BB_K_MISC);
}

/* colocate_Gf{x,y,z}: cell c (interior index, no ghosts) gets the mean of
 * the two faces of component `comp` that bound it; the geometry is read
 * from the device copy of the subdomain the caller passes */
__global__ void __launch_bounds__(256)
k_colocate(int comp, const dom_struct *__restrict__ d,
  const double *__restrict__ f, double *__restrict__ out)
{
  const grid_info &G = comp == 0 ? d->Gfx : comp == 1 ? d->Gfy : d->Gfz;
  const int nx = d->Gcc.in, ny = d->Gcc.jn, nz = d->Gcc.kn;
  const long n = (long) nx * ny * nz;
  const int step = comp == 0 ? 1 : comp == 1 ? G.s1b : G.s2b;
  for(long q = (long) blockIdx.x * 256 + threadIdx.x; q < n;
      q += (long) gridDim.x * 256) {
    const int i = (int) (q % nx), j = (int) ((q / nx) % ny);
    const int k = (int) (q / ((long) nx * ny));
    const int C = (i + DOM_BUF) + (j + DOM_BUF) * G.s1b + (k + DOM_BUF) * G.s2b;
    out[i + j * d->Gcc.s1 + k * d->Gcc.s2] = 0.5 * (f[C] + f[C + step]);
  }
}

static void colocate(int comp, real *f, real *out, dom_struct *d_dom)
{
  bb_ensure_context();
  if(f == NULL || out == NULL || d_dom == NULL) {
    fprintf(stderr, "cuda_colocate: null device pointer\n");
    exit(EXIT_FAILURE);
  }
  const long n = (long) dom[0].Gcc.in * dom[0].Gcc.jn * dom[0].Gcc.kn;
  long nb = (n + 255) / 256;
  if(nb > 4096) nb = 4096;
  if(nb < 1) nb = 1;
  k_colocate<<<(int) nb, 256, 0, g_bb.stream>>>(comp, d_dom, f, out);
  BB_CHECK_LAUNCH();
}

extern "C" void cuda_colocate_Gfx(real *_u_, real *_u_co, dom_struct *_dom_)
{
  colocate(0, _u_, _u_co, _dom_);
}

extern "C" void cuda_colocate_Gfy(real *_v_, real *_v_co, dom_struct *_dom_)
{
  colocate(1, _v_, _v_co, _dom_);
}

extern "C" void cuda_colocate_Gfz(real *_w_, real *_w_co, dom_struct *_dom_)
{
  colocate(2, _w_, _w_co, _dom_);
}
