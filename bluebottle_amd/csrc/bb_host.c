/*
 * bb_host.c -- driver side of the C ABI: the global state the reference's
 * driver defines (src/bluebottle.c:29-187, src/particle.c:29-75,
 * src/domain.c:27), the host set-up the reference does in domain_init /
 * parts_init / binDom_init (src/domain.c:1591-2347, src/particle.c:295-762),
 * the solver telemetry sink recorder_bicgstab (src/recorder.c:1927-1953),
 * and the per-step call sequence of src/bluebottle.c:445-545 and 598-711.
 *
 * The reference's own driver can replace this file unchanged: the HIP
 * library (libbluebottle_hip.so) only references these symbols.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "../../include/bluebottle.h"
#include "bb_host.h"

/* ---- src/bluebottle.c:29-187 ---- */
int dev_start;
int dev_end;
dom_struct *dom;
dom_struct **_dom;
dom_struct Dom;
real *p0, *p, *phi;
real **_p0, **_p, **_phi;
real *u, *u0, *v, *v0, *w, *w0;
real **_u, **_u0, **_v, **_v0, **_w, **_w0;
real *f_x, *f_y, *f_z;
real **_f_x, **_f_y, **_f_z;
real *diff0_u, *diff0_v, *diff0_w;
real **_diff0_u, **_diff0_v, **_diff0_w;
real *conv0_u, *conv0_v, *conv0_w;
real **_conv0_u, **_conv0_v, **_conv0_w;
real *diff_u, *diff_v, *diff_w;
real **_diff_u, **_diff_v, **_diff_w;
real *conv_u, *conv_v, *conv_w;
real **_conv_u, **_conv_v, **_conv_w;
real **_u_star, **_v_star, **_w_star;
real *u_star, *v_star, *w_star;
real *u_WE, *u_SN_S, *u_SN_N, *u_BT_B, *u_BT_T;
real *v_WE_W, *v_WE_E, *v_SN, *v_BT_B, *v_BT_T;
real *w_WE_W, *w_WE_E, *w_SN_S, *w_SN_N, *w_BT;
real **_u_WE, **_u_SN_S, **_u_SN_N, **_u_BT_B, **_u_BT_T;
real **_v_WE_W, **_v_WE_E, **_v_SN, **_v_BT_B, **_v_BT_T;
real **_w_WE_W, **_w_WE_E, **_w_SN_S, **_w_SN_N, **_w_BT;
real **_rhs_p;
real duration, ttime, vel_tDelay, p_tDelay, g_tDelay;
real dt, dt0, CFL;
int pp_max_iter;
real pp_residual;
int lamb_max_iter;
real lamb_residual, lamb_relax, lamb_cut;
int out_plane;
int stepnum;
g_struct g;
real rho_f, mu, nu;
BC bc;
int init_cond;
gradP_struct gradP;
long int cpumem, gpumem;
real pid_int, pid_back, Kp, Ki, Kd;
real turbA, turbl;
/* ---- src/domain.c:27 ---- */
int nsubdom;
/* ---- src/particle.c:29-75 ---- */
int *phase, *phase_shell;
int **_phase, **_phase_shell;
int *flag_u, *flag_v, *flag_w;
int **_flag_u, **_flag_v, **_flag_w;
int nparts;
real interactionLengthRatio;
part_struct *parts;
part_struct **_parts;
dom_struct binDom;
dom_struct *_binDom;
int coeff_stride;
real *pnm_re, *pnm_im, *phinm_re, *phinm_im, *chinm_re, *chinm_im;
real *pnm_re0, *pnm_im0, *phinm_re0, *phinm_im0, *chinm_re0, *chinm_im0;
real *pnm_re00, *pnm_im00, *phinm_re00, *phinm_im00, *chinm_re00, *chinm_im00;
real **_pnm_re, **_pnm_im, **_phinm_re, **_phinm_im, **_chinm_re, **_chinm_im;
real **_pnm_re0, **_pnm_im0, **_phinm_re0, **_phinm_im0, **_chinm_re0,
  **_chinm_im0;
real **_pnm_re00, **_pnm_im00, **_phinm_re00, **_phinm_im00, **_chinm_re00,
  **_chinm_im00;

/* ---- solver telemetry (src/recorder.c:1927-1953) ---- */
char bb_record_dir[FILE_NAME_SIZE] = "";
int bb_rec_niter = 0;
real bb_rec_resid = 0.;
int bb_rec_calls = 0;
long bb_rec_total_iters = 0;

void recorder_bicgstab(char *name, int niter, real resid)
{
  bb_rec_niter = niter;
  bb_rec_resid = resid;
  bb_rec_calls++;
  bb_rec_total_iters += niter;
  if(bb_record_dir[0] == '\0') return;
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/%s", bb_record_dir, name);
  FILE *rec = fopen(path, "r+");
  if(rec == NULL) {
    rec = fopen(path, "w");
    if(rec == NULL) return;
    fprintf(rec, "%-12s", "stepnum");
    fprintf(rec, "%-15s", "ttime");
    fprintf(rec, "%-15s", "dt");
    fprintf(rec, "%-8s", "niter");
    fprintf(rec, "%-15s", "resid");
  }
  fseek(rec, 0, SEEK_END);
  fprintf(rec, "\n");
  fprintf(rec, "%-12d", stepnum);
  fprintf(rec, "%-15e", ttime);
  fprintf(rec, "%-15e", dt);
  fprintf(rec, "%-8d", niter);
  fprintf(rec, "%-15e", resid);
  fclose(rec);
}

/* ---- grids: src/domain.c:1611-1725 and the device copies at 1777-1990 */
static void set_grid(grid_info *G, int nx, int ny, int nz)
{
  G->is = DOM_BUF; G->isb = G->is - DOM_BUF; G->in = nx;
  G->inb = G->in + 2 * DOM_BUF; G->ie = G->is + G->in; G->ieb = G->ie + DOM_BUF;
  G->js = DOM_BUF; G->jsb = G->js - DOM_BUF; G->jn = ny;
  G->jnb = G->jn + 2 * DOM_BUF; G->je = G->js + G->jn; G->jeb = G->je + DOM_BUF;
  G->ks = DOM_BUF; G->ksb = G->ks - DOM_BUF; G->kn = nz;
  G->knb = G->kn + 2 * DOM_BUF; G->ke = DOM_BUF + G->kn; G->keb = G->ke + DOM_BUF;
  G->s1 = G->in; G->s2 = G->s1 * G->jn; G->s3 = G->s2 * G->kn;
  G->s1b = G->inb; G->s2b = G->s1b * G->jnb; G->s3b = G->s2b * G->knb;
  G->_is = G->is; G->_isb = G->isb; G->_in = G->in; G->_inb = G->inb;
  G->_ie = G->ie; G->_ieb = G->ieb;
  G->_js = G->js; G->_jsb = G->jsb; G->_jn = G->jn; G->_jnb = G->jnb;
  G->_je = G->je; G->_jeb = G->jeb;
  G->_ks = G->ks; G->_ksb = G->ksb; G->_kn = G->kn; G->_knb = G->knb;
  G->_ke = G->ke; G->_keb = G->keb;
  G->_s1 = G->s1; G->_s2 = G->s2; G->_s3 = G->s3;
  G->_s1b = G->s1b; G->_s2b = G->s2b; G->_s3b = G->s3b;
}

static void *zalloc(size_t n, size_t sz)
{
  void *ptr = calloc(n ? n : 1, sz);
  if(!ptr) { fprintf(stderr, "bb_host: out of host memory\n"); exit(EXIT_FAILURE); }
  cpumem += (long) (n * sz);
  return ptr;
}

int bb_host_domain(real xs, real xe, int xn, real ys, real ye, int yn,
  real zs, real ze, int zn)
{
  memset(&Dom, 0, sizeof(Dom));
  Dom.xs = xs; Dom.xe = xe; Dom.xn = xn;
  Dom.ys = ys; Dom.ye = ye; Dom.yn = yn;
  Dom.zs = zs; Dom.ze = ze; Dom.zn = zn;
  Dom.xl = Dom.xe - Dom.xs; Dom.yl = Dom.ye - Dom.ys; Dom.zl = Dom.ze - Dom.zs;
  Dom.dx = Dom.xl / Dom.xn; Dom.dy = Dom.yl / Dom.yn; Dom.dz = Dom.zl / Dom.zn;
  set_grid(&Dom.Gcc, xn, yn, zn);
  set_grid(&Dom.Gfx, xn + 1, yn, zn);
  set_grid(&Dom.Gfy, xn, yn + 1, zn);
  set_grid(&Dom.Gfz, xn, yn, zn + 1);
  Dom.E = Dom.W = Dom.N = Dom.S = Dom.T = Dom.B = -1;
  nsubdom = 1;
  dev_start = dev_end = 0;
  dom = (dom_struct *) zalloc(1, sizeof(dom_struct));
  dom[0] = Dom;

  size_t cc = Dom.Gcc.s3b, fx = Dom.Gfx.s3b, fy = Dom.Gfy.s3b, fz = Dom.Gfz.s3b;
  p0 = zalloc(cc, sizeof(real)); p = zalloc(cc, sizeof(real));
  phi = zalloc(cc, sizeof(real));
  u = zalloc(fx, sizeof(real)); u0 = zalloc(fx, sizeof(real));
  v = zalloc(fy, sizeof(real)); v0 = zalloc(fy, sizeof(real));
  w = zalloc(fz, sizeof(real)); w0 = zalloc(fz, sizeof(real));
  f_x = zalloc(fx, sizeof(real)); f_y = zalloc(fy, sizeof(real));
  f_z = zalloc(fz, sizeof(real));
  diff0_u = zalloc(fx, sizeof(real)); diff0_v = zalloc(fy, sizeof(real));
  diff0_w = zalloc(fz, sizeof(real));
  conv0_u = zalloc(fx, sizeof(real)); conv0_v = zalloc(fy, sizeof(real));
  conv0_w = zalloc(fz, sizeof(real));
  diff_u = zalloc(fx, sizeof(real)); diff_v = zalloc(fy, sizeof(real));
  diff_w = zalloc(fz, sizeof(real));
  conv_u = zalloc(fx, sizeof(real)); conv_v = zalloc(fy, sizeof(real));
  conv_w = zalloc(fz, sizeof(real));
  u_star = zalloc(fx, sizeof(real)); v_star = zalloc(fy, sizeof(real));
  w_star = zalloc(fz, sizeof(real));
  /* the precursor inflow planes (src/domain.c:2058-2087) */
  {
    const grid_info *X = &Dom.Gfx, *Y = &Dom.Gfy, *Z = &Dom.Gfz;
    u_WE = zalloc((size_t) X->jnb * X->knb, sizeof(real));
    u_SN_S = zalloc((size_t) X->inb * X->knb, sizeof(real));
    u_SN_N = zalloc((size_t) X->inb * X->knb, sizeof(real));
    u_BT_B = zalloc((size_t) X->inb * X->jnb, sizeof(real));
    u_BT_T = zalloc((size_t) X->inb * X->jnb, sizeof(real));
    v_WE_W = zalloc((size_t) Y->jnb * Y->knb, sizeof(real));
    v_WE_E = zalloc((size_t) Y->jnb * Y->knb, sizeof(real));
    v_SN = zalloc((size_t) Y->inb * Y->knb, sizeof(real));
    v_BT_B = zalloc((size_t) Y->inb * Y->jnb, sizeof(real));
    v_BT_T = zalloc((size_t) Y->inb * Y->jnb, sizeof(real));
    w_WE_W = zalloc((size_t) Z->jnb * Z->knb, sizeof(real));
    w_WE_E = zalloc((size_t) Z->jnb * Z->knb, sizeof(real));
    w_SN_S = zalloc((size_t) Z->inb * Z->knb, sizeof(real));
    w_SN_N = zalloc((size_t) Z->inb * Z->knb, sizeof(real));
    w_BT = zalloc((size_t) Z->inb * Z->jnb, sizeof(real));
  }
  phase = zalloc(cc, sizeof(int)); phase_shell = zalloc(cc, sizeof(int));
  flag_u = zalloc(fx, sizeof(int)); flag_v = zalloc(fy, sizeof(int));
  flag_w = zalloc(fz, sizeof(int));
  /* src/domain.c:2320-2332: the initial dt estimate and Euler first step */
  real dx_min = Dom.dx;
  if(Dom.dy < dx_min) dx_min = Dom.dy;
  if(Dom.dz < dx_min) dx_min = Dom.dz;
  dt = 2. * nu / (dx_min * dx_min);
  dt += 2. * nu / (Dom.dy * Dom.dy);
  dt += 2. * nu / (Dom.dz * Dom.dz);
  dt = CFL / dt;
  dt0 = -1.;
  stepnum = 0;
  ttime = 0.;
  nparts = 0;
  parts = NULL;
  coeff_stride = 0;
  turbl = 0.;
  return 0;
}

int bb_host_particles(int n)
{
  nparts = n;
  free(parts);
  parts = (part_struct *) zalloc((size_t) (n > 0 ? n : 1), sizeof(part_struct));
  return 0;
}

/* src/particle.c:295-535 (QUIESCENT initial state), flags_reset
 * (src/particle.c:559-596), binDom_init (src/particle.c:598-762). */
int bb_host_particles_init(void)
{
  for(int k = Dom.Gcc.ks; k < Dom.Gcc.ke; k++)
  for(int j = Dom.Gcc.js; j < Dom.Gcc.je; j++)
  for(int i = Dom.Gcc.is; i < Dom.Gcc.ie; i++) {
    int C = i + j*Dom.Gcc.s1b + k*Dom.Gcc.s2b;
    phase[C] = 1;
    phase_shell[C] = 1;
  }
  for(int n = 0; n < Dom.Gfx.s3b; n++) flag_u[n] = 1;
  for(int n = 0; n < Dom.Gfy.s3b; n++) flag_v[n] = 1;
  for(int n = 0; n < Dom.Gfz.s3b; n++) flag_w[n] = 1;

  coeff_stride = 0;
  for(int i = 0; i < nparts; i++) {
    part_struct *P = &parts[i];
    memset(&P->cage, 0, sizeof(cage_struct));
    P->rs = P->rs * P->r;
    P->ncoeff = 0;
    for(int j = 0; j <= P->order; j++) P->ncoeff += j + 1;
    P->x0 = P->x; P->y0 = P->y; P->z0 = P->z;
    P->u = P->v = P->w = 0.; P->u0 = P->v0 = P->w0 = 0.;
    P->udot = P->vdot = P->wdot = 0.; P->udot0 = P->vdot0 = P->wdot0 = 0.;
    P->axx = 1.; P->axy = 0.; P->axz = 0.;
    P->ayx = 0.; P->ayy = 1.; P->ayz = 0.;
    P->azx = 0.; P->azy = 0.; P->azz = 1.;
    P->ox = P->oy = P->oz = 0.; P->ox0 = P->oy0 = P->oz0 = 0.;
    P->oxdot = P->oydot = P->ozdot = 0.;
    P->oxdot0 = P->oydot0 = P->ozdot0 = 0.;
    P->Fx = P->Fy = P->Fz = 0.; P->Lx = P->Ly = P->Lz = 0.;
    P->kFx = P->kFy = P->kFz = 0.;
    P->iFx = P->iFy = P->iFz = 0.; P->iLx = P->iLy = P->iLz = 0.;
    if(P->ncoeff > coeff_stride) coeff_stride = P->ncoeff;
    for(int j = 0; j < NNODES; j++) P->nodes[j] = -1;
    for(int j = 0; j < MAX_NEIGHBORS; j++) { P->St[j] = 0.; P->iSt[j] = -1; }
  }
  size_t nc = (size_t) coeff_stride * nparts;
  real **cf[18] = {&pnm_re, &pnm_im, &phinm_re, &phinm_im, &chinm_re, &chinm_im,
    &pnm_re0, &pnm_im0, &phinm_re0, &phinm_im0, &chinm_re0, &chinm_im0,
    &pnm_re00, &pnm_im00, &phinm_re00, &phinm_im00, &chinm_re00, &chinm_im00};
  for(int c = 0; c < 18; c++) { free(*cf[c]); *cf[c] = zalloc(nc, sizeof(real)); }

  /* binDom_init */
  real rmax = 0;
  for(int i = 0; i < nparts; i++) rmax = rmax + (parts[i].r > rmax)*(parts[i].r - rmax);
  memset(&binDom, 0, sizeof(binDom));
  /* with no particles rmax = 0 and the reference's floor(xl / 0) to int
   * is undefined; one bin per axis is used instead (no kernel reads the
   * bins when nparts = 0) */
  const real bin_w = rmax > 0 ? 2.*rmax + interactionLengthRatio*rmax : 2.*Dom.xl;
  binDom.xs = Dom.xs; binDom.xe = Dom.xe; binDom.xl = Dom.xl;
  binDom.xn = rmax > 0 ? floor(Dom.xl/bin_w) : 0;
  if(binDom.xn == 0) { binDom.xn = 1; binDom.dx = Dom.xl; }
  else binDom.dx = Dom.xl / binDom.xn;
  binDom.ys = Dom.ys; binDom.ye = Dom.ye; binDom.yl = Dom.yl;
  binDom.yn = rmax > 0 ? floor(Dom.yl/bin_w) : 0;
  if(binDom.yn == 0) { binDom.yn = 1; binDom.dy = Dom.yl; }
  else binDom.dy = Dom.yl / binDom.yn;
  binDom.zs = Dom.zs; binDom.ze = Dom.ze; binDom.zl = Dom.zl;
  binDom.zn = rmax > 0 ? floor(Dom.zl/bin_w) : 0;
  if(binDom.zn == 0) { binDom.zn = 1; binDom.dz = Dom.zl; }
  else binDom.dz = Dom.zl / binDom.zn;
  binDom.Gcc.is = DOM_BUF; binDom.Gcc.in = binDom.xn;
  binDom.Gcc.ie = binDom.Gcc.is + binDom.Gcc.in;
  binDom.Gcc.js = DOM_BUF; binDom.Gcc.jn = binDom.yn;
  binDom.Gcc.je = binDom.Gcc.js + binDom.Gcc.jn;
  binDom.Gcc.ks = DOM_BUF; binDom.Gcc.kn = binDom.zn;
  binDom.Gcc.ke = DOM_BUF + binDom.Gcc.kn;
  binDom.Gcc.s1 = binDom.Gcc.in;
  binDom.Gcc.s2 = binDom.Gcc.s1 * binDom.Gcc.jn;
  binDom.Gcc.s3 = binDom.Gcc.s2 * binDom.Gcc.kn;
  return 0;
}

void bb_host_free(void)
{
  void *ptrs[] = {p0, p, phi, u, u0, v, v0, w, w0, f_x, f_y, f_z, diff0_u,
    diff0_v, diff0_w, conv0_u, conv0_v, conv0_w, diff_u, diff_v, diff_w,
    conv_u, conv_v, conv_w, u_star, v_star, w_star, phase, phase_shell,
    flag_u, flag_v, flag_w, parts, dom, pnm_re, pnm_im, phinm_re, phinm_im,
    chinm_re, chinm_im, pnm_re0, pnm_im0, phinm_re0, phinm_im0, chinm_re0,
    chinm_im0, pnm_re00, pnm_im00, phinm_re00, phinm_im00, chinm_re00,
    chinm_im00, u_WE, u_SN_S, u_SN_N, u_BT_B, u_BT_T, v_WE_W, v_WE_E, v_SN,
    v_BT_B, v_BT_T, w_WE_W, w_WE_E, w_SN_S, w_SN_N, w_BT};
  for(size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
  u_WE = u_SN_S = u_SN_N = u_BT_B = u_BT_T = NULL;
  v_WE_W = v_WE_E = v_SN = v_BT_B = v_BT_T = NULL;
  w_WE_W = w_WE_E = w_SN_S = w_SN_N = w_BT = NULL;
  p0 = p = phi = u = u0 = v = v0 = w = w0 = NULL;
  f_x = f_y = f_z = NULL;
  diff0_u = diff0_v = diff0_w = conv0_u = conv0_v = conv0_w = NULL;
  diff_u = diff_v = diff_w = conv_u = conv_v = conv_w = NULL;
  u_star = v_star = w_star = NULL;
  phase = phase_shell = flag_u = flag_v = flag_w = NULL;
  parts = NULL; dom = NULL;
  pnm_re = pnm_im = phinm_re = phinm_im = chinm_re = chinm_im = NULL;
  pnm_re0 = pnm_im0 = phinm_re0 = phinm_im0 = chinm_re0 = chinm_im0 = NULL;
  pnm_re00 = pnm_im00 = phinm_re00 = phinm_im00 = chinm_re00 = chinm_im00 = NULL;
  nparts = 0;
}

/* src/domain.c:3144-3221 */
static real ramp(real delta, real a, real m)
{
  if(a == 0) return m;
  if(fabs(delta * a) > fabs(m)) return m;
  return delta * a;
}

void compute_vel_BC(void)
{
  real delta = ttime - vel_tDelay;
  if(delta < 0) return;
  bc.uWD = ramp(delta, bc.uWDa, bc.uWDm); bc.uED = ramp(delta, bc.uEDa, bc.uEDm);
  bc.uSD = ramp(delta, bc.uSDa, bc.uSDm); bc.uND = ramp(delta, bc.uNDa, bc.uNDm);
  bc.uBD = ramp(delta, bc.uBDa, bc.uBDm); bc.uTD = ramp(delta, bc.uTDa, bc.uTDm);
  bc.vWD = ramp(delta, bc.vWDa, bc.vWDm); bc.vED = ramp(delta, bc.vEDa, bc.vEDm);
  bc.vSD = ramp(delta, bc.vSDa, bc.vSDm); bc.vND = ramp(delta, bc.vNDa, bc.vNDm);
  bc.vBD = ramp(delta, bc.vBDa, bc.vBDm); bc.vTD = ramp(delta, bc.vTDa, bc.vTDm);
  bc.wWD = ramp(delta, bc.wWDa, bc.wWDm); bc.wED = ramp(delta, bc.wEDa, bc.wEDm);
  bc.wSD = ramp(delta, bc.wSDa, bc.wSDm); bc.wND = ramp(delta, bc.wNDa, bc.wNDm);
  bc.wBD = ramp(delta, bc.wBDa, bc.wBDm); bc.wTD = ramp(delta, bc.wTDa, bc.wTDm);
}

static void first_dt_and_bcs(void)
{
  dt = cuda_find_dt();
  if(nparts > 0) cuda_part_BC();
  cuda_dom_BC();
  cuda_parts_internal();
  cuda_dom_BC();
}

/* src/bluebottle.c:445-545 after the pushes: cages, first dt, BCs. */
int bb_host_prepare(void)
{
  cuda_build_cages();
  cuda_part_pull();
  first_dt_and_bcs();
  return 0;
}

/* src/bluebottle.c:449-545 with restart requested: the cages of the
 * configured particles, then the restart file replaces the host state and
 * is pushed, and dt is recomputed from the restored field before the
 * boundary conditions are applied. */
int bb_host_restart(const char *path)
{
  cuda_build_cages();
  cuda_part_pull();
  if(bb_in_restart(path) != 0) return -1;
  cuda_dom_push();
  cuda_part_push();
  first_dt_and_bcs();
  return 0;
}

/* src/bluebottle.c:598-711: one time step; returns the number of Lamb
 * sub-iterations taken, or -1 when the new dt fails the driver's blow-up
 * test (dt < 1e-10 or dt > 1e10, or not a number), after which the state
 * must not be stepped further. */
int bb_host_step(void)
{
  ttime += dt;
  stepnum++;
  cuda_compute_forcing(&pid_int, &pid_back, Kp, Ki, Kd);
  compute_vel_BC();
  cuda_build_cages();
  int iter = 0;
  real iter_err = FLT_MAX;
  while(iter_err > lamb_residual) {
    if(!bb_is_implicit()) cuda_U_star_2();
    else {
      cuda_ustar_helmholtz(0);
      cuda_vstar_helmholtz(0);
      cuda_wstar_helmholtz(0);
    }
    if(nparts > 0) cuda_part_BC_star();
    cuda_dom_BC_star();
    cuda_solvability();
    if(nparts > 0) cuda_part_BC_star();
    cuda_dom_BC_star();
    cuda_PP_bicgstab(0);
    cuda_dom_BC_phi();
    cuda_project();
    if(nparts > 0) cuda_part_BC();
    cuda_dom_BC();
    cuda_update_p();
    if(nparts > 0) {
      cuda_part_BC();
      cuda_part_p_fill();
    }
    cuda_dom_BC_p();
    cuda_move_parts_sub();
    cuda_Lamb();
    iter_err = cuda_lamb_err();
    iter++;
    if(iter == lamb_max_iter) break;
  }
  cuda_move_parts();
  cuda_parts_internal();
  cuda_dom_BC();
  cuda_store_u();
  if(nparts > 0) cuda_store_coeffs();
  dt0 = dt;
  dt = cuda_find_dt();
  /* the driver's blow-up test (src/bluebottle.c:800-804) */
  if(!(dt >= 1.e-10 && dt <= 1.e10)) {
    printf("The solution has diverged.  Ending simulation.\n");
    fflush(stdout);
    return -1;
  }
  return iter;
}

/* ---------------------------------------------------------------------
 * The precursor domain (the turbulence rank of src/bluebottle.c:850-1156)
 * ------------------------------------------------------------------- */

/* p[idx] where idx may lie past the cell array: the v correction below
 * indexes p with v's strides, which reach past p when zn > yn + 2 */
static real p_at(long idx)
{
  return idx < (long) Dom.Gcc.s3b ? p[idx] : 0.;
}

/* src/domain.c:2922-3051: the divergence of (u, v, w) over every cell,
 * ghosts included, divided by the cell count, into p; each interior face
 * value of u, v, w (and u0, v0, w0) less the mean of the two cells either
 * side; then each component less its mean over the faces the reference
 * averages, over the whole ghosted array */
int bb_host_turb_correct(void)
{
  const grid_info *X = &Dom.Gfx, *Y = &Dom.Gfy, *Z = &Dom.Gfz, *G = &Dom.Gcc;
  real *face[3][2] = {{u, u0}, {v, v0}, {w, w0}};
  const int n3[3] = {X->s3b, Y->s3b, Z->s3b};
  /* the divergence over every cell, ghosts included, over the cell count */
  const real vol = Dom.xn * Dom.yn * Dom.zn;
  for(int k = G->ksb; k < G->keb; k++)
    for(int j = G->jsb; j < G->jeb; j++)
      for(int i = G->isb; i < G->ieb; i++) {
        const int W = i + j * X->s1b + k * X->s2b;
        const int S = i + j * Y->s1b + k * Y->s2b;
        const int B = i + j * Z->s1b + k * Z->s2b;
        const int C = i + j * G->s1b + k * G->s2b;
        p[C] = (u[W + 1] - u[W]) / Dom.dx + (v[S + Y->s1b] - v[S]) / Dom.dy
          + (w[B + Z->s2b] - w[B]) / Dom.dz;
        p[C] = p[C] / vol;
      }
  for(int k = X->ks; k < X->ke; k++)
    for(int j = X->js; j < X->je; j++)
      for(int i = X->is; i < X->ie; i++) {
        const int C = i + j * X->s1b + k * X->s2b;
        const int E = i + j * G->s1b + k * G->s2b;
        u[C] = u[C] - 0.5 * (p[E - 1] + p[E]);
        u0[C] = u0[C] - 0.5 * (p[E - 1] + p[E]);
      }
  for(int k = Y->ks; k < Y->ke; k++)
    for(int j = Y->js; j < Y->je; j++)
      for(int i = Y->is; i < Y->ie; i++) {
        const long C = i + (long) j * Y->s1b + (long) k * Y->s2b;
        const real ps = p_at(C - Y->s1b), pn = p_at(C);
        v[C] = v[C] - 0.5 * (ps + pn);
        v0[C] = v0[C] - 0.5 * (ps + pn);
      }
  for(int k = Z->ks; k < Z->ke; k++)
    for(int j = Z->js; j < Z->je; j++)
      for(int i = Z->is; i < Z->ie; i++) {
        const int C = i + j * Z->s1b + k * Z->s2b;
        w[C] = w[C] - 0.5 * (p[C - Z->s2b] + p[C]);
        w0[C] = w0[C] - 0.5 * (p[C - Z->s2b] + p[C]);
      }

  /* means over the faces less the last one on each component's axis */
  real umean = 0., vmean = 0., wmean = 0.;
  for(int k = X->ks; k < X->ke; k++)
    for(int j = X->js; j < X->je; j++)
      for(int i = X->is; i < X->ie - 1; i++) umean += u[i + j * X->s1b + k * X->s2b];
  for(int k = Y->ks; k < Y->ke; k++)
    for(int j = Y->js; j < Y->je - 1; j++)
      for(int i = Y->is; i < Y->ie; i++) vmean += v[i + j * Y->s1b + k * Y->s2b];
  for(int k = Z->ks; k < Z->ke - 1; k++)
    for(int j = Z->js; j < Z->je; j++)
      for(int i = Z->is; i < Z->ie; i++) wmean += w[i + j * Z->s1b + k * Z->s2b];
  umean /= vol;
  vmean /= vol;
  wmean /= vol;
  const real mean[3] = {umean, vmean, wmean};
  for(int c = 0; c < 3; c++) {
    real *x = face[c][0], *x0 = face[c][1];
    for(int i = 0; i < n3[c]; i++) {
      x[i] = x[i] - mean[c];
      x0[i] = x0[i] - mean[c];
    }
  }

  return 0;
}

/* src/domain.c:2856-3093 after the grids and arrays exist (bb_host_domain):
 * uniform random velocities of amplitude 3 turbA l (l the mean side
 * length, the integral scale turbl), the reference's one-sweep divergence
 * correction and mean removal, zeroed inflow planes, and the first dt
 * from the viscous limit alone.  The generator is seeded with `seed`
 * (the reference uses the wall clock).  As there, the correction of v
 * indexes the cell divergence with v's own strides; values that would lie
 * past the cell array (zn > yn + 2) are taken as 0, where the reference
 * reads beyond it. */
int bb_host_init_turb(unsigned long long seed)
{
  const grid_info *X = &Dom.Gfx, *Y = &Dom.Gfy, *Z = &Dom.Gfz, *G = &Dom.Gcc;
  memset(p0, 0, sizeof(real) * G->s3b);
  memset(p, 0, sizeof(real) * G->s3b);
  memset(phi, 0, sizeof(real) * G->s3b);
  real *face[3][8] = {
    {u, diff0_u, diff_u, u0, conv0_u, conv_u, f_x, u_star},
    {v, diff0_v, diff_v, v0, conv0_v, conv_v, f_y, v_star},
    {w, diff0_w, diff_w, w0, conv0_w, conv_w, f_z, w_star}};
  const int n3[3] = {X->s3b, Y->s3b, Z->s3b};
  for(int c = 0; c < 3; c++)
    for(int a = 0; a < 8; a++) memset(face[c][a], 0, sizeof(real) * n3[c]);

  rng_init(seed);
  turbl = (Dom.xl + Dom.yl + Dom.zl) / 3.;
  const real urms = 3 * turbA * turbl;
  for(int c = 0; c < 3; c++) {
    real *x = face[c][0], *x0 = face[c][3];
    for(int i = 0; i < n3[c]; i++) {
      const real tmp = (rng_dbl() - 0.5) * urms;
      x[i] = tmp;
      x0[i] = tmp;
    }
  }

  bb_host_turb_correct();

  real *planes[15] = {u_WE, u_SN_S, u_SN_N, u_BT_B, u_BT_T, v_WE_W, v_WE_E,
    v_SN, v_BT_B, v_BT_T, w_WE_W, w_WE_E, w_SN_S, w_SN_N, w_BT};
  const size_t psz[15] = {(size_t) X->jnb * X->knb, (size_t) X->inb * X->knb,
    (size_t) X->inb * X->knb, (size_t) X->inb * X->jnb,
    (size_t) X->inb * X->jnb, (size_t) Y->jnb * Y->knb,
    (size_t) Y->jnb * Y->knb, (size_t) Y->inb * Y->knb,
    (size_t) Y->inb * Y->jnb, (size_t) Y->inb * Y->jnb,
    (size_t) Z->jnb * Z->knb, (size_t) Z->jnb * Z->knb,
    (size_t) Z->inb * Z->knb, (size_t) Z->inb * Z->knb,
    (size_t) Z->inb * Z->jnb};
  for(int q = 0; q < 15; q++) memset(planes[q], 0, sizeof(real) * psz[q]);

  real dx_min = Dom.dx;
  if(Dom.dy < dx_min) dx_min = Dom.dy;
  if(Dom.dz < dx_min) dx_min = Dom.dz;
  dt = 2. * nu / (dx_min * dx_min);
  dt = CFL / dt;
  dt0 = -1.;
  stepnum = 0;
  bb_reset_output_clocks();
  return 0;
}

/* src/domain.c:2349: the grids and arrays of the Dom that turb.config
 * gave, then the initial field seeded from the wall clock */
int domain_init_turb(void)
{
  if(nsubdom > dev_end - dev_start + 1) return EXIT_FAILURE;
  /* the driver reads part.config first (src/bluebottle.c:877-878); keep
   * its (empty) particle list */
  part_struct *keep = parts;
  const int nkeep = nparts;
  bb_host_domain(Dom.xs, Dom.xe, Dom.xn, Dom.ys, Dom.ye, Dom.yn, Dom.zs,
    Dom.ze, Dom.zn);
  parts = keep;
  nparts = nkeep;
  bb_host_init_turb((unsigned long long) time(NULL));
  return EXIT_SUCCESS;
}

/* src/bluebottle.c:945-1016 after the pushes: cages (of no particles),
 * the first dt from the random field, the BCs */
int bb_host_prepare_turb(void)
{
  cuda_build_cages();
  cuda_part_pull();
  dt = cuda_find_dt();
  cuda_dom_BC();
  return 0;
}

/* src/bluebottle.c:945-1016 with restart requested: the restart file
 * replaces the host state and is pushed before the first dt and BCs */
int bb_host_restart_turb(const char *path)
{
  cuda_build_cages();
  cuda_part_pull();
  if(bb_in_restart_turb(path) != 0) return -1;
  cuda_dom_push();
  cuda_part_push();
  dt = cuda_find_dt();
  cuda_dom_BC();
  return 0;
}

/* src/bluebottle.c:1036-1081: one precursor time step, without the
 * exchange with the main domain and the output; returns 0, or -1 when
 * the new dt fails the driver's blow-up test */
int bb_host_step_turb(void)
{
  ttime += dt;
  stepnum++;
  cuda_compute_forcing(&pid_int, &pid_back, Kp, Ki, Kd);
  cuda_compute_turb_forcing();
  compute_vel_BC();
  if(!bb_is_implicit()) cuda_U_star_2();
  else {
    cuda_ustar_helmholtz(1);
    cuda_vstar_helmholtz(1);
    cuda_wstar_helmholtz(1);
  }
  cuda_dom_BC_star();
  cuda_solvability();
  cuda_dom_BC_star();
  cuda_PP_bicgstab(1);
  cuda_dom_BC_phi();
  cuda_project();
  cuda_dom_BC();
  cuda_update_p();
  cuda_dom_BC_p();
  cuda_store_u();
  dt0 = dt;
  dt = cuda_find_dt();
  if(!(dt >= 1.e-10 && dt <= 1.e10)) {
    printf("The solution has diverged.  Ending simulation.\n");
    fflush(stdout);
    return -1;
  }
  return 0;
}
