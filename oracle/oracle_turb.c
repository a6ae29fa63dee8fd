/*
 * oracle_turb.c -- the precursor domain's linear forcing and initial field.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * orc_colocate:           src/bluebottle_kernel.cu:3849-3889
 * orc_compute_energy:     src/cuda_bluebottle.cu:3944-4017,
 *                         src/bluebottle_kernel.cu:3891-3906
 * orc_compute_turb_forcing: src/cuda_bluebottle.cu:3817-3942,
 *                         src/bluebottle_kernel.cu:2054-2131
 * orc_init_turb:          src/domain.c:2856-3093 (the field part),
 *                         with src/rng.c's generator (Ranq1, Numerical
 *                         Recipes 3rd ed. section 7.1.3)
 * orc_turb_correct:       src/domain.c:2922-3051, its correction sweep
 * The reductions the reference does with thrust::reduce are serial sums
 * here, in index order.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

void orc_colocate(const orc_state *s, int comp, real *out)
{
  const dom_struct *d = &s->dom;
  const grid_info *G = comp == 0 ? &d->Gfx : comp == 1 ? &d->Gfy : &d->Gfz;
  const real *f = comp == 0 ? s->u : comp == 1 ? s->v : s->w;
  if(comp == 0) {
    for(int tk = DOM_BUF; tk < G->knb - 1; tk++)
      for(int tj = DOM_BUF; tj < G->jnb - 1; tj++)
        for(int i = G->is; i < G->ie - 1; i++)
          out[(i - DOM_BUF) + (tj - DOM_BUF) * d->Gcc.s1
            + (tk - DOM_BUF) * d->Gcc.s2] = 0.5 * (f[i + tj * G->s1b
            + tk * G->s2b] + f[(i + 1) + tj * G->s1b + tk * G->s2b]);
  } else if(comp == 1) {
    for(int tk = DOM_BUF; tk < G->knb - 1; tk++)
      for(int ti = DOM_BUF; ti < G->inb - 1; ti++)
        for(int j = G->js; j < G->je - 1; j++)
          out[(ti - DOM_BUF) + (j - DOM_BUF) * d->Gcc.s1
            + (tk - DOM_BUF) * d->Gcc.s2] = 0.5 * (f[ti + j * G->s1b
            + tk * G->s2b] + f[ti + (j + 1) * G->s1b + tk * G->s2b]);
  } else {
    for(int ti = DOM_BUF; ti < G->inb - 1; ti++)
      for(int tj = DOM_BUF; tj < G->jnb - 1; tj++)
        for(int k = G->ks; k < G->ke - 1; k++)
          out[(ti - DOM_BUF) + (tj - DOM_BUF) * d->Gcc.s1
            + (k - DOM_BUF) * d->Gcc.s2] = 0.5 * (f[ti + tj * G->s1b
            + k * G->s2b] + f[ti + tj * G->s1b + (k + 1) * G->s2b]);
  }
}

static real sum(const real *x, int n)
{
  real a = 0.;
  for(int i = 0; i < n; i++) a += x[i];
  return a;
}

real orc_compute_energy(orc_state *s)
{
  const int n = s->dom.Gcc.s3;
  real *uc = malloc(sizeof(real) * n), *vc = malloc(sizeof(real) * n);
  real *wc = malloc(sizeof(real) * n);
  orc_colocate(s, 0, uc);
  orc_colocate(s, 1, vc);
  orc_colocate(s, 2, wc);
  for(int C = 0; C < n; C++) uc[C] = uc[C] * uc[C] + vc[C] * vc[C] + wc[C] * wc[C];
  real k = sum(uc, n);
  k = k / n * 0.5;
  free(uc); free(vc); free(wc);
  return k;
}

void orc_compute_turb_forcing(orc_state *s)
{
  const dom_struct *d = &s->dom;
  const real k = orc_compute_energy(s);
  const real turbA = s->turbA, turbl = s->turbl;
  const real k0 = 13.5 * turbA * turbA * 0.0361 * turbl * turbl;
  const int n = d->Gcc.s3;
  real *co = malloc(sizeof(real) * n);
  real mean[3];
  for(int c = 0; c < 3; c++) {
    orc_colocate(s, c, co);
    mean[c] = sum(co, n);
    mean[c] = mean[c] / n;
  }
  free(co);
  /* forcing_add_*_field over the whole ghosted arrays, then *_const */
  const real gain = turbA * k0 / k;
  real *f[3] = {s->f_x, s->f_y, s->f_z};
  const real *x[3] = {s->u, s->v, s->w};
  const int cnt[3] = {d->Gfx.s3b, d->Gfy.s3b, d->Gfz.s3b};
  for(int c = 0; c < 3; c++)
    for(int i = 0; i < cnt[c]; i++) f[c][i] += gain * x[c][i];
  for(int c = 0; c < 3; c++) {
    const real add = -turbA * k0 / k * mean[c];
    for(int i = 0; i < cnt[c]; i++) f[c][i] += add;
  }
}

/* src/rng.c: Ranq1 with the state carried in *v */
static unsigned long long ranq1(unsigned long long *v)
{
  *v ^= *v >> 21;
  *v ^= *v << 35;
  *v ^= *v >> 4;
  return *v * 2685821657736338717ULL;
}

unsigned long long orc_rng_seeded(unsigned long long state,
  unsigned long long seed)
{
  state ^= seed;
  return ranq1(&state);
}

double orc_rng_dbl(unsigned long long *state)
{
  return 5.42101086242752217e-20 * (double) ranq1(state);
}

/* src/domain.c:2922-3051: divergence into p, the correction of the
 * interior faces, the mean removal */
void orc_turb_correct(orc_state *s)
{
  const dom_struct *d = &s->dom;
  const grid_info *X = &d->Gfx, *Y = &d->Gfy, *Z = &d->Gfz, *G = &d->Gcc;
  real *p = s->p, *u = s->u, *v = s->v, *w = s->w;
  real *u0 = s->u0, *v0 = s->v0, *w0 = s->w0;
  const real vol = d->xn * d->yn * d->zn;
  for(int k = G->ksb; k < G->keb; k++)
    for(int j = G->jsb; j < G->jeb; j++)
      for(int i = G->isb; i < G->ieb; i++) {
        int W = i + j * X->s1b + k * X->s2b;
        int E = (i + 1) + j * X->s1b + k * X->s2b;
        int S = i + j * Y->s1b + k * Y->s2b;
        int N = i + (j + 1) * Y->s1b + k * Y->s2b;
        int B = i + j * Z->s1b + k * Z->s2b;
        int T = i + j * Z->s1b + (k + 1) * Z->s2b;
        int C = i + j * G->s1b + k * G->s2b;
        p[C] = (u[E] - u[W]) / d->dx + (v[N] - v[S]) / d->dy
          + (w[T] - w[B]) / d->dz;
        p[C] = p[C] / vol;
      }
  for(int k = X->ks; k < X->ke; k++)
    for(int j = X->js; j < X->je; j++)
      for(int i = X->is; i < X->ie; i++) {
        int C = i + j * X->s1b + k * X->s2b;
        int W = (i - 1) + j * G->s1b + k * G->s2b;
        int E = i + j * G->s1b + k * G->s2b;
        u[C] = u[C] - 0.5 * (p[W] + p[E]);
        u0[C] = u0[C] - 0.5 * (p[W] + p[E]);
      }
  /* the reference indexes p with v's strides here; past the end of p
   * (zn > yn + 2) the value is taken as 0 */
  for(int k = Y->ks; k < Y->ke; k++)
    for(int j = Y->js; j < Y->je; j++)
      for(int i = Y->is; i < Y->ie; i++) {
        int C = i + j * Y->s1b + k * Y->s2b;
        int S = i + (j - 1) * Y->s1b + k * Y->s2b;
        int N = i + j * Y->s1b + k * Y->s2b;
        real pS = S < G->s3b ? p[S] : 0., pN = N < G->s3b ? p[N] : 0.;
        v[C] = v[C] - 0.5 * (pS + pN);
        v0[C] = v0[C] - 0.5 * (pS + pN);
      }
  for(int k = Z->ks; k < Z->ke; k++)
    for(int j = Z->js; j < Z->je; j++)
      for(int i = Z->is; i < Z->ie; i++) {
        int C = i + j * Z->s1b + k * Z->s2b;
        int B = i + j * Z->s1b + (k - 1) * Z->s2b;
        int T = i + j * Z->s1b + k * Z->s2b;
        w[C] = w[C] - 0.5 * (p[B] + p[T]);
        w0[C] = w0[C] - 0.5 * (p[B] + p[T]);
      }
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
  for(int i = 0; i < X->s3b; i++) { u[i] = u[i] - umean; u0[i] = u0[i] - umean; }
  for(int i = 0; i < Y->s3b; i++) { v[i] = v[i] - vmean; v0[i] = v0[i] - vmean; }
  for(int i = 0; i < Z->s3b; i++) { w[i] = w[i] - wmean; w0[i] = w0[i] - wmean; }
}

void orc_init_turb(orc_state *s, unsigned long long *rng_state)
{
  const dom_struct *d = &s->dom;
  const grid_info *X = &d->Gfx, *Y = &d->Gfy, *Z = &d->Gfz, *G = &d->Gcc;
  memset(s->p0, 0, sizeof(real) * G->s3b);
  memset(s->p, 0, sizeof(real) * G->s3b);
  memset(s->phi, 0, sizeof(real) * G->s3b);
  real *arr[3][8] = {
    {s->u, s->diff0_u, s->diff_u, s->u0, s->conv0_u, s->conv_u, s->f_x, s->u_star},
    {s->v, s->diff0_v, s->diff_v, s->v0, s->conv0_v, s->conv_v, s->f_y, s->v_star},
    {s->w, s->diff0_w, s->diff_w, s->w0, s->conv0_w, s->conv_w, s->f_z, s->w_star}};
  const int cnt[3] = {X->s3b, Y->s3b, Z->s3b};
  for(int c = 0; c < 3; c++)
    for(int a = 0; a < 8; a++) memset(arr[c][a], 0, sizeof(real) * cnt[c]);

  s->turbl = (d->xl + d->yl + d->zl) / 3.;
  const real urms = 3 * s->turbA * s->turbl;
  for(int c = 0; c < 3; c++)
    for(int i = 0; i < cnt[c]; i++) {
      const real tmp = (orc_rng_dbl(rng_state) - 0.5) * urms;
      arr[c][0][i] = tmp;
      arr[c][3][i] = tmp;
    }

  orc_turb_correct(s);

  for(int q = 0; q < 15; q++) {
    size_t n = 0;
    int is_int = 0;
    real *pl = orc_array(s, orc_plane_names[q], &n, &is_int);
    if(pl) memset(pl, 0, sizeof(real) * n);
  }
  real dx_min = d->dx;
  if(d->dy < dx_min) dx_min = d->dy;
  if(d->dz < dx_min) dx_min = d->dz;
  s->dt = 2. * s->nu / (dx_min * dx_min);
  s->dt = s->CFL / s->dt;
  s->dt0 = -1.;
  s->stepnum = 0;
}
