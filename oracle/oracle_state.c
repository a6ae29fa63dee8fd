/*
 * oracle_state.c -- allocation of the oracle's copy of the reference host
 * state.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Array extents follow src/domain.c:1611-1725 (Gcc/Gfx/Gfy/Gfz s3b) and
 * src/particle.c:295-535 (nparts * coeff_stride coefficient arrays).
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

const char *orc_plane_names[15] = {
  "u_WE", "u_SN_S", "u_SN_N", "u_BT_B", "u_BT_T",
  "v_WE_W", "v_WE_E", "v_SN", "v_BT_B", "v_BT_T",
  "w_WE_W", "w_WE_E", "w_SN_S", "w_SN_N", "w_BT"
};

/* entries of plane q: the transverse ghosted extents of its grid */
static size_t plane_size(const dom_struct *d, int q)
{
  const grid_info *G = q < 5 ? &d->Gfx : q < 10 ? &d->Gfy : &d->Gfz;
  static const int axis[15] = {0, 1, 1, 2, 2, 0, 0, 1, 2, 2, 0, 0, 1, 1, 2};
  if(axis[q] == 0) return (size_t) G->jnb * G->knb;
  if(axis[q] == 1) return (size_t) G->inb * G->knb;
  return (size_t) G->inb * G->jnb;
}

const char *orc_coeff_names[18] = {
  "pnm_re", "pnm_im", "phinm_re", "phinm_im", "chinm_re", "chinm_im",
  "pnm_re0", "pnm_im0", "phinm_re0", "phinm_im0", "chinm_re0", "chinm_im0",
  "pnm_re00", "pnm_im00", "phinm_re00", "phinm_im00", "chinm_re00",
  "chinm_im00"
};

static void *zalloc(size_t n, size_t sz)
{
  void *ptr = calloc(n ? n : 1, sz);
  return ptr;
}

orc_state *orc_create(const dom_struct *Dom, int nparts, int coeff_stride)
{
  orc_state *s = (orc_state *) zalloc(1, sizeof(orc_state));
  s->dom = *Dom;
  size_t cc = (size_t) Dom->Gcc.s3b;
  size_t fx = (size_t) Dom->Gfx.s3b;
  size_t fy = (size_t) Dom->Gfy.s3b;
  size_t fz = (size_t) Dom->Gfz.s3b;
  s->p0 = zalloc(cc, sizeof(real));
  s->p = zalloc(cc, sizeof(real));
  s->phi = zalloc(cc, sizeof(real));
  s->phase = zalloc(cc, sizeof(int));
  s->phase_shell = zalloc(cc, sizeof(int));
  real **fxs[] = {&s->u, &s->u0, &s->f_x, &s->diff0_u, &s->conv0_u,
    &s->diff_u, &s->conv_u, &s->u_star};
  real **fys[] = {&s->v, &s->v0, &s->f_y, &s->diff0_v, &s->conv0_v,
    &s->diff_v, &s->conv_v, &s->v_star};
  real **fzs[] = {&s->w, &s->w0, &s->f_z, &s->diff0_w, &s->conv0_w,
    &s->diff_w, &s->conv_w, &s->w_star};
  for(int i = 0; i < 8; i++) {
    *fxs[i] = zalloc(fx, sizeof(real));
    *fys[i] = zalloc(fy, sizeof(real));
    *fzs[i] = zalloc(fz, sizeof(real));
  }
  s->flag_u = zalloc(fx, sizeof(int));
  s->flag_v = zalloc(fy, sizeof(int));
  s->flag_w = zalloc(fz, sizeof(int));
  s->rhs_p = zalloc((size_t) Dom->Gcc.s3, sizeof(real));
  s->nparts = nparts;
  s->coeff_stride = coeff_stride;
  s->parts = zalloc((size_t) nparts, sizeof(part_struct));
  for(int c = 0; c < 18; c++)
    s->coeff[c] = zalloc((size_t) nparts * coeff_stride, sizeof(real));
  for(int q = 0; q < 15; q++) s->plane[q] = zalloc(plane_size(Dom, q), sizeof(real));
  return s;
}

void orc_free(orc_state *s)
{
  if(!s) return;
  void *ptrs[] = {s->p0, s->p, s->phi, s->phase, s->phase_shell,
    s->u, s->u0, s->f_x, s->diff0_u, s->conv0_u, s->diff_u, s->conv_u,
    s->u_star, s->v, s->v0, s->f_y, s->diff0_v, s->conv0_v, s->diff_v,
    s->conv_v, s->v_star, s->w, s->w0, s->f_z, s->diff0_w, s->conv0_w,
    s->diff_w, s->conv_w, s->w_star, s->flag_u, s->flag_v, s->flag_w,
    s->rhs_p, s->parts};
  for(size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
  for(int c = 0; c < 18; c++) free(s->coeff[c]);
  for(int q = 0; q < 15; q++) free(s->plane[q]);
  free(s);
}

size_t orc_sizeof_state(void) { return sizeof(orc_state); }

void *orc_array(orc_state *s, const char *name, size_t *n, int *is_int)
{
  const dom_struct *d = &s->dom;
  size_t cc = (size_t) d->Gcc.s3b, fx = (size_t) d->Gfx.s3b;
  size_t fy = (size_t) d->Gfy.s3b, fz = (size_t) d->Gfz.s3b;
  struct { const char *nm; void *ptr; size_t n; int isint; } tab[] = {
    {"p0", s->p0, cc, 0}, {"p", s->p, cc, 0}, {"phi", s->phi, cc, 0},
    {"phase", s->phase, cc, 1}, {"phase_shell", s->phase_shell, cc, 1},
    {"u", s->u, fx, 0}, {"u0", s->u0, fx, 0}, {"f_x", s->f_x, fx, 0},
    {"diff0_u", s->diff0_u, fx, 0}, {"conv0_u", s->conv0_u, fx, 0},
    {"diff_u", s->diff_u, fx, 0}, {"conv_u", s->conv_u, fx, 0},
    {"u_star", s->u_star, fx, 0}, {"flag_u", s->flag_u, fx, 1},
    {"v", s->v, fy, 0}, {"v0", s->v0, fy, 0}, {"f_y", s->f_y, fy, 0},
    {"diff0_v", s->diff0_v, fy, 0}, {"conv0_v", s->conv0_v, fy, 0},
    {"diff_v", s->diff_v, fy, 0}, {"conv_v", s->conv_v, fy, 0},
    {"v_star", s->v_star, fy, 0}, {"flag_v", s->flag_v, fy, 1},
    {"w", s->w, fz, 0}, {"w0", s->w0, fz, 0}, {"f_z", s->f_z, fz, 0},
    {"diff0_w", s->diff0_w, fz, 0}, {"conv0_w", s->conv0_w, fz, 0},
    {"diff_w", s->diff_w, fz, 0}, {"conv_w", s->conv_w, fz, 0},
    {"w_star", s->w_star, fz, 0}, {"flag_w", s->flag_w, fz, 1},
    {"rhs_p", s->rhs_p, (size_t) d->Gcc.s3, 0},
  };
  for(size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); i++) {
    if(strcmp(tab[i].nm, name) == 0) {
      *n = tab[i].n;
      *is_int = tab[i].isint;
      return tab[i].ptr;
    }
  }
  for(int q = 0; q < 15; q++) {
    if(strcmp(orc_plane_names[q], name) == 0) {
      *n = plane_size(d, q);
      *is_int = 0;
      return s->plane[q];
    }
  }
  for(int c = 0; c < 18; c++) {
    if(strcmp(orc_coeff_names[c], name) == 0) {
      *n = (size_t) s->nparts * s->coeff_stride;
      *is_int = 0;
      return s->coeff[c];
    }
  }
  *n = 0;
  *is_int = 0;
  return NULL;
}

#ifdef _OPENMP
#include <omp.h>
void orc_set_threads(int n) { if(n > 0) omp_set_num_threads(n); }
#else
void orc_set_threads(int n) { (void) n; }
#endif
