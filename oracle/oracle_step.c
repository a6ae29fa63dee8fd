/*
 * oracle_step.c -- the driver's call sequence around the hot path, on the
 * oracle's state.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * orc_prepare: src/bluebottle.c:445-545 (cages, first dt, BCs, interior).
 * orc_step:    src/bluebottle.c:598-711 (one time step with the Lamb
 *              sub-iteration), plus compute_vel_BC (src/domain.c:3144-3221).
 * orc_prepare_turb, orc_step_turb: the precursor domain's set-up after the
 *              pushes and one step (src/bluebottle.c:945-1016, 1036-1081).
 */
#include <float.h>
#include <math.h>
#include "oracle.h"

static real ramp(real delta, real a, real m)
{
  if(a == 0) return m;
  if(fabs(delta * a) > fabs(m)) return m;
  return delta * a;
}

static void compute_vel_BC(orc_state *s)
{
  BC *b = &s->bc;
  real delta = s->ttime - s->vel_tDelay;
  if(delta < 0) return;
  b->uWD = ramp(delta, b->uWDa, b->uWDm); b->uED = ramp(delta, b->uEDa, b->uEDm);
  b->uSD = ramp(delta, b->uSDa, b->uSDm); b->uND = ramp(delta, b->uNDa, b->uNDm);
  b->uBD = ramp(delta, b->uBDa, b->uBDm); b->uTD = ramp(delta, b->uTDa, b->uTDm);
  b->vWD = ramp(delta, b->vWDa, b->vWDm); b->vED = ramp(delta, b->vEDa, b->vEDm);
  b->vSD = ramp(delta, b->vSDa, b->vSDm); b->vND = ramp(delta, b->vNDa, b->vNDm);
  b->vBD = ramp(delta, b->vBDa, b->vBDm); b->vTD = ramp(delta, b->vTDa, b->vTDm);
  b->wWD = ramp(delta, b->wWDa, b->wWDm); b->wED = ramp(delta, b->wEDa, b->wEDm);
  b->wSD = ramp(delta, b->wSDa, b->wSDm); b->wND = ramp(delta, b->wNDa, b->wNDm);
  b->wBD = ramp(delta, b->wBDa, b->wBDm); b->wTD = ramp(delta, b->wTDa, b->wTDm);
}

void orc_prepare(orc_state *s)
{
  orc_build_cages(s);
  s->dt = orc_find_dt(s);
  if(s->nparts > 0) orc_part_BC(s);
  orc_dom_BC(s);
  orc_parts_internal(s);
  orc_dom_BC(s);
}

int orc_step(orc_state *s)
{
  s->ttime += s->dt;
  s->stepnum++;
  orc_compute_forcing(s);
  compute_vel_BC(s);
  orc_build_cages(s);
  int iter = 0;
  real iter_err = FLT_MAX;
  while(iter_err > s->lamb_residual) {
    if(!s->implicit) orc_U_star_2(s);
    else {
      orc_star_helmholtz(s, 0);
      orc_star_helmholtz(s, 1);
      orc_star_helmholtz(s, 2);
    }
    if(s->nparts > 0) orc_part_BC_star(s);
    orc_dom_BC_star(s);
    orc_solvability(s);
    if(s->nparts > 0) orc_part_BC_star(s);
    orc_dom_BC_star(s);
    orc_PP_bicgstab(s);
    orc_dom_BC_phi(s);
    orc_project(s);
    if(s->nparts > 0) orc_part_BC(s);
    orc_dom_BC(s);
    orc_update_p(s);
    if(s->nparts > 0) { orc_part_BC(s); orc_part_p_fill(s); }
    orc_dom_BC_p(s);
    orc_move_parts_sub(s);
    orc_Lamb(s);
    iter_err = orc_lamb_err(s);
    iter++;
    if(iter == s->lamb_max_iter) break;
  }
  orc_move_parts(s);
  orc_parts_internal(s);
  orc_dom_BC(s);
  orc_store_u(s);
  if(s->nparts > 0) orc_store_coeffs(s);
  s->dt0 = s->dt;
  s->dt = orc_find_dt(s);
  return iter;
}

void orc_prepare_turb(orc_state *s)
{
  orc_build_cages(s);
  s->dt = orc_find_dt(s);
  orc_dom_BC(s);
}

int orc_step_turb(orc_state *s)
{
  s->ttime += s->dt;
  s->stepnum++;
  orc_compute_forcing(s);
  orc_compute_turb_forcing(s);
  compute_vel_BC(s);
  if(!s->implicit) orc_U_star_2(s);
  else {
    orc_star_helmholtz(s, 0);
    orc_star_helmholtz(s, 1);
    orc_star_helmholtz(s, 2);
  }
  orc_dom_BC_star(s);
  orc_solvability(s);
  orc_dom_BC_star(s);
  orc_PP_bicgstab(s);
  orc_dom_BC_phi(s);
  orc_project(s);
  orc_dom_BC(s);
  orc_update_p(s);
  orc_dom_BC_p(s);
  orc_store_u(s);
  s->dt0 = s->dt;
  s->dt = orc_find_dt(s);
  return 0;
}
