/*
 * bb_io.c -- the driver side's files, in the reference's formats:
 *   flow.config   the fixed-order input grammar of src/domain.c:30-730
 *   part.config   src/particle.c:79-159
 *   restart       the raw dump of src/domain.c:3223-3413
 *   lamb.rec      the Lamb-coefficient log of src/recorder.c:1912-2001
 *
 * The readers consume a file the way the reference's sequence of fscanf
 * calls does: each expected line is a scanf pattern tried in order, and a
 * line whose leading literal does not match is skipped over without
 * consuming input (so an absent optional line such as "vel_tDelay" leaves
 * its value unchanged, as in the reference).  One consequence kept as it
 * is: without a "g_tDelay" line the attempt consumes the "g" of "g.x", so
 * the g.x, g.y, g.z and gain lines that follow are skipped too and g keeps
 * its previous value.  The patterns are generated
 * from tables of faces, components and fields instead of one hand-written
 * block per line.  Extension (no reference counterpart): init_cond
 * TAYLOR_GREEN, the initial field of the first benchmark configuration,
 * which the Python set-up fills in.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/bluebottle.h"
#include "bb_host.h"

/* ---- driver globals the restart file carries (src/bluebottle.c:140-182) */
int rec_paraview_stepnum_out;
real rec_flow_field_ttime_out;
real rec_paraview_ttime_out;
real rec_particle_ttime_out;
real rec_restart_ttime_out;
real rec_prec_ttime_out;
real bc_plane_pos[9];

/* output settings of record.config (src/bluebottle.c:139-169) */
int rec_flow_field_stepnum_out, rec_particle_stepnum_out;
int rec_prec_stepnum_out, rec_prec_flow_field_stepnum_out;
real rec_flow_field_dt, rec_paraview_dt, rec_particle_dt, rec_restart_dt;
real rec_prec_dt, rec_prec_flow_field_dt, rec_prec_flow_field_ttime_out;
int rec_flow_field_vel, rec_flow_field_p, rec_flow_field_phase;
int rec_prec_flow_field_vel, rec_prec_flow_field_p, rec_prec_flow_field_phase;
int rec_restart_stop;
int rec_particle_pos, rec_particle_a, rec_particle_vel, rec_particle_omega;
int rec_particle_force, rec_particle_moment;

/* the output counters and clocks domain_init and domain_init_turb zero
 * (src/domain.c:2334-2343, 3081-3090) */
void bb_reset_output_clocks(void)
{
  rec_flow_field_stepnum_out = 0;
  rec_paraview_stepnum_out = 0;
  rec_particle_stepnum_out = 0;
  rec_prec_stepnum_out = 0;
  rec_flow_field_ttime_out = 0;
  rec_paraview_ttime_out = 0;
  rec_particle_ttime_out = 0;
  rec_restart_ttime_out = 0;
  rec_prec_ttime_out = 0;
}

/* the directory holding input/ and record/ (ROOT_DIR of the reference) */
char bb_root_dir[FILE_NAME_SIZE] = ".";

#define TAYLOR_GREEN_INIT 100

static int read_error(FILE *f, const char *what)
{
  fprintf(stderr, "%s read error.\n", what);
  if(f) fclose(f);
  return -1;
}

/* ---------------------------------------------------------------------
 * flow.config
 * ------------------------------------------------------------------- */
static const char FACES[6] = {'W', 'E', 'S', 'N', 'B', 'T'};

/* pointers to the type and value slots of field c ('p', 'u', 'v', 'w') on
 * face q of the BC struct, whose members run face by face:
 * pW..pT (int), pWD..pTD (real) for p; uW, uWDm, uWD, uWDa, ... per face
 * for the velocities */
static int *bc_type(char c, int q)
{
  switch(c) {
    case 'p': { int *t[6] = {&bc.pW, &bc.pE, &bc.pS, &bc.pN, &bc.pB, &bc.pT};
      return t[q]; }
    case 'u': { int *t[6] = {&bc.uW, &bc.uE, &bc.uS, &bc.uN, &bc.uB, &bc.uT};
      return t[q]; }
    case 'v': { int *t[6] = {&bc.vW, &bc.vE, &bc.vS, &bc.vN, &bc.vB, &bc.vT};
      return t[q]; }
    default: { int *t[6] = {&bc.wW, &bc.wE, &bc.wS, &bc.wN, &bc.wB, &bc.wT};
      return t[q]; }
  }
}

/* Dirichlet value slots (Dm: maximum, D: current, Da: ramp rate) */
static real *bc_val(char c, int q, int which)
{
  real *u[6][3] = {{&bc.uWDm, &bc.uWD, &bc.uWDa}, {&bc.uEDm, &bc.uED, &bc.uEDa},
    {&bc.uSDm, &bc.uSD, &bc.uSDa}, {&bc.uNDm, &bc.uND, &bc.uNDa},
    {&bc.uBDm, &bc.uBD, &bc.uBDa}, {&bc.uTDm, &bc.uTD, &bc.uTDa}};
  real *v[6][3] = {{&bc.vWDm, &bc.vWD, &bc.vWDa}, {&bc.vEDm, &bc.vED, &bc.vEDa},
    {&bc.vSDm, &bc.vSD, &bc.vSDa}, {&bc.vNDm, &bc.vND, &bc.vNDa},
    {&bc.vBDm, &bc.vBD, &bc.vBDa}, {&bc.vTDm, &bc.vTD, &bc.vTDa}};
  real *w[6][3] = {{&bc.wWDm, &bc.wWD, &bc.wWDa}, {&bc.wEDm, &bc.wED, &bc.wEDa},
    {&bc.wSDm, &bc.wSD, &bc.wSDa}, {&bc.wNDm, &bc.wND, &bc.wNDa},
    {&bc.wBDm, &bc.wBD, &bc.wBDa}, {&bc.wTDm, &bc.wTD, &bc.wTDa}};
  return c == 'u' ? u[q][which] : c == 'v' ? v[q][which] : w[q][which];
}

static real *bc_ds(int q)
{
  real *d[6] = {&bc.dsW, &bc.dsE, &bc.dsS, &bc.dsN, &bc.dsB, &bc.dsT};
  return d[q];
}

/* one "bc.<c><F> TYPE [values]" line; the precursor domain's own file
 * knows no PRECURSOR faces (src/domain.c:898-1219) */
static int read_bc_line(FILE *f, char c, int q, int allow_precursor)
{
  char pat[32], buf[CHAR_BUF_SIZE] = "";
  snprintf(pat, sizeof(pat), "bc.%c%c %%255s", c, FACES[q]);
  if(fscanf(f, pat, buf) != 1) return -1;
  int *type = bc_type(c, q);
  if(c == 'p') {
    if(strcmp(buf, "PERIODIC") == 0) {
      *type = PERIODIC;
      *bc_ds(q) = 0;
    } else if(strcmp(buf, "NEUMANN") == 0) {
      *type = NEUMANN;
      if(fscanf(f, "%lf", bc_ds(q)) != 1) *bc_ds(q) = 0;
    } else {
      return -1;
    }
  } else {
    real *dm = bc_val(c, q, 0), *d = bc_val(c, q, 1), *da = bc_val(c, q, 2);
    if(strcmp(buf, "PERIODIC") == 0) {
      *type = PERIODIC;
    } else if(strcmp(buf, "DIRICHLET") == 0 ||
              (allow_precursor && strcmp(buf, "PRECURSOR") == 0)) {
      *type = buf[0] == 'D' ? DIRICHLET : PRECURSOR;
      if(fscanf(f, "%lf %lf", dm, da) != 2) return -1;
      *d = 0;
    } else if(strcmp(buf, "NEUMANN") == 0) {
      *type = NEUMANN;
      *dm = 0; *da = 0; *d = 0;
    } else {
      return -1;
    }
  }
  if(fscanf(f, "\n") < 0) {}
  return 0;
}

/* a keyword from a list (the index, or -1) */
static int read_keyword(FILE *f, const char *pat, const char *const *names,
  const int *values, int n, int *out)
{
  char buf[CHAR_BUF_SIZE] = "";
  if(fscanf(f, pat, buf) != 1) return -1;
  for(int i = 0; i < n; i++)
    if(strcmp(buf, names[i]) == 0) { *out = values[i]; return 0; }
  return -1;
}

/* match the literal text `literal` (the file's fixed section headers and
 * blank lines); the patterns are this file's own constants */
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wformat-security"
static void skip(FILE *f, const char *literal)
{
  if(fscanf(f, literal) < 0) {}
}
#pragma GCC diagnostic pop

static void read_real(FILE *f, const char *pat, real *x)
{
  if(fscanf(f, pat, x) < 0) {}
}

static void read_int(FILE *f, const char *pat, int *x)
{
  if(fscanf(f, pat, x) < 0) {}
}

/* DOMAIN, GPU DOMAIN DECOMPOSITION and PHYSICAL PARAMETERS, shared by
 * flow.config and turb.config (src/domain.c:52-104, 754-802); 0 or -1
 * (the file is closed on error) */
static int read_domain_head(FILE *f, const char *what)
{
  cpumem = 0;
  gpumem = 0;
  skip(f, "DOMAIN\n");
  if(fscanf(f, "(Xs, Xe, Xn) %lf %lf %d\n", &Dom.xs, &Dom.xe, &Dom.xn) != 3 ||
     fscanf(f, "(Ys, Ye, Yn) %lf %lf %d\n", &Dom.ys, &Dom.ye, &Dom.yn) != 3 ||
     fscanf(f, "(Zs, Ze, Zn) %lf %lf %d\n", &Dom.zs, &Dom.ze, &Dom.zn) != 3)
    return read_error(f, what);
  skip(f, "\n");
  skip(f, "GPU DOMAIN DECOMPOSITION\n");
  if(fscanf(f, "DEV RANGE %d %d\n", &dev_start, &dev_end) != 2 ||
     fscanf(f, "n %d\n", &nsubdom) != 1 || nsubdom != 1) {
    fprintf(stderr, "%s: one subdomain (one GPU) expected\n", what);
    return read_error(f, what);
  }
  /* the subdomain record (src/domain.c:68-91) */
  dom_struct sub;
  memset(&sub, 0, sizeof(sub));
  if(fscanf(f, "(Xs, Xe, Xn) %lf %lf %d\n", &sub.xs, &sub.xe, &sub.xn) != 3 ||
     fscanf(f, "(Ys, Ye, Yn) %lf %lf %d\n", &sub.ys, &sub.ye, &sub.yn) != 3 ||
     fscanf(f, "(Zs, Ze, Zn) %lf %lf %d\n", &sub.zs, &sub.ze, &sub.zn) != 3 ||
     fscanf(f, "E %d W %d N %d S %d T %d B %d\n", &sub.E, &sub.W, &sub.N,
       &sub.S, &sub.T, &sub.B) != 6)
    return read_error(f, what);
  skip(f, "\n");
  skip(f, "PHYSICAL PARAMETERS\n");
  read_real(f, "rho_f %lf\n", &rho_f);
  read_real(f, "nu %lf\n", &nu);
  mu = nu * rho_f;
  skip(f, "\n");
  return 0;
}

/* the four BC groups (PRESSURE, X-, Y-, Z-VELOCITY); 0 or -1 */
static int read_bc_groups(FILE *f, const char *what, int allow_precursor)
{
  const char *groups[4] = {"PRESSURE\n", "X-VELOCITY\n", "Y-VELOCITY\n",
    "Z-VELOCITY\n"};
  const char comps[4] = {'p', 'u', 'v', 'w'};
  for(int g = 0; g < 4; g++) {
    skip(f, groups[g]);
    for(int q = 0; q < 6; q++)
      if(read_bc_line(f, comps[g], q, allow_precursor))
        return read_error(f, what);
    if(g == 0 && (bc.dsE < 0 || bc.dsW < 0 || bc.dsN < 0 || bc.dsS < 0 ||
                  bc.dsT < 0 || bc.dsB < 0)) {
      fprintf(stderr, "%s read error -- screen offsets must be >= 0\n", what);
      fclose(f);
      return -1;
    }
  }
  return 0;
}

static int read_out_plane(FILE *f, const char *what)
{
  const char *names[7] = {"WEST", "EAST", "SOUTH", "NORTH", "BOTTOM", "TOP",
    "HOMOGENEOUS"};
  const int vals[7] = {WEST, EAST, SOUTH, NORTH, BOTTOM, TOP, HOMOGENEOUS};
  skip(f, "SOLVABILITY ENFORCEMENT PLANE\n");
  if(read_keyword(f, "out_plane %255s", names, vals, 7, &out_plane))
    return read_error(f, what);
  skip(f, "\n");
  skip(f, "\n");
  return 0;
}

int bb_read_flow_config(const char *path)
{
  FILE *f = fopen(path, "r");
  if(f == NULL) {
    fprintf(stderr, "Could not open file %s\n", path);
    return -1;
  }
  if(read_domain_head(f, "flow.config")) return -1;
  skip(f, "SIMULATION PARAMETERS\n");
  read_real(f, "duration %lf\n", &duration);
  read_real(f, "CFL %lf\n", &CFL);
  read_int(f, "pp_max_iter %d\n", &pp_max_iter);
  read_real(f, "pp_residual %lf\n", &pp_residual);
  read_int(f, "lamb_max_iter %d\n", &lamb_max_iter);
  read_real(f, "lamb_residual %lf\n", &lamb_residual);
  read_real(f, "lamb_relax %lf\n", &lamb_relax);
  read_real(f, "lamb_cut %lf\n", &lamb_cut);
  skip(f, "\n");
  skip(f, "BOUNDARY CONDITIONS\n");
  read_real(f, "vel_tDelay %lf\n", &vel_tDelay);
  if(read_bc_groups(f, "flow.config", 1)) return -1;
  skip(f, "\n");
  skip(f, "INITIAL CONDITION\n");
  {
    const char *names[4] = {"QUIESCENT", "SHEAR", "CHANNEL", "TAYLOR_GREEN"};
    const int vals[4] = {QUIESCENT, SHEAR, CHANNEL, TAYLOR_GREEN_INIT};
    if(read_keyword(f, "init_cond %255s", names, vals, 4, &init_cond))
      return read_error(f, "flow.config");
  }
  skip(f, "\n");
  skip(f, "\n");
  if(read_out_plane(f, "flow.config")) return -1;
  skip(f, "SIMULATION DRIVING CONDITIONS\n");
  read_real(f, "p_tDelay %lf\n", &p_tDelay);
  if(fscanf(f, "gradP.x %lf %lf\n", &gradP.xm, &gradP.xa) < 0) {}
  if(fscanf(f, "gradP.y %lf %lf\n", &gradP.ym, &gradP.ya) < 0) {}
  if(fscanf(f, "gradP.z %lf %lf\n", &gradP.zm, &gradP.za) < 0) {}
  gradP.x = gradP.y = gradP.z = 0;
  read_real(f, "g_tDelay %lf\n", &g_tDelay);
  if(fscanf(f, "g.x %lf %lf\n", &g.xm, &g.xa) < 0) {}
  if(fscanf(f, "g.y %lf %lf\n", &g.ym, &g.ya) < 0) {}
  if(fscanf(f, "g.z %lf %lf\n", &g.zm, &g.za) < 0) {}
  g.x = g.y = g.z = 0;
  skip(f, "\n");
  skip(f, "PID CONTROLLER GAINS\n");
  read_real(f, "Kp %lf\n", &Kp);
  read_real(f, "Ki %lf\n", &Ki);
  read_real(f, "Kd %lf\n", &Kd);
  turbA = 0.;
  fclose(f);
  return 0;
}

/* src/domain.c:30 (reads ROOT_DIR/input/flow.config; exits on error) */
void domain_read_input(void)
{
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/input/flow.config", bb_root_dir);
  if(bb_read_flow_config(path)) exit(EXIT_FAILURE);
}

/* ---------------------------------------------------------------------
 * turb.config: the precursor domain's file (src/domain.c:732-1287).  The
 * flow.config grammar without the Lamb parameters, vel_tDelay, the
 * PRECURSOR face type, the initial condition and the PID gains; its
 * driving section holds only turbA.  The settings the precursor does not
 * use are zeroed as there.
 * ------------------------------------------------------------------- */
int bb_read_turb_config(const char *path)
{
  FILE *f = fopen(path, "r");
  if(f == NULL) {
    fprintf(stderr, "Could not open file %s\n", path);
    return -1;
  }
  if(read_domain_head(f, "turb.config")) return -1;
  skip(f, "SIMULATION PARAMETERS\n");
  read_real(f, "duration %lf\n", &duration);
  read_real(f, "CFL %lf\n", &CFL);
  read_int(f, "pp_max_iter %d\n", &pp_max_iter);
  read_real(f, "pp_residual %lf\n", &pp_residual);
  lamb_max_iter = 0;
  lamb_residual = 0.;
  lamb_relax = 0.;
  lamb_cut = 0.;
  skip(f, "\n");
  skip(f, "BOUNDARY CONDITIONS\n");
  vel_tDelay = 0;
  if(read_bc_groups(f, "turb.config", 0)) return -1;
  skip(f, "\n");
  if(read_out_plane(f, "turb.config")) return -1;
  skip(f, "SIMULATION DRIVING CONDITIONS\n");
  read_real(f, "turbA %lf\n", &turbA);
  p_tDelay = 0;
  memset(&gradP, 0, sizeof(gradP));
  g_tDelay = 0;
  memset(&g, 0, sizeof(g));
  fclose(f);
  return 0;
}

/* src/domain.c:732 (ROOT_DIR/input/turb.config; exits on error) */
void turb_read_input(void)
{
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/input/turb.config", bb_root_dir);
  if(bb_read_turb_config(path)) exit(EXIT_FAILURE);
}

/* ---------------------------------------------------------------------
 * part.config
 * ------------------------------------------------------------------- */
int bb_read_part_config(const char *path)
{
  FILE *f = fopen(path, "r");
  if(f == NULL) {
    fprintf(stderr, "Could not open file %s\n", path);
    return -1;
  }
  int n = 0;
  if(fscanf(f, "n %d\n", &n) != 1 || n < 0) return read_error(f, "part.config");
  read_real(f, "(l/a) %lf\n", &interactionLengthRatio);
  if(bb_host_particles(n)) return read_error(f, "part.config");
  for(int i = 0; i < n; i++) {
    part_struct *P = &parts[i];
    skip(f, "\n");
    /* the radius and the position are required; every other line may be
     * absent, as with the reference's unchecked fscanf sequence (older
     * files have no e_dry, coeff_fric or rs/r), and then stays 0 */
    int ok = fscanf(f, "r %lf\n", &P->r) == 1;
    ok &= fscanf(f, "(x, y, z) %lf %lf %lf\n", &P->x, &P->y, &P->z) == 3;
    if(!ok) {
      fprintf(stderr, "part.config: particle %d incomplete\n", i);
      return read_error(f, "part.config");
    }
    if(fscanf(f, "(aFx, aFy, aFz) %lf %lf %lf\n", &P->aFx, &P->aFy,
       &P->aFz) < 0) {}
    if(fscanf(f, "(aLx, aLy, aLz) %lf %lf %lf\n", &P->aLx, &P->aLy,
       &P->aLz) < 0) {}
    read_real(f, "rho %lf\n", &P->rho);
    read_real(f, "E %lf\n", &P->E);
    read_real(f, "sigma %lf\n", &P->sigma);
    read_real(f, "e_dry %lf\n", &P->e_dry);
    read_real(f, "coeff_fric %lf\n", &P->coeff_fric);
    read_int(f, "order %d\n", &P->order);
    read_real(f, "rs/r %lf\n", &P->rs);
    read_real(f, "spring_k %lf\n", &P->spring_k);
    if(fscanf(f, "spring (x, y, z) %lf %lf %lf\n", &P->spring_x,
       &P->spring_y, &P->spring_z) < 0) {}
    read_real(f, "spring_l %lf\n", &P->spring_l);
    read_int(f, "translating %d\n", &P->translating);
    read_int(f, "rotating %d\n", &P->rotating);
  }
  fclose(f);
  return 0;
}

/* src/particle.c:79 (ROOT_DIR/input/part.config; exits on error) */
void parts_read_input(int turb)
{
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/input/part.config", bb_root_dir);
  if(bb_read_part_config(path)) exit(EXIT_FAILURE);
  if(turb) bb_host_particles(0);
}

/* ---------------------------------------------------------------------
 * Restart dump (src/domain.c:3223-3413): the step counters, every field
 * of the three face grids and the cell grid, the integer flags, the
 * precursor plane positions, the raw particle records, the 18 Lamb
 * coefficient arrays and the recorder / controller state.  The writer
 * stores gradP.z with a count of 2 (gradP.z and gradP.zm, as the reference
 * writes it) and the reader takes 1, so a file ends one real after what
 * is read back.
 * ------------------------------------------------------------------- */
typedef struct { void *p; size_t size, count; } chunk_t;

static int face_chunks(chunk_t *c, int n, real *x, real *x0, real *d0,
  real *cv0, real *d, real *cv, real *xs, size_t cnt, int with_diff0)
{
  c[n++] = (chunk_t) {x, sizeof(real), cnt};
  c[n++] = (chunk_t) {x0, sizeof(real), cnt};
  if(with_diff0) c[n++] = (chunk_t) {d0, sizeof(real), cnt};
  c[n++] = (chunk_t) {cv0, sizeof(real), cnt};
  c[n++] = (chunk_t) {d, sizeof(real), cnt};
  c[n++] = (chunk_t) {cv, sizeof(real), cnt};
  c[n++] = (chunk_t) {xs, sizeof(real), cnt};
  return n;
}

/* the dump's layout, in order; returns the number of chunks */
static int restart_layout(chunk_t *c, int for_read)
{
  const int diff0 = !bb_is_implicit();  /* absent in the IMPLICIT variant */
  const size_t gx = Dom.Gfx.s3b, gy = Dom.Gfy.s3b, gz = Dom.Gfz.s3b;
  const size_t gc = Dom.Gcc.s3b;
  const size_t nc = (size_t) nparts * coeff_stride;
  int n = 0;
  c[n++] = (chunk_t) {&ttime, sizeof(real), 1};
  c[n++] = (chunk_t) {&dt0, sizeof(real), 1};
  c[n++] = (chunk_t) {&dt, sizeof(real), 1};
  c[n++] = (chunk_t) {&stepnum, sizeof(int), 1};
  c[n++] = (chunk_t) {&rec_paraview_stepnum_out, sizeof(int), 1};
  n = face_chunks(c, n, u, u0, diff0_u, conv0_u, diff_u, conv_u, u_star, gx,
    diff0);
  n = face_chunks(c, n, v, v0, diff0_v, conv0_v, diff_v, conv_v, v_star, gy,
    diff0);
  n = face_chunks(c, n, w, w0, diff0_w, conv0_w, diff_w, conv_w, w_star, gz,
    diff0);
  c[n++] = (chunk_t) {p, sizeof(real), gc};
  c[n++] = (chunk_t) {phi, sizeof(real), gc};
  c[n++] = (chunk_t) {p0, sizeof(real), gc};
  c[n++] = (chunk_t) {phase, sizeof(int), gc};
  c[n++] = (chunk_t) {phase_shell, sizeof(int), gc};
  c[n++] = (chunk_t) {flag_u, sizeof(int), gx};
  c[n++] = (chunk_t) {flag_v, sizeof(int), gy};
  c[n++] = (chunk_t) {flag_w, sizeof(int), gz};
  c[n++] = (chunk_t) {bc_plane_pos, sizeof(real), 9};
  c[n++] = (chunk_t) {parts, sizeof(part_struct), (size_t) nparts};
  c[n++] = (chunk_t) {&coeff_stride, sizeof(int), 1};
  real *coeffs[18] = {pnm_re, pnm_im, pnm_re0, pnm_im0, pnm_re00, pnm_im00,
    phinm_re, phinm_im, phinm_re0, phinm_im0, phinm_re00, phinm_im00,
    chinm_re, chinm_im, chinm_re0, chinm_im0, chinm_re00, chinm_im00};
  for(int a = 0; a < 18; a++) c[n++] = (chunk_t) {coeffs[a], sizeof(real), nc};
  c[n++] = (chunk_t) {&rec_flow_field_ttime_out, sizeof(real), 1};
  c[n++] = (chunk_t) {&rec_paraview_ttime_out, sizeof(real), 1};
  c[n++] = (chunk_t) {&rec_particle_ttime_out, sizeof(real), 1};
  c[n++] = (chunk_t) {&rec_restart_ttime_out, sizeof(real), 1};
  c[n++] = (chunk_t) {&rec_prec_ttime_out, sizeof(real), 1};
  c[n++] = (chunk_t) {&pid_int, sizeof(real), 1};
  c[n++] = (chunk_t) {&pid_back, sizeof(real), 1};
  c[n++] = (chunk_t) {&gradP.z, sizeof(real), for_read ? 1 : 2};
  return n;
}

/* the precursor domain's dump (src/domain.c:3415-3546): the same head
 * and field arrays, the plane positions, then its two output clocks; no
 * particles, coefficients or controller state */
static int restart_turb_layout(chunk_t *c)
{
  const int diff0 = !bb_is_implicit();
  const size_t gx = Dom.Gfx.s3b, gy = Dom.Gfy.s3b, gz = Dom.Gfz.s3b;
  const size_t gc = Dom.Gcc.s3b;
  int n = 0;
  c[n++] = (chunk_t) {&ttime, sizeof(real), 1};
  c[n++] = (chunk_t) {&dt0, sizeof(real), 1};
  c[n++] = (chunk_t) {&dt, sizeof(real), 1};
  c[n++] = (chunk_t) {&stepnum, sizeof(int), 1};
  c[n++] = (chunk_t) {&rec_paraview_stepnum_out, sizeof(int), 1};
  n = face_chunks(c, n, u, u0, diff0_u, conv0_u, diff_u, conv_u, u_star, gx,
    diff0);
  n = face_chunks(c, n, v, v0, diff0_v, conv0_v, diff_v, conv_v, v_star, gy,
    diff0);
  n = face_chunks(c, n, w, w0, diff0_w, conv0_w, diff_w, conv_w, w_star, gz,
    diff0);
  c[n++] = (chunk_t) {p, sizeof(real), gc};
  c[n++] = (chunk_t) {phi, sizeof(real), gc};
  c[n++] = (chunk_t) {p0, sizeof(real), gc};
  c[n++] = (chunk_t) {phase, sizeof(int), gc};
  c[n++] = (chunk_t) {phase_shell, sizeof(int), gc};
  c[n++] = (chunk_t) {flag_u, sizeof(int), gx};
  c[n++] = (chunk_t) {flag_v, sizeof(int), gy};
  c[n++] = (chunk_t) {flag_w, sizeof(int), gz};
  c[n++] = (chunk_t) {bc_plane_pos, sizeof(real), 9};
  c[n++] = (chunk_t) {&rec_prec_ttime_out, sizeof(real), 1};
  c[n++] = (chunk_t) {&rec_prec_flow_field_ttime_out, sizeof(real), 1};
  return n;
}

static int write_chunks(const char *path, const chunk_t *c, int n)
{
  FILE *f = fopen(path, "w");
  if(f == NULL) {
    fprintf(stderr, "Could not open file %s\n", path);
    return -1;
  }
  for(int i = 0; i < n; i++)
    if(c[i].count && fwrite(c[i].p, c[i].size, c[i].count, f) != c[i].count) {
      fclose(f);
      return -1;
    }
  fclose(f);
  return 0;
}

int bb_out_restart(const char *path)
{
  chunk_t c[96];
  return write_chunks(path, c, restart_layout(c, 0));
}

int bb_out_restart_turb(const char *path)
{
  chunk_t c[64];
  return write_chunks(path, c, restart_turb_layout(c));
}

int bb_in_restart_turb(const char *path)
{
  FILE *f = fopen(path, "r");
  if(f == NULL) {
    fprintf(stderr, "Could not open file %s\n", path);
    return -1;
  }
  chunk_t c[64];
  const int n = restart_turb_layout(c);
  for(int i = 0; i < n; i++)
    if(c[i].count && fread(c[i].p, c[i].size, c[i].count, f) != c[i].count) {
      fprintf(stderr, "restart: file ends early\n");
      fclose(f);
      return -1;
    }
  fclose(f);
  return 0;
}

int bb_in_restart(const char *path)
{
  FILE *f = fopen(path, "r");
  if(f == NULL) {
    fprintf(stderr, "Could not open file %s\n", path);
    return -1;
  }
  chunk_t c[96];
  int n = restart_layout(c, 1);
  for(int i = 0; i < n; i++) {
    if(c[i].p == (void *) &coeff_stride) {
      /* the coefficient arrays that follow are sized by this value */
      int cs = 0;
      if(fread(&cs, sizeof(int), 1, f) != 1 || cs != coeff_stride) {
        fprintf(stderr, "restart: coefficient stride %d, the set-up has "
          "%d\n", cs, coeff_stride);
        fclose(f);
        return -1;
      }
      continue;
    }
    if(c[i].count && fread(c[i].p, c[i].size, c[i].count, f) != c[i].count) {
      fprintf(stderr, "restart: file ends early\n");
      fclose(f);
      return -1;
    }
  }
  fclose(f);
  return 0;
}

/* src/domain.c:3223 and 3319 (ROOT_DIR/input/restart.config) */
void out_restart(void)
{
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/input/restart.config", bb_root_dir);
  if(bb_out_restart(path)) exit(EXIT_FAILURE);
}

void in_restart(void)
{
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/input/restart.config", bb_root_dir);
  if(bb_in_restart(path)) exit(EXIT_FAILURE);
}

/* src/domain.c:3415 and 3480 (ROOT_DIR/input/restart_turb.config) */
void out_restart_turb(void)
{
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/input/restart_turb.config", bb_root_dir);
  if(bb_out_restart_turb(path)) exit(EXIT_FAILURE);
}

void in_restart_turb(void)
{
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/input/restart_turb.config", bb_root_dir);
  if(bb_in_restart_turb(path)) exit(EXIT_FAILURE);
}

/* ---------------------------------------------------------------------
 * Lamb-coefficient log (src/recorder.c:1912-2001): per call one header
 * line, then per particle its coefficients n, m, p_nm, phi_nm, chi_nm
 * (real and imaginary parts) for the particle's own expansion length.
 * ------------------------------------------------------------------- */
void recorder_lamb(char *name, int iter)
{
  char path[2 * FILE_NAME_SIZE + 16];
  snprintf(path, sizeof(path), "%s/record/%s", bb_root_dir, name);
  FILE *rec = fopen(path, "a");
  if(rec == NULL) {
    fprintf(stderr, "Could not open file %s\n", name);
    exit(EXIT_FAILURE);
  }
  fprintf(rec, "ttime = %e, iter = %d\n", ttime, iter);
  for(int i = 0; i < nparts; i++) {
    fprintf(rec, "particle %d\n", i);
    fprintf(rec, "%4s%4s%12s%12s", "n", "m", "pnm_re", "pnm_im");
    fprintf(rec, "%12s%12s", "phinm_re", "phinm_im");
    fprintf(rec, "%12s%12s\n", "chinm_re", "chinm_im");
    int n = 0, m = 0;
    for(int c = 0; c < parts[i].ncoeff; c++) {
      const int k = coeff_stride * i + c;
      fprintf(rec, "%4d%4d", n, m);
      fprintf(rec, "%12.3e%12.3e", pnm_re[k], pnm_im[k]);
      fprintf(rec, "%12.3e%12.3e", phinm_re[k], phinm_im[k]);
      fprintf(rec, "%12.3e%12.3e\n", chinm_re[k], chinm_im[k]);
      if(++m > n) { n++; m = 0; }
    }
  }
  fclose(rec);
}

/* ---------------------------------------------------------------------
 * record.config (src/recorder.c:26-371): output intervals, one "TYPE dt"
 * line each (dt < 0 turns the output off); FLOW_FIELD and PARTICLE take
 * option lines up to the next blank line, every other type consumes the
 * line after it.  As in the reference, only "velocity" and "pressure"
 * are accepted after FLOW_FIELD and only "position" after PARTICLE, and
 * each type line is tested against the types in table order with the
 * buffer as the earlier types' handling left it.  Departure: an option
 * list that runs to the end of the file ends there (the reference keeps
 * re-reading its last line).
 * ------------------------------------------------------------------- */
static const char *REC_TYPES[7] = {"FLOW_FIELD", "PARAVIEW", "PARTICLE",
  "RESTART", "PRECURSOR_VTK", "RESTART_STOP", "PRECURSOR"};

/* option lines after FLOW_FIELD / PARTICLE; -1 on an unknown option */
static int rec_options(FILE *f, char *buf, int len, const char *type,
  const char *const *opts, int *const *flags, int nopt)
{
  if(fgets(buf, len, f) == NULL) return 0;
  while(strcmp("\n", buf) != 0) {
    size_t ln = strlen(buf);
    if(ln > 0 && buf[ln - 1] == '\n') buf[ln - 1] = '\0';
    int found = 0;
    for(int j = 0; j < nopt; j++)
      if(strcmp(opts[j], buf) == 0) { *flags[j] = 1; found = 1; }
    if(!found) {
      fprintf(stderr, "Unrecognized record.config %s option %s\n", type, buf);
      return -1;
    }
    if(fgets(buf, len, f) == NULL) break;
  }
  return 0;
}

int bb_read_record_config(const char *path)
{
  rec_flow_field_dt = -1; rec_flow_field_vel = 0; rec_flow_field_p = 0;
  rec_flow_field_phase = 0;
  rec_paraview_dt = -1;
  rec_particle_dt = -1; rec_particle_pos = 0; rec_particle_a = 0;
  rec_particle_vel = 0; rec_particle_omega = 0; rec_particle_force = 0;
  rec_particle_moment = 0;
  rec_restart_dt = -1; rec_prec_dt = -1; rec_restart_stop = 0;
  rec_prec_flow_field_dt = -1; rec_prec_flow_field_vel = 0;
  rec_prec_flow_field_p = 0; rec_prec_flow_field_phase = 0;
  FILE *f = fopen(path, "r");
  if(f == NULL) {
    fprintf(stderr, "Could not open file %s\n", path);
    return -1;
  }
  static const char *const flow_opts[2] = {"velocity", "pressure"};
  int *const flow_flags[2] = {&rec_flow_field_vel, &rec_flow_field_p};
  static const char *const part_opts[1] = {"position"};
  int *const part_flags[1] = {&rec_particle_pos};
  char buf[CHAR_BUF_SIZE] = "", word[CHAR_BUF_SIZE] = "";
  while(fgets(buf, CHAR_BUF_SIZE, f) != NULL) {
    int found = 0;
    for(int i = 0; i < 7; i++) {
      real dtout = 0;
      sscanf(buf, "%s %lf\n", word, &dtout);
      if(strcmp(REC_TYPES[i], word) != 0) continue;
      found = 1;
      int err = 0;
      switch(i) {
        case 0: rec_flow_field_dt = dtout;
          err = rec_options(f, buf, CHAR_BUF_SIZE, word, flow_opts, flow_flags, 2);
          break;
        case 2: rec_particle_dt = dtout;
          err = rec_options(f, buf, CHAR_BUF_SIZE, word, part_opts, part_flags, 1);
          break;
        default: {
          real *dst[7] = {NULL, &rec_paraview_dt, NULL, &rec_restart_dt,
            &rec_prec_dt, NULL, &rec_prec_flow_field_dt};
          if(i == 5) rec_restart_stop = (int) dtout;
          else *dst[i] = dtout;
          if(fgets(buf, CHAR_BUF_SIZE, f) == NULL) buf[0] = '\0';
          break;
        }
      }
      if(err) { fclose(f); return -1; }
    }
    if(!found) {
      fprintf(stderr, "Unrecognized record.config type %s\n", word);
      fclose(f);
      return -1;
    }
  }
  fclose(f);
  return 0;
}

void recorder_read_config(void)
{
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/input/record.config", bb_root_dir);
  if(bb_read_record_config(path)) exit(EXIT_FAILURE);
}

/* src/recorder.c:1891-1925: create record/<name> (the solver log with its
 * column header, the Lamb log empty) */
static FILE *record_file(const char *name)
{
  char path[2 * FILE_NAME_SIZE + 16];
  snprintf(path, sizeof(path), "%s/record/%s", bb_root_dir, name);
  FILE *rec = fopen(path, "w");
  if(rec == NULL) {
    fprintf(stderr, "Could not open file %s\n", name);
    exit(EXIT_FAILURE);
  }
  return rec;
}

void recorder_bicgstab_init(char *name)
{
  FILE *rec = record_file(name);
  fprintf(rec, "%-12s", "stepnum");
  fprintf(rec, "%-15s", "ttime");
  fprintf(rec, "%-15s", "dt");
  fprintf(rec, "%-8s", "niter");
  fprintf(rec, "%-15s", "resid");
  fclose(rec);
}

void recorder_lamb_init(char *name)
{
  fclose(record_file(name));
}
