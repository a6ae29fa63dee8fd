/*
 * bb_host.h -- set-up and stepping API of the driver-side library
 * (libbluebottle_host.so).  It stands in for the parts of the reference's
 * driver that are not on the hot path: domain_init / parts_init /
 * binDom_init and the main loop body of src/bluebottle.c.
 */
#ifndef BB_HOST_H
#define BB_HOST_H

#include "../../include/bluebottle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Directory that recorder_bicgstab appends solver_expd.rec rows to; empty
 * (the default) keeps only the last row in bb_rec_niter / bb_rec_resid. */
extern char bb_record_dir[FILE_NAME_SIZE];
extern int bb_rec_niter;
extern real bb_rec_resid;
extern int bb_rec_calls;
extern long bb_rec_total_iters;

/* Grid set-up and zeroed host arrays (src/domain.c:1591-2347).  The caller
 * sets rho_f, nu, mu, CFL and the BC/driving globals before calling, and
 * fills the initial velocity afterwards. */
int bb_host_domain(real xs, real xe, int xn, real ys, real ye, int yn,
  real zs, real ze, int zn);
/* Allocate n zeroed particles; the caller fills the parts.config fields. */
int bb_host_particles(int n);
/* Initial particle state, flags, coefficients and bins
 * (src/particle.c:295-762). */
int bb_host_particles_init(void);
void bb_host_free(void);
void compute_vel_BC(void);
/* src/bluebottle.c:445-545 after the device pushes. */
int bb_host_prepare(void);
int bb_host_restart(const char *path);
/* src/bluebottle.c:598-711; returns the Lamb sub-iteration count, or -1
 * when the solution has diverged (the driver's dt test). */
int bb_host_step(void);

/* ---- the precursor domain (src/bluebottle.c:850-1156) ---- */
/* the random initial field of src/domain.c:2856-3093 on the arrays that
 * bb_host_domain made, seeded with `seed`; sets turbl and the first dt */
int bb_host_init_turb(unsigned long long seed);
/* its correction and mean removal alone, on the current u, v, w (and u0,
 * v0, w0), leaving the scaled divergence in p (src/domain.c:2922-3051) */
int bb_host_turb_correct(void);
/* src/domain.c:2349 (grids, arrays, the field seeded from the clock);
 * EXIT_SUCCESS or EXIT_FAILURE */
int domain_init_turb(void);
/* src/bluebottle.c:945-1016 after the pushes */
int bb_host_prepare_turb(void);
/* the same with the precursor's restart file in place of its set-up */
int bb_host_restart_turb(const char *path);
/* src/bluebottle.c:1036-1081: one step; 0, or -1 when it diverged */
int bb_host_step_turb(void);
/* zero the output step counters and clocks */
void bb_reset_output_clocks(void);

/* ---- random numbers (bb_rng.c; src/rng.h) ---- */
extern unsigned long long int rng_j, rng_v;
void rng_init(unsigned long long int seed);
unsigned long long int rng_uint64(void);
unsigned int rng_uint32(void);
long long int rng_int64(void);
int rng_int32(void);
float rng_flt(void);
double rng_dbl(void);
double gaussrand(void);

/* ---- files in the reference's formats (bb_io.c) ---- */
/* directory holding input/ and record/ (the reference's ROOT_DIR, ".") */
extern char bb_root_dir[FILE_NAME_SIZE];
/* flow.config (src/domain.c:30-730) into the globals; 0 or -1 with the
 * reference's message.  Extension: init_cond TAYLOR_GREEN (100). */
int bb_read_flow_config(const char *path);
/* part.config (src/particle.c:79-159): allocates and fills parts[] */
int bb_read_part_config(const char *path);
/* the raw restart dump of src/domain.c:3223-3413 (host arrays) */
int bb_out_restart(const char *path);
int bb_in_restart(const char *path);
/* turb.config (src/domain.c:732-1287), the precursor domain's file */
int bb_read_turb_config(const char *path);
/* the precursor's dump (src/domain.c:3415-3546) */
int bb_out_restart_turb(const char *path);
int bb_in_restart_turb(const char *path);
/* the reference's names, on ROOT_DIR/input/...; they exit on error */
void domain_read_input(void);
void turb_read_input(void);
void out_restart_turb(void);
void in_restart_turb(void);
void parts_read_input(int turb);
void out_restart(void);
void in_restart(void);
/* src/recorder.c:1955-2001, appending to ROOT_DIR/record/<name> */
void recorder_lamb(char *name, int iter);
/* record.config (src/recorder.c:26-371) into the rec_* globals; 0 or -1 */
int bb_read_record_config(const char *path);
/* the reference's names: ROOT_DIR/input/record.config (exits on error),
 * and the creation of ROOT_DIR/record/<name> */
void recorder_read_config(void);
void recorder_bicgstab_init(char *name);
void recorder_lamb_init(char *name);

/* ---- ParaView output (bb_vtk.c, src/vtk.c) into directory `dir` ---- */
int bb_vtk_init(const char *dir);             /* empty out.pvd */
int bb_vtk_grid(const char *dir, int step);   /* out_N.pvtr, out_N_0.vtr */
int bb_vtk_parts(const char *dir, int step);  /* out_N.vtp */
int bb_vtk_nodes(const char *dir, int step);  /* out_nodes_N.vtp */
int bb_vtk_out(const char *dir, int step);    /* all of them + out.pvd */
/* the precursor domain's files (src/vtk.c:324-611) */
int bb_vtk_turb_init(const char *dir);        /* empty out_turb.pvd */
int bb_vtk_turb_grid(const char *dir, int step); /* out_turb_N.pvtr, _N_0.vtr */
int bb_vtk_turb_out(const char *dir, int step);  /* both + out_turb.pvd */
/* the debugging output with the ghost cells (src/vtk.c:612-933) */
int bb_vtk_ghost_init(const char *dir);        /* empty out_ghost.pvd */
int bb_vtk_ghost_grid(const char *dir, int step); /* out_ghost_N.pvtr, _N_0.vtr */
int bb_vtk_ghost_out(const char *dir, int step);  /* + out_N.vtp, out_ghost.pvd */
/* the reference's names, on ROOT_DIR/output; they exit on error */
void init_VTK(void);
void out_VTK(void);
void dom_out_VTK(void);
void part_out_VTK(void);
void quadnodes_out_VTK(void);
void init_VTK_turb(void);
void out_VTK_turb(void);
void dom_out_VTK_turb(void);
void init_VTK_ghost(void);
void out_VTK_ghost(void);
void dom_out_VTK_ghost(void);

#ifdef __cplusplus
}
#endif

#endif
