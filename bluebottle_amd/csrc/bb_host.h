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
/* src/bluebottle.c:598-711; returns the Lamb sub-iteration count. */
int bb_host_step(void);

#ifdef __cplusplus
}
#endif

#endif
