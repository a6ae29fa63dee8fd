/*
 * bb_device.h -- shared declarations of the HIP library: the per-process
 * device context (stream, Krylov and particle workspaces), the compact
 * grid geometry passed to kernels by value, error checking, block
 * reductions for 64-wide wavefronts, and per-kernel event timing.
 */
#ifndef BB_DEVICE_H
#define BB_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/bluebottle.h"

/* A driver built as the reference's IMPLICIT variant defines no diff0
 * history arrays (src/bluebottle.c:60-67): they are referenced weakly and
 * their absence selects that variant (bb_set_implicit). */
extern "C" {
#pragma weak diff0_u
#pragma weak diff0_v
#pragma weak diff0_w
#pragma weak _diff0_u
#pragma weak _diff0_v
#pragma weak _diff0_w
}

#define BB_EXIT_HIP 86
void bb_fatal(const char *what, hipError_t err, const char *file, int line);

#define BB_CHECK(call)                                                      \
  do {                                                                      \
    hipError_t bb_e_ = (call);                                              \
    if(bb_e_ != hipSuccess) bb_fatal(#call, bb_e_, __FILE__, __LINE__);     \
  } while(0)

#define BB_CHECK_LAUNCH() BB_CHECK(hipGetLastError())

/* Geometry of the single subdomain, in kernel-argument form.  Index maps:
 *   Gcc: i + j*cs1 + k*cs2      (i in [0, xn+2), ...)
 *   Gfx: i + j*us1 + k*us2      (i in [0, xn+3))
 *   Gfy: i + j*vs1 + k*vs2      (j in [0, yn+3))
 *   Gfz: i + j*ws1 + k*ws2      (k in [0, zn+3))
 * Solver vectors have no ghosts: (i-1) + (j-1)*xn + (k-1)*xn*yn. */
struct BBGeom {
  int xn, yn, zn;
  int cs1, cs2, us1, us2, vs1, vs2, ws1, ws2;
  int ccount, ucount, vcount, wcount;   /* s3b of each grid */
  double dx, dy, dz, xs, ys, zs, xe, ye, ze, xl, yl, zl;
};

BBGeom bb_geom(void);

/* Packed per-cell operator flags for the pressure Poisson matrix. */
enum {
  BB_FW = 1, BB_FE = 2, BB_FS = 4, BB_FN = 8, BB_FB = 16, BB_FT = 32,
  BB_SOLID = 64
};

/* Device-resident Krylov scalars (one copy, written by the finalise kernels
 * and read by the sweeps; the host polls `done`). */
struct KrylovScalars {
  double rho, alpha, omega, beta;
  double tol, resid, bnorm;
  double sum0, sum1, sum2;
  int iter, max_iter, done, s_conv, converged;
  int cur;                  /* two-sweep form: buffer set of the last iteration */
};

struct BBWorkspace {
  size_t n;                 /* Gcc.s3 */
  int nblk;                 /* reduction partial slots */
  double *x, *r, *rs, *p, *v, *t, *b;   /* FP64 Krylov vectors */
  double *p2, *v2;          /* second buffers of p and v (ping-pong) */
  float *xf, *rf, *rsf, *pf, *vf, *tf, *bf;  /* FP32 inner-solver vectors */
  float *pf2, *vf2, *rf2;   /* second buffers of the two-sweep form */
  float *phf, *shf;         /* preconditioned p and s of the flexible solve */
  uint8_t *flags;
  double *partials;
  KrylovScalars *sc;        /* device */
  KrylovScalars *sc_host;   /* pinned mirror */
  double *red;              /* general reduction scratch */
  double *red_host;         /* pinned */
};

struct BBContext {
  int ready;
  int device;
  hipStream_t stream;
  BBWorkspace ws;
  int solver_mode;
  bb_solve_stats stats;
  int exit_on_fail;         /* reference behaviour: exit(EXIT_FAILURE) */
  int implicit;             /* IMPLICIT variant, fixed at cuda_dom_malloc */
  /* particle workspace */
  int *face_list;           /* flagged faces, packed per component */
  int face_cap;
  int *face_count;          /* device, 3 counters */
  int *face_count_host;      /* pinned host copy of the 3 counters */
  int *bin_ind, *bin_bin, *bin_start, *bin_end;
  int bin_cap, nbins_cap;
  double *quad;             /* nparts x 26 x 4 node values */
  double *coeff_err;        /* nparts */
  double *node_t, *node_p;  /* device node tables (26 each) */
  double *host_err;         /* pinned */
  int part_cap;
};

extern BBContext g_bb;

/* Make sure the context exists (device selected, stream created). */
void bb_ensure_context(void);
/* Allocate / grow the Krylov workspace for the current Dom. */
void bb_ensure_workspace(void);
void bb_free_workspace(void);
/* learned preconditioner: pair store and network (bb_poisson.hip) */
void bb_learn_free(void);
/* Crank-Nicolson solver workspace (bb_helmholtz.hip) */
void bb_free_helmholtz(void);
/* precursor inflow planes (bb_precursor.hip), with cuda_dom_malloc/free */
void bb_alloc_planes(void);
void bb_free_planes(void);
/* Particle workspace sized for nparts and the current grid. */
void bb_ensure_particle_workspace(void);
void bb_invalidate_face_lists(void);
/* copies between the driver's (pageable) host arrays and device memory,
 * through the library's pinned staging buffer; both return when done */
void bb_h2d(void *dev, const void *host, size_t bytes);
void bb_d2h(void *host, const void *dev, size_t bytes);

/* ---- kernel timing (enabled by bb_prof_enable) ---- */
#define BB_PROF_MAX_KINDS 32
void bb_prof_begin(int kind);
void bb_prof_end(int kind);
enum {
  BB_K_USTAR = 0, BB_K_BC, BB_K_RHS, BB_K_PACK, BB_K_SPMV_P, BB_K_SPMV_S,
  BB_K_UPDATE_XR, BB_K_UPDATE_P, BB_K_FINAL, BB_K_RESID64, BB_K_CORRECT64,
  BB_K_SPMV_P32, BB_K_SPMV_S32, BB_K_UPDATE_XR32, BB_K_UPDATE_P32,
  BB_K_PROJECT, BB_K_UPDATEP, BB_K_FINDDT, BB_K_CAGE, BB_K_PARTBC,
  BB_K_QUAD, BB_K_MOVE, BB_K_MISC, BB_K_UNET, BB_K_UNET_CONV0, BB_K_FSPMV,
  BB_K_FVEC, BB_K_SWEEP_A32, BB_K_SWEEP_B32, BB_K_NKINDS
};
static_assert((int) BB_K_NKINDS == (int) BB_PROF_NKINDS &&
  (int) BB_K_SPMV_S32 == (int) BB_PROF_SPMV_S32 &&
  (int) BB_K_MISC == (int) BB_PROF_MISC, "kernel kinds follow bluebottle.h");

/* ---- block reductions over 256 threads (4 wavefronts of 64) ---- */
__device__ inline double bb_wave_sum(double v)
{
#pragma unroll
  for(int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

/* Sum `v` over the block; the result is valid in thread 0.  `sh` needs
 * blockDim/64 doubles. */
__device__ inline double bb_block_sum(double v, double *sh)
{
  int tid = threadIdx.x + threadIdx.y * blockDim.x + threadIdx.z * blockDim.x * blockDim.y;
  int nth = blockDim.x * blockDim.y * blockDim.z;
  v = bb_wave_sum(v);
  __syncthreads();
  if((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  double s = 0.;
  if(tid == 0) {
    for(int w = 0; w < (nth + 63) / 64; w++) s += sh[w];
  }
  __syncthreads();
  return s;
}

#endif
