/*
 * bb_helmholtz.hip -- Crank-Nicolson Helmholtz solves of the reference's
 * IMPLICIT build (src/cuda_bicgstab.cu:44-714): cuda_{u,v,w}star_helmholtz.
 *
 * The explicit build (cuda_U_star_2), which every benchmark configuration
 * uses, does not call these.  They are exported so that a driver built
 * with -DIMPLICIT links; until the matrix-free Jacobi-PCG version lands
 * they stop with a message instead of returning wrong fields.
 */
#include <stdio.h>
#include <stdlib.h>
#include "bb_device.h"

static void implicit_unavailable(const char *which)
{
  fprintf(stderr, "bluebottle_hip: %s (IMPLICIT build) is not available in "
    "this library; build the driver without -DIMPLICIT\n", which);
  fflush(stderr);
  exit(EXIT_FAILURE);
}

extern "C" void cuda_ustar_helmholtz(int rank)
{
  (void) rank;
  implicit_unavailable("cuda_ustar_helmholtz");
}

extern "C" void cuda_vstar_helmholtz(int rank)
{
  (void) rank;
  implicit_unavailable("cuda_vstar_helmholtz");
}

extern "C" void cuda_wstar_helmholtz(int rank)
{
  (void) rank;
  implicit_unavailable("cuda_wstar_helmholtz");
}
