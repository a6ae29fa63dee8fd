/*
 * bb_outside.hip -- entry points the reference driver links that lie
 * outside this library's path: the -DTEST checks of src/cuda_testing.cu
 * and a few helpers no caller uses.  They exist so
 * that the unchanged src/bluebottle.c links against this library; each one
 * prints which path it belongs to and exits, as a missing feature should
 * rather than run wrong.
 */
#include <stdio.h>
#include <stdlib.h>
#include "bb_device.h"

static void outside(const char *name, const char *what)
{
  fprintf(stderr, "bluebottle_hip: %s (%s) is outside this library's path\n",
    name, what);
  fflush(stderr);
  exit(EXIT_FAILURE);
}

#define TEST_PATH "-DTEST check"

/* src/bluebottle.h:2616, 2628, 2753, 2766, 2780, 2792 */
extern "C" void cuda_U_star_test_exp(void) { outside(__func__, TEST_PATH); }
extern "C" void cuda_U_star_test_cos(void) { outside(__func__, TEST_PATH); }
extern "C" void cuda_project_test(void) { outside(__func__, TEST_PATH); }
extern "C" void cuda_BC_test(void) { outside(__func__, TEST_PATH); }
extern "C" void cuda_quad_interp_test(void) { outside(__func__, TEST_PATH); }
extern "C" void cuda_lamb_test(void) { outside(__func__, TEST_PATH); }

/* src/bluebottle.h:2704, 2717, 2943: declared, not called on the path
 * (cuda_div_U's one call is commented out in bluebottle.c:700) */
extern "C" void cuda_PP_bicgstab_jacobi(void) { outside(__func__, "unused solver variant"); }
extern "C" void cuda_div_U(void) { outside(__func__, "divergence diagnostic"); }
extern "C" void cuda_collisions(void) { outside(__func__, "unused collision variant"); }
