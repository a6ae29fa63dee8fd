/*
 * The learned preconditioner's network (bb_unet.hip) uses the gfx950
 * matrix-core instructions, which this CPU emulation does not model: in
 * the emulation build the network reports itself unavailable, so the
 * learned solver form runs its Jacobi path.  Test infrastructure only.
 */
#include <stdio.h>
#include <stdlib.h>
#include "hip/hip_runtime.h"
#include "bb_device.h"
#include "bb_unet.h"

static void not_emulated(const char *what)
{
  fprintf(stderr, "bb_emu: %s is not emulated (gfx950 matrix cores)\n", what);
  abort();
}

UNetDinv bb_unet_dinv(void) { UNetDinv t = {}; return t; }
float bb_unet_h2(void) { return 0.f; }
int bb_unet_ready(void) { return 0; }
void bb_unet_apply_dev(const float *, const uint8_t *, float *, hipStream_t)
{ not_emulated("bb_unet_apply_dev"); }
double bb_unet_train_pair(const float *, const float *, const uint8_t *,
  float, int, hipStream_t) { not_emulated("bb_unet_train_pair"); return 0.; }
void bb_unet_free(void) {}
double bb_unet_pair_residual(const float *, const float *, const uint8_t *,
  hipStream_t) { not_emulated("bb_unet_pair_residual"); return 0.; }
extern "C" {
long bb_unet_num_params(void) { return 79920; }
real bb_unet_output_scale(void) { return 0.; }
void bb_unet_init(unsigned long) { not_emulated("bb_unet_init"); }
void bb_unet_set_params(const float *, long) { not_emulated("bb_unet_set_params"); }
void bb_unet_get_params(float *, long) { not_emulated("bb_unet_get_params"); }
void bb_unet_get_grads(float *, long) { not_emulated("bb_unet_get_grads"); }
void bb_unet_apply(const float *, const unsigned char *, float *)
{ not_emulated("bb_unet_apply"); }
real bb_unet_loss_grad(const float *, const float *, const unsigned char *, real)
{ not_emulated("bb_unet_loss_grad"); return 0.; }
}
