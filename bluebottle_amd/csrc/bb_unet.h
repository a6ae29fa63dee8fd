/*
 * bb_unet.h -- internal interface of the learned preconditioner
 * (bb_unet.hip) used by the flexible BiCGSTAB of bb_poisson.hip.
 */
#ifndef BB_UNET_H
#define BB_UNET_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define UNET_NLAYERS 11

/* 1/diag (d) and diag (g) of the Poisson operator per packed flag byte */
struct UNetDinv { float d[128]; float g[128]; };

UNetDinv bb_unet_dinv(void);
/* the scale of the network's output (the cell size squared) */
float bb_unet_h2(void);
/* allocate (or keep) the network for the current domain; 0 if the grid
 * does not suit it (sizes not divisible by 4) */
int bb_unet_ready(void);
/* out = M(in) on `st` (device vectors of the unghosted cell count) */
void bb_unet_apply_dev(const float *in, const uint8_t *flags, float *out,
  hipStream_t st);
/* loss and gradients of one (r, e) pair; one Adam step if `step`;
 * returns the loss (synchronises `st`) */
double bb_unet_train_pair(const float *r, const float *e,
  const uint8_t *flags, float lr, int step, hipStream_t st);
void bb_unet_free(void);
/* |A e - r| / |r| over fluid rows (synchronises `st`) */
double bb_unet_pair_residual(const float *r, const float *e,
  const uint8_t *flags, hipStream_t st);

#endif
