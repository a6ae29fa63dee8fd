/*
 * bb_rng.c -- the random numbers of the precursor's initial field, under
 * the reference's names (src/rng.h, src/rng.c).
 *
 * The generator is Ranq1 of Press, Teukolsky, Vetterling and Flannery,
 * Numerical Recipes, 3rd ed. (2007), section 7.1.3: a 64-bit xorshift
 * (shifts 21 right, 35 left, 4 right) whose state is returned multiplied by
 * 2685821657736338717 mod 2^64, seeded by XOR-ing the seed into the state
 * 4101842887655102017 and stepping once; a double in [0, 1) is the 64-bit
 * output times 2^-64.  Seeding XORs into the current state, so a second
 * rng_init continues from the first rather than restarting, as the
 * reference's does.  gaussrand is Marsaglia's polar method (G. Marsaglia
 * and T. A. Bray, SIAM Review 6 (1964) 260), returning the second variate
 * of each accepted pair on the next call.
 */
#include <math.h>
#include "bb_host.h"

unsigned long long int rng_j = 1;
unsigned long long int rng_v = 4101842887655102017ULL;

static unsigned long long next_state(unsigned long long x)
{
  x ^= x >> 21;
  x ^= x << 35;
  x ^= x >> 4;
  return x;
}

unsigned long long int rng_uint64(void)
{
  rng_v = next_state(rng_v);
  return rng_v * 2685821657736338717ULL;
}

void rng_init(unsigned long long int seed)
{
  rng_j = seed;
  rng_v ^= rng_j;
  rng_v = rng_uint64();
}

unsigned int rng_uint32(void) { return (unsigned int) rng_uint64(); }
long long int rng_int64(void) { return (long long int) rng_uint64(); }
int rng_int32(void) { return (int) rng_uint64(); }

double rng_dbl(void)
{
  /* 2^-64 */
  return 5.42101086242752217e-20 * (double) rng_uint64();
}

float rng_flt(void) { return (float) rng_dbl(); }

double gaussrand(void)
{
  static double y1, y2, s;
  static int second = 0;
  if(second) {
    second = 0;
    return y2 * sqrt(-2. * log(s) / s);
  }
  do {
    y1 = 2. * rng_dbl() - 1.;
    y2 = 2. * rng_dbl() - 1.;
    s = y1 * y1 + y2 * y2;
  } while(s >= 1. || s == 0.);
  second = 1;
  return y1 * sqrt(-2. * log(s) / s);
}
