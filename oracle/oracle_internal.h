/*
 * oracle_internal.h -- helpers shared between the oracle's source files.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#ifndef BB_ORACLE_INTERNAL_H
#define BB_ORACLE_INTERNAL_H

#include "oracle.h"

#define OIDX(G, i, j, k) ((i) + (j) * (G).s1b + (k) * (G).s2b)

real orc_X(int n, real theta, real phi, const real *re, const real *im,
  int pp, int stride);
real orc_Y(int n, real theta, real phi, const real *re, const real *im,
  int pp, int stride);
real orc_Z(int n, real theta, real phi, const real *re, const real *im,
  int pp, int stride);

#endif
