/*
 * oracle_lamb.c -- Lamb's-solution basis functions: normalisation N_nm,
 * associated Legendre functions P_nm (Condon-Shortley phase, n <= 5),
 * Cartesian -> spherical conversion and the X/Y/Z vector-harmonic sums,
 * and the analytic Lamb velocity.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates src/particle_kernel.cu:1020-1369 (Nnm, Pnm, xyz2rtp, X_pn,
 * X_phin, Y_pn, Y_phin, Y_chin, Z_pn, Z_phin, Z_chin, lamb_vel).  The
 * quadrature kernels use an identical copy (src/quadrature_kernel.cu:
 * 305-375), so one restatement serves both.
 */
#include <math.h>
#include "oracle.h"
#include "oracle_internal.h"

real orc_Nnm(int n, int m)
{
  real fact_top = 1;
  real fact_bot = 1;
  for(int i = 1; i <= (n-m); i++) fact_top *= (real)i;
  for(int i = 1; i <= (n+m); i++) fact_bot *= (real)i;
  return sqrt((2.*n+1.) / 4. / PI * fact_top / fact_bot);
}

real orc_Pnm(int n, int m, real theta)
{
  real x = cos(theta);
  real y = sin(theta);
  switch(n) {
    case 0: return 1;
    case 1:
      if(m == 0) return x;
      if(m == 1) return -y;
      break;
    case 2:
      if(m == 0) return 0.5*(3.*x*x - 1.);
      if(m == 1) return -3.*x*y;
      if(m == 2) return 3.*y*y;
      break;
    case 3:
      if(m == 0) return 0.5*x*(5.*x*x - 3.);
      if(m == 1) return -1.5*(5.*x*x - 1.)*y;
      if(m == 2) return 15.*x*y*y;
      if(m == 3) return -15.*y*y*y;
      break;
    case 4:
      if(m == 0) return 0.125*(35.*x*x*x*x - 30.*x*x + 3.);
      if(m == 1) return -2.5*(7.*x*x - 3.)*x*y;
      if(m == 2) return 7.5*(7.*x*x - 1.)*y*y;
      if(m == 3) return -105.*x*y*y*y;
      if(m == 4) return 105.*y*y*y*y;
      break;
    case 5:
      if(m == 0) return 0.125*x*(63.*x*x*x*x - 70.*x*x + 15.);
      if(m == 1) return -1.875*y*(21.*x*x*x*x - 14.*x*x + 1.);
      if(m == 2) return 52.5*x*y*y*(3.*x*x - 1.);
      if(m == 3) return -52.5*y*y*y*(9.*x*x - 1.);
      if(m == 4) return 945.*x*y*y*y*y;
      if(m == 5) return -945.*y*y*y*y*y;
      break;
  }
  return 0;
}

void orc_xyz2rtp(real x, real y, real z, real *r, real *theta, real *phi)
{
  real XY = x*x + y*y;
  real XYZ = XY + z*z;
  if(XYZ >= 0 && XYZ < DIV_ST) XYZ = DIV_ST;
  *r = sqrt(XYZ);
  *theta = acos(z / *r);
  if(XY >= 0 && XY < DIV_ST) XY = DIV_ST;
  *phi = acos(x / sqrt(XY));
  if(y < 0.) *phi = 2.*PI - *phi;
}

static int coeff_base(int n, int pp, int stride)
{
  int coeff = 0;
  for(int j = 0; j < n; j++) coeff += j+1;
  return coeff + pp*stride;
}

/* X_pn / X_phin share one form (src/particle_kernel.cu:1107-1145). */
real orc_X(int n, real theta, real phi, const real *re, const real *im,
  int pp, int stride)
{
  int coeff = coeff_base(n, pp, stride);
  int m = 0;
  real sum = orc_Nnm(n,m)*orc_Pnm(n,m,theta)*re[coeff];
  for(m = 1; m <= n; m++) {
    coeff++;
    sum += 2.*orc_Nnm(n,m)*orc_Pnm(n,m,theta)
      *(re[coeff]*cos(m*phi) - im[coeff]*sin(m*phi));
  }
  return sum;
}

/* Y_pn / Y_phin / Y_chin (src/particle_kernel.cu:1147-1239). */
real orc_Y(int n, real theta, real phi, const real *re, const real *im,
  int pp, int stride)
{
  real ct = cos(theta);
  real st = sin(theta);
  if(st >= 0 && st < DIV_ST) st = DIV_ST;
  else if(st < 0 && st > -DIV_ST) st = -DIV_ST;
  int coeff = coeff_base(n, pp, stride);
  int m = 0;
  real sum = orc_Nnm(n,m)
    *(-(n+1)*ct/st*orc_Pnm(n,m,theta)+(n-m+1)/st*orc_Pnm(n+1,m,theta))
    *re[coeff];
  for(m = 1; m <= n; m++) {
    coeff++;
    sum += 2.*orc_Nnm(n,m)
      *(-(n+1)*ct/st*orc_Pnm(n,m,theta)+(n-m+1)/st*orc_Pnm(n+1,m,theta))
      *(re[coeff]*cos(m*phi) - im[coeff]*sin(m*phi));
  }
  return sum;
}

/* Z_pn / Z_phin / Z_chin (src/particle_kernel.cu:1241-1309). */
real orc_Z(int n, real theta, real phi, const real *re, const real *im,
  int pp, int stride)
{
  real st = sin(theta);
  if(st >= 0 && st < DIV_ST) st = DIV_ST;
  else if(st < 0 && st > -DIV_ST) st = -DIV_ST;
  int coeff = coeff_base(n, pp, stride);
  int m = 0;
  real sum = 0.;
  for(m = 1; m <= n; m++) {
    coeff++;
    sum += -2.*m/st*orc_Nnm(n,m)*orc_Pnm(n,m,theta)
      *(re[coeff]*sin(m*phi) + im[coeff]*cos(m*phi));
  }
  return sum;
}

/* src/particle_kernel.cu:1311-1369 */
void orc_lamb_vel(const orc_state *s, int order, real a, real r, real theta,
  real phi, int P, real *Ux, real *Uy, real *Uz)
{
  const int stride = s->coeff_stride;
  const real *pre = s->coeff[0], *pim = s->coeff[1];
  const real *fre = s->coeff[2], *fim = s->coeff[3];
  const real *cre = s->coeff[4], *cim = s->coeff[5];
  real nu = s->nu;
  real ar = a / r;
  real ra = r / a;
  real ur = 0.;
  real ut = 0.5*ra*orc_Y(0, theta, phi, pre, pim, P, stride);
  real up = 0.5*ra*orc_Z(0, theta, phi, pre, pim, P, stride);
  for(int n = 1; n <= order; n++) {
    real powranp1 = pow(ra,n+1.);
    real powranm1 = pow(ra,n-1.);
    real powarnp2 = pow(ar,n+2.);
    real powarnp1 = pow(ar,n+1.);
    real powarn = pow(ar,n);
    real od2np3 = 1./(2.*n+3.);
    real odnp1 = 1./(n+1.);
    ur += (0.5*n*od2np3*powranp1
      + 0.25*n*((2.*n+1.)*od2np3*ar*ar-1.)*powarn)
      * orc_X(n, theta, phi, pre, pim, P, stride);
    ur += (n*powranm1
      + 0.5*n*(2.*n-1.-(2.*n+1.)*ra*ra)*pow(ar,n+2.))
      * orc_X(n, theta, phi, fre, fim, P, stride);
    ut += (0.5*(n+3.)*odnp1*od2np3*powranp1
      + 0.25*odnp1*(n-2.-n*(2.*n+1.)*od2np3*ar*ar)*powarn)
      * orc_Y(n, theta, phi, pre, pim, P, stride);
    ut += (powranm1
      + 0.5*odnp1*((n-2.)*(2.*n+1.)*ra*ra-n*(2.*n-1.))*powarnp2)
      * orc_Y(n, theta, phi, fre, fim, P, stride);
    ut += (powranm1 - powarnp1)
      * orc_Z(n, theta, phi, cre, cim, P, stride);
    up += (0.5*(n+3.)*odnp1*od2np3*powranp1
      + 0.25*odnp1*(n-2.-n*(2.*n+1.)*od2np3*ar*ar)*powarn)
      * orc_Z(n, theta, phi, pre, pim, P, stride);
    up += (powranm1
      + 0.5*odnp1*((n-2.)*(2.*n+1.)*ra*ra-n*(2.*n-1.))*powarnp2)
      * orc_Z(n, theta, phi, fre, fim, P, stride);
    up += (-powranm1 + powarnp1)
      * orc_Y(n, theta, phi, cre, cim, P, stride);
  }
  ur *= nu / a;
  ut *= nu / a;
  up *= nu / a;
  *Ux = ur*sin(theta)*cos(phi)+ut*cos(theta)*cos(phi)-up*sin(phi);
  *Uy = ur*sin(theta)*sin(phi)+ut*cos(theta)*sin(phi)+up*cos(phi);
  *Uz = ur*cos(theta)-ut*sin(theta);
}
