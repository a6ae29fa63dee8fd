/*
 * bb_lamb.h -- device-side Lamb's-solution basis shared by the particle
 * boundary conditions and the quadrature: N_nm, P_nm (Condon-Shortley
 * phase, n <= 5), Cartesian -> spherical, the X/Y/Z vector-harmonic sums
 * and the analytic velocity.  Same mathematics as
 * src/particle_kernel.cu:1020-1369 and src/quadrature_kernel.cu:305-375;
 * cos(m phi) / sin(m phi) are evaluated once per point and shared by all
 * orders instead of once per term.
 */
#ifndef BB_LAMB_H
#define BB_LAMB_H

#include "bb_device.h"

/* N_nm for n <= 5 (21 entries, index n*(n+1)/2 + m), filled on the host
 * with sqrt((2n+1)/(4 pi) (n-m)!/(n+m)!). */
struct LambNorm { double N[21]; };

LambNorm bb_lamb_norm(void);

__device__ inline double lamb_P(int n, int m, double x, double y)
{
  /* x = cos(theta), y = sin(theta) */
  switch(n) {
    case 0: return 1.;
    case 1: return m == 0 ? x : -y;
    case 2:
      if(m == 0) return 0.5*(3.*x*x - 1.);
      if(m == 1) return -3.*x*y;
      return 3.*y*y;
    case 3:
      if(m == 0) return 0.5*x*(5.*x*x - 3.);
      if(m == 1) return -1.5*(5.*x*x - 1.)*y;
      if(m == 2) return 15.*x*y*y;
      return -15.*y*y*y;
    case 4:
      if(m == 0) return 0.125*(35.*x*x*x*x - 30.*x*x + 3.);
      if(m == 1) return -2.5*(7.*x*x - 3.)*x*y;
      if(m == 2) return 7.5*(7.*x*x - 1.)*y*y;
      if(m == 3) return -105.*x*y*y*y;
      return 105.*y*y*y*y;
    default:
      if(m == 0) return 0.125*x*(63.*x*x*x*x - 70.*x*x + 15.);
      if(m == 1) return -1.875*y*(21.*x*x*x*x - 14.*x*x + 1.);
      if(m == 2) return 52.5*x*y*y*(3.*x*x - 1.);
      if(m == 3) return -52.5*y*y*y*(9.*x*x - 1.);
      if(m == 4) return 945.*x*y*y*y*y;
      return -945.*y*y*y*y*y;
  }
}

__device__ inline void lamb_xyz2rtp(double x, double y, double z, double *r,
  double *theta, double *phi)
{
  double XY = x*x + y*y;
  double XYZ = XY + z*z;
  if(XYZ >= 0 && XYZ < DIV_ST) XYZ = DIV_ST;
  *r = sqrt(XYZ);
  *theta = acos(z / *r);
  if(XY >= 0 && XY < DIV_ST) XY = DIV_ST;
  *phi = acos(x / sqrt(XY));
  if(y < 0.) *phi = 2.*PI - *phi;
}

/* Angular data of one evaluation point. */
struct LambPoint {
  double ct, st, sts;   /* cos, sin, sin clamped away from 0 */
  double cm[6], sm[6];  /* cos(m phi), sin(m phi) */
};

__device__ inline void lamb_point(double theta, double phi, LambPoint *q)
{
  q->ct = cos(theta);
  q->st = sin(theta);
  double st = q->st;
  if(st >= 0 && st < DIV_ST) st = DIV_ST;
  else if(st < 0 && st > -DIV_ST) st = -DIV_ST;
  q->sts = st;
  for(int m = 0; m < 6; m++) { q->cm[m] = cos(m*phi); q->sm[m] = sin(m*phi); }
}

__device__ inline double lamb_X(const LambNorm &L, int n, const LambPoint &q,
  const double *re, const double *im, int base)
{
  int c = base + n*(n+1)/2;
  double sum = L.N[n*(n+1)/2] * lamb_P(n, 0, q.ct, q.st) * re[c];
  for(int m = 1; m <= n; m++) {
    c++;
    sum += 2.*L.N[n*(n+1)/2 + m]*lamb_P(n, m, q.ct, q.st)
      *(re[c]*q.cm[m] - im[c]*q.sm[m]);
  }
  return sum;
}

__device__ inline double lamb_Y(const LambNorm &L, int n, const LambPoint &q,
  const double *re, const double *im, int base)
{
  int c = base + n*(n+1)/2;
  double st = q.sts, ct = q.ct;
  int m = 0;
  double sum = L.N[n*(n+1)/2]
    *(-(n+1)*ct/st*lamb_P(n, m, q.ct, q.st)+(n-m+1)/st*lamb_P(n+1, m, q.ct, q.st))
    *re[c];
  for(m = 1; m <= n; m++) {
    c++;
    sum += 2.*L.N[n*(n+1)/2 + m]
      *(-(n+1)*ct/st*lamb_P(n, m, q.ct, q.st)+(n-m+1)/st*lamb_P(n+1, m, q.ct, q.st))
      *(re[c]*q.cm[m] - im[c]*q.sm[m]);
  }
  return sum;
}

__device__ inline double lamb_Z(const LambNorm &L, int n, const LambPoint &q,
  const double *re, const double *im, int base)
{
  int c = base + n*(n+1)/2;
  double st = q.sts;
  double sum = 0.;
  for(int m = 1; m <= n; m++) {
    c++;
    sum += -2.*m/st*L.N[n*(n+1)/2 + m]*lamb_P(n, m, q.ct, q.st)
      *(re[c]*q.sm[m] + im[c]*q.cm[m]);
  }
  return sum;
}

struct LambCoeffs {
  const double *pre, *pim, *fre, *fim, *cre, *cim;
};

/* Velocity of Lamb's solution at (r, theta, phi) about a particle of
 * radius a whose coefficients start at index `base`. */
__device__ inline void lamb_velocity(const LambNorm &L, const LambCoeffs &K,
  int order, double a, double r, double theta, double phi, double nu,
  int base, double *Ux, double *Uy, double *Uz)
{
  LambPoint q;
  lamb_point(theta, phi, &q);
  double ar = a / r, ra = r / a;
  double ur = 0.;
  double ut = 0.5*ra*lamb_Y(L, 0, q, K.pre, K.pim, base);
  double up = 0.5*ra*lamb_Z(L, 0, q, K.pre, K.pim, base);
  for(int n = 1; n <= order; n++) {
    double powranp1 = pow(ra, n+1.);
    double powranm1 = pow(ra, n-1.);
    double powarnp2 = pow(ar, n+2.);
    double powarnp1 = pow(ar, n+1.);
    double powarn = pow(ar, (double) n);
    double od2np3 = 1./(2.*n+3.);
    double odnp1 = 1./(n+1.);
    double Xp = lamb_X(L, n, q, K.pre, K.pim, base);
    double Xf = lamb_X(L, n, q, K.fre, K.fim, base);
    double Yp = lamb_Y(L, n, q, K.pre, K.pim, base);
    double Yf = lamb_Y(L, n, q, K.fre, K.fim, base);
    double Yc = lamb_Y(L, n, q, K.cre, K.cim, base);
    double Zp = lamb_Z(L, n, q, K.pre, K.pim, base);
    double Zf = lamb_Z(L, n, q, K.fre, K.fim, base);
    double Zc = lamb_Z(L, n, q, K.cre, K.cim, base);
    ur += (0.5*n*od2np3*powranp1
      + 0.25*n*((2.*n+1.)*od2np3*ar*ar-1.)*powarn) * Xp;
    ur += (n*powranm1
      + 0.5*n*(2.*n-1.-(2.*n+1.)*ra*ra)*powarnp2) * Xf;
    ut += (0.5*(n+3.)*odnp1*od2np3*powranp1
      + 0.25*odnp1*(n-2.-n*(2.*n+1.)*od2np3*ar*ar)*powarn) * Yp;
    ut += (powranm1
      + 0.5*odnp1*((n-2.)*(2.*n+1.)*ra*ra-n*(2.*n-1.))*powarnp2) * Yf;
    ut += (powranm1 - powarnp1) * Zc;
    up += (0.5*(n+3.)*odnp1*od2np3*powranp1
      + 0.25*odnp1*(n-2.-n*(2.*n+1.)*od2np3*ar*ar)*powarn) * Zp;
    up += (powranm1
      + 0.5*odnp1*((n-2.)*(2.*n+1.)*ra*ra-n*(2.*n-1.))*powarnp2) * Zf;
    up += (-powranm1 + powarnp1) * Yc;
  }
  ur *= nu / a;
  ut *= nu / a;
  up *= nu / a;
  *Ux = ur*q.st*q.cm[1]+ut*q.ct*q.cm[1]-up*q.sm[1];
  *Uy = ur*q.st*q.sm[1]+ut*q.ct*q.sm[1]+up*q.cm[1];
  *Uz = ur*q.ct-ut*q.st;
}

#endif
