/*
 * field_util.c -- the host-array helpers of the reference's field.h and
 * myComplex.h: incident-wave sources added to a caller's array, the PML and
 * NS-PML coefficient formulas, the MPI subdomain getters, the text writers and
 * the complex-array allocation and interpolation helpers.
 *
 * The solvers of this library do not call the sources (their kernels form the
 * same terms on the GPU); these exist so that code written against field.h /
 * myComplex.h links and behaves as with the reference.  Each function follows
 * the reference's expression and operation order (cited per function).
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mpifdtd_field.h"

#ifndef M_PI
#define M_PI 3.1415926535897932384626433832795
#endif

/* field.c:58-64: this rank's subdomain; the single-rank state leaves it unset */
int field_getOffsetX(void) { return field_getSubFieldInfo_S().OFFSET_X; }
int field_getOffsetY(void) { return field_getSubFieldInfo_S().OFFSET_Y; }
int field_getSubNx(void) { return field_getSubFieldInfo_S().SUB_N_X; }
int field_getSubNy(void) { return field_getSubFieldInfo_S().SUB_N_Y; }
int field_getSubNpx(void) { return field_getSubFieldInfo_S().SUB_N_PX; }
int field_getSubNpy(void) { return field_getSubFieldInfo_S().SUB_N_PY; }
int field_getSubNcell(void) { return field_getSubFieldInfo_S().SUB_N_CELL; }

int ind(const int i, const int j) { return i * N_PY + j; } /* field.c:74 */

/* field.c:288-295 */
double field_pmlCoef(double ep_mu, double sig) { return (1.0 - sig / ep_mu) / (1.0 + sig / ep_mu); }
double field_pmlCoef_LXY(double ep_mu, double sig) { return 1.0 / (ep_mu + sig); }

/* field.c:298-309: beta and beta* of the NS PML from tanh of alpha and alpha* */
double field_ns_beta(double alpha, double alpha_aster)
{
  const double ta = tanh(alpha), tb = tanh(alpha_aster);
  return ta / (1 + ta * tb);
}
double field_ns_beta_aster(double alpha, double alpha_aster)
{
  const double ta = tanh(alpha), tb = tanh(alpha_aster);
  return tb / (1 + ta * tb);
}

/* the plane wave's phase k.r at (i + gapX, j + gapY): k cos / k sin formed once */
typedef struct { double t, w, rc, kc, ks; } wave_now;
static wave_now wave_state(double dot)
{
  wave_now s;
  const double rad = field_getWaveAngle() * M_PI / 180.0, k = field_getK();
  s.t = field_getTime();
  s.w = field_getOmega();
  s.rc = field_getRayCoef() * dot;
  s.kc = cos(rad) * k;
  s.ks = sin(rad) * k;
  return s;
}

/* field.c:202-218: the CW scattered-wave source of the UPML solvers,
   ray_coef (eps0/eps - 1) e^{i(k.r - w t)} over the interior cells */
void field_scatteredWave(fdtd_dcomplex *p, double *eps, double gapX, double gapY)
{
  const wave_now s = wave_state(1.0);
  for (int i = 1; i < N_PX - 1; i++)
    for (int j = 1; j < N_PY - 1; j++) {
      const int k = field_index(i, j);
      const double kr = (i + gapX) * s.kc + (j + gapY) * s.ks;
      p[k] += s.rc * (1.0 / eps[k] - 1.0) * cexp(I * (kr - s.w * s.t));
    }
}

/* field.c:179-199: the split-field PML solvers' source, the difference of the
   wave at t + 1/2 and t - 1/2 */
void field_scatteredWaveNotUPML(fdtd_dcomplex *p, double *eps, double gapX, double gapY)
{
  const wave_now s = wave_state(1.0);
  for (int i = 1; i < N_PX - 1; i++)
    for (int j = 1; j < N_PY - 1; j++) {
      const int k = field_index(i, j);
      const double kr = (i + gapX) * s.kc + (j + gapY) * s.ks;
      p[k] += s.rc * (1.0 / eps[k] - 1.0) * (cexp(I * (kr - s.w * (s.t + 0.5))) - cexp(I * (kr - s.w * (s.t - 0.5))));
    }
}

/* field.c:155-177: the NS-FDTD solvers' source, the factor 1/(_n n) - 1 with the
   NS dispersion ratio _n = u0/u1, phases t + 1 and t, projected by dot */
void field_nsScatteredWaveNotUPML(fdtd_dcomplex *p, double *eps, double gapX, double gapY, double dot)
{
  const wave_now s = wave_state(dot);
  const double k = field_getK();
  for (int i = 1; i < N_PX - 1; i++)
    for (int j = 1; j < N_PY - 1; j++) {
      const int c = field_index(i, j);
      const double kr = (i + gapX) * s.kc + (j + gapY) * s.ks;
      const double n = sqrt(eps[c] / 1.0);
      const double u0 = sin(s.w * 0.5) / sin(k * 0.5), u1 = sin(s.w * 0.5) / sin(n * k * 0.5);
      const double _n = u0 / u1;
      p[c] += s.rc * (1.0 / (_n * n) - 1.0) * (cexp(I * (kr - s.w * (s.t + 1.0))) - cexp(I * (kr - s.w * (s.t))));
    }
}

/* field.c:224-256: the plain UPML solvers' Gaussian pulse (width 50 cells),
   timed to peak at the grid centre at t = 500, vacuum cells skipped */
void field_scatteredPulse(fdtd_dcomplex *p, double *eps, double gapX, double gapY, double dot)
{
  const double t = field_getTime(), w = field_getOmega(), rad = field_getWaveAngle() * M_PI / 180.0;
  const double c0 = field_getLightSpeed();
  const double cpc = cos(rad) / c0, spc = sin(rad) / c0, width = 50;
  const double center_peak = (N_PX / 2.0 + gapX) * cpc + (N_PY / 2 + gapY) * spc; /* N_PY / 2 in ints, as field.c:236 */
  const double t0 = -center_peak + 500;
  for (int i = 1; i < N_PX - 1; i++)
    for (int j = 1; j < N_PY - 1; j++) {
      const int k = field_index(i, j);
      if (1.0 == eps[k]) continue;
      const double r = (i + gapX) * cpc + (j + gapY) * spc - (t - t0);
      const double g = exp(-pow(r / width, 2));
      p[k] += dot * g * (1.0 / eps[k] - 1) * cexp(I * r * w);
    }
}

/* myComplex.c:3-38 */
double *newDouble(int size)
{
  double *a = (double *)malloc(sizeof(double) * size);
  memset(a, 0, sizeof(double) * size);
  return a;
}
void freeDouble(double *array) { free(array); }
fdtd_dcomplex *newDComplex(int size)
{
  fdtd_dcomplex *a = (fdtd_dcomplex *)malloc(sizeof(fdtd_dcomplex) * size);
  memset(a, 0, sizeof(fdtd_dcomplex) * size);
  return a;
}
void freeDComplex(fdtd_dcomplex *array) { free(array); }

/* myComplex.c:33-38: |c|^2 */
double cnorm(fdtd_dcomplex c)
{
  const double re = creal(c), im = cimag(c);
  return re * re + im * im;
}

/* myComplex.c:40-51: bilinear interpolation of a width x height array (index
   i * height + j) at (x, y); the caller keeps (x, y) at least one cell inside */
fdtd_dcomplex cbilinear(fdtd_dcomplex *p, double x, double y, int width, int height)
{
  (void)width;
  const int i = (int)floor(x), j = (int)floor(y);
  const double dx = x - i, dy = y - j;
  const int c = i * height + j;
  return p[c] * (1.0 - dx) * (1.0 - dy) + p[c + height] * dx * (1.0 - dy) + p[c + 1] * (1.0 - dx) * dy +
         p[c + height + 1] * dx * dy;
}

/* field.c:322-341: |data|^2 on the circle of radius 1.2 lambda around the grid
   centre, 181 lines "%d %lf".  Where the circle leaves the grid (a side under
   2.4 lambda) the reference reads outside its array; 0 is written instead. */
void field_outputElliptic(const char *fileName, fdtd_dcomplex *data)
{
  printf("output start\n");
  FILE *fp = fopen(fileName, "w");
  if (!fp) {
    printf("cannot open file %s \n", fileName);
    exit(2);
  }
  const double lam = field_getLambda();
  for (int ang = 180; ang >= 0; ang--) {
    const double rad = ang * M_PI / 180.0;
    const double x = 1.2 * lam * cos(rad) + N_PX / 2.0, y = 1.2 * lam * sin(rad) + N_PY / 2.0;
    const int i = (int)floor(x), j = (int)floor(y);
    const int inside = i >= 0 && j >= 0 && i + 1 < N_PX && j + 1 < N_PY;
    fprintf(fp, "%d %lf \n", 180 - ang, inside ? cnorm(cbilinear(data, x, y, N_PX, N_PY)) : 0.0);
  }
  fclose(fp);
  printf("output to %s end\n", fileName);
}

/* field.c:346-366: |data|^2 of every cell, one "%lf" line each */
void field_outputAllDataComplex(const char *fileName, fdtd_dcomplex *data)
{
  printf("output all data start\n");
  FILE *fp = fopen(fileName, "w");
  if (!fp) {
    printf("cannot open file %s \n", fileName);
    exit(2);
  }
  for (int i = 0; i < N_PX; i++)
    for (int j = 0; j < N_PY; j++) fprintf(fp, "%lf \n", cnorm(data[ind(i, j)]));
  fclose(fp);
  printf("output all data to %s end\n", fileName);
}

/* field.c:368-389: every cell's value other than 1.0, one "%lf" line each */
void field_outputAllDataDouble(const char *fileName, double *data)
{
  printf("output all data double start\n");
  FILE *fp = fopen(fileName, "w");
  if (!fp) {
    printf("cannot open file %s \n", fileName);
    exit(2);
  }
  for (int i = 0; i < N_PX; i++)
    for (int j = 0; j < N_PY; j++)
      if (data[ind(i, j)] != 1.0) fprintf(fp, "%lf \n", data[ind(i, j)]);
  fclose(fp);
  printf("output all data to %s end\n", fileName);
}
