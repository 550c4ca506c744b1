/*
 * field.c -- grid, time and incident-wave state of the running simulation.
 *
 * Behaviour of the reference's field.c (field.c:39-319): cell-unit conversion,
 * NTFF contour, the wave's k/omega, the ray_coef ramp, sigma profiles and the
 * point source.  One simulation per process, as in the reference; the
 * batched multi-instance API lives in fdtd_batch.
 */
#include <math.h>
#include <string.h>
#include "mpifdtd_field.h"

#ifndef M_PI
#define M_PI 3.1415926535897932384626433832795
#endif

int N_X, N_Y, N_CELL, N_PML, N_PX, N_PY;

static double c0_ = 0.7071;   /* field.h:70-71 */
static int ntff_inset_ = 5;   /* field.c:135-138 */
static int point_source_ = 0;

static double time_, ray_coef_, lambda_s_, k_s_, w_s_, T_s_, maxTime_;
static NTFFInfo ntff_info_;
static FieldInfo fieldInfo_;
static FieldInfo_S fieldInfo_s_;
static WaveInfo_S waveInfo_s_;
static SubFieldInfo_S subFieldInfo_s_;

void field_setLightSpeed(double c0) { c0_ = c0 > 0 ? c0 : 0.7071; }
double field_getLightSpeed(void) { return c0_; }
double field_getMu0(void) { return 1.0 / c0_ / c0_; }
double field_getZ0(void) { return c0_ == 0.7071 ? 1.41422712488 : 1.0 / c0_; }
void field_setNTFFInset(int cells) { ntff_inset_ = cells > 0 ? cells : 5; }
void field_setPointSource(int on) { point_source_ = on ? 1 : 0; }
int field_getPointSource(void) { return point_source_; }

void field_setWaveAngle(int deg)
{
  fieldInfo_.angle_deg = deg;
  waveInfo_s_.Angle_deg = deg;
}

double field_getT(void) { return waveInfo_s_.T_s; }
double field_getK(void) { return waveInfo_s_.K_s; }
double field_getRayCoef(void) { return ray_coef_; }
double field_getOmega(void) { return waveInfo_s_.Omega_s; }
double field_getLambda(void) { return lambda_s_; }
double field_getWaveAngle(void) { return waveInfo_s_.Angle_deg; }
double field_getTime(void) { return time_; }
double field_getMaxTime(void) { return maxTime_; }
NTFFInfo field_getNTFFInfo(void) { return ntff_info_; }
WaveInfo_S field_getWaveInfo_S(void) { return waveInfo_s_; }
SubFieldInfo_S field_getSubFieldInfo_S(void) { return subFieldInfo_s_; }
FieldInfo_S field_getFieldInfo_S(void) { return fieldInfo_s_; }
FieldInfo field_getFieldInfo(void) { return fieldInfo_; }

int field_subIndex(int i, int j) { return i * subFieldInfo_s_.SUB_N_PY + j; }
int field_index(int i, int j) { return i * fieldInfo_s_.N_PY + j; }

double field_toCellUnit(const double nm) { return nm / fieldInfo_.h_u_nm; }
double field_toPhisycalUnit(const double cells) { return cells * fieldInfo_.h_u_nm; }

void field_reset(void)
{
  time_ = 0;
  ray_coef_ = 0;
}

void field_init(FieldInfo fi)
{
  fieldInfo_ = fi;
  fieldInfo_s_.N_PX = (int)field_toCellUnit(fi.width_nm);
  fieldInfo_s_.N_PY = (int)field_toCellUnit(fi.height_nm);
  fieldInfo_s_.N_PML = fi.pml;
  fieldInfo_s_.N_X = fieldInfo_s_.N_PX - 2 * fieldInfo_s_.N_PML;
  fieldInfo_s_.N_Y = fieldInfo_s_.N_PY - 2 * fieldInfo_s_.N_PML;
  fieldInfo_s_.N_CELL = fieldInfo_s_.N_PY * fieldInfo_s_.N_PX;
  fieldInfo_s_.DX = fieldInfo_s_.N_PY;
  fieldInfo_s_.DY = 1;

  waveInfo_s_.Lambda_s = field_toCellUnit(fi.lambda_nm);
  waveInfo_s_.T_s = waveInfo_s_.Lambda_s / c0_;
  waveInfo_s_.K_s = 2 * M_PI / waveInfo_s_.Lambda_s;
  waveInfo_s_.K_0_s = 2 * M_PI / waveInfo_s_.Lambda_s;
  waveInfo_s_.Omega_s = c0_ * waveInfo_s_.K_s;
  waveInfo_s_.Angle_deg = fi.angle_deg;

  /* the subdomain record stays unset, as the reference's field_init leaves it
     (its mpiSplit call is commented out, field.c:112); the MPI solvers keep
     their own subdomain sizes (mpi_fdtdTM_upml_getSubNx ...) */
  memset(&subFieldInfo_s_, 0, sizeof subFieldInfo_s_);

  N_PX = fieldInfo_s_.N_PX;
  N_PY = fieldInfo_s_.N_PY;
  N_PML = fi.pml;
  N_X = N_PX - 2 * N_PML;
  N_Y = N_PY - 2 * N_PML;
  N_CELL = N_PX * N_PY;
  time_ = 0;
  maxTime_ = fi.stepNum;

  lambda_s_ = field_toCellUnit(fi.lambda_nm);
  k_s_ = 2 * M_PI / lambda_s_;
  w_s_ = c0_ * k_s_;
  T_s_ = 2 * M_PI / w_s_;
  ray_coef_ = 0;

  ntff_info_.cx = N_PX / 2;
  ntff_info_.cy = N_PY / 2;
  ntff_info_.top = N_PY - N_PML - ntff_inset_;
  ntff_info_.bottom = N_PML + ntff_inset_;
  ntff_info_.left = N_PML + ntff_inset_;
  ntff_info_.right = N_PX - N_PML - ntff_inset_;
  double len = (ntff_info_.top - ntff_info_.bottom) / 2;  /* int division, as field.c:140 */
  ntff_info_.RFperC = len * 2;
  ntff_info_.arraySize = (int)(maxTime_ + 2 * ntff_info_.RFperC);
}

fdtd_dcomplex field_pointLight(void)
{
  /* ray_coef * cexp(i w t), written as cos + i sin of the same argument */
  double a = waveInfo_s_.Omega_s * time_;
  return ray_coef_ * cos(a) + I * (ray_coef_ * sin(a));
}

/* M = 2 polynomial grading (field.c:259-283) */
double field_sigmaX(const double x, const double y)
{
  (void)y;
  const int M = 2;
  double v;
  if (x < N_PML) {
    v = 1.0 * (N_PML - x) / N_PML;
    return pow(v, M);
  } else if (N_PML <= x && x < (N_X + N_PML)) {
    return 0;
  }
  v = 1.0 * (x - (N_PX - N_PML - 1)) / N_PML;
  return pow(v, M);
}

double field_sigmaY(const double x, double y)
{
  (void)x;
  const int M = 2;
  double v;
  if (y < N_PML) {
    v = 1.0 * (N_PML - y) / N_PML;
    return pow(v, M);
  } else if (y >= N_PML && y < (N_Y + N_PML)) {
    return 0.0;
  }
  v = 1.0 * (y - (N_PY - N_PML - 1)) / N_PML;
  return pow(v, M);
}

void field_nextStep(void)
{
  time_ += 1.0;
  ray_coef_ = 1.0 - exp(-pow(0.01 * time_, 2));
}

int field_isFinish(void) { return time_ >= maxTime_; }
