/*
 * mpifdtd_field.h -- grid / time / wave state of a simulation (host side).
 *
 * Replaces the reference's field.h (field.h:20-162) with the same types,
 * globals and entry points; the solvers pull every per-step scalar from these
 * getters, exactly as the reference's solvers do.  `bool` is the reference's
 * `int` (bool.h:5-7).
 */
#ifndef MPIFDTD_FIELD_H
#define MPIFDTD_FIELD_H

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> fdtd_dcomplex;
extern "C" {
#else
#include <complex.h>
typedef double _Complex fdtd_dcomplex;
#endif

/* field.h:20-28 */
typedef struct FieldInfo {
  int width_nm, height_nm;  /* domain size including the PML (nm) */
  int h_u_nm;               /* cell size (nm) */
  int pml;                  /* PML thickness (cells) */
  int lambda_nm;            /* wavelength (nm) */
  int angle_deg;            /* incidence angle (deg) */
  int stepNum;              /* number of time steps */
} FieldInfo;

/* field.h:31-38 */
typedef struct FieldInfo_S {
  int N_X, N_Y;
  int N_PX, N_PY;
  int N_CELL;
  int N_PML;
  int DX, DY;
} FieldInfo_S;

/* field.h:41-49; filled for a single rank (offsets 0, SUB_N_P* = N_P* + 2) */
typedef struct SubFieldInfo_S {
  int OFFSET_X, OFFSET_Y;
  int SUB_N_X, SUB_N_Y;
  int SUB_N_PX, SUB_N_PY;
  int SUB_N_CELL;
  int Rank;
  int RtRank, LtRank, TpRank, BmRank;
} SubFieldInfo_S;

/* field.h:52-60 */
typedef struct WaveInfo_S {
  double Lambda_s, T_s, Omega_s, K_s, K_0_s, Angle_deg;
} WaveInfo_S;

/* field.h:62-68 */
typedef struct NFFInfo {
  int top, bottom, left, right;
  int cx, cy;
  double RFperC;
  int arraySize;
} NTFFInfo;

/* field.h:77-82 */
extern int N_X, N_Y, N_CELL, N_PML, N_PX, N_PY;

void field_init(FieldInfo field_info);               /* field.c:89-143 */
void field_reset(void);                              /* field.c:83-87 */
double field_sigmaX(double x, double y);             /* field.c:259-270 */
double field_sigmaY(double x, double y);             /* field.c:272-283 */
double field_toCellUnit(const double nm);            /* field.c:76-78 */
double field_toPhisycalUnit(const double cells);     /* field.c:79-81 */
void field_nextStep(void);                           /* field.c:312-315 */
int field_isFinish(void);                            /* field.c:317-319 */
void field_setWaveAngle(int deg);                    /* field.c:39-42 */
double field_getT(void);                             /* field.c:44-51 */
double field_getK(void);
double field_getRayCoef(void);
double field_getOmega(void);
double field_getLambda(void);
double field_getWaveAngle(void);
double field_getTime(void);
double field_getMaxTime(void);
NTFFInfo field_getNTFFInfo(void);                    /* field.c:52-56 */
WaveInfo_S field_getWaveInfo_S(void);
SubFieldInfo_S field_getSubFieldInfo_S(void);
FieldInfo_S field_getFieldInfo_S(void);
FieldInfo field_getFieldInfo(void);
fdtd_dcomplex field_pointLight(void);                /* field.c:145-152 */
int field_index(int i, int j);                       /* field.c:70-72 */
int field_subIndex(int i, int j);                    /* field.c:66-68 */
int field_getOffsetX(void);                          /* field.c:58-64 */
int field_getOffsetY(void);
int field_getSubNx(void);
int field_getSubNy(void);
int field_getSubNpx(void);
int field_getSubNpy(void);
int field_getSubNcell(void);
int ind(const int i, const int j);                   /* field.c:74 */
double field_pmlCoef(double ep_mu, double sig);      /* field.c:288-291 */
double field_pmlCoef_LXY(double ep_mu, double sig);  /* field.c:292-295 */
double field_ns_beta(double alpha, double alpha_aster);        /* field.c:298-303 */
double field_ns_beta_aster(double alpha, double alpha_aster);  /* field.c:304-309 */
/* sources added to a caller's host array over the interior cells (the solvers
 * here form the same terms on the GPU and do not call these) */
void field_scatteredWave(fdtd_dcomplex *p, double *eps, double gapX, double gapY);          /* field.c:202-218 */
void field_scatteredWaveNotUPML(fdtd_dcomplex *p, double *eps, double gapX, double gapY);   /* field.c:179-199 */
void field_nsScatteredWaveNotUPML(fdtd_dcomplex *p, double *eps, double gapX, double gapY,
                                  double dot);                                             /* field.c:155-177 */
void field_scatteredPulse(fdtd_dcomplex *p, double *eps, double gapX, double gapY, double dot); /* field.c:224-256 */
/* text writers (field.c:322-389); field_outputElliptic writes 0 where its circle
 * leaves the grid, where the reference reads outside its array */
void field_outputElliptic(const char *fileName, fdtd_dcomplex *data);
void field_outputAllDataComplex(const char *fileName, fdtd_dcomplex *data);
void field_outputAllDataDouble(const char *fileName, double *data);

/* myComplex.h:7-14 (myComplex.c:3-51) */
double *newDouble(int size);
void freeDouble(double *array);
fdtd_dcomplex *newDComplex(int size);
void freeDComplex(fdtd_dcomplex *array);
double cnorm(fdtd_dcomplex c);
fdtd_dcomplex cbilinear(fdtd_dcomplex *p, double x, double y, int width, int height);

/* --- extensions (no counterpart in the reference) --- */
/* light speed in cell units (dt = dx = 1, so this is the Courant number);
 * the reference hard-wires 0.7071 (field.h:70-71).  Takes effect at field_init. */
void field_setLightSpeed(double c0);
double field_getLightSpeed(void);
double field_getMu0(void);   /* 1/c0^2 (field.h:74) */
double field_getZ0(void);    /* sqrt(mu/eps); the reference's literal for c0 = 0.7071 (field.h:75) */
/* NTFF contour inset from the PML in cells; reference 5 (field.c:135-138) */
void field_setNTFFInset(int cells);
/* 1: the MPI TM step adds field_pointLight() at (N_PX/2, N_PY/2) after its
 * scattered-wave source (the reference defines field_pointLight but never calls it) */
void field_setPointSource(int on);
int field_getPointSource(void);

#ifdef __cplusplus
}
#endif
#endif
