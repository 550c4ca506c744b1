/*
 * fdtd_oracle.h -- CPU restatement of rennone/mpiFDTD's UPML solvers.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library (mpifdtd_amd) never links, loads or calls it.
 *
 * It restates, serially and in FP64 (double complex, like the reference), the
 * single-rank behaviour of:
 *   field.c        field_init / field_nextStep / sigma / sources      (field.c:89-319)
 *   mpiTM_UPML.c   MPI TMz UPML step + per-step NTFF                  (mpiTM_UPML.c:196-530, 663-716, 840-1037)
 *   mpiTE_UPML.c   MPI TEz UPML step + per-step NTFF                  (mpiTE_UPML.c:250-399, 489-536, 602-794)
 *   fdtdTM_upml.c  plain TMz UPML step (H first, Gaussian pulse)      (fdtdTM_upml.c:54-274)
 *   fdtdTE_upml.c  plain TEz UPML step                                (fdtdTE_upml.c:168-412)
 *   ntffTM.c / ntffTE.c / cfft.c / ntff.c  NTFF transforms + spectra
 *   models.c and the noModel / circle / multiLayer / zigzag / morphoScale eps callbacks.
 *
 * Unlike the reference (file-static globals, one instance per process) every
 * call takes an explicit orc_sim handle, and every compile-time model constant
 * is a runtime field of orc_model_params whose defaults equal the reference's
 * #defines.
 */
#ifndef FDTD_ORACLE_H
#define FDTD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* enum SOLVER values (simulator.h:8-18) */
/* enum SOLVER (simulator.h:8-18); TM_2D / TE_2D are the split-field PML solvers (fdtdTM.c, fdtdTE.c) */
enum { ORC_TM_2D = 0, ORC_TE_2D = 1, ORC_TM_UPML = 2, ORC_TE_UPML = 3, ORC_MPI_TM_UPML = 4, ORC_MPI_TE_UPML = 5,
       ORC_NS_TM_2D = 6 /* non-standard FDTD TM (nsFdtdTM.c) */, ORC_NS_TE_2D = 7 /* nsFdtdTE.c */ };
/* enum MODEL values (models.h:5-14) */
enum { ORC_NO_MODEL = 0, ORC_MIE_CYLINDER = 1, ORC_LAYER = 2, ORC_MORPHO_SCALE = 3, ORC_ZIGZAG = 5 };
/* enum MODE values (models.h:16-21) */
enum { ORC_D_X = 0, ORC_D_Y = 1, ORC_D_XY = 2 };

#define ORC_MAX_LAYERS 64

/* Model parameters.  Defaults (orc_default_model_params) are the reference's
 * start-of-sweep #define values. */
typedef struct orc_model_params {
  /* circleModel.c:6-10 */
  double circle_radius_nm;   /* ST_RADIUS_NM 500 */
  double circle_n;           /* N 1.6 */
  /* multiLayerModel.c:6-56 */
  double ml_n0, ml_n1, ml_n0x, ml_n1x;
  int ml_uniaxial, ml_asymmetry, ml_use_gap;
  int ml_width_nm0, ml_width_nm1;
  int ml_thick_nm0, ml_thick_nm1;
  int ml_layer_num, ml_branch_nm, ml_left_gap_nm;
  double ml_edge_rate, ml_curve;
  /* zigzagModel.c:6-30 */
  int zz_width_nm, zz_thick_nm, zz_layer_num, zz_degree;
  double zz_n;
  /* morphoScaleModel.c:10-99 */
  int mo_width_nm, mo_thick_nm0, mo_thick_nm1, mo_layer_num, mo_branch_nm, mo_left_gap_nm;
  double mo_edge_rate, mo_curve, mo_n, mo_nx;
  int mo_uniaxial, mo_asymmetry, mo_randomness, mo_seed;
} orc_model_params;

typedef struct orc_config {
  /* FieldInfo (field.h:20-28) */
  int width_nm, height_nm, h_u_nm, pml, lambda_nm, angle_deg, stepNum;
  int solver;        /* ORC_*_UPML */
  int model;         /* ORC_* model */
  double c0;         /* light speed in cell units; the reference hard-wires 0.7071 (field.h:70-71) */
  int ntff_inset;    /* NTFF contour inset from the PML; reference 5 (field.c:135-138) */
  int point_source;  /* 1: add field_pointLight() (field.c:145-152) to Ez at global (N_PX/2, N_PY/2)
                        right after the scattered-wave source of the MPI TM step; the reference never
                        calls it, so a NoModel run would otherwise stay identically zero */
  int ntff_time;     /* 1: run the per-step time-domain NTFF accumulation (part of every reference step) */
  orc_model_params mp;
} orc_config;

typedef struct orc_sim orc_sim;

void orc_default_model_params(orc_model_params *mp);
void orc_default_config(orc_config *cfg);

orc_sim *orc_create(const orc_config *cfg);
void orc_destroy(orc_sim *s);

/* one simulator_calc(): solver update then field_nextStep (simulator.c:215-219) */
void orc_step(orc_sim *s);
void orc_run(orc_sim *s, int nsteps);
/* simulator_reset semantics without the spectra file output: zero state, time=0 */
void orc_reset(orc_sim *s);
void orc_set_angle(orc_sim *s, int deg);  /* field_setWaveAngle (field.c:39-42) */

/* geometry getters */
void orc_dims(const orc_sim *s, int out[8]); /* N_PX, N_PY, NPXa, NPYa, off, N_PML, arraySize, maxTime */
void orc_ntff_info(const orc_sim *s, int out[6], double *rf_per_c); /* top,bottom,left,right,cx,cy */
double orc_time(const orc_sim *s);
double orc_ray_coef(const orc_sim *s);
double orc_omega(const orc_sim *s);
double orc_k(const orc_sim *s);

/* Field access.  which: 0..8 = the nine state arrays in the order
 *   TM: Ez Jz Dz Hx Mx Bx Hy My By     TE: Ex Jx Dx Ey Jy Dy Hz Mz Bz
 * out receives NPXa*NPYa complex doubles (re,im interleaved) in the solver's
 * own layout (MPI solvers: padded (N_PX+2)x(N_PY+2); plain: N_PX x N_PY). */
void orc_get_field(const orc_sim *s, int which, double *out);
void orc_set_field(orc_sim *s, int which, const double *in);
/* eps maps: which 0 = EPS_EZ (TM) / EPS_EX (TE), 1 = EPS_HX / EPS_EY, 2 = EPS_HY / EPS_HZ */
void orc_get_eps(const orc_sim *s, int which, double *out);
/* time-domain NTFF accumulators: which 0,1,2 = Ux,Uy,Wz (TM) / Wx,Wy,Uz (TE), 360*arraySize complex */
void orc_get_ntff_time(const orc_sim *s, int which, double *out);

/* frequency-domain far field of the current fields, 360 complex values.
 * TM: ntffTM_Frequency (ntffTM.c:72-158) on the fields in N_PX x N_PY layout
 *     (MPI fields are repacked as mpiTM_UPML.c:542-544 does), R0 from ntffTM_init.
 * TE: single-rank restatement of ntffTE_FrequencySplit (ntffTE.c:341-470) with
 *     R0 from ntffTE_init (ntffTE.c:14-17); ntffTE_Frequency itself reads out of
 *     bounds in a serial run and is not restated. */
void orc_ntff_frequency(const orc_sim *s, double *out360);

/* time-domain far field + spectra, as ntffTM_TimeOutput / ntffTE_TimeOutput
 * (ntffTM.c:197-233, ntffTE.c:160-196) without the file output:
 * out[(lambda-380)*360 + ang] for lambda = 380..700 nm (321 x 360). */
void orc_time_spectra(const orc_sim *s, double *out);
/* TimeTranslate only: Eth/Eph as 360*arraySize complex (first maxTime used). */
void orc_time_translate(const orc_sim *s, double *eth, double *eph);
/* field_outputElliptic (field.c:322-341) of the TM_2D / NS_TM_2D solvers' Ez or the TE_2D / NS_TE_2D solvers' Ey
   (their reset() outputs): 181 values, index 180 - ang
   for ang = 180 .. 0, the squared magnitude of Ez interpolated bilinearly (myComplex.c:40-51) on
   the circle of radius 1.2 lambda around the grid centre */
void orc_elliptic(const orc_sim *s, double *out181);

/* models_eps (models.c:166-184) for the configured model; valid after orc_create */
double orc_model_eps(const orc_sim *s, double x, double y, int mode);
/* eps map over the N_PX x N_PY cells: out[x*N_PY+y] = models_eps(x+dx, y+dy, mode) */
void orc_model_eps_map(const orc_sim *s, int mode, double dx, double dy, double *out);

/* sigma profiles field_sigmaX/Y (field.c:259-283) */
double orc_sigma_x(const orc_sim *s, double x, double y);
double orc_sigma_y(const orc_sim *s, double x, double y);

/* cfft (cfft.c:304-307 -> fft1x flag 0): in-place, unnormalised, twiddle e^{+i pi k/xp2} */
void orc_cfft(double *data, int n);

#ifdef __cplusplus
}
#endif
#endif
