/*
 * mpifdtd_models.h -- dielectric model interface (host side).
 *
 * Replaces the reference's models.h (models.h:5-31) for the five models the
 * UPML path uses.  The reference fixes each model's parameters with
 * compile-time #defines and steps them with models_isFinish(); here the same
 * parameters are runtime fields (fdtd_model_params) whose defaults are the
 * reference's start-of-sweep values, and models_isFinish() steps them with
 * the reference's sweep rules.
 */
#ifndef MPIFDTD_MODELS_H
#define MPIFDTD_MODELS_H

#include "fdtd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* models.h:5-14 */
enum MODEL { NO_MODEL, MIE_CYLINDER, LAYER, MORPHO_SCALE, CONCENTRIC_CIRCLE, ZIGZAG, TRACE_IMAGE };
/* models.h:16-21 */
enum MODE { D_X, D_Y, D_XY };

typedef struct fdtd_model_params {
  /* circleModel.c:6-10 */
  double circle_radius_nm, circle_n;
  /* multiLayerModel.c:6-56 */
  double ml_n0, ml_n1, ml_n0x, ml_n1x;
  int ml_uniaxial, ml_asymmetry, ml_use_gap;
  int ml_width_nm0, ml_width_nm1, ml_thick_nm0, ml_thick_nm1;
  int ml_layer_num, ml_branch_nm, ml_left_gap_nm;
  double ml_edge_rate, ml_curve;
  /* zigzagModel.c:6-30 */
  int zz_width_nm, zz_thick_nm, zz_layer_num, zz_degree;
  double zz_n;
  /* morphoScaleModel.c:10-99 */
  int mo_width_nm, mo_thick_nm0, mo_thick_nm1, mo_layer_num, mo_branch_nm, mo_left_gap_nm;
  double mo_edge_rate, mo_curve, mo_n, mo_nx;
  int mo_uniaxial, mo_asymmetry, mo_randomness, mo_seed;
} fdtd_model_params;

void models_setModel(enum MODEL model);                 /* models.c:139-164 */
double models_eps(double x, double y, enum MODE mode);  /* models.c:166-184 */
int models_isFinish(void);                              /* models.c:126-129 (bool) */
void models_moveDirectory(void);                        /* models.c:132-137 */
void models_needSize(int *x_nm, int *y_nm);             /* models.c:186-189 */
void models_initModel(void);                            /* models.c:191-194 */

/* --- extensions --- */
void models_defaultParams(fdtd_model_params *p);
void models_setParams(const fdtd_model_params *p);
void models_getParams(fdtd_model_params *p);
enum MODEL models_getModel(void);
/* derived state in cell units, valid after models_initModel(); consumed by the
 * device eps-fill kernel */
const fdtd_model_state *models_getState(void);

#ifdef __cplusplus
}
#endif
#endif
