/* batch.c -- sweeps of many MPI-solver simulations advanced together on one GPU
 * (the single-GPU form of the reference's rank-per-simulation sweep,
 * main.c:114-138). */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fdtd_hip.h"
#include "host_internal.h"
#include "mpifdtd_batch.h"
#include "mpifdtd_field.h"
#include "mpifdtd_models.h"
#include "mpifdtd_simulator.h"

struct fdtd_batch {
  int n;
  fdtd_grid **g;
  double *k, *scale;
  fdtd_step_params *p0;
  double complex *coef;
  int *steps;
  int (*dims)[4];
  fdtd_hip_batch *hb;
};

static void batch_free(fdtd_batch *b)
{
  if (!b) return;
  if (b->hb) fdtd_hip_batch_destroy(b->hb);
  for (int i = 0; i < b->n; i++)
    if (b->g && b->g[i]) fdtd_hip_grid_destroy(b->g[i]);
  free(b->g);
  free(b->p0);
  free(b->k);
  free(b->scale);
  free(b->coef);
  free(b->steps);
  free(b->dims);
  free(b);
}

int fdtd_batch_create(const fdtd_batch_item *items, int n, int T, int R, fdtd_batch **out)
{
  if (!items || !out || n < 1) return FDTD_EINVAL;
  *out = NULL;
  fdtd_batch *b = (fdtd_batch *)calloc(1, sizeof *b);
  if (!b) return FDTD_ENOMEM;
  b->g = (fdtd_grid **)calloc(n, sizeof *b->g);
  b->p0 = (fdtd_step_params *)calloc(n, sizeof(fdtd_step_params));
  b->k = (double *)calloc(n, sizeof(double));
  b->scale = (double *)calloc(n, sizeof(double));
  b->coef = (double complex *)calloc(n, sizeof(double complex));
  b->steps = (int *)calloc(n, sizeof(int));
  b->dims = (int (*)[4])calloc(n, sizeof *b->dims);
  if (!b->g || !b->p0 || !b->k || !b->scale || !b->coef || !b->steps || !b->dims) {
    batch_free(b);
    return FDTD_ENOMEM;
  }
  b->n = n;
  int rc = FDTD_OK;
  for (int i = 0; i < n && rc == FDTD_OK; i++) {
    const fdtd_batch_item *it = &items[i];
    int kind;
    if (it->solver == MPI_TM_UPML_2D) kind = SLV_MPI_TM;
    else if (it->solver == MPI_TE_UPML_2D) kind = SLV_MPI_TE;
    else {
      rc = FDTD_EINVAL;
      break;
    }
    /* the per-instance part of simulator_init (simulator.c:215-219) */
    models_setModel((enum MODEL)it->model);
    models_setParams(&it->params);
    field_init(it->field);
    models_initModel();
    rc = solver_make_grid(kind, FDTD_FP32, 0, T, R, &b->g[i]);
    if (rc != FDTD_OK) break;
    b->p0[i].omega = field_getOmega();
    b->p0[i].k = field_getK();
    b->p0[i].angle_deg = field_getWaveAngle();
    b->k[i] = field_getK();
    b->coef[i] = solver_ff_coef();
    b->scale[i] = sqrt((double)it->field.h_u_nm);
    b->steps[i] = (int)field_getMaxTime();
    rc = fdtd_hip_grid_dims(b->g[i], b->dims[i]);
  }
  if (rc == FDTD_OK) rc = fdtd_hip_batch_create(b->g, b->p0, b->steps, n, T, R, &b->hb);
  if (rc != FDTD_OK) {
    batch_free(b);
    return rc;
  }
  *out = b;
  return FDTD_OK;
}

int fdtd_batch_run(fdtd_batch *b) { return b ? fdtd_hip_batch_run(b->hb, -1) : FDTD_EINVAL; }
int fdtd_batch_run_blocks(fdtd_batch *b, int blocks)
{
  return b && blocks >= 0 ? fdtd_hip_batch_run(b->hb, blocks) : FDTD_EINVAL;
}
int fdtd_batch_sync(fdtd_batch *b, int *steps_done) { return b ? fdtd_hip_batch_sync(b->hb, steps_done) : FDTD_EINVAL; }
int fdtd_batch_size(const fdtd_batch *b) { return b ? b->n : 0; }
int fdtd_batch_set_launch_mode(fdtd_batch *b, int mode)
{
  return b ? fdtd_hip_batch_set_launch_mode(b->hb, mode) : FDTD_EINVAL;
}

int fdtd_batch_dims(const fdtd_batch *b, int i, int out[4])
{
  if (!b || i < 0 || i >= b->n || !out) return FDTD_EINVAL;
  memcpy(out, b->dims[i], sizeof b->dims[i]);
  return FDTD_OK;
}

int fdtd_batch_field(fdtd_batch *b, int i, int which, double *out)
{
  if (!b || i < 0 || i >= b->n || !out) return FDTD_EINVAL;
  return fdtd_hip_grid_get_field(b->g[i], which, out);
}

int fdtd_batch_far_field(fdtd_batch *b, int i, double *out360)
{
  if (!b || i < 0 || i >= b->n || !out360) return FDTD_EINVAL;
  return fdtd_hip_grid_far_field(b->g[i], b->k[i], creal(b->coef[i]), cimag(b->coef[i]), b->scale[i], out360);
}

void fdtd_batch_destroy(fdtd_batch *b) { batch_free(b); }

/* main.c:71-88 (calcFieldSize) and main.c:114-138 (nextSimulation): the angles of
 * one structure, then the next structure, each rank's stride being 1 here */
int fdtd_batch_sweep(FieldInfo base, int solver, int model, int start_deg, int end_deg, int delta_deg, int fit_size,
                     fdtd_batch_item *items, int max)
{
  if (solver != MPI_TM_UPML_2D && solver != MPI_TE_UPML_2D) return FDTD_EINVAL;
  if (model != NO_MODEL && model != MIE_CYLINDER && model != LAYER && model != MORPHO_SCALE && model != ZIGZAG)
    return FDTD_EINVAL;
  if (max < 0 || (max > 0 && !items)) return FDTD_EINVAL;
  if (end_deg < start_deg || (delta_deg <= 0 && end_deg > start_deg)) return FDTD_EINVAL;
  const int per_structure = delta_deg > 0 ? (end_deg - start_deg) / delta_deg + 1 : 1;
  fdtd_model_params p;
  models_setModel((enum MODEL)model);  /* keeps the model's current parameters */
  int n = 0;
  for (int guard = 0; guard < 1000000; guard++) {
    models_getParams(&p);
    FieldInfo fi = base;
    if (fit_size) {
      int x_nm, y_nm;
      models_needSize(&x_nm, &y_nm);
      fi.width_nm = x_nm + fi.h_u_nm * (fi.pml + 5) * 2 + 200;
      fi.height_nm = y_nm + fi.h_u_nm * (fi.pml + 5) * 2 + 200;
    }
    for (int a = 0; a < per_structure; a++, n++) {
      if (n >= max) continue;
      fi.angle_deg = start_deg + a * (delta_deg > 0 ? delta_deg : 0);
      items[n].field = fi;
      items[n].solver = solver;
      items[n].model = model;
      items[n].params = p;
    }
    if (models_isFinish()) return n;
  }
  return FDTD_EINVAL;  /* a model whose sweep never ends */
}
