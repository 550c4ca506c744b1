/*
 * mpifdtd_simulator.h -- the simulator dispatcher (host side).
 *
 * Replaces the reference's simulator.h (simulator.h:8-31) and ntff.h
 * (ntff.h:4-9).  Only the four UPML solvers are backed here; selecting another
 * enum SOLVER value prints an error and exits(2), as the reference does for
 * unknown solvers (simulator.c:202-205).
 */
#ifndef MPIFDTD_SIMULATOR_H
#define MPIFDTD_SIMULATOR_H

#include "mpifdtd_field.h"
#include "mpifdtd_models.h"

#ifdef __cplusplus
extern "C" {
#endif

/* simulator.h:8-18 */
enum SOLVER { TM_2D, TE_2D, TM_UPML_2D, TE_UPML_2D, MPI_TM_UPML_2D, MPI_TE_UPML_2D, NS_TM_2D, NS_TE_2D };

void simulator_init(FieldInfo field_info);         /* simulator.c:226-234 */
void simulator_calc(void);                         /* simulator.c:215-219 */
int simulator_isFinish(void);                      /* simulator.c:270-273 (bool) */
void simulator_finish(void);                       /* simulator.c:257-264 */
void simulator_reset(void);                        /* simulator.c:243-250 */
void simulator_solverInit(void);                   /* simulator.c:236-241 */
void simulator_changeModelAndRestart(void);        /* simulator.c:252-255 */
fdtd_dcomplex *simulator_getDrawingData(void);     /* simulator.c:266-268 */
double *simulator_getEps(void);                    /* simulator.c:276-279 */
void simulator_moveDirectory(void);                /* simulator.c:209-213 */
void simulator_setSolver(enum SOLVER solverType);  /* simulator.c:221-224 */

/* ntff.h:4-9 */
#define LAMBDA_ST_NM 380
#define LAMBDA_EN_NM 700
#define NTFF_NUM 8192
void ntff_outputEnormTxt(double **e_norm, const char *file_name);   /* ntff.c:6-19 */
void ntff_outputEnormBin(double **e_norm, const char *file_name);   /* ntff.c:21-32 */

/* --- extensions --- */
/* configSample.txt reader (main.c:317-366 order: width, height, h_u, pml,
 * lambda, step, start/end/delta angle, model, solver; '#' comments,
 * parser.c:3-20).  Returns 0 on success. */
typedef struct fdtd_run_config {
  FieldInfo field_info;
  int startAngle, endAngle, deltaAngle;
  int model;
  int solver;
} fdtd_run_config;
int fdtd_readConfig(const char *path, fdtd_run_config *out);
/* wall-clock seconds since the last simulator_init / simulator_reset */
double simulator_elapsed(void);
/* nsteps x simulator_calc() in one call */
void fdtd_simulator_run(int nsteps);

#ifdef __cplusplus
}
#endif
#endif
