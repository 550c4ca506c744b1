/*
 * simulator.c -- solver dispatcher (simulator.c:19-279 of the reference).
 *
 * Selects one of the four UPML solvers, the two split-field PML solvers or
 * the two non-standard FDTD solvers by enum SOLVER, wires its methods into the
 * update / init / finish / reset table and drives field_nextStep after every
 * update.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <unistd.h>
#include "mpifdtd_field.h"
#include "mpifdtd_models.h"
#include "mpifdtd_simulator.h"
#include "mpifdtd_solvers.h"

static fdtd_method update_ = NULL, init_ = NULL, finish_ = NULL, reset_ = NULL;
static fdtd_dcomplex *(*draw_)(void) = NULL;
static double *(*eps_)(void) = NULL;
static const char *solverDir_ = "";
static struct timeval t1_;

void simulator_setSolver(enum SOLVER solver)
{
  switch (solver) {
  case TM_2D:  /* setTM (simulator.c:32-47) */
    update_ = fdtdTM_getUpdate(); init_ = fdtdTM_getInit();
    finish_ = fdtdTM_getFinish(); reset_ = fdtdTM_getReset();
    eps_ = fdtdTM_getEps; draw_ = fdtdTM_getEz;
    printf("TM mode \n");
    solverDir_ = "TM";
    break;
  case TE_2D:  /* setTE (simulator.c:49-64) */
    update_ = fdtdTE_getUpdate(); init_ = fdtdTE_getInit();
    finish_ = fdtdTE_getFinish(); reset_ = fdtdTE_getReset();
    eps_ = fdtdTE_getEps; draw_ = fdtdTE_getEy;
    printf("TE mode \n");
    solverDir_ = "TE";
    break;
  case NS_TM_2D:  /* setNsTM (simulator.c:66-81) */
    update_ = nsFdtdTM_getUpdate(); init_ = nsFdtdTM_getInit();
    finish_ = nsFdtdTM_getFinish(); reset_ = nsFdtdTM_getReset();
    eps_ = nsFdtdTM_getEps; draw_ = nsFdtdTM_getEz;
    printf("NS TM mode \n");
    solverDir_ = "NS_TM";
    break;
  case NS_TE_2D:  /* setNsTE (simulator.c:83-99) */
    update_ = nsFdtdTE_getUpdate(); init_ = nsFdtdTE_getInit();
    finish_ = nsFdtdTE_getFinish(); reset_ = nsFdtdTE_getReset();
    eps_ = nsFdtdTE_getEps; draw_ = nsFdtdTE_getEy;
    printf("NS TE mode \n");
    solverDir_ = "NS_TE";
    break;
  case TM_UPML_2D:
    update_ = fdtdTM_upml_getUpdate(); init_ = fdtdTM_upml_getInit();
    finish_ = fdtdTM_upml_getFinish(); reset_ = fdtdTM_upml_getReset();
    eps_ = fdtdTM_upml_getEps; draw_ = fdtdTM_upml_getEz;
    printf("TM UPML mode \n");
    solverDir_ = "TM_UPML";
    break;
  case TE_UPML_2D:
    update_ = fdtdTE_upml_getUpdate(); init_ = fdtdTE_upml_getInit();
    finish_ = fdtdTE_upml_getFinish(); reset_ = fdtdTE_upml_getReset();
    eps_ = fdtdTE_upml_getEps; draw_ = fdtdTE_upml_getEy;
    printf("TE UPML mode \n");
    solverDir_ = "TE_UPML";
    break;
  case MPI_TM_UPML_2D:
    update_ = mpi_fdtdTM_upml_getUpdate(); init_ = mpi_fdtdTM_upml_getInit();
    finish_ = mpi_fdtdTM_upml_getFinish(); reset_ = mpi_fdtdTM_upml_getReset();
    eps_ = mpi_fdtdTM_upml_getEps; draw_ = mpi_fdtdTM_upml_getEz;
    printf("MPI TM UPML mode \n");
    solverDir_ = "MPI_TM_UPML";
    break;
  case MPI_TE_UPML_2D:
    update_ = mpi_fdtdTE_upml_getUpdate(); init_ = mpi_fdtdTE_upml_getInit();
    finish_ = mpi_fdtdTE_upml_getFinish(); reset_ = mpi_fdtdTE_upml_getReset();
    eps_ = mpi_fdtdTE_upml_getEps; draw_ = mpi_fdtdTE_upml_getEy;
    printf("MPI TE UPML mode \n");
    solverDir_ = "MPI_TE_UPML";
    break;
  default:
    printf("error, not implement simulator (simulator.c)\n");
    exit(2);
  }
}

void simulator_moveDirectory(void)
{
  mkdir(solverDir_, 0775);
  if (chdir(solverDir_) != 0) printf("cannot move to %s\n", solverDir_);
}

void simulator_calc(void)
{
  (*update_)();
  field_nextStep();
}

void simulator_init(FieldInfo field_info)
{
  if (!init_) {
    printf("simulator_init: no solver selected\n");
    exit(2);
  }
  field_init(field_info);
  models_initModel();
  (*init_)();
  gettimeofday(&t1_, NULL);
}

void simulator_solverInit(void)
{
  simulator_moveDirectory();
  (*init_)();
}

void simulator_reset(void)
{
  printf("simulator_reset \n");
  field_reset();
  (*reset_)();
  gettimeofday(&t1_, NULL);
}

void simulator_changeModelAndRestart(void)
{
  if (chdir("..") != 0) printf("cannot move to ..\n");
}

double simulator_elapsed(void)
{
  struct timeval t2;
  gettimeofday(&t2, NULL);
  return t2.tv_sec - t1_.tv_sec + (t2.tv_usec - t1_.tv_usec) * 1e-6;
}

void simulator_finish(void)
{
  printf("simulator_finish at %d step \n", (int)field_getTime());
  printf("time = %lf \n", simulator_elapsed());
  (*finish_)();
}

fdtd_dcomplex *simulator_getDrawingData(void) { return (*draw_)(); }
int simulator_isFinish(void) { return field_isFinish(); }
double *simulator_getEps(void) { return (*eps_)(); }

/* n x simulator_calc() without returning to the caller in between */
void fdtd_simulator_run(int nsteps)
{
  for (int n = 0; n < nsteps; n++) {
    (*update_)();
    field_nextStep();
  }
}
