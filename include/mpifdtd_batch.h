/* mpifdtd_batch.h -- many simulations of the MPI UPML solvers on one GPU.
 *
 * No single counterpart in the reference: its solvers keep one instance per
 * process in file-static state (mpiTM_UPML.c:38-57), and a parameter sweep runs
 * one simulation per MPI rank (main.c:114-138, nextSimulation).  Here a batch
 * holds n instances of uneven size, each described by what simulator_init and
 * the model setters take for one run, and advances them together in FP32 with
 * the fused kernels.  Creating a batch sets the field and model globals once
 * per instance (as simulator_init does); it does not touch the active solver of
 * the single-instance API.  Errors return FDTD_* codes (fdtd_hip.h) with
 * fdtd_hip_last_error() for the message. */
#ifndef MPIFDTD_BATCH_H
#define MPIFDTD_BATCH_H

#include "mpifdtd_field.h"
#include "mpifdtd_models.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  FieldInfo field;          /* grid, wavelength, angle and stepNum of this instance */
  int solver;               /* MPI_TM_UPML_2D or MPI_TE_UPML_2D (simulator.h) */
  int model;                /* enum MODEL */
  fdtd_model_params params; /* model parameters (models_defaultParams for the reference's) */
} fdtd_batch_item;

typedef struct fdtd_batch fdtd_batch;

/* T steps per launch, R rows per wavefront of the interior tiles (4, 8, 16) */
int fdtd_batch_create(const fdtd_batch_item *items, int n, int T, int R, fdtd_batch **out);
/* every instance to its stepNum, then synchronised */
int fdtd_batch_run(fdtd_batch *b);
/* queue `blocks` blocks of T steps (asynchronous); read nothing before fdtd_batch_sync */
int fdtd_batch_run_blocks(fdtd_batch *b, int blocks);
int fdtd_batch_sync(fdtd_batch *b, int *steps_done);
/* launch shape of every block (FAST_LAUNCH_* of fdtd_hip.h; TWO_STREAMS_PML8 by default).
 * FDTD_EINVAL for values outside the enum and for the *_P16 shapes, whose
 * 16-wavefront UPML tiles exist for single grids only. */
int fdtd_batch_set_launch_mode(fdtd_batch *b, int mode);
int fdtd_batch_size(const fdtd_batch *b);
/* instance i: NPXa, NPYa, off (1: padded MPI layout), arraySize */
int fdtd_batch_dims(const fdtd_batch *b, int i, int out[4]);
/* instance i's state slot `which` (TM Ez Jz Dz Hx Mx Bx Hy My By, TE likewise) as double complex in
 * the solver layout */
int fdtd_batch_field(fdtd_batch *b, int i, int which, double *out);
/* instance i's 360-angle far field (ntffTM_Frequency / ntffTE_FrequencySplit), double complex */
int fdtd_batch_far_field(fdtd_batch *b, int i, double *out360);
void fdtd_batch_destroy(fdtd_batch *b);

/* The reference's parameter sweep as batch items (main.c:114-138 nextSimulation,
 * main.c:71-88 calcFieldSize): for the current parameters of `model`
 * (models_setParams, or the reference's start values), one item per incidence
 * angle from start_deg to end_deg by delta_deg; then models_isFinish() steps the
 * model to its next structure, until it reports the sweep finished.  With
 * fit_size, each structure's field is the model's models_needSize() plus
 * 2 h (pml + 5) + 200 nm per axis, as calcFieldSize sets it; otherwise base's
 * width_nm / height_nm.  Writes at most max items (items may be NULL with max 0)
 * and returns how many the sweep has, or an FDTD_* error (< 0).  The model is
 * left selected, with its parameters at the end of the sweep, as after the
 * reference's loop. */
int fdtd_batch_sweep(FieldInfo base, int solver, int model, int start_deg, int end_deg, int delta_deg, int fit_size,
                     fdtd_batch_item *items, int max);

#ifdef __cplusplus
}
#endif
#endif
