/* host_internal.h -- functions shared between the host files of libmpifdtd_amd
 * (not exported). */
#ifndef MPIFDTD_HOST_INTERNAL_H
#define MPIFDTD_HOST_INTERNAL_H

#include <complex.h>
#include "fdtd_hip.h"

#define FDTD_HIDDEN __attribute__((visibility("hidden")))

/* solver kinds of solvers.c, in the order of their entry points */
enum { SLV_MPI_TM, SLV_MPI_TE, SLV_TM, SLV_TE, SLV_SPLIT_TM, SLV_SPLIT_TE, SLV_NS_TM, SLV_NS_TE, SLV_N };

/* grid of solver `kind` for the current field and model globals */
FDTD_HIDDEN int solver_make_grid(int kind, int precision, int ntff_time, int T, int R, fdtd_grid **out);
/* fdtd_hip_grid_far_field without any device allocation (fdtd_grid.hip) */
FDTD_HIDDEN int fdtd_hip_grid_far_field_noalloc(fdtd_grid *g, double k, double coef_re, double coef_im, double scale,
                                                double *out360);
/* far-field coefficient sqrt(i k / (8 pi R0)) e^{i k R0} of the current field globals */
FDTD_HIDDEN double complex solver_ff_coef(void);

#endif
