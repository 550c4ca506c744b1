/*
 * mpifdtd_solvers.h -- the four UPML solvers', the two split-field PML
 * solvers' and the two non-standard FDTD solvers' entry points, backed by HIP.
 *
 * Same symbols and signatures as the reference's solver headers:
 *   mpiTM_UPML.h:5-18   (mpiTM_UPML.c:88-136)   MPI TMz UPML, run as one rank on one GPU
 *   mpiTE_UPML.h:5-19   (mpiTE_UPML.c:72-126)   MPI TEz UPML, run as one rank on one GPU
 *   fdtdTM_upml.h:5-14  (fdtdTM_upml.c:117-152) plain TMz UPML
 *   fdtdTE_upml.h:5-14                          plain TEz UPML
 *   fdtdTM.h:5-16       (fdtdTM.c:39-78)        split-field PML TMz (FP64 only; its
 *                                               reset() writes tm_<h>nm.txt)
 *   fdtdTE.h:5-16       (fdtdTE.c:29-66)        split-field PML TEz (FP64 only; its
 *                                               reset() writes te_<h>nm.txt)
 *   nsFdtdTM.h:6-17     (nsFdtdTM.c:32-42)      non-standard FDTD TMz (FP64 only; its
 *                                               reset() writes ns_tm_<h>nm.txt;
 *                                               getEzx / getEzy, declared but not
 *                                               defined by the reference, are defined)
 *   nsFdtdTE.h:6-17     (nsFdtdTE.c:29-39)      non-standard FDTD TEz (FP64 only; its
 *                                               reset() writes ns_te_<h>nm.txt; getEps
 *                                               returns EPS_EY with the reference's
 *                                               zero border; getHzx / getHzy defined)
 * get{Update,Init,Finish,Reset} return the solver's void(void) methods; the
 * field getters return borrowed host pointers (valid from init to finish) that
 * are refreshed from device memory on every getter call; getEps returns the
 * same eps map the reference does (MPI TM: EPS_EZ, MPI TE: EPS_EY, plain TM:
 * EPS_EZ, plain TE: EPS_EX).  The MPI variants keep the reference's padded
 * (N_PX+2)x(N_PY+2) layout; none of them calls MPI_Finalize.
 */
#ifndef MPIFDTD_SOLVERS_H
#define MPIFDTD_SOLVERS_H

#include "mpifdtd_field.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*fdtd_method)(void);

/* mpiTM_UPML.h:5-18 */
fdtd_method mpi_fdtdTM_upml_getUpdate(void);
fdtd_method mpi_fdtdTM_upml_getFinish(void);
fdtd_method mpi_fdtdTM_upml_getReset(void);
fdtd_method mpi_fdtdTM_upml_getInit(void);
fdtd_dcomplex *mpi_fdtdTM_upml_getHx(void);
fdtd_dcomplex *mpi_fdtdTM_upml_getHy(void);
fdtd_dcomplex *mpi_fdtdTM_upml_getEz(void);
double *mpi_fdtdTM_upml_getEps(void);
int mpi_fdtdTM_upml_getSubNx(void);
int mpi_fdtdTM_upml_getSubNy(void);
int mpi_fdtdTM_upml_getSubNpx(void);
int mpi_fdtdTM_upml_getSubNpy(void);
int mpi_fdtdTM_upml_getSubNcell(void);

/* mpiTE_UPML.h:5-19 */
fdtd_method mpi_fdtdTE_upml_getUpdate(void);
fdtd_method mpi_fdtdTE_upml_getFinish(void);
fdtd_method mpi_fdtdTE_upml_getReset(void);
fdtd_method mpi_fdtdTE_upml_getInit(void);
fdtd_dcomplex *mpi_fdtdTE_upml_getEx(void);
fdtd_dcomplex *mpi_fdtdTE_upml_getEy(void);
fdtd_dcomplex *mpi_fdtdTE_upml_getHz(void);
double *mpi_fdtdTE_upml_getEps(void);
int mpi_fdtdTE_upml_getSubNx(void);
int mpi_fdtdTE_upml_getSubNy(void);
int mpi_fdtdTE_upml_getSubNpx(void);
int mpi_fdtdTE_upml_getSubNpy(void);
int mpi_fdtdTE_upml_getSubNcell(void);

/* fdtdTM_upml.h:5-14 */
fdtd_method fdtdTM_upml_getUpdate(void);
fdtd_method fdtdTM_upml_getFinish(void);
fdtd_method fdtdTM_upml_getReset(void);
fdtd_method fdtdTM_upml_getInit(void);
fdtd_dcomplex *fdtdTM_upml_getHx(void);
fdtd_dcomplex *fdtdTM_upml_getHy(void);
fdtd_dcomplex *fdtdTM_upml_getEz(void);
double *fdtdTM_upml_getEps(void);

/* fdtdTM.h:5-16 (split-field PML TMz; getEps returns EPS_EZ) */
fdtd_method fdtdTM_getUpdate(void);
fdtd_method fdtdTM_getFinish(void);
fdtd_method fdtdTM_getReset(void);
fdtd_method fdtdTM_getInit(void);
fdtd_dcomplex *fdtdTM_getEzx(void);
fdtd_dcomplex *fdtdTM_getEzy(void);
fdtd_dcomplex *fdtdTM_getEz(void);
fdtd_dcomplex *fdtdTM_getHy(void);
fdtd_dcomplex *fdtdTM_getHx(void);
double *fdtdTM_getEps(void);

/* fdtdTE.h:5-16 (split-field PML TEz; getEps returns EPS_EY) */
fdtd_method fdtdTE_getUpdate(void);
fdtd_method fdtdTE_getFinish(void);
fdtd_method fdtdTE_getReset(void);
fdtd_method fdtdTE_getInit(void);
fdtd_dcomplex *fdtdTE_getEx(void);
fdtd_dcomplex *fdtdTE_getEy(void);
fdtd_dcomplex *fdtdTE_getHzx(void);
fdtd_dcomplex *fdtdTE_getHzy(void);
fdtd_dcomplex *fdtdTE_getHz(void);
double *fdtdTE_getEps(void);

/* nsFdtdTM.h:6-19 (non-standard FDTD TMz; getEps returns EPS_EZ). */
fdtd_method nsFdtdTM_getUpdate(void);
fdtd_method nsFdtdTM_getFinish(void);
fdtd_method nsFdtdTM_getReset(void);
fdtd_method nsFdtdTM_getInit(void);
fdtd_dcomplex *nsFdtdTM_getEzx(void);
fdtd_dcomplex *nsFdtdTM_getEzy(void);
fdtd_dcomplex *nsFdtdTM_getEz(void);
fdtd_dcomplex *nsFdtdTM_getHy(void);
fdtd_dcomplex *nsFdtdTM_getHx(void);
double *nsFdtdTM_getEps(void);
/* solver.h:7-20's struct Solver (ten function pointers), here incomplete so
 * that the reference's own typedef stays compatible.  As nsFdtdTM.c:45-61 and
 * nsFdtdTE.c:42-58 do, the getters return a static Solver whose fields are all
 * unset (null). */
struct Solver;
struct Solver *nsFdtdTM_getSolver(void);
struct Solver *nsFdtdTE_getSolver(void);
/* nsFdtdTM.c:35-37 / nsFdtdTE.c:33-35 define these without a header declaration:
 * TM X EPS_HX, Y EPS_HY, Z EPS_EZ; TE X EPS_EX, Y EPS_EY, Z EPS_HZ (the TE maps
 * with the reference's zero border).  Each is a host copy of its own. */
double *nsFdtdTM_getEpsX(void);
double *nsFdtdTM_getEpsY(void);
double *nsFdtdTM_getEpsZ(void);
double *nsFdtdTE_getEpsX(void);
double *nsFdtdTE_getEpsY(void);
double *nsFdtdTE_getEpsZ(void);

/* nsFdtdTE.h:6-19 (non-standard FDTD TEz; getEps returns EPS_EY; nsFdtdTE_getSolver
 * is declared with nsFdtdTM_getSolver above). */
fdtd_method nsFdtdTE_getUpdate(void);
fdtd_method nsFdtdTE_getFinish(void);
fdtd_method nsFdtdTE_getReset(void);
fdtd_method nsFdtdTE_getInit(void);
fdtd_dcomplex *nsFdtdTE_getHzx(void);
fdtd_dcomplex *nsFdtdTE_getHzy(void);
fdtd_dcomplex *nsFdtdTE_getHz(void);
fdtd_dcomplex *nsFdtdTE_getEy(void);
fdtd_dcomplex *nsFdtdTE_getEx(void);
double *nsFdtdTE_getEps(void);

/* fdtdTE_upml.h:5-14 */
fdtd_method fdtdTE_upml_getUpdate(void);
fdtd_method fdtdTE_upml_getFinish(void);
fdtd_method fdtdTE_upml_getReset(void);
fdtd_method fdtdTE_upml_getInit(void);
fdtd_dcomplex *fdtdTE_upml_getEx(void);
fdtd_dcomplex *fdtdTE_upml_getEy(void);
fdtd_dcomplex *fdtdTE_upml_getHz(void);
double *fdtdTE_upml_getEps(void);

/* --- extensions on the active solver (no counterpart in the reference) --- */
/* which 0..8 in the solver's slot order (TM Ez Jz Dz Hx Mx Bx Hy My By,
 * TE Ex Jx Dx Ey Jy Dy Hz Mz Bz); NPXa*NPYa complex doubles.  FP32 solvers
 * store the UPML auxiliaries (slots 1 2 4 5 7 8) only outside the box that
 * holds no UPML cell: there getState reads zero and setState drops the values */
int fdtd_solver_getState(int which, double *out);
int fdtd_solver_setState(int which, const double *in);
/* time-domain NTFF accumulators (which 0..2: Ux Uy Wz / Wx Wy Uz), 360*arraySize complex */
int fdtd_solver_getNTFFTime(int which, double *out);
/* frequency-domain far field of the current fields, 360 complex (re,im):
 * TM ntffTM_Frequency (ntffTM.c:72-158); TE single-rank ntffTE_FrequencySplit
 * (ntffTE.c:341-470) with R0 of ntffTE_init */
int fdtd_solver_farField(double *out360);
/* fdtd_solver_farField without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The fields are read in place (FP32
 * solvers in FP32, without the FP64 export buffers fdtd_solver_farField
 * allocates on its first call) and the partial sums live in a file-scope
 * device array of 1.11 MB per GPU, shared by every call on that GPU: calls that
 * overlap on one GPU race.  FDTD_EINVAL for an NTFF contour of more than 65536
 * cells (sides up to ~16000 cells), FDTD_ESTATE while FP32 steps are queued
 * (any getter launches them).  Same values as fdtd_solver_farField up to the
 * order of the FP64 sums. */
int fdtd_solver_farFieldNoAlloc(double *out360);
/* time-domain far-field spectra 380..700 nm x 360 angles, as the plain
 * solvers' ntff*_TimeOutput computes them (ntffTM.c:197-233, ntffTE.c:160-196) */
int fdtd_solver_timeSpectra(double *out321x360);
/* running DFT of the NTFF contour at n wavelengths (nm) of the active solver,
 * from the next step on (n = 0 turns it off).  No counterpart in the
 * reference: the spectra of its pulse solvers come from the time-domain
 * accumulators and an 8192-point FFT (ntffTM.c:197-233); the running DFT gives
 * the same transform at exact frequencies without the [360][arraySize] bins. */
int fdtd_solver_setDFT(const double *lambda_nm, int n);
/* out[n][3][360] complex: sum_m U_a[m] e^{i w m} of the three time-domain
 * accumulator rows (Ux Uy Wz / Wx Wy Uz) of every angle at each frequency */
int fdtd_solver_getDFT(double *out);
/* out[n][360]: |E|^2 / 8192 with E the time translation of fdtd_solver_timeSpectra
 * applied to fdtd_solver_getDFT at each wavelength */
int fdtd_solver_dftSpectra(double *out);
/* 0: skip per-step time-domain NTFF accumulation (default 1, as the reference); set before init */
void fdtd_solver_setNTFFTime(int on);
/* 0: reset()/finish() write no output files (default 1, as the reference) */
void fdtd_solver_setFileOutput(int on);
/* dims of the active solver: NPXa, NPYa, off, arraySize */
int fdtd_solver_dims(int out[4]);
/* FDTD_FP64 (default): reference-order FP64 kernels; FDTD_FP32: fused FP32
 * kernels with temporal blocking (set before init).  Getters return double
 * complex in both cases. */
void fdtd_solver_setPrecision(int precision);
int fdtd_solver_getPrecision(void);
/* FP32: steps per fused launch T and rows per wavefront R of the active / next solver */
int fdtd_solver_setBlocking(int T, int R);
/* FP32: launch shape of the active solver's UPML and interior tiles
 * (FAST_LAUNCH_* of fdtd_hip.h); FDTD_ESTATE for FP64 solvers */
int fdtd_solver_setLaunchMode(int mode);
/* FP32 with contour samples: the sample packing, binning and running DFT of a
 * block beside the next block on a stream of their own (1), in order after it
 * (0), or by grid size (-1, the default; fdtd_hip_grid_set_ntff_overlap) */
int fdtd_solver_setNTFFOverlap(int on);
/* the setting, or with AUTO the shape of the last launch */
int fdtd_solver_getLaunchMode(int *mode);
/* blocking of the active solver; FDTD_ESTATE when it does not queue steps (FP64) */
int fdtd_solver_getBlocking(int *T, int *R);
/* Non-finite field detection (no counterpart in the reference, whose runs carry
 * NaN or Inf on to the end).  out[3]: values with a NaN or Inf part in each
 * physical field of the active solver (TM Ez Hx Hy / TE Ex Ey Hz), after its
 * queued steps; FDTD_ESTATE without an active solver. */
int fdtd_solver_checkFinite(long long out[3]);
/* the same check inside the step loop every `every` steps (0: off, the
 * default): the first detection since init / reset prints the count and the
 * step time and is kept for fdtd_solver_finiteStatus; with abort_on_detect it
 * then exits with status 2, the reference's failure exit.  FP32 solvers launch
 * their queued steps for each check: make `every` a multiple of the block depth. */
int fdtd_solver_setFiniteCheck(int every, int abort_on_detect);
/* 1 and the first detection's step time and count, or 0 (step -1, count 0) */
int fdtd_solver_finiteStatus(int *step, long long *count);

#ifdef __cplusplus
}
#endif
#endif
