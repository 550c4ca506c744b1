/*
 * fdtd_hip.h -- thin C ABI between the C host (field / models / simulator /
 * solver entry points) and the gfx950 HIP kernels of libmpifdtd_amd.so.
 *
 * Plain pointers, sizes and PODs only.  All host pointers are ordinary host
 * memory; device memory is owned by the grid / batch objects.  Every call
 * returns 0 on success or a negative FDTD_E* code; fdtd_hip_last_error()
 * returns a message for the last failure on the calling thread.
 *
 * This layer has no counterpart in the reference (which is CPU-only); the
 * reference-facing entry points are declared in mpifdtd_solvers.h,
 * mpifdtd_field.h, mpifdtd_models.h and mpifdtd_simulator.h.
 */
#ifndef FDTD_HIP_H
#define FDTD_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FDTD_OK = 0, FDTD_EINVAL = -1, FDTD_EHIP = -2, FDTD_ENOMEM = -3, FDTD_ESTATE = -4 };

enum { FDTD_MODE_TM = 0, FDTD_MODE_TE = 1 };
/* MPI flavour: padded (N_PX+2)x(N_PY+2) arrays, E half first, CW scattered wave
 *              (mpiTM_UPML.c / mpiTE_UPML.c single rank).
 * plain flavour: N_PX x N_PY arrays, H half first, Gaussian pulse
 *              (fdtdTM_upml.c / fdtdTE_upml.c).
 * split flavour: the split-field PML solvers (fdtdTM.c, fdtdTE.c): N_PX x N_PY
 *              arrays, CW scattered wave; FP64, no per-step NTFF.  TM: H half
 *              first, the wave on Ezx, slots Ez 0, Ezx 1, Ezy 2, Hx 3, Hy 6;
 *              TE: E half first, the wave on Ey, slots Ex 0, Ey 3, Hz 6, Hzx 7,
 *              Hzy 8.
 * ns flavour: the non-standard FDTD solvers (nsFdtdTM.c, nsFdtdTE.c): as the
 *              split flavour (N_PX x N_PY, FP64, no per-step NTFF, H half first,
 *              the same slots), with the non-standard second differences in
 *              the half that reads the split field's sum.  TM: the NS scattered
 *              wave on Ezy; coefficients from fdtd_hip_grid_set_coefs (1-D
 *              decay tables) and fdtd_hip_grid_set_ns_coefs (per cell).  TE: the
 *              NS wave on Ex and Ey at gap (0, 1/2); every coefficient per cell
 *              (fdtd_hip_grid_set_ns_coefs); the eps maps sample Hz at D_XY and
 *              keep a zero border, as nsFdtdTE.c:116-121 leaves them. */
enum { FDTD_FLAVOUR_MPI = 0, FDTD_FLAVOUR_PLAIN = 1, FDTD_FLAVOUR_SPLIT = 2, FDTD_FLAVOUR_NS = 3 };
enum { FDTD_FP64 = 0, FDTD_FP32 = 1 };

/* Model ids: enum MODEL (models.h:5-14) values */
enum { FDTD_NO_MODEL = 0, FDTD_MIE_CYLINDER = 1, FDTD_LAYER = 2, FDTD_MORPHO_SCALE = 3, FDTD_ZIGZAG = 5 };

#define FDTD_MAX_LAYERS 32

/* Derived model state in cell units (what each model's *_init computes),
 * shared by the host eps callback and the device eps-fill kernel. */
typedef struct fdtd_model_state {
  int model;
  int N_PX, N_PY, N_X, N_Y, N_PML;
  double eps0;
  /* circle (circleModel.c:119-126) */
  double ci_radius_s, ci_eps_s, ci_posx, ci_posy;
  /* multilayer (multiLayerModel.c:331-348) */
  int ml_layer_num, ml_uniaxial, ml_asymmetry, ml_use_gap;
  double ml_width_s[2], ml_thick_s[2], ml_branch_s, ml_gap_s, ml_ep[2], ml_epx[2], ml_c0, ml_c1, ml_edge_rate;
  /* zigzag (zigzagModel.c:146-160) */
  int zz_layer_num;
  double zz_rad, zz_width_s, zz_thick_s, zz_height_s, zz_ox, zz_oy, zz_exw, zz_exh, zz_ep;
  /* morpho scale (morphoScaleModel.c:379-422) */
  int mo_layer_num, mo_uniaxial;
  double mo_width_s, mo_thick_s[2], mo_branch_s, mo_gap_s, mo_height_s, mo_ox, mo_oy, mo_ep, mo_epx;
  double mo_edge_rate, mo_curve;
  double mo_rand[2][FDTD_MAX_LAYERS];
  double mo_lam_a[2][FDTD_MAX_LAYERS], mo_lam_b1[2][FDTD_MAX_LAYERS], mo_lam_b2[2][FDTD_MAX_LAYERS];
} fdtd_model_state;

/* Geometry and constants of one simulation instance (from field.c's state). */
typedef struct fdtd_grid_desc {
  int mode, flavour, precision;
  int N_PX, N_PY, N_PML;
  /* NTFF contour (field.c:133-142) */
  int ntff_top, ntff_bottom, ntff_left, ntff_right, ntff_cx, ntff_cy;
  double rf_per_c;
  int array_size, max_time;
  double c0, mu0, z0, eps0;
  int ntff_time;     /* 1: per-step time-domain NTFF accumulation */
  int point_source;  /* 1: soft point source at (N_PX/2, N_PY/2) in the MPI TM step */
} fdtd_grid_desc;

/* UPML coefficient tables, indexed by physical cell coordinate.  Built on the
 * host from field_sigmaX/Y exactly as setCoefficient does
 * (mpiTM_UPML.c:663-716, mpiTE_UPML.c:489-536, fdtdTM_upml.c:224-274,
 * fdtdTE_upml.c:361-412); only the two coefficients that mix sigma_x and
 * sigma_y are formed on the device, from the raw sigma tables sx / sy.
 *   TM: x0 C_JZ  x1 C_JZHXHY  x2 C_BXMX1  x3 C_BXMX0  x4 C_BY      sx = sig_hy_x
 *       y0 C_DZ  y1 C_DZJZ1(=C_DZJZ0)  y2 C_MX  y3 C_MXEZ         sy = sig_hy_y
 *   TE: x0 C_DXJX1 x1 C_DXJX0 x2 C_DY  x3 C_MZ   x4 C_MZEXEY     sx = sig_ey_x
 *       y0 C_JX  y1 C_JXHZ  y2 C_BZ  y3 C_BZMZ1(=C_BZMZ0)         sy = sig_ey_y
 * Split flavour (fdtdTM.c:184-243, fdtdTE.c:200-241; the coefficients are formed
 * per cell on the device from these conductivities and eps, as field_pmlCoef /
 * _LXY do; TM and TE use the same four tables):
 *       x0 sig_max sigma_x(i)  (TM sig_ez_x, TE sig_ey_x)
 *       x1 mu0/eps0 sig_max sigma_x(i + 1/2)  (TM sig_hy_xx, TE sig_hz_xx)
 *       y0 sig_max sigma_y(j)  (TM sig_ez_y, TE sig_ex_y)
 *       y1 mu0/eps0 sig_max sigma_y(j + 1/2)  (TM sig_hx_yy, TE sig_hz_yy)
 * NS flavour (nsFdtdTM.c:226-306, coef1 (1-b)/(1+b) of the simplified beta b):
 *       x0 C_EZX(i)  x1 C_HY(i + 1/2)   y0 C_EZY(j)  y1 C_HX(j + 1/2)
 * Each x table has N_PX entries, each y table N_PY. */
typedef struct fdtd_coef_tables {
  const double *x[5];
  const double *sx;
  const double *y[4];
  const double *sy;
} fdtd_coef_tables;

/* Per-step scalars, read from field.c's getters by the host each step. */
typedef struct fdtd_step_params {
  double time;       /* field_getTime() */
  double ray_coef;   /* field_getRayCoef() */
  double omega;      /* field_getOmega() */
  double k;          /* field_getK() */
  double angle_deg;  /* field_getWaveAngle() */
  /* point source value ray_coef * cexp(i omega time) (field.c:145-152) */
  double ps_re, ps_im;
  /* plain TE: projection of the pulse onto Ex / Ey (fdtdTE_upml.c:182-189) */
  double dot_x, dot_y;
  int use_x, use_y;
} fdtd_step_params;

typedef struct fdtd_grid fdtd_grid;

const char *fdtd_hip_last_error(void);
int fdtd_hip_device_count(void);
/* device memory and synchronisation through the library's own HIP runtime, for
 * callers that hold no other GPU runtime (no counterpart in the reference) */
int fdtd_hip_alloc(size_t bytes, void **out);
int fdtd_hip_free(void *p);
int fdtd_hip_upload(void *dst, const void *src, size_t bytes);
int fdtd_hip_download(void *dst, const void *src, size_t bytes);
int fdtd_hip_synchronize(void);
/* the solver objects keep freed device blocks for reuse by later objects of
 * the same sizes; this returns them to the runtime */
int fdtd_hip_trim(void);

int fdtd_hip_grid_create(const fdtd_grid_desc *d, fdtd_grid **out);
int fdtd_hip_grid_destroy(fdtd_grid *g);
int fdtd_hip_grid_set_coefs(fdtd_grid *g, const fdtd_coef_tables *t);
/* eps maps on the device from the model state (3 maps, see fdtd_hip_grid_get_eps) */
int fdtd_hip_grid_fill_eps(fdtd_grid *g, const fdtd_model_state *m);
/* which: TM 0 EPS_EZ 1 EPS_HX 2 EPS_HY ; TE 0 EPS_EX 1 EPS_EY 2 EPS_HZ; solver layout */
int fdtd_hip_grid_set_eps(fdtd_grid *g, int which, const double *host);
int fdtd_hip_grid_get_eps(fdtd_grid *g, int which, double *host);
/* NS flavour: the per-cell coefficients as N_PX*N_PY maps in this order.
 *   TM, 5 maps: C_EZXLX, C_EZYLY, C_HXLY, C_HYLX (nsFdtdTM.c:276-304) and the
 *       factor 1/(_n n) - 1 of the NS scattered wave (field.c:155-177) of EPS_EZ;
 *   TE, 10 maps: C_EX, C_EXLY, C_EY, C_EYLX, C_HZX, C_HZXLX, C_HZY, C_HZYLY
 *       (nsFdtdTE.c:145-178) and the wave factors of EPS_EX and of EPS_EY.
 * r_half is r/2 of the second differences, r = 1/6 + k^2/180 - k^4/23040
 * (nsFdtdTM.c:117-118, nsFdtdTE.c:255-256) */
int fdtd_hip_grid_set_ns_coefs(fdtd_grid *g, const double *cells, double r_half);
/* one reference update(): E/H halves, source, per-step NTFF */
int fdtd_hip_grid_step(fdtd_grid *g, const fdtd_step_params *p);
/* which 0..8: TM Ez Jz Dz Hx Mx Bx Hy My By / TE Ex Jx Dx Ey Jy Dy Hz Mz Bz;
 * host buffers hold NPXa*NPYa complex doubles (re,im) in the solver layout */
int fdtd_hip_grid_get_field(fdtd_grid *g, int which, double *host);
int fdtd_hip_grid_set_field(fdtd_grid *g, int which, const double *host);
/* time-domain NTFF accumulators (which 0..2), 360*array_size complex */
int fdtd_hip_grid_get_ntff_time(fdtd_grid *g, int which, double *host);
/* zero fields, aux fields and NTFF accumulators */
int fdtd_hip_grid_reset(fdtd_grid *g);
/* frequency-domain far field of the current fields (360 complex, re/im pairs):
 * TM: ntffTM_Frequency semantics on the physical cells (ntffTM.c:72-158);
 * TE: single-rank ntffTE_FrequencySplit semantics (ntffTE.c:341-470).
 * coef = sqrt(i k/(8 pi R0)) exp(i k R0) is passed in as (re,im). */
int fdtd_hip_grid_far_field(fdtd_grid *g, double k, double coef_re, double coef_im, double scale, double *out360);
/* time-domain far field + spectra (ntffTM.c:161-233, ntffTE.c:20-55,160-196):
 * translation coefficient (re,im), 4096 FFT twiddles e^{i pi k/4096} (re,im
 * pairs), h_u_nm; out[321*360] lambda-major. */
int fdtd_hip_grid_time_spectra(fdtd_grid *g, double coef_re, double coef_im, const double *twiddles,
                               int h_u_nm, double *out);
/* running DFT of the NTFF contour at nf angular frequencies (rad per step):
 * every later step adds its contour samples times e^{i w tb} (tb = time - 1,
 * the binning's base time) into FP64 per-cell accumulators; nf = 0 turns it off.
 * Without the per-step time-domain NTFF the samples of up to 128 consecutive
 * steps are staged on the device and added in one launch (get_dft and reset
 * take the staged steps into account).  No counterpart in the reference
 * (ntffTM.c:279-371 bins in the time domain). */
int fdtd_hip_grid_set_dft(fdtd_grid *g, const double *omega, int nf);
/* out[nf][3][360] complex: sum_m U_a[m] e^{i w m} over the rows of the three
 * time-domain accumulators as the binning would fill them */
int fdtd_hip_grid_get_dft(fdtd_grid *g, double *out);
/* non-finite values (NaN or Inf in either part) of the three physical fields
 * (TM Ez Hx Hy / TE Ex Ey Hz) of the current state, after the queued steps;
 * no counterpart in the reference, whose runs go on with NaN fields */
int fdtd_hip_grid_count_nonfinite(fdtd_grid *g, long long out[3]);
/* dimensions: out[0..3] = NPXa, NPYa, off, array_size */
int fdtd_hip_grid_dims(const fdtd_grid *g, int out[4]);
int fdtd_hip_grid_sync(fdtd_grid *g);

/* FP32 grids: temporal block depth T (steps per fused launch, 1..15) and rows
 * per wavefront R (4, 8, 16, 18, 20) of the interior tiles, which are 4R rows
 * high (R = 18 and 20 only with the 8-wavefront interior shapes, *_W8 and *_P16;
 * a launch in another shape fails with FDTD_EINVAL).  Steps are queued by
 * fdtd_hip_grid_step and launched T at a time; every read of the state
 * (get_field, far_field, ...) first launches the queued steps. */
int fdtd_hip_grid_set_blocking(fdtd_grid *g, int T, int R);
int fdtd_hip_grid_get_blocking(const fdtd_grid *g, int *T, int *R);
/* FP32 grids: how a block's UPML and interior tiles are launched.  AUTO picks
 * SPLIT for large grids (over 4096 interior workgroups, 8 per UPML one), else
 * MERGED; MERGED is one launch of both
 * (one register budget, the UPML body's); SPLIT is two launches in a row, each
 * at its own occupancy; TWO_STREAMS runs the two launches side by side on two
 * streams. */
enum {
  FAST_LAUNCH_AUTO = 0,
  FAST_LAUNCH_MERGED = 1,
  FAST_LAUNCH_SPLIT = 2,
  FAST_LAUNCH_TWO_STREAMS = 3,
  /* SPLIT / TWO_STREAMS with the wide UPML tiles: blocks of 48 rows (27 core
   * rows at T = 10, against 11 for the 32-row blocks of the other shapes), as
   * 12 wavefronts of 4 rows (TM) or 8 of 6 (TE).  (Named for the first form of
   * this body, 8 wavefronts of 4 rows of a 32-row block.) */
  FAST_LAUNCH_SPLIT_PML8 = 4,
  FAST_LAUNCH_TWO_STREAMS_PML8 = 5,
  /* SPLIT_PML8 / TWO_STREAMS_PML8 with the interior tiles (4R rows) as 8
   * wavefronts of R/2 rows: twice the wavefronts per SIMD (R = 8 or 16; with
   * R = 4 the interior tiles keep 4 wavefronts) */
  FAST_LAUNCH_SPLIT_W8 = 6,
  FAST_LAUNCH_TWO_STREAMS_W8 = 7,
  /* SPLIT_W8 / TWO_STREAMS_W8 with the wide UPML tiles as 16 wavefronts of 3
   * rows (1024 threads, four wavefronts per SIMD): for grids whose UPML band has
   * fewer tiles than the chip has CUs.  Single grids only (not batches). */
  FAST_LAUNCH_SPLIT_P16 = 8,
  FAST_LAUNCH_TWO_STREAMS_P16 = 9
};
int fdtd_hip_grid_set_launch_mode(fdtd_grid *g, int mode);
/* FP32 grids with contour samples: the sample packing, binning and running DFT
 * of a block run beside the next block on a stream of their own (1) or after
 * it on the step stream (0); -1 (default) chooses by size: beside it on grids
 * up to 8 Mi cells, whose step kernel leaves the GPU partly idle */
int fdtd_hip_grid_set_ntff_overlap(fdtd_grid *g, int on);
/* the setting; with AUTO, the shape of the last launch (AUTO before the first) */
int fdtd_hip_grid_get_launch_mode(const fdtd_grid *g, int *mode);
int fdtd_hip_grid_flush(fdtd_grid *g);
/* event timing of the FP32 fused launches since the last enable (one launch
 * advances the UPML and the interior tiles together by T steps).  on = k > 0
 * brackets every k-th launch with a pair of events (the first launch after the
 * enable is timed); 0 turns it off.  A pair costs the stream a few microseconds,
 * so short launches are sampled rather than all timed.
 * out = {ms of the timed launches, 0 (reserved), timed launches, algorithmic
 *        bytes of their interior tiles, of their UPML tiles, their cell updates} */
void fdtd_hip_profile_enable(int on);
void fdtd_hip_profile_read(double out[6]);
/* the achievable HBM rate (no counterpart in the reference): a streaming
 * device-to-device copy of `bytes` repeated `reps` times with event timing.
 * out = {best GB/s, mean GB/s (bytes read + written), mean us per launch,
 *        bytes moved per launch} */
int fdtd_hip_copy_rate(size_t bytes, int reps, double out[4]);

/* Batched instances (no counterpart in the reference, whose solvers hold one
 * instance per process; its sweeps run one instance per MPI rank, main.c:114-138).
 * Several FP32, MPI-flavour grids without the per-step NTFF advance together:
 * each block of T steps is one launch per mode whose workgroups pull the tiles
 * of every unfinished instance from device queues, and the loop of blocks is a
 * replayed hipGraph.  Instance i stops after steps[i] steps in total; p0[i]
 * gives its wave (omega, k, angle_deg; the time and ray_coef fields are not
 * used: the device forms every step's source from omega and the step count).
 * The grids stay owned by the caller and must not be used between
 * fdtd_hip_batch_run and fdtd_hip_batch_sync, which hands each grid its
 * progress back. */
typedef struct fdtd_hip_batch fdtd_hip_batch;
int fdtd_hip_batch_create(fdtd_grid *const *grids, const fdtd_step_params *p0, const int *steps, int n, int T, int R,
                          fdtd_hip_batch **out);
/* blocks >= 0: queue that many blocks (asynchronous); blocks < 0: run every
 * instance to its step count and synchronise */
int fdtd_hip_batch_run(fdtd_hip_batch *b, int blocks);
/* the interior tiles of a block are launched, then the UPML tiles
 * (FAST_LAUNCH_SPLIT; AUTO and MERGED mean the same here), or the two side by
 * side on two streams (FAST_LAUNCH_TWO_STREAMS); the *_PML8 shapes run the
 * UPML tiles as 8 wavefronts, the *_W8 shapes the interior tiles as well
 * (R >= 8).  Default: FAST_LAUNCH_TWO_STREAMS_PML8. */
int fdtd_hip_batch_set_launch_mode(fdtd_hip_batch *b, int mode);
int fdtd_hip_batch_sync(fdtd_hip_batch *b, int *steps_done);
int fdtd_hip_batch_destroy(fdtd_hip_batch *b);

#ifdef __cplusplus
}
#endif
#endif
