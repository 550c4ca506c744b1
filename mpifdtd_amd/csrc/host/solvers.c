/*
 * solvers.c -- the four UPML solvers' and the two split-field PML solvers'
 * entry points on top of the HIP grid.
 *
 * Each solver keeps the reference's lifecycle (init -> update* -> reset /
 * finish) and getter ABI:
 *   MPI TM  mpiTM_UPML.c:88-250      MPI TE  mpiTE_UPML.c:72-335
 *   TM      fdtdTM_upml.c:54-152     TE      fdtdTE_upml.c:101-192
 *   split-field PML TM  fdtdTM.c:1-327, TE  fdtdTE.c:1-321 (FP64 only)
 * but the state lives in device memory.  setCoefficient (mpiTM_UPML.c:663-716,
 * mpiTE_UPML.c:489-536, fdtdTM_upml.c:224-274, fdtdTE_upml.c:361-412) is
 * restated as 1-D tables per axis, because every UPML coefficient depends on
 * sigma_x(x) or sigma_y(y) alone except the two built on the device from both.
 * Errors print the HIP message and exit(2), as the reference does on failure.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include "fdtd_hip.h"
#include "mpifdtd_field.h"
#include "mpifdtd_models.h"
#include "mpifdtd_simulator.h"
#include "host_internal.h"
#include "mpifdtd_solvers.h"

#ifndef M_PI
#define M_PI 3.1415926535897932384626433832795
#endif


typedef struct {
  fdtd_grid *g;
  int NPXa, NPYa, off, arraySize;
  double complex *mirror[3];  /* TM: Ez Hx Hy   TE: Ex Ey Hz */
  double complex *split_mirror[2];  /* split-field and NS TM: Ezx Ezy   TE: Hzx Hzy */
  double *eps_mirror;
  double *eps_maps[3];  /* the NS solvers' getEpsX / Y / Z: one host copy per map */
  int blocked;    /* FP32 grid: steps are queued and run T per launch */
  int ntff_time;  /* the per-step time-domain NTFF runs (its outputs exist) */
  int ndft;       /* frequencies of the running DFT (fdtd_solver_setDFT) */
} solver_t;

static solver_t slv_[SLV_N];
static int active_ = -1;
static int ntff_time_ = 1;
static int file_output_ = 1;
static int precision_ = FDTD_FP64;
static int block_T_ = 0, block_R_ = 16;  /* 0: default depth */
/* non-finite field check (fdtd_solver_setFiniteCheck): every finite_every_ steps,
   0 off; the first detection since init / reset (step time and count) */
static int finite_every_ = 0, finite_abort_ = 0, finite_since_ = 0;
static int finite_step_ = -1;
static long long finite_count_ = 0;

static const int slot_of_[3] = {0, 3, 6};

static void die(const char *what)
{
  fprintf(stderr, "mpifdtd_amd: %s failed: %s\n", what, fdtd_hip_last_error());
  exit(2);
}
#define CHECK(call) do { if ((call) != FDTD_OK) die(#call); } while (0)

static int is_tm(int kind) { return kind == SLV_MPI_TM || kind == SLV_TM || kind == SLV_SPLIT_TM || kind == SLV_NS_TM; }
static int is_mpi(int kind) { return kind == SLV_MPI_TM || kind == SLV_MPI_TE; }
/* the solvers with split components (FP64, no NTFF, elliptic output on reset):
   the split-field PML TM / TE and the non-standard FDTD TM / TE */
static int is_split(int kind)
{
  return kind == SLV_SPLIT_TM || kind == SLV_SPLIT_TE || kind == SLV_NS_TM || kind == SLV_NS_TE;
}
static int is_ns(int kind) { return kind == SLV_NS_TM || kind == SLV_NS_TE; }

void fdtd_solver_setNTFFTime(int on) { ntff_time_ = on ? 1 : 0; }
void fdtd_solver_setFileOutput(int on) { file_output_ = on ? 1 : 0; }
void fdtd_solver_setPrecision(int precision) { precision_ = precision == FDTD_FP32 ? FDTD_FP32 : FDTD_FP64; }
int fdtd_solver_getPrecision(void) { return precision_; }
int fdtd_solver_setBlocking(int T, int R)
{
  block_T_ = T;
  block_R_ = R;
  /* the active solver's own precision decides, not the setting for the next init */
  if (active_ >= 0 && slv_[active_].blocked) return fdtd_hip_grid_set_blocking(slv_[active_].g, T, R);
  return FDTD_OK;
}

int fdtd_solver_setLaunchMode(int mode)
{
  if (active_ < 0 || !slv_[active_].blocked) return FDTD_ESTATE;
  return fdtd_hip_grid_set_launch_mode(slv_[active_].g, mode);
}

int fdtd_solver_setNTFFOverlap(int on)
{
  if (active_ < 0 || !slv_[active_].blocked) return FDTD_ESTATE;
  return fdtd_hip_grid_set_ntff_overlap(slv_[active_].g, on);
}

int fdtd_solver_getLaunchMode(int *mode)
{
  if (active_ < 0 || !slv_[active_].blocked) return FDTD_ESTATE;
  return fdtd_hip_grid_get_launch_mode(slv_[active_].g, mode);
}

int fdtd_solver_getBlocking(int *T, int *R)
{
  if (active_ < 0 || !slv_[active_].blocked) return FDTD_ESTATE;
  return fdtd_hip_grid_get_blocking(slv_[active_].g, T, R);
}

int fdtd_solver_checkFinite(long long out[3])
{
  if (!out) return FDTD_EINVAL;
  if (active_ < 0 || !slv_[active_].g) return FDTD_ESTATE;
  return fdtd_hip_grid_count_nonfinite(slv_[active_].g, out);
}

int fdtd_solver_setFiniteCheck(int every, int abort_on_detect)
{
  if (every < 0) return FDTD_EINVAL;
  finite_every_ = every;
  finite_abort_ = abort_on_detect ? 1 : 0;
  finite_since_ = 0;
  return FDTD_OK;
}

int fdtd_solver_finiteStatus(int *step, long long *count)
{
  if (step) *step = finite_step_;
  if (count) *count = finite_count_;
  return finite_step_ >= 0 ? 1 : 0;
}

/* the simplified beta and coef1 of nsFdtdTM.c:221-229 */
static double ns_beta(double alpha) { return tanh(alpha) / (1 + pow(tanh(alpha), 2)); }
static double ns_coef1(double beta) { return (1 - beta) / (1 + beta); }

/* the factor 1/(_n n) - 1 of field_nsScatteredWaveNotUPML (field.c:155-177),
   exactly 0 where n = 1 */
static double ns_wave_factor(double eps, double eps0, double w_s, double k_s)
{
  const double nn = sqrt(eps / eps0);
  const double u0 = sin(w_s * 0.5) / sin(k_s * 0.5), u1 = sin(w_s * 0.5) / sin(nn * k_s * 0.5);
  const double _n = u0 / u1;
  return 1.0 / (_n * nn) - 1.0;
}

/* nsFdtdTE.c:100-181: all ten TE maps from EPS_EX, EPS_HZ and EPS_EY, over the
   interior cells the reference fills (the border stays zero) */
static void ns_te_cells(const double *eps, double *cell, size_t n)
{
  const int nx = N_PX, ny = N_PY;
  const double R = 1.0e-8, M = 2.0, eps0 = 1.0, C = field_getLightSpeed(), mu0 = field_getMu0();
  const double sig_max = -(M + 1.0) * eps0 * C / N_PML * log(R);
  const double w_s = field_getOmega(), k_s = field_getK();
  memset(cell, 0, 10 * n * sizeof(double));
  for (int i = 1; i < nx - 1; i++) {
    for (int j = 1; j < ny - 1; j++) {
      const size_t k = (size_t)i * ny + j;
      const double e_ex = eps[k], e_ey = eps[n + k], e_hz = eps[2 * n + k];
      const double sig_ex_y = sig_max * field_sigmaY(i + 0.5, j), sig_ey_x = sig_max * field_sigmaX(i, j + 0.5);
      const double sig_hz_x = sig_max * field_sigmaX(i + 0.5, j + 0.5), sig_hz_y = sig_max * field_sigmaY(i + 0.5, j + 0.5);
      const double b_ex_y = ns_beta(sig_ex_y / (2 * e_ex)), b_ey_x = ns_beta(sig_ey_x / (2 * e_ey));
      const double b_hz_x = ns_beta(sig_hz_x / (2 * e_hz)), b_hz_y = ns_beta(sig_hz_y / (2 * e_hz));
      const double z_hz = sqrt(mu0 / e_hz), n_hz = sqrt(e_hz / eps0);
      const double u_hz = sin(w_s * 0.5) / sin(k_s * n_hz * 0.5);
      cell[4 * n + k] = ns_coef1(b_hz_x);                  /* C_HZX */
      cell[5 * n + k] = u_hz / z_hz / (1.0 + b_hz_x);      /* C_HZXLX */
      cell[6 * n + k] = ns_coef1(b_hz_y);                  /* C_HZY */
      cell[7 * n + k] = u_hz / z_hz / (1.0 + b_hz_y);      /* C_HZYLY */
      const double z_ex = sqrt(mu0 / e_ex), n_ex = sqrt(e_ex / eps0);
      const double u_ex = sin(w_s * 0.5) / sin(k_s * n_ex * 0.5);
      cell[k] = ns_coef1(b_ex_y);                          /* C_EX */
      cell[n + k] = u_ex * z_ex / (1.0 + b_ex_y);          /* C_EXLY */
      const double z_ey = sqrt(mu0 / e_ey), n_ey = sqrt(e_ey / eps0);
      const double u_ey = sin(w_s * 0.5) / sin(k_s * n_ey * 0.5);
      cell[2 * n + k] = ns_coef1(b_ey_x);                  /* C_EY */
      cell[3 * n + k] = u_ey * z_ey / (1.0 + b_ey_x);      /* C_EYLX */
      cell[8 * n + k] = ns_wave_factor(e_ex, eps0, w_s, k_s);
      cell[9 * n + k] = ns_wave_factor(e_ey, eps0, w_s, k_s);
    }
  }
}

/* nsFdtdTM.c:226-306 / nsFdtdTE.c:100-181 and field.c:155-177: the per-cell
   coefficients of the NS solvers from the eps maps the device filled (TM EPS_EZ,
   EPS_HX, EPS_HY; TE EPS_EX, EPS_EY, EPS_HZ), each formed with the reference's
   expression and operation order */
static int build_ns_coefs(int kind, fdtd_grid *g)
{
  const int nx = N_PX, ny = N_PY;
  const size_t n = (size_t)nx * ny;
  const int nmaps = is_tm(kind) ? 5 : 10;
  double *eps = (double *)malloc(3 * n * sizeof(double)), *cell = (double *)malloc(nmaps * n * sizeof(double));
  if (!eps || !cell) {
    free(eps);
    free(cell);
    return FDTD_ENOMEM;
  }
  int rc = FDTD_OK;
  for (int q = 0; q < 3 && rc == FDTD_OK; q++) rc = fdtd_hip_grid_get_eps(g, q, eps + q * n);
  const double R = 1.0e-8, M = 2.0, eps0 = 1.0, C = field_getLightSpeed(), mu0 = field_getMu0();
  const double sig_max = -(M + 1.0) * eps0 * C / N_PML * log(R);
  const double w_s = field_getOmega(), k_s = field_getK();
  if (rc == FDTD_OK && !is_tm(kind)) ns_te_cells(eps, cell, n);
  for (int i = 0; i < nx && rc == FDTD_OK && is_tm(kind); i++) {
    const double b_ez_x = ns_beta(sig_max * field_sigmaX(i, 0) / (2 * eps0));
    const double b_hy_x = ns_beta(sig_max * field_sigmaX(i + 0.5, 0) / (2 * eps0));
    for (int j = 0; j < ny; j++) {
      const size_t k = (size_t)i * ny + j;
      const double b_hx_y = ns_beta(sig_max * field_sigmaY(0, j + 0.5) / (2 * eps0));
      const double e_ez = eps[k], e_hx = eps[n + k], e_hy = eps[2 * n + k];
      const double z_ez = sqrt(mu0 / e_ez), n_ez = sqrt(e_ez / eps0);
      const double u_ez = sin(w_s * 0.5) / sin(k_s * n_ez * 0.5);
      cell[k] = u_ez * z_ez / (1 + b_ez_x);          /* C_EZXLX */
      cell[n + k] = u_ez * z_ez / (1.0 + b_ez_x);    /* C_EZYLY: b_ez_x, as nsFdtdTM.c:282 has it */
      const double z_hx = sqrt(mu0 / e_hx), n_hx = sqrt(e_hx / eps0);
      const double u_hx = sin(w_s * 0.5) / sin(k_s * n_hx * 0.5);
      cell[2 * n + k] = u_hx / z_hx / (1.0 + b_hx_y);  /* C_HXLY */
      const double z_hy = sqrt(mu0 / e_hy), n_hy = sqrt(e_hy / eps0);
      const double u_hy = sin(w_s * 0.5) / sin(k_s * n_hy * 0.5);
      cell[3 * n + k] = u_hy / z_hy / (1.0 + b_hy_x);  /* C_HYLX */
      cell[4 * n + k] = ns_wave_factor(e_ez, eps0, w_s, k_s);
    }
  }
  const double r = 1.0 / 6.0 + k_s * k_s / 180.0 - pow(k_s, 4) / 23040;  /* nsFdtdTM.c:117, nsFdtdTE.c:255 */
  if (rc == FDTD_OK) rc = fdtd_hip_grid_set_ns_coefs(g, cell, r / 2.0);
  free(eps);
  free(cell);
  return rc;
}

static void build_coefs(int kind, fdtd_grid *g)
{
  const int nx = N_PX, ny = N_PY;
  const double R = 1.0e-8, M = 2.0, eps = 1.0, C = field_getLightSpeed();
  double *buf = (double *)calloc((size_t)6 * nx + (size_t)5 * ny, sizeof(double));
  double *X[5], *SX, *Y[4], *SY;
  for (int q = 0; q < 5; q++) X[q] = buf + (size_t)q * nx;
  SX = buf + (size_t)5 * nx;
  for (int q = 0; q < 4; q++) Y[q] = buf + (size_t)6 * nx + (size_t)q * ny;
  SY = buf + (size_t)6 * nx + (size_t)4 * ny;
  const double sig_z = 0;
  if (is_ns(kind)) {
    /* nsFdtdTM.c:226-306: the TM decay factors coef1 of the simplified beta, which
       depend on one axis only; the eps-dependent ones, and every TE one (its alpha
       divides by the cell's eps), follow in build_ns_coefs */
    const double sig_max = -(M + 1.0) * eps * C / N_PML * log(R);
    for (int x = 0; x < nx && is_tm(kind); x++) {
      X[0][x] = ns_coef1(ns_beta(sig_max * field_sigmaX(x, 0) / (2 * eps)));        /* C_EZX */
      X[1][x] = ns_coef1(ns_beta(sig_max * field_sigmaX(x + 0.5, 0) / (2 * eps)));  /* C_HY */
    }
    for (int y = 0; y < ny && is_tm(kind); y++) {
      Y[0][y] = ns_coef1(ns_beta(sig_max * field_sigmaY(0, y) / (2 * eps)));        /* C_EZY */
      Y[1][y] = ns_coef1(ns_beta(sig_max * field_sigmaY(0, y + 0.5) / (2 * eps)));  /* C_HX */
    }
  } else if (is_split(kind)) {
    /* fdtdTM.c:184-243, fdtdTE.c:200-241: the conductivities (the same four tables
       for TM and TE); the coefficients are formed per cell on the device with eps
       (field_pmlCoef / field_pmlCoef_LXY of the component's eps map and MU_0_S) */
    const double sig_max = -(M + 1.0) * eps * C / 2.0 / N_PML * log(R);
    const double mu_eps = field_getMu0() / eps;  /* MU_0_S / EPSILON_0_S */
    for (int x = 0; x < nx; x++) {
      X[0][x] = sig_max * field_sigmaX(x, 0);                 /* TM sig_ez_x, TE sig_ey_x */
      X[1][x] = mu_eps * (sig_max * field_sigmaX(x + 0.5, 0)); /* TM sig_hy_xx, TE sig_hz_xx */
    }
    for (int y = 0; y < ny; y++) {
      Y[0][y] = sig_max * field_sigmaY(0, y);                 /* TM sig_ez_y, TE sig_ex_y */
      Y[1][y] = mu_eps * (sig_max * field_sigmaY(0, y + 0.5)); /* TM sig_hx_yy, TE sig_hz_yy */
    }
  } else if (is_tm(kind)) {
    const double sig_max = -(M + 1.0) * eps * C / 2.0 / N_PML / cos(M_PI / 3) * log(R);
    for (int x = 0; x < nx; x++) {
      double sig_ez_x = sig_max * field_sigmaX(x, 0);
      double sig_hx_x = sig_max * field_sigmaX(x, 0.5);
      double sig_hy_x = sig_max * field_sigmaX(x + 0.5, 0);
      X[0][x] = (2 * eps - sig_ez_x) / (2 * eps + sig_ez_x);   /* C_JZ */
      X[1][x] = (2 * eps) / (2 * eps + sig_ez_x);              /* C_JZHXHY */
      X[2][x] = (2 * eps + sig_hx_x) / (2 * eps + sig_z);      /* C_BXMX1 */
      X[3][x] = (2 * eps - sig_hx_x) / (2 * eps + sig_z);      /* C_BXMX0 */
      X[4][x] = (2 * eps - sig_hy_x) / (2 * eps + sig_hy_x);   /* C_BY */
      SX[x] = sig_hy_x;
    }
    for (int y = 0; y < ny; y++) {
      double sig_ez_y = sig_max * field_sigmaY(0, y);
      double sig_hx_y = sig_max * field_sigmaY(0, y + 0.5);
      double sig_hy_y = sig_max * field_sigmaY(0.5, y);
      Y[0][y] = (2 * eps - sig_ez_y) / (2 * eps + sig_ez_y);   /* C_DZ */
      Y[1][y] = (2 * eps + sig_z) / (2 * eps + sig_ez_y);      /* C_DZJZ1 = C_DZJZ0 */
      Y[2][y] = (2 * eps - sig_hx_y) / (2 * eps + sig_hx_y);   /* C_MX */
      Y[3][y] = (2 * eps) / (2 * eps + sig_hx_y);              /* C_MXEZ */
      SY[y] = sig_hy_y;
    }
  } else {
    const double sig_max = -(M + 1.0) * eps * C / 2.0 / N_PML * log(R);
    for (int x = 0; x < nx; x++) {
      double sig_ex_x = sig_max * field_sigmaX(x + 0.5, 0);
      double sig_ey_x = sig_max * field_sigmaX(x, 0.5);
      double sig_hz_x = sig_max * field_sigmaX(x + 0.5, 0.5);
      X[0][x] = (2 * eps + sig_ex_x) / (2 * eps + sig_z);      /* C_DXJX1 */
      X[1][x] = (2 * eps - sig_ex_x) / (2 * eps + sig_z);      /* C_DXJX0 */
      X[2][x] = (2 * eps - sig_ey_x) / (2 * eps + sig_ey_x);   /* C_DY */
      X[3][x] = (2 * eps - sig_hz_x) / (2 * eps + sig_hz_x);   /* C_MZ */
      X[4][x] = (2 * eps) / (2 * eps + sig_hz_x);              /* C_MZEXEY */
      SX[x] = sig_ey_x;
    }
    for (int y = 0; y < ny; y++) {
      double sig_ex_y = sig_max * field_sigmaY(0.5, y);
      double sig_ey_y = sig_max * field_sigmaY(0, y + 0.5);
      double sig_hz_y = sig_max * field_sigmaY(0.5, y + 0.5);
      Y[0][y] = (2 * eps - sig_ex_y) / (2 * eps + sig_ex_y);   /* C_JX */
      Y[1][y] = (2 * eps) / (2 * eps + sig_ex_y);              /* C_JXHZ */
      Y[2][y] = (2 * eps - sig_hz_y) / (2 * eps + sig_hz_y);   /* C_BZ */
      Y[3][y] = (2 * eps + sig_z) / (2 * eps + sig_hz_y);      /* C_BZMZ1 = C_BZMZ0 */
      SY[y] = sig_ey_y;
    }
  }
  fdtd_coef_tables t;
  for (int q = 0; q < 5; q++) t.x[q] = X[q];
  t.sx = SX;
  for (int q = 0; q < 4; q++) t.y[q] = Y[q];
  t.sy = SY;
  CHECK(fdtd_hip_grid_set_coefs(g, &t));
  free(buf);
}

static void solver_free(int kind)
{
  solver_t *s = &slv_[kind];
  if (s->g) fdtd_hip_grid_destroy(s->g);
  for (int q = 0; q < 3; q++) free(s->mirror[q]);
  for (int q = 0; q < 2; q++) free(s->split_mirror[q]);
  free(s->eps_mirror);
  for (int q = 0; q < 3; q++) free(s->eps_maps[q]);
  memset(s, 0, sizeof *s);
  if (active_ == kind) active_ = -1;
}

/* The grid of solver `kind` for the current field/model globals: descriptor,
 * UPML coefficient tables, eps fill and, for FP32 without the per-step NTFF,
 * the temporal blocking.  Shared by solver init and the batch builder. */
int solver_make_grid(int kind, int precision, int ntff_time, int T, int R, fdtd_grid **out)
{
  FieldInfo_S fs = field_getFieldInfo_S();
  NTFFInfo ni = field_getNTFFInfo();
  fdtd_grid_desc d;
  memset(&d, 0, sizeof d);
  d.mode = is_tm(kind) ? FDTD_MODE_TM : FDTD_MODE_TE;
  d.flavour = is_ns(kind) ? FDTD_FLAVOUR_NS
                          : (is_split(kind) ? FDTD_FLAVOUR_SPLIT : (is_mpi(kind) ? FDTD_FLAVOUR_MPI : FDTD_FLAVOUR_PLAIN));
  if (is_split(kind)) {  /* FP64 only; the reference's fdtdTM.c / fdtdTE.c / nsFdtdTM.c have no NTFF */
    precision = FDTD_FP64;
    ntff_time = 0;
  }
  d.precision = precision;
  d.N_PX = fs.N_PX;
  d.N_PY = fs.N_PY;
  d.N_PML = fs.N_PML;
  d.ntff_top = ni.top;
  d.ntff_bottom = ni.bottom;
  d.ntff_left = ni.left;
  d.ntff_right = ni.right;
  d.ntff_cx = ni.cx;
  d.ntff_cy = ni.cy;
  d.rf_per_c = ni.RFperC;
  d.array_size = ni.arraySize;
  d.max_time = (int)field_getMaxTime();
  d.c0 = field_getLightSpeed();
  d.mu0 = field_getMu0();
  d.z0 = field_getZ0();
  d.eps0 = 1.0;
  d.ntff_time = ntff_time;
  d.point_source = field_getPointSource();
  fdtd_grid *g = NULL;
  int rc = fdtd_hip_grid_create(&d, &g);
  if (rc != FDTD_OK) return rc;
  build_coefs(kind, g);
  if ((rc = fdtd_hip_grid_fill_eps(g, models_getState())) == FDTD_OK && precision == FDTD_FP32)
    rc = fdtd_hip_grid_set_blocking(g, T > 0 ? T : 4, R);
  if (rc == FDTD_OK && is_ns(kind)) rc = build_ns_coefs(kind, g);
  if (rc != FDTD_OK) {
    fdtd_hip_grid_destroy(g);
    return rc;
  }
  *out = g;
  return FDTD_OK;
}

static void solver_init(int kind)
{
  solver_t *s = &slv_[kind];
  if (s->g) solver_free(kind);
  CHECK(solver_make_grid(kind, precision_, ntff_time_, block_T_, block_R_, &s->g));
  const int split = is_split(kind);
  int dims[4];
  CHECK(fdtd_hip_grid_dims(s->g, dims));
  s->NPXa = dims[0];
  s->NPYa = dims[1];
  s->off = dims[2];
  s->arraySize = dims[3];
  s->blocked = !split && precision_ == FDTD_FP32;
  s->ntff_time = split ? 0 : ntff_time_;
  s->ndft = 0;
  size_t n = (size_t)s->NPXa * s->NPYa;
  for (int q = 0; q < 3; q++) s->mirror[q] = (double complex *)calloc(n, sizeof(double complex));
  for (int q = 0; q < 2 && split; q++) s->split_mirror[q] = (double complex *)calloc(n, sizeof(double complex));
  s->eps_mirror = (double *)calloc(n, sizeof(double));
  active_ = kind;
  finite_since_ = 0;
  finite_step_ = -1;
  finite_count_ = 0;
}

static void solver_update(int kind)
{
  solver_t *s = &slv_[kind];
  if (!s->g) { fprintf(stderr, "mpifdtd_amd: update before init\n"); exit(2); }
  fdtd_step_params p;
  memset(&p, 0, sizeof p);
  p.time = field_getTime();
  p.ray_coef = field_getRayCoef();
  p.omega = field_getOmega();
  p.k = field_getK();
  p.angle_deg = field_getWaveAngle();
  fdtd_dcomplex ps = field_pointLight();
  p.ps_re = creal(ps);
  p.ps_im = cimag(ps);
  if (kind == SLV_TE || kind == SLV_NS_TE) {
    /* the pulse / wave is projected on the direction 90 deg from the wave vector
       (fdtdTE_upml.c:182-189, nsFdtdTE.c:250-256) */
    WaveInfo_S w = field_getWaveInfo_S();
    double co = cos((w.Angle_deg + 90) * M_PI / 180.0);
    double si = sin((w.Angle_deg + 90) * M_PI / 180.0);
    p.dot_x = co;
    p.dot_y = si;
    p.use_x = co != 0.0;
    p.use_y = si != 0.0;
  }
  CHECK(fdtd_hip_grid_step(s->g, &p));
  if (finite_every_ > 0 && ++finite_since_ >= finite_every_) {
    finite_since_ = 0;
    long long c[3];
    CHECK(fdtd_hip_grid_count_nonfinite(s->g, c));
    const long long n = c[0] + c[1] + c[2];
    if (n > 0 && finite_step_ < 0) {
      finite_step_ = (int)p.time;
      finite_count_ = n;
      printf("mpifdtd_amd: %lld non-finite field values after the step of time %d\n", n, finite_step_);
      fflush(stdout);
      if (finite_abort_) exit(2);
    }
  }
}

static double complex *solver_field(int kind, int q)
{
  solver_t *s = &slv_[kind];
  if (!s->g) return NULL;
  CHECK(fdtd_hip_grid_get_field(s->g, slot_of_[q], (double *)s->mirror[q]));
  return s->mirror[q];
}

static double *solver_eps(int kind)
{
  solver_t *s = &slv_[kind];
  if (!s->g) return NULL;
  /* MPI TM: EPS_EZ, MPI TE: EPS_EY, plain TM: EPS_EZ, plain TE: EPS_EX, split TM: EPS_EZ,
     split TE: EPS_EY (fdtdTE.c:62-65) */
  int which = (kind == SLV_MPI_TE || kind == SLV_SPLIT_TE || kind == SLV_NS_TE) ? 1 : 0;
  CHECK(fdtd_hip_grid_get_eps(s->g, which, s->eps_mirror));
  return s->eps_mirror;
}

/* ---------------------------------------------------------------- NTFF outputs */

double complex solver_ff_coef(void)
{
  /* ntffTM_Frequency / ntffTE_FrequencySplit with R0 of ntffTM_init / ntffTE_init */
  const double R0 = 1.0e6 * field_toCellUnit(500);
  const double k_s = field_getK();
  return csqrt(I * k_s / (8 * M_PI * R0)) * cexp(I * k_s * R0);
}

int fdtd_solver_farField(double *out360)
{
  if (active_ < 0) return FDTD_ESTATE;
  double complex coef = solver_ff_coef();
  FieldInfo fi = field_getFieldInfo();
  return fdtd_hip_grid_far_field(slv_[active_].g, field_getK(), creal(coef), cimag(coef), sqrt(fi.h_u_nm), out360);
}

int fdtd_solver_farFieldNoAlloc(double *out360)
{
  if (active_ < 0) return FDTD_ESTATE;
  double complex coef = solver_ff_coef();
  FieldInfo fi = field_getFieldInfo();
  return fdtd_hip_grid_far_field_noalloc(slv_[active_].g, field_getK(), creal(coef), cimag(coef), sqrt(fi.h_u_nm),
                                         out360);
}

static void twiddles(double *tw)
{
  /* e^{+i pi k/4096} with the argument formed as fft1x does (cfft.c:138-141) */
  const double w = M_PI / 4096;
  for (int k = 0; k < 4096; k++) {
    double th = w * k;
    tw[2 * k] = cos(th);
    tw[2 * k + 1] = sin(th);
  }
}

int fdtd_solver_timeSpectra(double *out)
{
  if (active_ < 0) return FDTD_ESTATE;
  const double C = field_getLightSpeed(), w_s = field_getOmega();
  /* ntffTM_TimeTranslate / ntffTE_TimeTranslate (ntffTM.c:165, ntffTE.c:25);
     the MPI solvers' ntffOutput omits 1/(4 pi c) (mpiTM_UPML.c:802, mpiTE_UPML.c:843) */
  double complex coef = is_mpi(active_) ? csqrt(2 * M_PI * C / (I * w_s))
                                        : 1.0 / (4 * M_PI * C) * csqrt(2 * M_PI * C / (I * w_s));
  double tw[8192];
  twiddles(tw);
  FieldInfo fi = field_getFieldInfo();
  return fdtd_hip_grid_time_spectra(slv_[active_].g, creal(coef), cimag(coef), tw, fi.h_u_nm, out);
}

int fdtd_solver_setDFT(const double *lambda_nm, int n)
{
  if (active_ < 0) return FDTD_ESTATE;
  if (n < 0 || (n > 0 && !lambda_nm)) return FDTD_EINVAL;
  /* w_s = C_0_S 2 pi / lambda_s, lambda_s = lambda / h (field.c:89-143) */
  const double C = field_getLightSpeed();
  FieldInfo fi = field_getFieldInfo();
  double *w = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  for (int f = 0; f < n; f++) w[f] = C * 2 * M_PI / (lambda_nm[f] / fi.h_u_nm);
  int rc = fdtd_hip_grid_set_dft(slv_[active_].g, w, n);
  free(w);
  if (rc == FDTD_OK) slv_[active_].ndft = n;
  return rc;
}

int fdtd_solver_getDFT(double *out)
{
  if (active_ < 0) return FDTD_ESTATE;
  return fdtd_hip_grid_get_dft(slv_[active_].g, out);
}

int fdtd_solver_dftSpectra(double *out)
{
  if (active_ < 0) return FDTD_ESTATE;
  const int n = slv_[active_].ndft;
  if (n <= 0) return FDTD_ESTATE;
  double complex *U = (double complex *)malloc((size_t)n * 3 * 360 * sizeof(double complex));
  int rc = fdtd_hip_grid_get_dft(slv_[active_].g, (double *)U);
  if (rc) {
    free(U);
    return rc;
  }
  /* the translation and normalisation of fdtd_solver_timeSpectra, on the DFT of
     the accumulator rows at each frequency instead of the interpolated FFT */
  const double C = field_getLightSpeed(), w_s = field_getOmega(), Z0 = field_getZ0();
  const double complex coef = is_mpi(active_) ? csqrt(2 * M_PI * C / (I * w_s))
                                              : 1.0 / (4 * M_PI * C) * csqrt(2 * M_PI * C / (I * w_s));
  const int tm = is_tm(active_);
  for (int f = 0; f < n; f++)
    for (int ang = 0; ang < 360; ang++) {
      const double phi = ang * M_PI / 180.0, sp = sin(phi), cp = cos(phi);
      const double complex u0 = U[((size_t)f * 3 + 0) * 360 + ang], u1 = U[((size_t)f * 3 + 1) * 360 + ang],
                           u2 = U[((size_t)f * 3 + 2) * 360 + ang];
      double complex e;
      if (tm) /* Eth = coef (-Z0 WTH - UPH), WTH = -Wz, UPH = -Ux sin + Uy cos */
        e = coef * (-Z0 * (-u2) - (-u0 * sp + u1 * cp));
      else /* Eph = coef (-Z0 WPH + UTH), WPH = -Wx sin + Wy cos, UTH = -Uz */
        e = coef * (-Z0 * (-u0 * sp + u1 * cp) + (-u2));
      out[(size_t)f * 360 + ang] = (creal(e) * creal(e) + cimag(e) * cimag(e)) / NTFF_NUM;
    }
  free(U);
  return FDTD_OK;
}

static void write_spectra(void)
{
  double *sp = (double *)malloc(321 * 360 * sizeof(double));
  CHECK(fdtd_solver_timeSpectra(sp));
  double *rows[321];
  for (int l = 0; l < 321; l++) rows[l] = sp + (size_t)l * 360;
  char buf[256];
  snprintf(buf, sizeof buf, "%d[deg].txt", (int)field_getWaveAngle());
  ntff_outputEnormTxt(rows, buf);
  snprintf(buf, sizeof buf, "%d[deg]_%dnm_%dnm_b.dat", (int)field_getWaveAngle(), LAMBDA_ST_NM, LAMBDA_EN_NM);
  ntff_outputEnormBin(rows, buf);
  free(sp);
}

/* ntffOutput of the MPI TE solver (mpiTE_UPML.c:796-879): Eth / Eph time series per angle */
static void mpi_te_ntff_output(void)
{
  solver_t *s = &slv_[SLV_MPI_TE];
  const int maxTime = (int)field_getMaxTime();
  const size_t n = (size_t)360 * s->arraySize;
  double complex *W[3];
  for (int a = 0; a < 3; a++) {
    W[a] = (double complex *)malloc(n * sizeof(double complex));
    CHECK(fdtd_hip_grid_get_ntff_time(s->g, a, (double *)W[a]));
  }
  const double w_s = field_getOmega(), Z0 = field_getZ0();
  const double complex coef = csqrt(2 * M_PI * field_getLightSpeed() / (I * w_s));
  mkdir("MPI_TE_UPML", 0775);
  const char *names[4] = {"MPI_TE_UPML/Eph_r.txt", "MPI_TE_UPML/Eph_i.txt", "MPI_TE_UPML/Eth_r.txt", "MPI_TE_UPML/Eth_i.txt"};
  FILE *fp[4];
  for (int f = 0; f < 4; f++) {
    fp[f] = fopen(names[f], "w");
    if (!fp[f]) { printf("cannot open file %s \n", names[f]); exit(2); }
  }
  for (int ang = 0; ang < 360; ang++) {
    double phi = ang * M_PI / 180.0;
    double sx = cos(phi), sy = sin(phi), sz = -1.0, px = -sin(phi), py = cos(phi);
    size_t k = (size_t)ang * s->arraySize;
    for (int i = 0; i < maxTime; i++) {
      double complex WTH = W[0][k + i] * sx + W[1][k + i] * sy + 0;
      double complex WPH = W[0][k + i] * px + W[1][k + i] * py;
      double complex UTH = 0 + 0 + W[2][k + i] * sz;
      double complex ETH = coef * (-Z0 * WTH - 0);
      double complex EPH = coef * (-Z0 * WPH + UTH);
      fprintf(fp[0], "%.20lf ", creal(EPH));
      fprintf(fp[1], "%.20lf ", cimag(EPH));
      fprintf(fp[2], "%.20lf ", creal(ETH));
      fprintf(fp[3], "%.20lf ", cimag(ETH));
    }
    for (int f = 0; f < 4; f++) fprintf(fp[f], "\n");
  }
  for (int f = 0; f < 4; f++) fclose(fp[f]);
  for (int a = 0; a < 3; a++) free(W[a]);
  /* ntffStrTE.txt: |E_phi|^2 per angle; the reference's split version prints NaN here (R0 unset) */
  double ff[720];
  CHECK(fdtd_solver_farField(ff));
  FILE *f = fopen("ntffStrTE.txt", "w");
  if (f) {
    for (int ang = 0; ang < 360; ang++) fprintf(f, "%.18lf\n", ff[2 * ang] * ff[2 * ang] + ff[2 * ang + 1] * ff[2 * ang + 1]);
    fclose(f);
  }
}

/* the split-field and NS solvers' reset output (fdtdTM.c:167-173, fdtdTE.c:269-274,
   nsFdtdTM.c:160-164, nsFdtdTE.c:231-235): field_outputElliptic of Ez (TM) or Ey (TE) */
static void write_elliptic(int kind)
{
  FieldInfo fi = field_getFieldInfo();
  char name[128];
  const char *prefix = kind == SLV_NS_TM ? "ns_tm" : kind == SLV_NS_TE ? "ns_te" : kind == SLV_SPLIT_TM ? "tm" : "te";
  snprintf(name, sizeof name, "%s_%dnm.txt", prefix, fi.h_u_nm);
  field_outputElliptic(name, (fdtd_dcomplex *)solver_field(kind, is_tm(kind) ? 0 : 1));
}

static void solver_reset(int kind)
{
  solver_t *s = &slv_[kind];
  if (!s->g) return;
  finite_since_ = 0;
  finite_step_ = -1;
  finite_count_ = 0;
  if (is_split(kind) && file_output_) write_elliptic(kind);  /* fdtdTM.c:167-173, fdtdTE.c:269-274, nsFdtdTM.c:160-164 */
  if (!is_mpi(kind) && file_output_ && s->ntff_time) write_spectra();  /* fdtdTM_upml.c:87, fdtdTE_upml.c:117 */
  CHECK(fdtd_hip_grid_reset(s->g));
}

static void solver_finish(int kind)
{
  solver_t *s = &slv_[kind];
  if (!s->g) return;
  if (!is_mpi(kind)) solver_reset(kind);                               /* fdtdTM_upml.c:78-82 */
  else if (kind == SLV_MPI_TE && file_output_ && s->ntff_time) mpi_te_ntff_output();  /* mpiTE_UPML.c:306-318 */
  solver_free(kind);
}

/* ---------------------------------------------------------------- extensions */
int fdtd_solver_getState(int which, double *out)
{
  if (active_ < 0) return FDTD_ESTATE;
  return fdtd_hip_grid_get_field(slv_[active_].g, which, out);
}
int fdtd_solver_setState(int which, const double *in)
{
  if (active_ < 0) return FDTD_ESTATE;
  return fdtd_hip_grid_set_field(slv_[active_].g, which, in);
}
int fdtd_solver_getNTFFTime(int which, double *out)
{
  if (active_ < 0) return FDTD_ESTATE;
  return fdtd_hip_grid_get_ntff_time(slv_[active_].g, which, out);
}
int fdtd_solver_dims(int out[4])
{
  if (active_ < 0) return FDTD_ESTATE;
  out[0] = slv_[active_].NPXa;
  out[1] = slv_[active_].NPYa;
  out[2] = slv_[active_].off;
  out[3] = slv_[active_].arraySize;
  return FDTD_OK;
}

/* ---------------------------------------------------------------- MPI TM */
static void mpi_tm_init(void) { solver_init(SLV_MPI_TM); }
static void mpi_tm_update(void) { solver_update(SLV_MPI_TM); }
static void mpi_tm_reset(void) { solver_reset(SLV_MPI_TM); }
static void mpi_tm_finish(void) { solver_finish(SLV_MPI_TM); }
fdtd_method mpi_fdtdTM_upml_getUpdate(void) { return mpi_tm_update; }
fdtd_method mpi_fdtdTM_upml_getFinish(void) { return mpi_tm_finish; }
fdtd_method mpi_fdtdTM_upml_getReset(void) { return mpi_tm_reset; }
fdtd_method mpi_fdtdTM_upml_getInit(void) { return mpi_tm_init; }
fdtd_dcomplex *mpi_fdtdTM_upml_getEz(void) { return (fdtd_dcomplex *)solver_field(SLV_MPI_TM, 0); }
fdtd_dcomplex *mpi_fdtdTM_upml_getHx(void) { return (fdtd_dcomplex *)solver_field(SLV_MPI_TM, 1); }
fdtd_dcomplex *mpi_fdtdTM_upml_getHy(void) { return (fdtd_dcomplex *)solver_field(SLV_MPI_TM, 2); }
double *mpi_fdtdTM_upml_getEps(void) { return solver_eps(SLV_MPI_TM); }
int mpi_fdtdTM_upml_getSubNx(void) { return N_PX; }
int mpi_fdtdTM_upml_getSubNy(void) { return N_PY; }
int mpi_fdtdTM_upml_getSubNpx(void) { return N_PX + 2; }
int mpi_fdtdTM_upml_getSubNpy(void) { return N_PY + 2; }
int mpi_fdtdTM_upml_getSubNcell(void) { return (N_PX + 2) * (N_PY + 2); }

/* ---------------------------------------------------------------- MPI TE */
static void mpi_te_init(void) { solver_init(SLV_MPI_TE); }
static void mpi_te_update(void) { solver_update(SLV_MPI_TE); }
static void mpi_te_reset(void) { solver_reset(SLV_MPI_TE); }
static void mpi_te_finish(void) { solver_finish(SLV_MPI_TE); }
fdtd_method mpi_fdtdTE_upml_getUpdate(void) { return mpi_te_update; }
fdtd_method mpi_fdtdTE_upml_getFinish(void) { return mpi_te_finish; }
fdtd_method mpi_fdtdTE_upml_getReset(void) { return mpi_te_reset; }
fdtd_method mpi_fdtdTE_upml_getInit(void) { return mpi_te_init; }
fdtd_dcomplex *mpi_fdtdTE_upml_getEx(void) { return (fdtd_dcomplex *)solver_field(SLV_MPI_TE, 0); }
fdtd_dcomplex *mpi_fdtdTE_upml_getEy(void) { return (fdtd_dcomplex *)solver_field(SLV_MPI_TE, 1); }
fdtd_dcomplex *mpi_fdtdTE_upml_getHz(void) { return (fdtd_dcomplex *)solver_field(SLV_MPI_TE, 2); }
double *mpi_fdtdTE_upml_getEps(void) { return solver_eps(SLV_MPI_TE); }
int mpi_fdtdTE_upml_getSubNx(void) { return N_PX; }
int mpi_fdtdTE_upml_getSubNy(void) { return N_PY; }
int mpi_fdtdTE_upml_getSubNpx(void) { return N_PX + 2; }
int mpi_fdtdTE_upml_getSubNpy(void) { return N_PY + 2; }
int mpi_fdtdTE_upml_getSubNcell(void) { return (N_PX + 2) * (N_PY + 2); }

/* ---------------------------------------------------------------- plain TM */
static void tm_init(void) { solver_init(SLV_TM); }
static void tm_update(void) { solver_update(SLV_TM); }
static void tm_reset(void) { solver_reset(SLV_TM); }
static void tm_finish(void) { solver_finish(SLV_TM); }
fdtd_method fdtdTM_upml_getUpdate(void) { return tm_update; }
fdtd_method fdtdTM_upml_getFinish(void) { return tm_finish; }
fdtd_method fdtdTM_upml_getReset(void) { return tm_reset; }
fdtd_method fdtdTM_upml_getInit(void) { return tm_init; }
fdtd_dcomplex *fdtdTM_upml_getEz(void) { return (fdtd_dcomplex *)solver_field(SLV_TM, 0); }
fdtd_dcomplex *fdtdTM_upml_getHx(void) { return (fdtd_dcomplex *)solver_field(SLV_TM, 1); }
fdtd_dcomplex *fdtdTM_upml_getHy(void) { return (fdtd_dcomplex *)solver_field(SLV_TM, 2); }
double *fdtdTM_upml_getEps(void) { return solver_eps(SLV_TM); }

/* ---------------------------------------------------------------- split-field PML TM (fdtdTM.c) */
static void split_tm_init(void) { solver_init(SLV_SPLIT_TM); }
static void split_tm_update(void) { solver_update(SLV_SPLIT_TM); }
static void split_tm_reset(void) { solver_reset(SLV_SPLIT_TM); }
static void split_tm_finish(void) { solver_finish(SLV_SPLIT_TM); }
/* the split components: TM q 0 Ezx (slot 1), 1 Ezy (slot 2); TE q 0 Hzx (slot 7), 1 Hzy (slot 8) */
static double complex *split_field(int kind, int q)
{
  solver_t *s = &slv_[kind];
  if (!s->g) return NULL;
  CHECK(fdtd_hip_grid_get_field(s->g, (is_tm(kind) ? 1 : 7) + q, (double *)s->split_mirror[q]));
  return s->split_mirror[q];
}
fdtd_method fdtdTM_getUpdate(void) { return split_tm_update; }
fdtd_method fdtdTM_getFinish(void) { return split_tm_finish; }
fdtd_method fdtdTM_getReset(void) { return split_tm_reset; }
fdtd_method fdtdTM_getInit(void) { return split_tm_init; }
fdtd_dcomplex *fdtdTM_getEzx(void) { return (fdtd_dcomplex *)split_field(SLV_SPLIT_TM, 0); }
fdtd_dcomplex *fdtdTM_getEzy(void) { return (fdtd_dcomplex *)split_field(SLV_SPLIT_TM, 1); }
fdtd_dcomplex *fdtdTM_getEz(void) { return (fdtd_dcomplex *)solver_field(SLV_SPLIT_TM, 0); }
fdtd_dcomplex *fdtdTM_getHy(void) { return (fdtd_dcomplex *)solver_field(SLV_SPLIT_TM, 2); }
fdtd_dcomplex *fdtdTM_getHx(void) { return (fdtd_dcomplex *)solver_field(SLV_SPLIT_TM, 1); }
double *fdtdTM_getEps(void) { return solver_eps(SLV_SPLIT_TM); }

/* ---------------------------------------------------------------- split-field PML TE (fdtdTE.c) */
static void split_te_init(void) { solver_init(SLV_SPLIT_TE); }
static void split_te_update(void) { solver_update(SLV_SPLIT_TE); }
static void split_te_reset(void) { solver_reset(SLV_SPLIT_TE); }
static void split_te_finish(void) { solver_finish(SLV_SPLIT_TE); }
fdtd_method fdtdTE_getUpdate(void) { return split_te_update; }
fdtd_method fdtdTE_getFinish(void) { return split_te_finish; }
fdtd_method fdtdTE_getReset(void) { return split_te_reset; }
fdtd_method fdtdTE_getInit(void) { return split_te_init; }
fdtd_dcomplex *fdtdTE_getEx(void) { return (fdtd_dcomplex *)solver_field(SLV_SPLIT_TE, 0); }
fdtd_dcomplex *fdtdTE_getEy(void) { return (fdtd_dcomplex *)solver_field(SLV_SPLIT_TE, 1); }
fdtd_dcomplex *fdtdTE_getHzx(void) { return (fdtd_dcomplex *)split_field(SLV_SPLIT_TE, 0); }
fdtd_dcomplex *fdtdTE_getHzy(void) { return (fdtd_dcomplex *)split_field(SLV_SPLIT_TE, 1); }
fdtd_dcomplex *fdtdTE_getHz(void) { return (fdtd_dcomplex *)solver_field(SLV_SPLIT_TE, 2); }
double *fdtdTE_getEps(void) { return solver_eps(SLV_SPLIT_TE); }

/* ---------------------------------------------------------------- non-standard FDTD TM (nsFdtdTM.c) */
static void ns_tm_init(void) { solver_init(SLV_NS_TM); }
static void ns_tm_update(void) { solver_update(SLV_NS_TM); }
static void ns_tm_reset(void) { solver_reset(SLV_NS_TM); }
static void ns_tm_finish(void) { solver_finish(SLV_NS_TM); }
fdtd_method nsFdtdTM_getUpdate(void) { return ns_tm_update; }
fdtd_method nsFdtdTM_getFinish(void) { return ns_tm_finish; }
fdtd_method nsFdtdTM_getReset(void) { return ns_tm_reset; }
fdtd_method nsFdtdTM_getInit(void) { return ns_tm_init; }
fdtd_dcomplex *nsFdtdTM_getEzx(void) { return (fdtd_dcomplex *)split_field(SLV_NS_TM, 0); }
fdtd_dcomplex *nsFdtdTM_getEzy(void) { return (fdtd_dcomplex *)split_field(SLV_NS_TM, 1); }
fdtd_dcomplex *nsFdtdTM_getEz(void) { return (fdtd_dcomplex *)solver_field(SLV_NS_TM, 0); }
fdtd_dcomplex *nsFdtdTM_getHy(void) { return (fdtd_dcomplex *)solver_field(SLV_NS_TM, 2); }
fdtd_dcomplex *nsFdtdTM_getHx(void) { return (fdtd_dcomplex *)solver_field(SLV_NS_TM, 1); }
double *nsFdtdTM_getEps(void) { return solver_eps(SLV_NS_TM); }

/* solver.h:7-20: ten function pointers; the NS getters return one left unset,
   as nsFdtdTM.c:45-61 and nsFdtdTE.c:42-58 do */
struct Solver {
  void (*fn[10])(void);
};
struct Solver *nsFdtdTM_getSolver(void)
{
  static struct Solver solver;
  return &solver;
}
struct Solver *nsFdtdTE_getSolver(void)
{
  static struct Solver solver;
  return &solver;
}

/* ---------------------------------------------------------------- non-standard FDTD TE (nsFdtdTE.c) */
static void ns_te_init(void) { solver_init(SLV_NS_TE); }
static void ns_te_update(void) { solver_update(SLV_NS_TE); }
static void ns_te_reset(void) { solver_reset(SLV_NS_TE); }
static void ns_te_finish(void) { solver_finish(SLV_NS_TE); }
fdtd_method nsFdtdTE_getUpdate(void) { return ns_te_update; }
fdtd_method nsFdtdTE_getFinish(void) { return ns_te_finish; }
fdtd_method nsFdtdTE_getReset(void) { return ns_te_reset; }
fdtd_method nsFdtdTE_getInit(void) { return ns_te_init; }
fdtd_dcomplex *nsFdtdTE_getHzx(void) { return (fdtd_dcomplex *)split_field(SLV_NS_TE, 0); }
fdtd_dcomplex *nsFdtdTE_getHzy(void) { return (fdtd_dcomplex *)split_field(SLV_NS_TE, 1); }
fdtd_dcomplex *nsFdtdTE_getHz(void) { return (fdtd_dcomplex *)solver_field(SLV_NS_TE, 2); }
fdtd_dcomplex *nsFdtdTE_getEy(void) { return (fdtd_dcomplex *)solver_field(SLV_NS_TE, 1); }
fdtd_dcomplex *nsFdtdTE_getEx(void) { return (fdtd_dcomplex *)solver_field(SLV_NS_TE, 0); }
double *nsFdtdTE_getEps(void) { return solver_eps(SLV_NS_TE); }

/* one eps map of an NS solver in a host copy of its own (TM 0 EPS_EZ, 1 EPS_HX,
   2 EPS_HY; TE 0 EPS_EX, 1 EPS_EY, 2 EPS_HZ), refreshed on every call */
static double *ns_eps_map(int kind, int which)
{
  solver_t *s = &slv_[kind];
  if (!s->g) return NULL;
  if (!s->eps_maps[which]) s->eps_maps[which] = (double *)calloc((size_t)s->NPXa * s->NPYa, sizeof(double));
  CHECK(fdtd_hip_grid_get_eps(s->g, which, s->eps_maps[which]));
  return s->eps_maps[which];
}
/* nsFdtdTM.c:35-37 and nsFdtdTE.c:33-35 define these without declaring them in
   their headers */
double *nsFdtdTM_getEpsX(void) { return ns_eps_map(SLV_NS_TM, 1); }
double *nsFdtdTM_getEpsY(void) { return ns_eps_map(SLV_NS_TM, 2); }
double *nsFdtdTM_getEpsZ(void) { return ns_eps_map(SLV_NS_TM, 0); }
double *nsFdtdTE_getEpsX(void) { return ns_eps_map(SLV_NS_TE, 0); }
double *nsFdtdTE_getEpsY(void) { return ns_eps_map(SLV_NS_TE, 1); }
double *nsFdtdTE_getEpsZ(void) { return ns_eps_map(SLV_NS_TE, 2); }

/* ---------------------------------------------------------------- plain TE */
static void te_init(void) { solver_init(SLV_TE); }
static void te_update(void) { solver_update(SLV_TE); }
static void te_reset(void) { solver_reset(SLV_TE); }
static void te_finish(void) { solver_finish(SLV_TE); }
fdtd_method fdtdTE_upml_getUpdate(void) { return te_update; }
fdtd_method fdtdTE_upml_getFinish(void) { return te_finish; }
fdtd_method fdtdTE_upml_getReset(void) { return te_reset; }
fdtd_method fdtdTE_upml_getInit(void) { return te_init; }
fdtd_dcomplex *fdtdTE_upml_getEx(void) { return (fdtd_dcomplex *)solver_field(SLV_TE, 0); }
fdtd_dcomplex *fdtdTE_upml_getEy(void) { return (fdtd_dcomplex *)solver_field(SLV_TE, 1); }
fdtd_dcomplex *fdtdTE_upml_getHz(void) { return (fdtd_dcomplex *)solver_field(SLV_TE, 2); }
double *fdtdTE_upml_getEps(void) { return solver_eps(SLV_TE); }
