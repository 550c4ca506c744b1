/*
 * fdtd_oracle.c -- serial FP64 restatement of rennone/mpiFDTD's UPML path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle).  Nothing in the product links
 * or calls this file; tests/, smoke() and bench.py's cpu_baseline leg do.
 *
 * The arithmetic is written to reproduce the reference's operation order
 * (double complex, real x complex products, no FMA) so that the restatement
 * can be checked bit-for-bit against the compiled reference (oracle/_ref).
 * Each function cites the reference file:line it restates.
 *
 * How each part is pinned (tests/):
 *   steps of all eight solvers, eps maps of the five models, ntffTM_Frequency,
 *   the plain solvers' spectra, the MPI-TE Eth/Eph output, the field.h
 *   helpers: bit-exact (spectra 1e-12, text files to their printed
 *   resolution) against golden runs of the compiled reference
 *   (test_oracle_golden.py, test_field_util.py);
 *   the TE far field (no working reference routine): the Mie series of the
 *   dielectric cylinder (test_oracle_mie.py);
 *   mpi_tm_ntff (accumulators the reference never writes out): the CW
 *   steady-state identity with the golden-pinned frequency-domain NTFF, 7.5e-5
 *   (test_oracle_ntff_pin.py);
 *   the point source and the Courant setting (never used by the reference):
 *   closed forms of the first steps, mirror symmetry, the front speed
 *   (test_oracle_ntff_pin.py);
 *   the whole file: clean under ASan / UBSan (san_driver.c,
 *   test_oracle_sanitized.py).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>
#include "fdtd_oracle.h"

#ifndef M_PI
#define M_PI 3.1415926535897932384626433832795
#endif

#define NTFF_NUM 8192      /* ntff.h:6 */
#define LAMBDA_ST_NM 380   /* ntff.h:4 */
#define LAMBDA_EN_NM 700   /* ntff.h:5 */
#define NT_GUARD 4         /* guard bins around the flat [360][arraySize] accumulators (Q10 writes at -2 / +1) */

#define IMAX(a, b) ((a) > (b) ? (a) : (b))
#define IMIN(a, b) ((a) < (b) ? (a) : (b))

typedef double complex dcx;

enum { TM_EZ, TM_JZ, TM_DZ, TM_HX, TM_MX, TM_BX, TM_HY, TM_MY, TM_BY };
enum { TE_EX, TE_JX, TE_DX, TE_EY, TE_JY, TE_DY, TE_HZ, TE_MZ, TE_BZ };
/* TM coefficient slots (mpiTM_UPML.c:50-55) */
enum { C_JZ, C_JZHXHY, C_DZ, C_DZJZ1, C_DZJZ0, C_MX, C_MXEZ, C_BX, C_BXMX1, C_BXMX0,
       C_MY, C_MYEZ, C_BY, C_BYMY1, C_BYMY0 };
/* TE coefficient slots (mpiTE_UPML.c:37-42) */
enum { C_JX, C_JXHZ, C_DX, C_DXJX1, C_DXJX0, C_JY, C_JYHZ, C_DY, C_DYJY1, C_DYJY0,
       C_MZ, C_MZEXEY, C_BZ, C_BZMZ1, C_BZMZ0 };

typedef struct { double a, b1, b2; int id; } lamela_t;  /* morphoScaleModel.c:67-71 */

struct orc_sim {
  orc_config cfg;
  /* field.c statics */
  int N_X, N_Y, N_PX, N_PY, N_CELL, N_PML;
  double time, ray_coef, lambda_s, k_s, w_s, T_s, maxTime;
  double K_s, Omega_s, Angle_deg;
  double C0, MU0, Z0, EPS0;
  int top, bottom, left, right, cx, cy, arraySize;
  double RFperC;
  /* array layout */
  int NPXa, NPYa, off;
  bool mpi, tm, split;
  dcx *F[9];
  double *C[15];
  double *EPS[3];
  dcx *NTb[3];  /* allocation base */
  dcx *NT[3];   /* NTb + NT_GUARD */
  double R0;
  /* model state (cell units) */
  double ci_radius_s, ci_eps_s, ci_posx, ci_posy;
  double ml_width_s[2], ml_thick_s[2], ml_branch_s, ml_gap_s, ml_ep[2], ml_epx[2], ml_c0, ml_c1;
  double zz_rad, zz_width_s, zz_thick_s, zz_height_s, zz_ox, zz_oy, zz_exw, zz_exh, zz_ep;
  double mo_width_s, mo_thick_s[2], mo_branch_s, mo_gap_s, mo_height_s, mo_ox, mo_oy, mo_ep, mo_epx;
  double mo_rand[2][ORC_MAX_LAYERS];
  lamela_t mo_lam[2][ORC_MAX_LAYERS];
};

/* ------------------------------------------------------------------ defaults */
void orc_default_model_params(orc_model_params *mp)
{
  memset(mp, 0, sizeof(*mp));
  mp->circle_radius_nm = 500; mp->circle_n = 1.6;                       /* circleModel.c:6-10 */
  mp->ml_n0 = 1.0; mp->ml_n1 = 1.56; mp->ml_n0x = 1.0; mp->ml_n1x = 1.1;  /* multiLayerModel.c:7-15 */
  mp->ml_width_nm0 = 300; mp->ml_width_nm1 = 300;                       /* :18, :58 */
  mp->ml_thick_nm0 = 100; mp->ml_thick_nm1 = 90;                        /* :23-30, :61 */
  mp->ml_layer_num = 4; mp->ml_branch_nm = 0; mp->ml_left_gap_nm = 0;   /* :33, :46, :79 */
  mp->ml_edge_rate = 0.0; mp->ml_curve = 0.0;                           /* :51, :56 */
  mp->zz_width_nm = 300; mp->zz_thick_nm = 80; mp->zz_layer_num = 5;    /* zigzagModel.c:7-19 */
  mp->zz_degree = 10; mp->zz_n = 1.56;                                  /* :22-24 */
  mp->mo_width_nm = 300; mp->mo_thick_nm0 = 90; mp->mo_thick_nm1 = 90;  /* morphoScaleModel.c:18-30 */
  mp->mo_layer_num = 6; mp->mo_branch_nm = 0; mp->mo_left_gap_nm = 0;   /* :33, :45, :96 */
  mp->mo_edge_rate = 0.0; mp->mo_curve = 0.0; mp->mo_n = 1.56; mp->mo_nx = 1.1; /* :50-60, :15 */
  mp->mo_randomness = 0; mp->mo_seed = 0;                               /* :63-64 */
}

void orc_default_config(orc_config *cfg)
{
  memset(cfg, 0, sizeof(*cfg));
  cfg->width_nm = 1000; cfg->height_nm = 1000; cfg->h_u_nm = 10; cfg->pml = 10;
  cfg->lambda_nm = 500; cfg->angle_deg = 0; cfg->stepNum = 1000;
  cfg->solver = ORC_MPI_TM_UPML; cfg->model = ORC_NO_MODEL;
  cfg->c0 = 0.7071; cfg->ntff_inset = 5; cfg->point_source = 0; cfg->ntff_time = 1;
  orc_default_model_params(&cfg->mp);
}

/* ------------------------------------------------------------------ field.c */
static double to_cell(const orc_sim *s, double nm) { return nm / s->cfg.h_u_nm; } /* field.c:76-78 */

/* field_init (field.c:89-143) */
static void field_init(orc_sim *s)
{
  const orc_config *c = &s->cfg;
  s->C0 = c->c0;
  s->EPS0 = 1.0;                         /* field.h:73 */
  s->MU0 = 1.0 / s->C0 / s->C0;          /* field.h:74 */
  /* field.h:75 hard-codes sqrt(mu/eps) for C0 = 0.7071 */
  s->Z0 = (s->C0 == 0.7071) ? 1.41422712488 : 1.0 / s->C0;

  s->N_PX = (int)to_cell(s, c->width_nm);
  s->N_PY = (int)to_cell(s, c->height_nm);
  s->N_PML = c->pml;
  s->N_X = s->N_PX - 2 * s->N_PML;
  s->N_Y = s->N_PY - 2 * s->N_PML;
  s->N_CELL = s->N_PX * s->N_PY;

  /* waveInfo_s (field.c:105-110) */
  double Lambda_s = to_cell(s, c->lambda_nm);
  s->K_s = 2 * M_PI / Lambda_s;
  s->Omega_s = s->C0 * s->K_s;
  s->Angle_deg = c->angle_deg;

  s->time = 0;
  s->maxTime = c->stepNum;
  s->lambda_s = to_cell(s, c->lambda_nm);
  s->k_s = 2 * M_PI / s->lambda_s;
  s->w_s = s->C0 * s->k_s;
  s->T_s = 2 * M_PI / s->w_s;
  s->ray_coef = 0;

  /* NTFF contour (field.c:133-142) */
  int inset = c->ntff_inset;
  s->cx = s->N_PX / 2;
  s->cy = s->N_PY / 2;
  s->top = s->N_PY - s->N_PML - inset;
  s->bottom = s->N_PML + inset;
  s->left = s->N_PML + inset;
  s->right = s->N_PX - s->N_PML - inset;
  double len = (s->top - s->bottom) / 2;   /* int division, then double */
  s->RFperC = len * 2;
  s->arraySize = (int)(s->maxTime + 2 * s->RFperC);
}

/* field_nextStep (field.c:312-315) */
static void field_next_step(orc_sim *s)
{
  s->time += 1.0;
  s->ray_coef = 1.0 - exp(-pow(0.01 * s->time, 2));
}

/* field_sigmaX / field_sigmaY (field.c:259-283) */
static double sigma_x(const orc_sim *s, double x)
{
  const int M = 2;
  if (x < s->N_PML) return pow(1.0 * (s->N_PML - x) / s->N_PML, M);
  else if (s->N_PML <= x && x < (s->N_X + s->N_PML)) return 0;
  else return pow(1.0 * (x - (s->N_PX - s->N_PML - 1)) / s->N_PML, M);
}
static double sigma_y(const orc_sim *s, double y)
{
  const int M = 2;
  if (y < s->N_PML) return pow(1.0 * (s->N_PML - y) / s->N_PML, M);
  else if (y >= s->N_PML && y < (s->N_Y + s->N_PML)) return 0.0;
  else return pow(1.0 * (y - (s->N_PY - s->N_PML - 1)) / s->N_PML, M);
}
double orc_sigma_x(const orc_sim *s, double x, double y) { (void)y; return sigma_x(s, x); }
double orc_sigma_y(const orc_sim *s, double x, double y) { (void)x; return sigma_y(s, y); }

/* ------------------------------------------------------------------ models */
/* circleModel.c:72-103 */
static double eps_circle(const orc_sim *s, double x, double y, int col, int row)
{
  if (x < s->N_PML || y < s->N_PML || x > s->N_X + s->N_PML || y > s->N_Y + s->N_PML)
    return s->EPS0;
  double dx = x - s->ci_posx, dy = y - s->ci_posy;
  double len = dx * dx + dy * dy;
  double r = s->ci_radius_s;
  if (len >= (r + 1) * (r + 1)) return s->EPS0;
  if (len <= (r - 1) * (r - 1)) return s->ci_eps_s;
  double split = 10, half = split / 2, sum = 0;
  for (double i = -half + 0.5; i < half; i += 1)
    for (double j = -half + 0.5; j < half; j += 1)
      if (pow(dx + col * i / split, 2.0) + pow(dy + row * j / split, 2.0) <= r * r) sum += 1;
  sum /= split * split;
  return s->ci_eps_s * sum + s->EPS0 * (1 - sum);
}

/* multiLayerModel.c:87-103 calc_width */
static double ml_calc_width(const orc_sim *s, double sx, double sy, double wid, double hei, double modY, int k)
{
  const orc_model_params *mp = &s->cfg.mp;
  double p = 1 - sy / hei;
  double new_wid = (wid + s->ml_branch_s) * (p + (1 - p) * mp->ml_edge_rate);
  double dh = k == 0 ? modY : modY - s->ml_thick_s[0];
  double c = k == 0 ? s->ml_c0 : s->ml_c1;
  if (mp->ml_asymmetry && sx < 0) dh = (k == 1 ? modY : modY - s->ml_thick_s[1]);
  return c * pow((dh - s->ml_thick_s[k] / 2), 2) + new_wid;
}

/* multiLayerModel.c:107-185 */
static double eps_multilayer(const orc_sim *s, double x, double y, int col, int row)
{
  const orc_model_params *mp = &s->cfg.mp;
  double width = s->ml_width_s[0] > s->ml_width_s[1] ? s->ml_width_s[0] : s->ml_width_s[1];
  double thick = s->ml_thick_s[0] + s->ml_thick_s[1];
  double height = thick * mp->ml_layer_num + s->ml_gap_s;
  int oy = s->N_PY / 2 - height / 2;
  int ox = s->N_PX / 2;
  double _x = x - ox, _y = y - oy;
  if (fabs(_x) > (width / 2 + 0.5) || _y < -0.5 || _y > height + 0.5) return s->EPS0;
  height = thick * mp->ml_layer_num;
  double sm[2] = {0, 0};
  double split = 10, half = split / 2;
  for (double i = -half + 0.5; i < half; i += 1) {
    for (double j = -half + 0.5; j < half; j += 1) {
      double sx = _x + col * i / split;
      double sy = _y + row * j / split;
      if (sx < 0 && mp->ml_use_gap) sy -= s->ml_gap_s;
      if (sy < 0 || sy > height) continue;
      double p = 1 - sy / height;
      if (fabs(sx) < s->ml_branch_s * (p + (1 - p) * mp->ml_edge_rate)) { sm[1] += 1; continue; }
      double modY = sy - floor(sy / thick) * thick;
      if (modY == s->ml_thick_s[0]) {
        sm[0] += 0.5 * (fabs(sx) < s->ml_width_s[0] / 2);
        sm[1] += 0.5 * (fabs(sx) < s->ml_width_s[1] / 2);
        continue;
      }
      int k;
      if (sx < 0 && mp->ml_asymmetry) k = (modY < s->ml_thick_s[1]);
      else k = (modY > s->ml_thick_s[0]);
      double wid = ml_calc_width(s, sx, sy, s->ml_width_s[k], height, modY, k);
      if (fabs(sx) < wid / 2) sm[k] += 1;
    }
  }
  sm[0] /= split * split;
  sm[1] /= split * split;
  if (mp->ml_uniaxial && col == 0 && row == 1)
    return s->EPS0 * (1 - sm[0] - sm[1]) + s->ml_epx[0] * sm[0] + s->ml_epx[1] * sm[1];
  return s->EPS0 * (1 - sm[0] - sm[1]) + s->ml_ep[0] * sm[0] + s->ml_ep[1] * sm[1];
}

/* zigzagModel.c:42-78 */
static double eps_zigzag(const orc_sim *s, double _x, double _y, int col, int row)
{
  const orc_model_params *mp = &s->cfg.mp;
  double x = _x - s->zz_ox, y = _y - s->zz_oy;
  double rad = s->zz_rad, ws = s->zz_width_s, ts = s->zz_thick_s, hs = s->zz_height_s;
  if (fabs(x) > ws / 2.0 + s->zz_exw + 0.5 || y < -s->zz_exh - 0.5 || y > hs + s->zz_exh + 0.5)
    return s->EPS0;
  double sm = 0, split = 10, half = split / 2;
  for (double i = -half + 0.5; i < half; i += 1) {
    for (double j = -half + 0.5; j < half; j += 1) {
      double sx = x + col * i / split;
      double sy = y + row * j / split;
      if (fabs(sx) > ws / 2 + s->zz_exw || sy < -s->zz_exh || sy > hs + s->zz_exh) continue;
      /* k = min(layerNum-1, max(0, q)) with the reference's function-like macros:
         the conditional expressions are evaluated in double, then truncated to int */
      double q = sy / (ws * sin(rad) + cos(rad) * ts / 2);
      double mx = (0 > q ? 0 : q);
      int k = (int)((mp->zz_layer_num - 1) < mx ? (double)(mp->zz_layer_num - 1) : mx);
      double px = -ws * cos(rad) / 2.0;
      double py = floor((k + 1) / 2) * 2 * (ws * sin(rad)) + k * cos(rad) * ts / 2;
      double vx = (1 - ((k & 1) << 1)) * cos(rad);
      double vy = sin(rad);
      double sqrDist = pow(sx - px, 2) + pow(sy - py, 2) - pow((sx - px) * vx + (sy - py) * vy, 2);
      if (sqrDist <= ts * ts) sm += 1;
    }
  }
  sm /= split * split;
  return s->EPS0 * (1 - sm) + s->zz_ep * sm;
}

/* morphoScaleModel.c:101-104 */
static bool in_lamela(const lamela_t *l, double x, double y) { return (l->a * x + l->b1 <= y && y <= l->a * x + l->b2); }

/* morphoScaleModel.c:107-125 */
static double mo_get_width(const orc_sim *s, double sx, double sy, const lamela_t *lam)
{
  const orc_model_params *mp = &s->cfg.mp;
  double b = (sy - sx * lam->a);
  int signX = sx < 0 ? 0 : 1;
  double rad = s->mo_rand[signX][lam->id];
  double offset = (signX == 0 ? -1 : 1) * (0.5 * (lam->b2 + lam->b1) - b) * sin(rad);
  double p = 1 - b / s->mo_height_s;
  double width_p = (p + (1 - p) * mp->mo_edge_rate);
  double yy = (b - lam->b1) / (lam->b2 - lam->b1) - 0.5;
  double half_width = s->mo_width_s / 2.0;
  double c0 = 4.0 * half_width * mp->mo_curve;
  return width_p * (half_width * (1 - mp->mo_curve)) + c0 * (0.25 - yy * yy) + offset;
}

/* morphoScaleModel.c:128-145 */
static void mo_calc_lamela(const orc_sim *s, double lft, double rht, double btm, double top, lamela_t *lam)
{
  lam->id = -1;
  int signX = lft < 0 ? 0 : 1;
  for (int i = 0; i < s->cfg.mp.mo_layer_num; i++) {
    lam->a = s->mo_lam[signX][i].a;
    lam->b1 = s->mo_lam[signX][i].b1;
    lam->b2 = s->mo_lam[signX][i].b2;
    if (in_lamela(lam, lft, btm) || in_lamela(lam, rht, btm) || in_lamela(lam, lft, top) || in_lamela(lam, rht, top)) {
      lam->id = i;
      return;
    }
  }
}

/* morphoScaleModel.c:147-253 */
static double eps_morpho(const orc_sim *s, double x, double y, int col, int row)
{
  const orc_model_params *mp = &s->cfg.mp;
  double _x = x - s->mo_ox, _y = y - s->mo_oy;
  if (fabs(_x) > 0.5 * sqrt(pow(s->mo_width_s, 2) + pow(s->mo_thick_s[0], 2)) + 0.5) return s->EPS0;
  lamela_t lams[2];
  lams[0].id = lams[1].id = -1;
  double gap = s->mo_gap_s;
  if (fabs(_x) > 0.5) {
    if (_x < 0) mo_calc_lamela(s, _x - 0.5, _x + 0.5, _y - gap - 0.5, _y - gap + 0.5, &lams[0]);
    else mo_calc_lamela(s, _x - 0.5, _x + 0.5, _y - 0.5, _y + 0.5, &lams[1]);
  } else {
    mo_calc_lamela(s, _x - 0.5, 0, _y - gap - 0.5, _y - gap + 0.5, &lams[0]);
    mo_calc_lamela(s, 0, _x + 0.5, _y - 0.5, _y + 0.5, &lams[1]);
  }
  double sm = 0, split = 10, half = split / 2;
  for (double i = -half + 0.5; i < half; i += 1) {
    for (double j = -half + 0.5; j < half; j += 1) {
      double sx = _x + col * i / split;
      double sy = _y + row * j / split;
      if (sx < 0) sy -= gap;
      if (sy >= 0 && sy <= s->mo_height_s) {
        double p = 1 - sy / s->mo_height_s;
        double width_p = (p + (1 - p) * mp->mo_edge_rate);
        if (fabs(sx) < s->mo_branch_s / 2.0 * width_p) { sm += 1; continue; }
      }
      int signX = sx < 0 ? 0 : 1;
      const lamela_t *lam = &lams[signX];
      if (lam->id < 0) continue;
      lam = &s->mo_lam[signX][lam->id];
      if (in_lamela(lam, sx, sy)) {
        double sqrWidth = pow(sx, 2) * (1 + pow(lam->a, 2));
        double newWidth = mo_get_width(s, sx, sy, lam);
        if (sqrWidth < pow(newWidth, 2)) { sm += 1; continue; }
      }
      for (int k = -1; k < 2; k += 2) {
        int id = lam->id + k;
        if (id < 0 || id >= mp->mo_layer_num) continue;
        const lamela_t *l2 = &s->mo_lam[signX][id];
        if (in_lamela(l2, sx, sy)) {
          double sqrWidth = pow(sx, 2) * (1 + pow(l2->a, 2));
          double newWidth = mo_get_width(s, sx, sy, l2);
          if (sqrWidth < pow(newWidth, 2)) { sm += 1; break; }
        }
      }
    }
  }
  sm /= split * split;
  if (mp->mo_uniaxial && col == 0 && row == 1) return s->EPS0 * (1 - sm) + s->mo_epx * sm;
  return s->EPS0 * (1 - sm) + s->mo_ep * sm;
}

/* model init functions: circleModel.c:119-126, multiLayerModel.c:331-348,
   zigzagModel.c:146-160, morphoScaleModel.c:379-422 */
static void models_init(orc_sim *s)
{
  const orc_model_params *mp = &s->cfg.mp;
  switch (s->cfg.model) {
  case ORC_MIE_CYLINDER:
    s->ci_radius_s = to_cell(s, mp->circle_radius_nm);
    s->ci_posx = s->N_PX / 2;
    s->ci_posy = s->N_PY / 2;
    s->ci_eps_s = mp->circle_n * mp->circle_n * s->EPS0;
    break;
  case ORC_LAYER:
    s->ml_width_s[0] = to_cell(s, mp->ml_width_nm0);
    s->ml_width_s[1] = to_cell(s, mp->ml_width_nm1);
    s->ml_thick_s[0] = to_cell(s, mp->ml_thick_nm0);
    s->ml_thick_s[1] = to_cell(s, mp->ml_thick_nm1);
    s->ml_gap_s = to_cell(s, mp->ml_left_gap_nm);
    s->ml_ep[0] = mp->ml_n0 * mp->ml_n0 * s->EPS0;
    s->ml_ep[1] = mp->ml_n1 * mp->ml_n1 * s->EPS0;
    s->ml_epx[0] = mp->ml_n0x * mp->ml_n0x * s->EPS0;
    s->ml_epx[1] = mp->ml_n1x * mp->ml_n1x * s->EPS0;
    s->ml_branch_s = to_cell(s, mp->ml_branch_nm);
    s->ml_c0 = -4 * s->ml_width_s[0] * mp->ml_curve / s->ml_thick_s[0] / s->ml_thick_s[0];
    s->ml_c1 = -4 * s->ml_width_s[1] * mp->ml_curve / s->ml_thick_s[1] / s->ml_thick_s[1];
    break;
  case ORC_ZIGZAG:
    s->zz_rad = mp->zz_degree * M_PI / 180.0;
    s->zz_width_s = to_cell(s, mp->zz_width_nm);
    s->zz_thick_s = to_cell(s, mp->zz_thick_nm);
    s->zz_height_s = (s->zz_width_s * sin(s->zz_rad)) * mp->zz_layer_num + mp->zz_layer_num * s->zz_thick_s / 2 * cos(s->zz_rad);
    s->zz_exw = s->zz_thick_s / 2.0 * sin(s->zz_rad);
    s->zz_exh = s->zz_thick_s / 2.0 * cos(s->zz_rad);
    s->zz_ox = s->N_PX / 2.0;
    s->zz_oy = (s->N_PY - s->zz_height_s) / 2.0;
    s->zz_ep = mp->zz_n * mp->zz_n * s->EPS0;
    break;
  case ORC_MORPHO_SCALE: {
    s->mo_width_s = to_cell(s, mp->mo_width_nm);
    s->mo_thick_s[0] = to_cell(s, mp->mo_thick_nm0);
    s->mo_thick_s[1] = to_cell(s, mp->mo_thick_nm1);
    s->mo_branch_s = to_cell(s, mp->mo_branch_nm);
    s->mo_gap_s = to_cell(s, mp->mo_left_gap_nm);
    s->mo_height_s = (s->mo_thick_s[0] + s->mo_thick_s[1]) * mp->mo_layer_num + (s->mo_thick_s[0] + s->mo_thick_s[1]);
    s->mo_ep = mp->mo_n * mp->mo_n * s->EPS0;
    s->mo_epx = mp->mo_nx * mp->mo_nx * s->EPS0;
    s->mo_oy = s->N_PY / 2 - s->mo_height_s / 2;
    s->mo_ox = s->N_PX / 2;
    srand(mp->mo_seed);
    double TO_RAD = M_PI / 180.0;
    int R = mp->mo_randomness;
    for (int i = 0; i < mp->mo_layer_num; i++) {
      s->mo_rand[1][i] = (rand() % (R + 1) - R / 2) * TO_RAD;
      s->mo_rand[0][i] = (rand() % (R + 1) - R / 2) * TO_RAD;
    }
    double thick = s->mo_thick_s[0] + s->mo_thick_s[1];
    double half_thick = 0.5 * s->mo_thick_s[0];
    double off_l = mp->mo_asymmetry ? s->mo_thick_s[1] + half_thick : half_thick;
    double off_r = half_thick;
    for (int i = 0; i < mp->mo_layer_num; i++) {
      s->mo_lam[0][i].a = tan(s->mo_rand[0][i]);
      s->mo_lam[0][i].b1 = thick * i + off_l - half_thick / cos(s->mo_rand[0][i]);
      s->mo_lam[0][i].b2 = thick * i + off_l + half_thick / cos(s->mo_rand[0][i]);
      s->mo_lam[0][i].id = i;
      s->mo_lam[1][i].a = tan(s->mo_rand[1][i]);
      s->mo_lam[1][i].b1 = thick * i + off_r - half_thick / cos(s->mo_rand[1][i]);
      s->mo_lam[1][i].b2 = thick * i + off_r + half_thick / cos(s->mo_rand[1][i]);
      s->mo_lam[1][i].id = i;
    }
    break;
  }
  default:
    break;
  }
}

/* models_eps (models.c:166-184) with the per-model callbacks; noModel.c:5-8 returns eps0 */
static double model_eps_cr(const orc_sim *s, double x, double y, int col, int row)
{
  switch (s->cfg.model) {
  case ORC_MIE_CYLINDER: return eps_circle(s, x, y, col, row);
  case ORC_LAYER: return eps_multilayer(s, x, y, col, row);
  case ORC_ZIGZAG: return eps_zigzag(s, x, y, col, row);
  case ORC_MORPHO_SCALE: return eps_morpho(s, x, y, col, row);
  default: return s->EPS0;
  }
}
static double models_eps(const orc_sim *s, double x, double y, int mode)
{
  switch (mode) {
  case ORC_D_X: return model_eps_cr(s, x, y, 1, 0);
  case ORC_D_Y: return model_eps_cr(s, x, y, 0, 1);
  default: return model_eps_cr(s, x, y, 1, 1);
  }
}
double orc_model_eps(const orc_sim *s, double x, double y, int mode) { return models_eps(s, x, y, mode); }

/* ------------------------------------------------------------------ coefficients */
/* TM: mpiTM_UPML.c:663-716 (MPI, x=i-1) and fdtdTM_upml.c:224-274 (plain, x=i) */
static void set_coef_tm(orc_sim *s)
{
  const double R = 1.0e-8, M = 2.0;
  const double sig_max = -(M + 1.0) * s->EPS0 * s->C0 / 2.0 / s->N_PML / cos(M_PI / 3) * log(R);
  int i0 = s->mpi ? 1 : 0, i1 = s->mpi ? s->NPXa - 1 : s->NPXa;
  int j0 = s->mpi ? 1 : 0, j1 = s->mpi ? s->NPYa - 1 : s->NPYa;
  for (int i = i0; i < i1; i++) {
    for (int j = j0; j < j1; j++) {
      int k = i * s->NPYa + j;
      int x = i - s->off, y = j - s->off;
      s->EPS[0][k] = models_eps(s, x, y, ORC_D_XY);
      if (s->mpi) {
        s->EPS[1][k] = models_eps(s, x, y + 0.5, ORC_D_Y);
        s->EPS[2][k] = models_eps(s, x + 0.5, y, ORC_D_X);
      } else { /* fdtdTM_upml.c:238-239 samples these with D_XY (Q19) */
        s->EPS[1][k] = models_eps(s, x, y + 0.5, ORC_D_XY);
        s->EPS[2][k] = models_eps(s, x + 0.5, y, ORC_D_XY);
      }
      double sig_ez_x = sig_max * sigma_x(s, x), sig_ez_y = sig_max * sigma_y(s, y);
      double sig_hx_x = sig_max * sigma_x(s, x), sig_hx_y = sig_max * sigma_y(s, y + 0.5);
      double sig_hy_x = sig_max * sigma_x(s, x + 0.5), sig_hy_y = sig_max * sigma_y(s, y);
      double sig_z = 0, eps = s->EPS0;
      s->C[C_JZ][k] = (2 * eps - sig_ez_x) / (2 * eps + sig_ez_x);
      s->C[C_JZHXHY][k] = (2 * eps) / (2 * eps + sig_ez_x);
      s->C[C_DZ][k] = (2 * eps - sig_ez_y) / (2 * eps + sig_ez_y);
      s->C[C_DZJZ1][k] = (2 * eps + sig_z) / (2 * eps + sig_ez_y);
      s->C[C_DZJZ0][k] = (2 * eps - sig_z) / (2 * eps + sig_ez_y);
      s->C[C_MX][k] = (2 * eps - sig_hx_y) / (2 * eps + sig_hx_y);
      s->C[C_MXEZ][k] = (2 * eps) / (2 * eps + sig_hx_y);
      s->C[C_BX][k] = (2 * eps - sig_z) / (2 * eps + sig_z);
      s->C[C_BXMX1][k] = (2 * eps + sig_hx_x) / (2 * eps + sig_z);
      s->C[C_BXMX0][k] = (2 * eps - sig_hx_x) / (2 * eps + sig_z);
      s->C[C_MY][k] = (2 * eps - sig_z) / (2 * eps + sig_z);
      s->C[C_MYEZ][k] = (2 * eps) / (2 * eps + sig_z);
      s->C[C_BY][k] = (2 * eps - sig_hy_x) / (2 * eps + sig_hy_x);
      s->C[C_BYMY1][k] = (2 * eps + sig_hy_y) / (2 * eps + sig_hy_x);
      s->C[C_BYMY0][k] = (2 * eps - sig_hy_y) / (2 * eps + sig_hy_x);
    }
  }
}

/* Split-field PML TM (fdtdTM.c:184-243): per-cell coefficients from eps and the
   x / y conductivities, C[0] C_EZX, C[1] C_EZXLX, C[2] C_EZY, C[3] C_EZYLY,
   C[4] C_HX, C[5] C_HXLY, C[6] C_HY, C[7] C_HYLX; eps maps EPS_EZ (D_XY),
   EPS_HX (j + 0.5, D_Y), EPS_HY (i + 0.5, D_X) */
static double pml_coef(double ep_mu, double sig) { return (1.0 - sig / ep_mu) / (1.0 + sig / ep_mu); } /* field.c:288-291 */
static double pml_coef_lxy(double ep_mu, double sig) { return 1.0 / (ep_mu + sig); }                 /* field.c:292-295 */
static void set_coef_tm_split(orc_sim *s)
{
  const double R = 1.0e-8, M = 2.0;
  const double sig_max = -(M + 1.0) * s->EPS0 * s->C0 / 2.0 / s->N_PML * log(R);
  for (int i = 0; i < s->NPXa; i++) {
    for (int j = 0; j < s->NPYa; j++) {
      int k = i * s->NPYa + j;
      s->EPS[0][k] = models_eps(s, i, j, ORC_D_XY);
      s->EPS[1][k] = models_eps(s, i, j + 0.5, ORC_D_Y);
      s->EPS[2][k] = models_eps(s, i + 0.5, j, ORC_D_X);
      double sig_ez_x = sig_max * sigma_x(s, i);
      double sig_ez_y = sig_max * sigma_y(s, j);
      double sig_hx_yy = s->MU0 / s->EPS0 * (sig_max * sigma_y(s, j + 0.5));
      double sig_hy_xx = s->MU0 / s->EPS0 * (sig_max * sigma_x(s, i + 0.5));
      s->C[0][k] = pml_coef(s->EPS[0][k], sig_ez_x);
      s->C[1][k] = pml_coef_lxy(s->EPS[0][k], sig_ez_x);
      s->C[2][k] = pml_coef(s->EPS[0][k], sig_ez_y);
      s->C[3][k] = pml_coef_lxy(s->EPS[0][k], sig_ez_y);
      s->C[4][k] = pml_coef(s->MU0, sig_hx_yy);
      s->C[5][k] = pml_coef_lxy(s->MU0, sig_hx_yy);
      s->C[6][k] = pml_coef(s->MU0, sig_hy_xx);
      s->C[7][k] = pml_coef_lxy(s->MU0, sig_hy_xx);
    }
  }
}

/* Split-field PML TE (fdtdTE.c:200-241): C[0] C_EX, C[1] C_EXLY, C[2] C_EY,
   C[3] C_EYLX, C[4] C_HZX, C[5] C_HZXLX, C[6] C_HZY, C[7] C_HZYLY; eps maps
   EPS_EX (i + 0.5, D_Y), EPS_EY (j + 0.5, D_X), EPS_HZ (mean of D_X and D_Y
   at (i + 0.5, j + 0.5)) */
static void set_coef_te_split(orc_sim *s)
{
  const double R = 1.0e-8, M = 2.0;
  const double sig_max = -(M + 1.0) * s->EPS0 * s->C0 / 2.0 / s->N_PML * log(R);
  for (int i = 0; i < s->NPXa; i++) {
    for (int j = 0; j < s->NPYa; j++) {
      int k = i * s->NPYa + j;
      s->EPS[0][k] = models_eps(s, i + 0.5, j, ORC_D_Y);
      s->EPS[1][k] = models_eps(s, i, j + 0.5, ORC_D_X);
      s->EPS[2][k] = 0.5 * (models_eps(s, i + 0.5, j + 0.5, ORC_D_X) + models_eps(s, i + 0.5, j + 0.5, ORC_D_Y));
      double sig_ex_y = sig_max * sigma_y(s, j);
      double sig_ey_x = sig_max * sigma_x(s, i);
      double sig_hz_xx = s->MU0 / s->EPS0 * (sig_max * sigma_x(s, i + 0.5));
      double sig_hz_yy = s->MU0 / s->EPS0 * (sig_max * sigma_y(s, j + 0.5));
      s->C[0][k] = pml_coef(s->EPS[0][k], sig_ex_y);
      s->C[1][k] = pml_coef_lxy(s->EPS[0][k], sig_ex_y);
      s->C[2][k] = pml_coef(s->EPS[1][k], sig_ey_x);
      s->C[3][k] = pml_coef_lxy(s->EPS[1][k], sig_ey_x);
      s->C[4][k] = pml_coef(s->MU0, sig_hz_xx);
      s->C[5][k] = pml_coef_lxy(s->MU0, sig_hz_xx);
      s->C[6][k] = pml_coef(s->MU0, sig_hz_yy);
      s->C[7][k] = pml_coef_lxy(s->MU0, sig_hz_yy);
    }
  }
}

/* Non-standard FDTD TM (nsFdtdTM.c:222-306): C[0] C_EZX, C[1] C_EZXLX, C[2] C_EZY,
   C[3] C_EZYLY, C[4] C_HX, C[5] C_HXLY, C[6] C_HY, C[7] C_HYLX, and C[8] the
   scattered-wave factor 1/(_n n) - 1 of field_nsScatteredWaveNotUPML
   (field.c:155-177) with EPS_EZ; eps maps as the split-field TM solver's */
static double ns_beta(double alpha) { return tanh(alpha) / (1 + pow(tanh(alpha), 2)); } /* nsFdtdTM.c:221-224 */
static double ns_coef1(double alpha, double beta) { (void)alpha; return (1 - beta) / (1 + beta); }
static void set_coef_ns_tm(orc_sim *s)
{
  const double R = 1.0e-8, M = 2.0;
  const double sig_max = -(M + 1.0) * s->EPS0 * s->C0 / s->N_PML * log(R);
  const double w_s = s->Omega_s, k_s = s->K_s;
  for (int i = 0; i < s->NPXa; i++) {
    for (int j = 0; j < s->NPYa; j++) {
      int k = i * s->NPYa + j;
      s->EPS[0][k] = models_eps(s, i, j, ORC_D_XY);
      s->EPS[1][k] = models_eps(s, i, j + 0.5, ORC_D_Y);
      s->EPS[2][k] = models_eps(s, i + 0.5, j, ORC_D_X);
      double sig_ez_x = sig_max * sigma_x(s, i), sig_ez_y = sig_max * sigma_y(s, j);
      double sig_hx_y = sig_max * sigma_y(s, j + 0.5), sig_hy_x = sig_max * sigma_x(s, i + 0.5);
      double a_hx_y = sig_hx_y / (2 * s->EPS0), a_hy_x = sig_hy_x / (2 * s->EPS0);
      double a_ez_x = sig_ez_x / (2 * s->EPS0), a_ez_y = sig_ez_y / (2 * s->EPS0);
      double b_hx_y = ns_beta(a_hx_y), b_hy_x = ns_beta(a_hy_x), b_ez_x = ns_beta(a_ez_x), b_ez_y = ns_beta(a_ez_y);
      double z_ez = sqrt(s->MU0 / s->EPS[0][k]);
      double n_ez = sqrt(s->EPS[0][k] / s->EPS0);
      double k_ez_s = k_s * n_ez;
      double u_ez = sin(w_s * 0.5) / sin(k_ez_s * 0.5);
      s->C[0][k] = ns_coef1(a_ez_x, b_ez_x);
      s->C[1][k] = u_ez * z_ez / (1 + b_ez_x);
      s->C[2][k] = ns_coef1(a_ez_y, b_ez_y);
      s->C[3][k] = u_ez * z_ez / (1.0 + b_ez_x); /* b_ez_x, as nsFdtdTM.c:282 has it */
      double z_hx = sqrt(s->MU0 / s->EPS[1][k]);
      double n_hx = sqrt(s->EPS[1][k] / s->EPS0);
      double k_hx_s = k_s * n_hx;
      double u_hx = sin(w_s * 0.5) / sin(k_hx_s * 0.5);
      s->C[4][k] = ns_coef1(a_hx_y, b_hx_y);
      s->C[5][k] = u_hx / z_hx / (1.0 + b_hx_y);
      double z_hy = sqrt(s->MU0 / s->EPS[2][k]);
      double n_hy = sqrt(s->EPS[2][k] / s->EPS0);
      double k_hy_s = k_s * n_hy;
      double u_hy = sin(w_s * 0.5) / sin(k_hy_s * 0.5);
      s->C[6][k] = ns_coef1(a_hy_x, b_hy_x);
      s->C[7][k] = u_hy / z_hy / (1.0 + b_hy_x);
      /* field.c:168-174 */
      double n = sqrt(s->EPS[0][k] / s->EPS0);
      double u0 = sin(w_s * 0.5) / sin(k_s * 0.5);
      double u1 = sin(w_s * 0.5) / sin(n * k_s * 0.5);
      double _n = u0 / u1;
      s->C[8][k] = 1.0 / (_n * n) - 1.0;
    }
  }
}

/* Non-standard FDTD TE (nsFdtdTE.c:100-181): over the interior cells only (the
   border of every map stays zero, as the reference's zeroed arrays do), with
   alpha = sigma / (2 eps) of the component's own eps.  C[0] C_EX, C[1] C_EXLY,
   C[2] C_EY, C[3] C_EYLX, C[4] C_HZX, C[5] C_HZXLX, C[6] C_HZY, C[7] C_HZYLY, and
   the scattered-wave factors 1/(_n n) - 1 (field.c:155-177) of EPS_EX in C[8] and
   of EPS_EY in C[9]; EPS[0] EPS_EX, EPS[1] EPS_EY, EPS[2] EPS_HZ */
static double ns_wave_factor(const orc_sim *s, double eps)
{
  const double w_s = s->Omega_s, k_s = s->K_s;
  double n = sqrt(eps / s->EPS0);
  double u0 = sin(w_s * 0.5) / sin(k_s * 0.5);
  double u1 = sin(w_s * 0.5) / sin(n * k_s * 0.5);
  double _n = u0 / u1;
  return 1.0 / (_n * n) - 1.0;
}
static void set_coef_ns_te(orc_sim *s)
{
  const double R = 1.0e-8, M = 2.0;
  const double sig_max = -(M + 1.0) * s->EPS0 * s->C0 / s->N_PML * log(R);
  const double w_s = s->Omega_s, k_s = s->K_s;
  for (int i = 1; i < s->NPXa - 1; i++) {
    for (int j = 1; j < s->NPYa - 1; j++) {
      int k = i * s->NPYa + j;
      double e_ex = s->EPS[0][k] = models_eps(s, i + 0.5, j, ORC_D_Y);
      double e_ey = s->EPS[1][k] = models_eps(s, i, j + 0.5, ORC_D_X);
      double e_hz = s->EPS[2][k] = models_eps(s, i + 0.5, j + 0.5, ORC_D_XY);
      double sig_ex_y = sig_max * sigma_y(s, j), sig_ey_x = sig_max * sigma_x(s, i);
      double sig_hz_x = sig_max * sigma_x(s, i + 0.5), sig_hz_y = sig_max * sigma_y(s, j + 0.5);
      double b_ex_y = ns_beta(sig_ex_y / (2 * e_ex)), b_ey_x = ns_beta(sig_ey_x / (2 * e_ey));
      double b_hz_x = ns_beta(sig_hz_x / (2 * e_hz)), b_hz_y = ns_beta(sig_hz_y / (2 * e_hz));
      double z_hz = sqrt(s->MU0 / e_hz), n_hz = sqrt(e_hz / s->EPS0);
      double u_hz = sin(w_s * 0.5) / sin(k_s * n_hz * 0.5);
      s->C[4][k] = ns_coef1(0, b_hz_x);
      s->C[5][k] = u_hz / z_hz / (1.0 + b_hz_x);
      s->C[6][k] = ns_coef1(0, b_hz_y);
      s->C[7][k] = u_hz / z_hz / (1.0 + b_hz_y);
      double z_ex = sqrt(s->MU0 / e_ex), n_ex = sqrt(e_ex / s->EPS0);
      double u_ex = sin(w_s * 0.5) / sin(k_s * n_ex * 0.5);
      s->C[0][k] = ns_coef1(0, b_ex_y);
      s->C[1][k] = u_ex * z_ex / (1.0 + b_ex_y);
      double z_ey = sqrt(s->MU0 / e_ey), n_ey = sqrt(e_ey / s->EPS0);
      double u_ey = sin(w_s * 0.5) / sin(k_s * n_ey * 0.5);
      s->C[2][k] = ns_coef1(0, b_ey_x);
      s->C[3][k] = u_ey * z_ey / (1.0 + b_ey_x);
      s->C[8][k] = ns_wave_factor(s, e_ex);
      s->C[9][k] = ns_wave_factor(s, e_ey);
    }
  }
}

/* TE: mpiTE_UPML.c:489-536 (MPI) and fdtdTE_upml.c:361-412 (plain); no /cos(pi/3) (Q7) */
static void set_coef_te(orc_sim *s)
{
  const double R = 1.0e-8, M = 2.0;
  const double sig_max = -(M + 1.0) * s->EPS0 * s->C0 / 2.0 / s->N_PML * log(R);
  int i0 = s->mpi ? 1 : 0, i1 = s->mpi ? s->NPXa - 1 : s->NPXa;
  int j0 = s->mpi ? 1 : 0, j1 = s->mpi ? s->NPYa - 1 : s->NPYa;
  for (int i = i0; i < i1; i++) {
    for (int j = j0; j < j1; j++) {
      int k = i * s->NPYa + j;
      int x = i - s->off, y = j - s->off;
      s->EPS[0][k] = models_eps(s, x + 0.5, y, ORC_D_Y);
      s->EPS[1][k] = models_eps(s, x, y + 0.5, ORC_D_X);
      s->EPS[2][k] = 0.5 * (models_eps(s, x + 0.5, y + 0.5, ORC_D_X) + models_eps(s, x + 0.5, y + 0.5, ORC_D_Y));
      double sig_ex_x = sig_max * sigma_x(s, x + 0.5), sig_ex_y = sig_max * sigma_y(s, y);
      double sig_ey_x = sig_max * sigma_x(s, x), sig_ey_y = sig_max * sigma_y(s, y + 0.5);
      double sig_hz_x = sig_max * sigma_x(s, x + 0.5), sig_hz_y = sig_max * sigma_y(s, y + 0.5);
      double eps = s->EPS0, sig_z = 0;
      s->C[C_JX][k] = (2 * eps - sig_ex_y) / (2 * eps + sig_ex_y);
      s->C[C_JXHZ][k] = (2 * eps) / (2 * eps + sig_ex_y);
      s->C[C_DX][k] = (2 * eps - sig_z) / (2 * eps + sig_z);
      s->C[C_DXJX1][k] = (2 * eps + sig_ex_x) / (2 * eps + sig_z);
      s->C[C_DXJX0][k] = (2 * eps - sig_ex_x) / (2 * eps + sig_z);
      s->C[C_JY][k] = (2 * eps - sig_z) / (2 * eps + sig_z);
      s->C[C_JYHZ][k] = (2 * eps) / (2 * eps + sig_z);
      s->C[C_DY][k] = (2 * eps - sig_ey_x) / (2 * eps + sig_ey_x);
      s->C[C_DYJY1][k] = (2 * eps + sig_ey_y) / (2 * eps + sig_ey_x);
      s->C[C_DYJY0][k] = (2 * eps - sig_ey_y) / (2 * eps + sig_ey_x);
      s->C[C_MZ][k] = (2 * eps - sig_hz_x) / (2 * eps + sig_hz_x);
      s->C[C_MZEXEY][k] = (2 * eps) / (2 * eps + sig_hz_x);
      s->C[C_BZ][k] = (2 * eps - sig_hz_y) / (2 * eps + sig_hz_y);
      s->C[C_BZMZ1][k] = (2 * eps + sig_z) / (2 * eps + sig_hz_y);
      s->C[C_BZMZ0][k] = (2 * eps - sig_z) / (2 * eps + sig_hz_y);
    }
  }
}

/* ------------------------------------------------------------------ create */
orc_sim *orc_create(const orc_config *cfg)
{
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  s->cfg = *cfg;
  if (s->cfg.c0 == 0) s->cfg.c0 = 0.7071;
  if (s->cfg.ntff_inset == 0) s->cfg.ntff_inset = 5;
  field_init(s);
  models_init(s);  /* models_initModel runs after field_init (simulator.c:229-231) */
  s->mpi = (cfg->solver == ORC_MPI_TM_UPML || cfg->solver == ORC_MPI_TE_UPML);
  s->tm = (cfg->solver == ORC_MPI_TM_UPML || cfg->solver == ORC_TM_UPML || cfg->solver == ORC_TM_2D);
  s->split = (cfg->solver == ORC_TM_2D || cfg->solver == ORC_TE_2D || cfg->solver == ORC_NS_TM_2D ||
              cfg->solver == ORC_NS_TE_2D);
  if (cfg->solver == ORC_NS_TM_2D) s->tm = true;
  /* MPI single rank: SUB_N_PX = N_PX + 2 with a ghost ring (mpiTM_UPML.c:737-743) */
  s->off = s->mpi ? 1 : 0;
  s->NPXa = s->N_PX + 2 * s->off;
  s->NPYa = s->N_PY + 2 * s->off;
  size_t n = (size_t)s->NPXa * s->NPYa;
  for (int f = 0; f < 9; f++) s->F[f] = (dcx *)calloc(n, sizeof(dcx));
  for (int c = 0; c < 15; c++) s->C[c] = (double *)calloc(n, sizeof(double));
  for (int e = 0; e < 3; e++) s->EPS[e] = (double *)calloc(n, sizeof(double));
  size_t nn = (size_t)360 * s->arraySize + 2 * NT_GUARD;
  for (int a = 0; a < 3; a++) {
    s->NTb[a] = (dcx *)calloc(nn, sizeof(dcx));
    s->NT[a] = s->NTb[a] + NT_GUARD;
  }
  if (cfg->solver == ORC_TM_2D) set_coef_tm_split(s);
  else if (cfg->solver == ORC_TE_2D) set_coef_te_split(s);
  else if (cfg->solver == ORC_NS_TM_2D) set_coef_ns_tm(s);
  else if (cfg->solver == ORC_NS_TE_2D) set_coef_ns_te(s);
  else if (s->tm) set_coef_tm(s);
  else set_coef_te(s);
  s->R0 = 1.0e6 * to_cell(s, 500);  /* ntffTM_init (ntffTM.c:31), ntffTE_init (ntffTE.c:16) */
  return s;
}

void orc_destroy(orc_sim *s)
{
  if (!s) return;
  for (int f = 0; f < 9; f++) free(s->F[f]);
  for (int c = 0; c < 15; c++) free(s->C[c]);
  for (int e = 0; e < 3; e++) free(s->EPS[e]);
  for (int a = 0; a < 3; a++) free(s->NTb[a]);
  free(s);
}

/* ------------------------------------------------------------------ TM kernels */
#define IDX(i, j) ((i) * NPY + (j))

/* calcJD + calcE (mpiTM_UPML.c:431-458, fdtdTM_upml.c:155-178) */
static void tm_calc_jd_e(orc_sim *s)
{
  const int NPX = s->NPXa, NPY = s->NPYa;
  dcx *Ez = s->F[TM_EZ], *Jz = s->F[TM_JZ], *Dz = s->F[TM_DZ], *Hx = s->F[TM_HX], *Hy = s->F[TM_HY];
  double **C = s->C;
  for (int i = 1; i < NPX - 1; i++) {
    for (int j = 1; j < NPY - 1; j++) {
      const int k = IDX(i, j);
      const dcx nowJz = Jz[k];
      Jz[k] = C[C_JZ][k] * Jz[k] + C[C_JZHXHY][k] * (+Hy[k] - Hy[k - NPY] - Hx[k] + Hx[k - 1]);
      Dz[k] = C[C_DZ][k] * Dz[k] + C[C_DZJZ1][k] * Jz[k] - C[C_DZJZ0][k] * nowJz;
    }
  }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      const int k = IDX(i, j);
      Ez[k] = Dz[k] / s->EPS[0][k];
    }
}

/* calcMB + calcH (mpiTM_UPML.c:491-530, fdtdTM_upml.c:181-219) */
static void tm_calc_mb_h(orc_sim *s)
{
  const int NPX = s->NPXa, NPY = s->NPYa;
  dcx *Ez = s->F[TM_EZ], *Hx = s->F[TM_HX], *Mx = s->F[TM_MX], *Bx = s->F[TM_BX];
  dcx *Hy = s->F[TM_HY], *My = s->F[TM_MY], *By = s->F[TM_BY];
  double **C = s->C;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      const int k = IDX(i, j);
      dcx nowMx = Mx[k];
      Mx[k] = C[C_MX][k] * Mx[k] - C[C_MXEZ][k] * (Ez[k + 1] - Ez[k]);
      Bx[k] = C[C_BX][k] * Bx[k] + C[C_BXMX1][k] * Mx[k] - C[C_BXMX0][k] * nowMx;
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      const int k = IDX(i, j);
      dcx nowMy = My[k];
      My[k] = C[C_MY][k] * My[k] - C[C_MYEZ][k] * (-Ez[k + NPY] + Ez[k]);
      By[k] = C[C_BY][k] * By[k] + C[C_BYMY1][k] * My[k] - C[C_BYMY0][k] * nowMy;
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      const int k = IDX(i, j);
      Hx[k] = Bx[k] / s->MU0;
      Hy[k] = By[k] / s->MU0;
    }
}

/* ------------------------------------------------------------------ TE kernels */
/* calcJD + calcE (mpiTE_UPML.c:337-375, fdtdTE_upml.c:252-291) */
static void te_calc_jd_e(orc_sim *s)
{
  const int NPX = s->NPXa, NPY = s->NPYa;
  dcx *Ex = s->F[TE_EX], *Jx = s->F[TE_JX], *Dx = s->F[TE_DX];
  dcx *Ey = s->F[TE_EY], *Jy = s->F[TE_JY], *Dy = s->F[TE_DY], *Hz = s->F[TE_HZ];
  double **C = s->C;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      dcx nowJx = Jx[k];
      Jx[k] = C[C_JX][k] * Jx[k] + C[C_JXHZ][k] * (Hz[k] - Hz[k - 1]);
      Dx[k] = C[C_DX][k] * Dx[k] + C[C_DXJX1][k] * Jx[k] - C[C_DXJX0][k] * nowJx;
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      dcx nowJy = Jy[k];
      Jy[k] = C[C_JY][k] * Jy[k] + C[C_JYHZ][k] * (-Hz[k] + Hz[k - NPY]);
      Dy[k] = C[C_DY][k] * Dy[k] + C[C_DYJY1][k] * Jy[k] - C[C_DYJY0][k] * nowJy;
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      const int k = IDX(i, j);
      Ex[k] = Dx[k] / s->EPS[0][k];
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      const int k = IDX(i, j);
      Ey[k] = Dy[k] / s->EPS[1][k];
    }
}

/* calcMB + calcH (mpiTE_UPML.c:377-399, fdtdTE_upml.c:293-314) */
static void te_calc_mb_h(orc_sim *s)
{
  const int NPX = s->NPXa, NPY = s->NPYa;
  dcx *Ex = s->F[TE_EX], *Ey = s->F[TE_EY], *Hz = s->F[TE_HZ], *Mz = s->F[TE_MZ], *Bz = s->F[TE_BZ];
  double **C = s->C;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      dcx nowMz = Mz[k];
      Mz[k] = C[C_MZ][k] * Mz[k] - C[C_MZEXEY][k] * (Ey[k + NPY] - Ey[k] - Ex[k + 1] + Ex[k]);
      Bz[k] = C[C_BZ][k] * Bz[k] + C[C_BZMZ1][k] * Mz[k] - C[C_BZMZ0][k] * nowMz;
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      const int k = IDX(i, j);
      Hz[k] = Bz[k] / s->MU0;
    }
}

/* ------------------------------------------------------------------ sources */
/* scatteredWave of the MPI TM solver (mpiTM_UPML.c:337-374): one cexp per cell */
static void mpi_tm_scattered_wave(orc_sim *s, dcx *p, const double *eps)
{
  const int NPX = s->NPXa, NPY = s->NPYa;
  double time = s->time, w_s = s->Omega_s, ray_coef = s->ray_coef, k_s = s->K_s;
  double rad = s->Angle_deg * M_PI / 180;
  double ks_cos = cos(rad) * k_s, ks_sin = sin(rad) * k_s;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      int x = i - 1, y = j - 1;
      double kr = x * ks_cos + y * ks_sin;
      p[k] += ray_coef * (s->EPS0 / eps[k] - 1.0) * cexp(I * (kr - w_s * time));
    }
}

/* scatteredWave of the MPI TE solver (mpiTE_UPML.c:250-281), applied to Ey only */
static void mpi_te_scattered_wave(orc_sim *s, dcx *p, const double *eps)
{
  const int NPX = s->NPXa, NPY = s->NPYa;
  double time = s->time, w_s = s->Omega_s, ray_coef = s->ray_coef, k_s = s->K_s;
  double rad = s->Angle_deg * M_PI / 180.0;
  double ks_cos = cos(rad) * k_s, ks_sin = sin(rad) * k_s;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      int x = i - 1, y = j - 1;
      double ikx = x * ks_cos + y * ks_sin;
      p[k] += ray_coef * (s->EPS0 / eps[k] - 1) * (cos(ikx - w_s * time) + I * sin(ikx - w_s * time));
    }
}

/* field_scatteredPulse (field.c:224-256): Gaussian pulse, skips vacuum cells */
static void scattered_pulse(orc_sim *s, dcx *p, const double *eps, double gapX, double gapY, double dot)
{
  const int NPY = s->NPYa;
  double time = s->time, w_s = s->Omega_s;
  double rad = s->Angle_deg * M_PI / 180.0;
  double cos_per_c = cos(rad) / s->C0, sin_per_c = sin(rad) / s->C0;
  const double beam_width = 50;
  const double center_peak = (s->N_PX / 2.0 + gapX) * cos_per_c + (s->N_PY / 2 + gapY) * sin_per_c;
  const double t0 = -center_peak + 500;
  for (int i = 1; i < s->N_PX - 1; i++)
    for (int j = 1; j < s->N_PY - 1; j++) {
      int k = IDX(i, j);
      if (s->EPS0 == eps[k]) continue;
      const double r = (i + gapX) * cos_per_c + (j + gapY) * sin_per_c - (time - t0);
      const double gaussian_coef = exp(-pow(r / beam_width, 2));
      p[k] += dot * gaussian_coef * (s->EPS0 / eps[k] - 1) * cexp(I * r * w_s);
    }
}

/* ------------------------------------------------------------------ NTFF per step */
/* Accumulate into the flat [360][arraySize] array.  Indices outside the
   allocation (plus NT_GUARD bins each side) are dropped: the reference writes
   them out of bounds on strongly non-square grids (RFperC comes from the y
   extent only, field.c:140-142), which is undefined behaviour there.
   `a -= v` is written `a += -v`, which rounds identically. */
#define NT_ADD(slot, gidx, val)                                                   \
  do {                                                                            \
    long g_ = (long)(gidx);                                                       \
    if (g_ >= -NT_GUARD && g_ < 360L * s->arraySize + NT_GUARD) s->NT[slot][g_] += (val); \
  } while (0)

/* ntffCoef (mpiTM_UPML.c:840-847, mpiTE_UPML.c:602-609) */
static inline void ntff_coef(double time, double timeShift, int *m, double *a, double *b, double *ab)
{
  double t = time + timeShift;
  *m = floor(t + 0.5);
  *a = (0.5 + t - *m);
  *b = 1.0 - *a;
  *ab = *a - *b;
}

/* ntff() of the MPI TM solver (mpiTM_UPML.c:849-1037), single rank */
static void mpi_tm_ntff(orc_sim *s)
{
  const int SPX = s->NPXa, SPY = s->NPYa, NPY = SPY;
  const double C = s->C0, R = 1.0e6;
  const double coef = 1.0 / (4 * M_PI * C * R);
  const int cx = s->N_PX / 2, cy = s->N_PY / 2;
  double timeE = s->time - 1, timeH = s->time - 0.5;
  int tp = s->top, bm = s->bottom, rt = s->right, lt = s->left;
  dcx *Ez = s->F[TM_EZ], *Hx = s->F[TM_HX], *Hy = s->F[TM_HY];
  int m_e, m_h;
  double a_e, b_e, ab_e, a_h, b_h, ab_h;
  for (int ang = 0; ang < 360; ang++) {
    double rad = ang * M_PI / 180.0;
    double r1x = cos(rad), r1y = sin(rad);
    const int stp = ang * s->arraySize;
    if (0 < bm && bm < SPY - 1) {
      const int L = IMAX(1, lt), Rr = IMIN(SPX, rt);
      for (int i = L; i < Rr; i++) {
        const double r2x = i - cx, r2y = bm - cy;
        const double timeShift = -(r1x * r2x + r1y * r2y) / C + s->RFperC;
        ntff_coef(timeE, timeShift, &m_e, &a_e, &b_e, &ab_e);
        ntff_coef(timeH, timeShift, &m_h, &a_h, &b_h, &ab_h);
        int k = IDX(i, bm);
        const dcx ez = Ez[k];
        const dcx hx = 0.5 * (Hx[k] + Hx[k - 1]);
        NT_ADD(0, stp + m_e - 1, (ez * b_e * coef));
        NT_ADD(2, stp + m_h - 1, (hx * b_h * coef));
        NT_ADD(0, stp + m_e, (ez * ab_e * coef));
        NT_ADD(2, stp + m_h, (hx * ab_h * coef));
        NT_ADD(0, stp + m_e + 1, -(ez * a_e * coef));
        NT_ADD(2, stp + m_h + 1, -(hx * a_h * coef));
      }
    }
    if (0 < rt && rt < SPX - 1) {
      int T = IMIN(SPY, tp), B = IMAX(1, bm);
      for (int j = B; j < T; j++) {
        const double r2x = rt - cx, r2y = j - cy;
        const double timeShift = -(r1x * r2x + r1y * r2y) / C + s->RFperC;
        ntff_coef(timeE, timeShift, &m_e, &a_e, &b_e, &ab_e);
        ntff_coef(timeH, timeShift, &m_h, &a_h, &b_h, &ab_h);
        int k = IDX(rt, j);
        const dcx ez = Ez[k];
        const dcx hy = 0.5 * (Hy[k] + Hy[k - SPY]);
        NT_ADD(1, stp + m_e - 1, (ez * b_e * coef));
        NT_ADD(2, stp + m_h - 1, (hy * b_h * coef));
        NT_ADD(1, stp + m_e, (ez * ab_e * coef));
        NT_ADD(2, stp + m_h, (hy * ab_h * coef));
        NT_ADD(1, stp + m_e + 1, -(ez * a_e * coef));
        NT_ADD(2, stp + m_h + 1, -(hy * a_h * coef));
      }
    }
    if (0 < tp && tp < SPY - 1) {
      int L = IMAX(1, lt), Rr = IMIN(SPX, rt);
      for (int i = L; i < Rr; i++) {
        const double r2x = i - cx, r2y = tp - cy;
        const double timeShift = -(r1x * r2x + r1y * r2y) / C + s->RFperC;
        ntff_coef(timeE, timeShift, &m_e, &a_e, &b_e, &ab_e);
        ntff_coef(timeH, timeShift, &m_h, &a_h, &b_h, &ab_h);
        int k = IDX(i, tp);
        const dcx ez = -Ez[k];
        const dcx hx = -0.5 * (Hx[k] + Hx[k - 1]);
        NT_ADD(0, stp + m_e - 1, (ez * b_e * coef));
        NT_ADD(2, stp + m_h - 1, (hx * b_h * coef));
        NT_ADD(0, stp + m_e, (ez * ab_e * coef));
        NT_ADD(2, stp + m_h, (hx * ab_h * coef));
        NT_ADD(0, stp + m_e + 1, -(ez * a_e * coef));
        NT_ADD(2, stp + m_h + 1, -(hx * a_h * coef));
      }
    }
    if (0 < lt && lt < SPX) {
      int T = IMIN(SPY, tp), B = IMAX(1, bm);
      for (int j = B; j < T; j++) {
        const double r2x = lt - cx, r2y = j - cy;
        const double timeShift = -(r1x * r2x + r1y * r2y) / C + s->RFperC;
        ntff_coef(timeE, timeShift, &m_e, &a_e, &b_e, &ab_e);
        ntff_coef(timeH, timeShift, &m_h, &a_h, &b_h, &ab_h);
        int k = IDX(lt, j);
        const dcx ez = -Ez[k];
        const dcx hy = -0.5 * (Hy[k] + Hy[k - SPY]);
        NT_ADD(1, stp + m_e - 1, (ez * b_e * coef));
        NT_ADD(2, stp + m_h - 1, (hy * b_h * coef));
        NT_ADD(1, stp + m_e, (ez * ab_e * coef));
        NT_ADD(2, stp + m_h, (hy * ab_h * coef));
        NT_ADD(1, stp + m_e + 1, -(ez * a_e * coef));
        NT_ADD(2, stp + m_h + 1, -(hy * a_h * coef));
      }
    }
  }
}

/* ntff() of the MPI TE solver (mpiTE_UPML.c:611-794), single rank */
static void mpi_te_ntff(orc_sim *s)
{
  const int SPX = s->NPXa, SPY = s->NPYa, NPY = SPY;
  const double R = 1.0e6;
  const double coef = 1.0 / (4 * M_PI * s->C0 * R);
  const int cx = s->N_PX / 2, cy = s->N_PY / 2;
  double timeE = s->time - 1, timeH = s->time - 0.5;
  int tp = s->top, bm = s->bottom, rt = s->right, lt = s->left;
  dcx *Ex = s->F[TE_EX], *Ey = s->F[TE_EY], *Hz = s->F[TE_HZ];
  int m_e, m_h;
  double a_e, b_e, ab_e, a_h, b_h, ab_h;
  int ang, stp;
  for (ang = 0, stp = 0; ang < 360; ang++, stp += s->arraySize) {
    double rad = ang * M_PI / 180.0;
    double r1x = cos(rad), r1y = sin(rad);
    if (0 < bm && bm < SPY - 1) {
      const int L = IMAX(1, lt), Rr = IMIN(SPX, rt);
      for (int i = L; i < Rr; i++) {
        const double r2x = i - cx + 0.5, r2y = bm - cy;
        double timeShift = -(r1x * r2x + r1y * r2y) / s->C0 + s->RFperC;
        ntff_coef(timeE, timeShift, &m_e, &a_e, &b_e, &ab_e);
        ntff_coef(timeH, timeShift, &m_h, &a_h, &b_h, &ab_h);
        int k = IDX(i, bm);
        dcx ex = -Ex[k];
        dcx hz = -0.5 * (Hz[k] + Hz[k - 1]);
        NT_ADD(2, stp + m_e - 1, (ex * b_e * coef));
        NT_ADD(0, stp + m_h - 1, (hz * b_h * coef));
        NT_ADD(2, stp + m_e, (ex * ab_e * coef));
        NT_ADD(0, stp + m_h, (hz * ab_h * coef));
        NT_ADD(2, stp + m_e + 1, -(ex * a_e * coef));
        NT_ADD(0, stp + m_h + 1, -(hz * a_h * coef));
      }
    }
    if (0 < rt && rt < SPX - 1) {
      int T = IMIN(SPY, tp), B = IMAX(1, bm);
      for (int j = B; j < T; j++) {
        double r2x = rt - cx, r2y = j - cy + 0.5;
        double timeShift = -(r1x * r2x + r1y * r2y) / s->C0 + s->RFperC;
        ntff_coef(timeE, timeShift, &m_e, &a_e, &b_e, &ab_e);
        ntff_coef(timeH, timeShift, &m_h, &a_h, &b_h, &ab_h);
        int k = IDX(rt, j);
        dcx ey = -Ey[k];
        dcx hz = -0.5 * (Hz[k] + Hz[k - SPY]);
        NT_ADD(2, stp + m_e - 1, (ey * b_e * coef));
        NT_ADD(1, stp + m_h - 1, (hz * b_h * coef));
        NT_ADD(2, stp + m_e, (ey * ab_e * coef));
        NT_ADD(1, stp + m_h, (hz * ab_h * coef));
        NT_ADD(2, stp + m_e + 1, -(ey * a_e * coef));
        NT_ADD(1, stp + m_h + 1, -(hz * a_h * coef));
      }
    }
    if (0 < tp && tp < SPY - 1) {
      int L = IMAX(1, lt), Rr = IMIN(SPX, rt);
      for (int i = L; i < Rr; i++) {
        const double r2x = i - cx + 0.5, r2y = tp - cy;
        double timeShift = -(r1x * r2x + r1y * r2y) / s->C0 + s->RFperC;
        ntff_coef(timeE, timeShift, &m_e, &a_e, &b_e, &ab_e);
        ntff_coef(timeH, timeShift, &m_h, &a_h, &b_h, &ab_h);
        int k = IDX(i, tp);
        dcx ex = Ex[k];
        dcx hz = 0.5 * (Hz[k] + Hz[k - 1]);
        NT_ADD(2, stp + m_e - 1, (ex * b_e * coef));
        NT_ADD(0, stp + m_h - 1, (hz * b_h * coef));
        NT_ADD(2, stp + m_e, (ex * ab_e * coef));
        NT_ADD(0, stp + m_h, (hz * ab_h * coef));
        NT_ADD(2, stp + m_e + 1, -(ex * a_e * coef));
        NT_ADD(0, stp + m_h + 1, -(hz * a_h * coef));
      }
    }
    if (0 < lt && lt < SPX) {
      int T = IMIN(SPY, tp), B = IMAX(1, bm);
      for (int j = B; j < T; j++) {
        double r2x = lt - cx, r2y = j - cy + 0.5;
        double timeShift = -(r1x * r2x + r1y * r2y) / s->C0 + s->RFperC;
        ntff_coef(timeE, timeShift, &m_e, &a_e, &b_e, &ab_e);
        ntff_coef(timeH, timeShift, &m_h, &a_h, &b_h, &ab_h);
        int k = IDX(lt, j);
        dcx ey = Ey[k];
        dcx hz = 0.5 * (Hz[k] + Hz[k - SPY]);
        NT_ADD(2, stp + m_e - 1, (ey * b_e * coef));
        NT_ADD(1, stp + m_h - 1, (hz * b_h * coef));
        NT_ADD(2, stp + m_e, (ey * ab_e * coef));
        NT_ADD(1, stp + m_h, (hz * ab_h * coef));
        NT_ADD(2, stp + m_e + 1, -(ey * a_e * coef));
        NT_ADD(1, stp + m_h + 1, -(hz * a_h * coef));
      }
    }
  }
}

/* calc() of ntffTM.c:279-288 / ntffTE.c:57-66 (no coef) */
static inline void ntff_calc_at(orc_sim *s, int slot, long row, double tts, dcx eh)
{
  int m = floor(tts + 0.5);
  double a = (0.5 + tts - m);
  double b = 1.0 - a;
  double ab = a - b;
  NT_ADD(slot, row + m - 1, eh * b);
  NT_ADD(slot, row + m, eh * ab);
  NT_ADD(slot, row + m + 1, -(eh * a));
}

/* ntffTM_TimeCalc (ntffTM.c:293-371): the along-side time shift is accumulated
   by repeated subtraction, as the reference does */
static void plain_tm_ntff(orc_sim *s)
{
  const int NPY = s->NPYa;
  double timeE = s->time - 1, timeH = s->time - 0.5;
  double lt_cx = s->left - s->cx, rt_cx = s->right - s->cx;
  double bm_cy = s->bottom - s->cy, tp_cy = s->top - s->cy;
  int lb = IDX(s->left, s->bottom), ltI = IDX(s->left, s->top);
  int rb = IDX(s->right, s->bottom), rtI = IDX(s->right, s->top);
  dcx *Ez = s->F[TM_EZ], *Hx = s->F[TM_HX], *Hy = s->F[TM_HY];
  double ToRad = M_PI / 180.0;
  int index_ang = 0;
  for (int ang = 0; ang < 360; ang++, index_ang += s->arraySize) {
    double rad = ang * ToRad;
    double r1x_per_c = cos(rad) / s->C0, r1y_per_c = sin(rad) / s->C0;
    {
      const double r2x = lt_cx, r2y = bm_cy;
      double ts = -(r1x_per_c * r2x + r1y_per_c * r2y) + s->RFperC;
      for (int k = lb; k < rb; k += NPY) {
        ntff_calc_at(s, 0, index_ang, timeE + ts, Ez[k]);
        ntff_calc_at(s, 2, index_ang, timeH + ts, 0.5 * (Hx[k] + Hx[k - 1]));
        ts -= r1x_per_c;
      }
    }
    {
      const double r2x = rt_cx; double r2y = bm_cy;
      double ts = -(r1x_per_c * r2x + r1y_per_c * r2y) + s->RFperC;
      for (int k = rb; k < rtI; k++) {
        ntff_calc_at(s, 1, index_ang, timeE + ts, Ez[k]);
        ntff_calc_at(s, 2, index_ang, timeH + ts, 0.5 * (Hy[k] + Hy[k - NPY]));
        ts -= r1y_per_c;
      }
    }
    {
      double r2x = lt_cx; const double r2y = tp_cy;
      double ts = -(r1x_per_c * r2x + r1y_per_c * r2y) + s->RFperC;
      for (int k = ltI; k < rtI; k += NPY) {
        ntff_calc_at(s, 0, index_ang, timeE + ts, -Ez[k]);
        ntff_calc_at(s, 2, index_ang, timeH + ts, -0.5 * (Hx[k] + Hx[k - 1]));
        ts -= r1x_per_c;
      }
    }
    {
      const double r2x = lt_cx; double r2y = bm_cy;
      double ts = -(r1x_per_c * r2x + r1y_per_c * r2y) + s->RFperC;
      for (int k = lb; k < ltI; k++) {
        ntff_calc_at(s, 1, index_ang, timeE + ts, -Ez[k]);
        ntff_calc_at(s, 2, index_ang, timeH + ts, -0.5 * (Hy[k] + Hy[k - NPY]));
        ts -= r1y_per_c;
      }
    }
  }
}

/* ntffTE_TimeCalc (ntffTE.c:68-157) */
static void plain_te_ntff(orc_sim *s)
{
  const int NPY = s->NPYa;
  double timeE = s->time - 1, timeH = s->time - 0.5;
  double lt_cx = s->left - s->cx, rt_cx = s->right - s->cx;
  double bm_cy = s->bottom - s->cy, tp_cy = s->top - s->cy;
  int lb = IDX(s->left, s->bottom), ltI = IDX(s->left, s->top);
  int rb = IDX(s->right, s->bottom), rtI = IDX(s->right, s->top);
  dcx *Ex = s->F[TE_EX], *Ey = s->F[TE_EY], *Hz = s->F[TE_HZ];
  double ToRad = M_PI / 180.0;
  for (int ang = 0, stp = 0; ang < 360; ang++, stp += s->arraySize) {
    double rad = ang * ToRad;
    double r1x_per_c = cos(rad) / s->C0, r1y_per_c = sin(rad) / s->C0;
    {
      const double r2x = lt_cx + 0.5, r2y = bm_cy;
      double ts = -(r1x_per_c * r2x + r1y_per_c * r2y) + s->RFperC;
      for (int k = lb; k < rb; k += NPY) {
        dcx ex = -Ex[k];
        dcx hz = -0.5 * (Hz[k] + Hz[k - 1]);
        ntff_calc_at(s, 2, stp, timeE + ts, ex);
        ntff_calc_at(s, 0, stp, timeH + ts, hz);
        ts -= r1x_per_c;
      }
    }
    {
      const double r2x = rt_cx, r2y = bm_cy + 0.5;
      double ts = -(r1x_per_c * r2x + r1y_per_c * r2y) + s->RFperC;
      for (int k = rb; k < rtI; k++) {
        dcx ey = -Ey[k];
        dcx hz = -0.5 * (Hz[k] + Hz[k - NPY]);
        ntff_calc_at(s, 2, stp, timeE + ts, ey);
        ntff_calc_at(s, 1, stp, timeH + ts, hz);
        ts -= r1y_per_c;
      }
    }
    {
      const double r2x = lt_cx + 0.5, r2y = tp_cy;
      double ts = -(r1x_per_c * r2x + r1y_per_c * r2y) + s->RFperC;
      for (int k = ltI; k < rtI; k += NPY) {
        dcx ex = Ex[k];
        dcx hz = 0.5 * (Hz[k] + Hz[k - 1]);
        ntff_calc_at(s, 2, stp, timeE + ts, ex);
        ntff_calc_at(s, 0, stp, timeH + ts, hz);
        ts -= r1x_per_c;
      }
    }
    {
      const double r2x = lt_cx, r2y = bm_cy + 0.5;
      double ts = -(r1x_per_c * r2x + r1y_per_c * r2y) + s->RFperC;
      for (int k = lb; k < ltI; k++) {
        dcx ey = Ey[k];
        dcx hz = 0.5 * (Hz[k] + Hz[k - NPY]);
        ntff_calc_at(s, 2, stp, timeE + ts, ey);
        ntff_calc_at(s, 1, stp, timeH + ts, hz);
        ts -= r1y_per_c;
      }
    }
  }
}

/* ------------------------------------------------------------------ steps */
/* Split-field PML TM step (fdtdTM.c:269-327): calcH, calcE (Ezx, Ezy, then
   Ez = Ezx + Ezy), then the scattered wave on Ezx (field.c:179-199, gap 0);
   slots F[0] Ez, F[1] Ezx, F[2] Ezy, F[3] Hx, F[6] Hy */
static void split_tm_step(orc_sim *s)
{
  const int NPX = s->NPXa, NPY = s->NPYa;
  dcx *Ez = s->F[0], *Ezx = s->F[1], *Ezy = s->F[2], *Hx = s->F[3], *Hy = s->F[6];
  double **C = s->C;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Hx[k] = C[4][k] * Hx[k] - C[5][k] * (Ezx[k + 1] - Ezx[k] + Ezy[k + 1] - Ezy[k]);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Hy[k] = C[6][k] * Hy[k] + C[7][k] * (Ezx[k + NPY] - Ezx[k] + Ezy[k + NPY] - Ezy[k]);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Ezx[k] = C[0][k] * Ezx[k] + C[1][k] * (Hy[k] - Hy[k - NPY]);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Ezy[k] = C[2][k] * Ezy[k] - C[3][k] * (Hx[k] - Hx[k - 1]);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Ez[k] = Ezx[k] + Ezy[k];
    }
  double time = s->time, w_s = s->Omega_s, ray_coef = s->ray_coef, k_s = s->K_s;
  double rad = s->Angle_deg * M_PI / 180.0;
  double ks_cos = cos(rad) * k_s, ks_sin = sin(rad) * k_s;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      double kr = (i + 0.0) * ks_cos + (j + 0.0) * ks_sin;
      Ezx[k] += ray_coef * (s->EPS0 / s->EPS[0][k] - 1.0) *
                (cexp(I * (kr - w_s * (time + 0.5))) - cexp(I * (kr - w_s * (time - 0.5))));
    }
}

/* Split-field PML TE step (fdtdTE.c:281-321): calcE (Ex, Ey), the scattered wave
   on Ey (field.c:179-199, gap (0, 0.5)), calcH (Hzx, Hzy, then Hz = Hzx + Hzy);
   slots F[0] Ex, F[3] Ey, F[6] Hz, F[7] Hzx, F[8] Hzy */
static void split_te_step(orc_sim *s)
{
  const int NPX = s->NPXa, NPY = s->NPYa;
  dcx *Ex = s->F[0], *Ey = s->F[3], *Hz = s->F[6], *Hzx = s->F[7], *Hzy = s->F[8];
  double **C = s->C;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Ex[k] = C[0][k] * Ex[k] + C[1][k] * (Hzx[k] - Hzx[k - 1] + Hzy[k] - Hzy[k - 1]);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Ey[k] = C[2][k] * Ey[k] - C[3][k] * (Hzx[k] - Hzx[k - NPY] + Hzy[k] - Hzy[k - NPY]);
    }
  double time = s->time, w_s = s->Omega_s, ray_coef = s->ray_coef, k_s = s->K_s;
  double rad = s->Angle_deg * M_PI / 180.0;
  double ks_cos = cos(rad) * k_s, ks_sin = sin(rad) * k_s;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      double kr = (i + 0.0) * ks_cos + (j + 0.5) * ks_sin;
      Ey[k] += ray_coef * (s->EPS0 / s->EPS[1][k] - 1.0) *
               (cexp(I * (kr - w_s * (time + 0.5))) - cexp(I * (kr - w_s * (time - 0.5))));
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Hzx[k] = C[4][k] * Hzx[k] - C[5][k] * (Ey[k + NPY] - Ey[k]);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Hzy[k] = C[6][k] * Hzy[k] + C[7][k] * (Ex[k + 1] - Ex[k]);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Hz[k] = Hzx[k] + Hzy[k];
    }
}

/* Non-standard FDTD TM step (nsFdtdTM.c:68-151): calcH with the non-standard
   second differences (weight r/2, r = 1/6 + k^2/180 - k^4/23040), calcE (Ezx,
   Ezy), the scattered wave on Ezy (field.c:155-177, gap 0, dot 1), then
   Ez = Ezx + Ezy; slots as the split-field TM solver's */
static void ns_tm_step(orc_sim *s)
{
  const int NPX = s->NPXa, NPY = s->NPYa, dx = NPY, dy = 1;
  dcx *Ez = s->F[0], *Ezx = s->F[1], *Ezy = s->F[2], *Hx = s->F[3], *Hy = s->F[6];
  double **C = s->C;
  const double k_s = s->K_s;
  const double r = 1.0 / 6.0 + k_s * k_s / 180.0 - pow(k_s, 4) / 23040;
  const double r_2 = r / 2.0;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      dcx ns = r_2 * ((Ez[k + dy + dx] + Ez[k + dy - dx] - 2 * Ez[k + dy]) - (Ez[k + dx] + Ez[k - dx] - 2 * Ez[k]));
      Hx[k] = C[4][k] * Hx[k] - C[5][k] * (Ez[k + dy] - Ez[k] + ns);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      dcx ns = r_2 * ((Ez[k + dx + dy] + Ez[k + dx - dy] - 2 * Ez[k + dx]) - (Ez[k + dy] + Ez[k - dy] - 2 * Ez[k]));
      Hy[k] = C[6][k] * Hy[k] + C[7][k] * (Ez[k + dx] - Ez[k] + ns);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Ezx[k] = C[0][k] * Ezx[k] + C[1][k] * (Hy[k] - Hy[k - dx]);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Ezy[k] = C[2][k] * Ezy[k] - C[3][k] * (Hx[k] - Hx[k - dy]);
    }
  double time = s->time, w_s = s->Omega_s, ray_coef = s->ray_coef * 1.0;
  double rad = s->Angle_deg * M_PI / 180.0;
  double ks_cos = cos(rad) * k_s, ks_sin = sin(rad) * k_s;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      double kr = (i + 0) * ks_cos + (j + 0) * ks_sin;
      Ezy[k] += ray_coef * C[8][k] * (cexp(I * (kr - w_s * (time + 1.0))) - cexp(I * (kr - w_s * (time))));
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Ez[k] = Ezx[k] + Ezy[k];
    }
}

/* Non-standard FDTD TE step (nsFdtdTE.c:242-308): calcH (Hzx, Hzy), Hz = Hzx + Hzy,
   calcE with the non-standard second differences, then the NS scattered wave
   (field.c:155-177) on Ex with dot cos(angle + 90) and on Ey with sin(angle + 90),
   both at gap (0, 0.5); slots as the split-field TE solver's */
static void ns_te_step(orc_sim *s)
{
  const int NPX = s->NPXa, NPY = s->NPYa, dx = NPY, dy = 1;
  dcx *Ex = s->F[0], *Ey = s->F[3], *Hz = s->F[6], *Hzx = s->F[7], *Hzy = s->F[8];
  double **C = s->C;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Hzx[k] = C[4][k] * Hzx[k] - C[5][k] * (Ey[k + dx] - Ey[k]);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Hzy[k] = C[6][k] * Hzy[k] + C[7][k] * (Ex[k + dy] - Ex[k]);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      Hz[k] = Hzx[k] + Hzy[k];
    }
  const double k_s = s->K_s;
  const double r = 1.0 / 6.0 + k_s * k_s / 180.0 - pow(k_s, 4) / 23040;
  const double r_2 = r / 2.0;
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      dcx ns = r_2 * ((Hz[k + dx] + Hz[k - dx] - 2 * Hz[k]) - (Hz[k - dy + dx] + Hz[k - dy - dx] - 2 * Hz[k - dy]));
      Ex[k] = C[0][k] * Ex[k] + C[1][k] * (Hz[k] - Hz[k - dy] + ns);
    }
  for (int i = 1; i < NPX - 1; i++)
    for (int j = 1; j < NPY - 1; j++) {
      int k = IDX(i, j);
      dcx ns = r_2 * ((Hz[k + dy] + Hz[k - dy] - 2 * Hz[k]) - (Hz[k - dx + dy] + Hz[k - dx - dy] - 2 * Hz[k - dx]));
      Ey[k] = C[2][k] * Ey[k] - C[3][k] * (Hz[k] - Hz[k - dx] + ns);
    }
  double co = cos((s->Angle_deg + 90) * M_PI / 180.0);
  double si = sin((s->Angle_deg + 90) * M_PI / 180.0);
  double time = s->time, w_s = s->Omega_s;
  double rad = s->Angle_deg * M_PI / 180.0;
  double ks_cos = cos(rad) * k_s, ks_sin = sin(rad) * k_s;
  for (int c = 0; c < 2; c++) {
    double dot = c == 0 ? co : si;
    if (dot == 0.0) continue;
    dcx *p = c == 0 ? Ex : Ey;
    const double *fac = C[8 + c];
    double ray_coef = s->ray_coef * dot;
    for (int i = 1; i < NPX - 1; i++)
      for (int j = 1; j < NPY - 1; j++) {
        int k = IDX(i, j);
        double kr = (i + 0) * ks_cos + (j + 0.5) * ks_sin;
        p[k] += ray_coef * fac[k] * (cexp(I * (kr - w_s * (time + 1.0))) - cexp(I * (kr - w_s * (time))));
      }
  }
}

void orc_elliptic(const orc_sim *s, double *out181)
{
  /* TE_2D / NS_TE_2D: Ey (fdtdTE.c:269-274, nsFdtdTE.c:234-238) */
  const dcx *Ez = (s->cfg.solver == ORC_TE_2D || s->cfg.solver == ORC_NS_TE_2D) ? s->F[3] : s->F[0];
  const int NPY = s->NPYa;
  for (int ang = 180; ang >= 0; ang--) {
    double rad = ang * M_PI / 180.0;
    double x = 1.2 * s->lambda_s * cos(rad) + s->N_PX / 2.0;
    double y = 1.2 * s->lambda_s * sin(rad) + s->N_PY / 2.0;
    int i = (int)floor(x), j = (int)floor(y);
    if (i < 0 || j < 0 || i + 1 >= s->NPXa || j + 1 >= NPY) {
      out181[180 - ang] = 0.0;  /* the reference reads outside the array here (grids under 2.4 lambda) */
      continue;
    }
    double dx = x - i, dy = y - j;
    int k = i * NPY + j;
    dcx v = Ez[k] * (1.0 - dx) * (1.0 - dy) + Ez[k + NPY] * dx * (1.0 - dy) + Ez[k + 1] * (1.0 - dx) * dy +
            Ez[k + NPY + 1] * dx * dy;
    out181[180 - ang] = creal(v) * creal(v) + cimag(v) * cimag(v);
  }
}

void orc_step(orc_sim *s)
{
  switch (s->cfg.solver) {
  case ORC_TM_2D:
    split_tm_step(s);
    break;
  case ORC_TE_2D:
    split_te_step(s);
    break;
  case ORC_NS_TM_2D:
    ns_tm_step(s);
    break;
  case ORC_NS_TE_2D:
    ns_te_step(s);
    break;
  case ORC_MPI_TM_UPML: /* update (mpiTM_UPML.c:196-217) */
    tm_calc_jd_e(s);
    mpi_tm_scattered_wave(s, s->F[TM_EZ], s->EPS[0]);
    if (s->cfg.point_source) {
      /* field_pointLight (field.c:145-152) at global (N_PX/2, N_PY/2) */
      int k = (s->N_PX / 2 + 1) * s->NPYa + (s->N_PY / 2 + 1);
      s->F[TM_EZ][k] += s->ray_coef * cexp(I * s->Omega_s * s->time);
    }
    tm_calc_mb_h(s);
    if (s->cfg.ntff_time) mpi_tm_ntff(s);
    break;
  case ORC_MPI_TE_UPML: /* update (mpiTE_UPML.c:290-304) */
    te_calc_jd_e(s);
    mpi_te_scattered_wave(s, s->F[TE_EY], s->EPS[1]);
    te_calc_mb_h(s);
    if (s->cfg.ntff_time) mpi_te_ntff(s);
    break;
  case ORC_TM_UPML: /* update (fdtdTM_upml.c:54-66) */
    tm_calc_mb_h(s);
    tm_calc_jd_e(s);
    scattered_pulse(s, s->F[TM_EZ], s->EPS[0], 0, 0, 1.0);
    if (s->cfg.ntff_time) plain_tm_ntff(s);
    break;
  case ORC_TE_UPML: { /* update (fdtdTE_upml.c:168-192) */
    te_calc_mb_h(s);
    te_calc_jd_e(s);
    double co = cos((s->Angle_deg + 90) * M_PI / 180.0);
    double si = sin((s->Angle_deg + 90) * M_PI / 180.0);
    if (co != 0.0) scattered_pulse(s, s->F[TE_EX], s->EPS[0], 0.5, 0.0, co);
    if (si != 0.0) scattered_pulse(s, s->F[TE_EY], s->EPS[1], 0.0, 0.5, si);
    if (s->cfg.ntff_time) plain_te_ntff(s);
    break;
  }
  default:
    break;
  }
  field_next_step(s);
}

void orc_run(orc_sim *s, int nsteps)
{
  for (int n = 0; n < nsteps; n++) orc_step(s);
}

void orc_reset(orc_sim *s)
{
  size_t n = (size_t)s->NPXa * s->NPYa;
  for (int f = 0; f < 9; f++) memset(s->F[f], 0, n * sizeof(dcx));
  size_t nn = (size_t)360 * s->arraySize + 2 * NT_GUARD;
  for (int a = 0; a < 3; a++) memset(s->NTb[a], 0, nn * sizeof(dcx));
  s->time = 0;       /* field_reset (field.c:83-87) */
  s->ray_coef = 0;
}

void orc_set_angle(orc_sim *s, int deg) { s->cfg.angle_deg = deg; s->Angle_deg = deg; }

/* ------------------------------------------------------------------ getters */
void orc_dims(const orc_sim *s, int out[8])
{
  out[0] = s->N_PX; out[1] = s->N_PY; out[2] = s->NPXa; out[3] = s->NPYa;
  out[4] = s->off; out[5] = s->N_PML; out[6] = s->arraySize; out[7] = (int)s->maxTime;
}
void orc_ntff_info(const orc_sim *s, int out[6], double *rf)
{
  out[0] = s->top; out[1] = s->bottom; out[2] = s->left; out[3] = s->right; out[4] = s->cx; out[5] = s->cy;
  if (rf) *rf = s->RFperC;
}
double orc_time(const orc_sim *s) { return s->time; }
double orc_ray_coef(const orc_sim *s) { return s->ray_coef; }
double orc_omega(const orc_sim *s) { return s->Omega_s; }
double orc_k(const orc_sim *s) { return s->K_s; }

void orc_get_field(const orc_sim *s, int which, double *out)
{
  memcpy(out, s->F[which], (size_t)s->NPXa * s->NPYa * sizeof(dcx));
}
void orc_set_field(orc_sim *s, int which, const double *in)
{
  memcpy(s->F[which], in, (size_t)s->NPXa * s->NPYa * sizeof(dcx));
}
void orc_get_eps(const orc_sim *s, int which, double *out)
{
  memcpy(out, s->EPS[which], (size_t)s->NPXa * s->NPYa * sizeof(double));
}
void orc_get_ntff_time(const orc_sim *s, int which, double *out)
{
  memcpy(out, s->NT[which], (size_t)360 * s->arraySize * sizeof(dcx));
}

/* ------------------------------------------------------------------ far field */
/* ntffTM_Frequency (ntffTM.c:72-158) on N_PX x N_PY arrays */
static void ntff_frequency_tm_plain(const orc_sim *s, const dcx *Hx, const dcx *Hy, const dcx *Ez, dcx *res)
{
  const int NPY = s->N_PY;
  double cx = s->cx, cy = s->cy, k_s = s->K_s;
  dcx coef = csqrt(I * k_s / (8 * M_PI * s->R0)) * cexp(I * k_s * s->R0);
  int tp = s->top, bm = s->bottom, rt = s->right, lt = s->left;
  for (int ang = 0; ang < 360; ang++) {
    double rad = ang * M_PI / 180.0;
    double rx = cos(rad), ry = sin(rad), r2x, r2y;
    dcx Nz = 0, Lx = 0, Ly = 0, C_EZ, C_HX, C_HY;
    int i, j;
    for (i = lt; i < rt; i++) {
      r2x = i - cx; r2y = bm - cy;
      int k = IDX(i, bm);
      C_EZ = Ez[k];
      C_HX = 0.5 * (Hx[k] + Hx[k - 1]);
      double ip = rx * r2x + ry * r2y;
      Nz += C_HX * cexp(I * k_s * ip);
      Lx += C_EZ * cexp(I * k_s * ip);
    }
    for (j = bm; j < tp; j++) {
      r2x = rt - cx; r2y = j - cy;
      int k = IDX(rt, j);
      C_EZ = Ez[k];
      C_HY = 0.5 * (Hy[k] + Hy[k - NPY]);
      double ip = rx * r2x + ry * r2y;
      Nz += C_HY * cexp(I * k_s * ip);
      Ly += C_EZ * cexp(I * k_s * ip);
    }
    for (i = lt; i < rt; i++) {
      r2x = i - cx; r2y = tp - cy;
      int k = IDX(i, tp);
      C_EZ = Ez[k];
      C_HX = 0.5 * (Hx[k] + Hx[k - 1]);
      double ip = rx * r2x + ry * r2y;
      Nz -= C_HX * cexp(I * k_s * ip);
      Lx -= C_EZ * cexp(I * k_s * ip);
    }
    for (j = bm; j < tp; j++) {
      r2x = lt - cx; r2y = j - cy;
      int k = IDX(lt, j);
      C_EZ = Ez[k];
      C_HY = 0.5 * (Hy[k] + Hy[k - NPY]);
      double ip = rx * r2x + ry * r2y;
      Nz -= C_HY * cexp(I * k_s * ip);
      Ly -= C_EZ * cexp(I * k_s * ip);
    }
    dcx Lphi = -Lx * sin(rad) + Ly * cos(rad);
    res[ang] = coef * (s->Z0 * Nz + Lphi) * sqrt(s->cfg.h_u_nm);
  }
}

/* single-rank restatement of ntffTE_FrequencySplit (ntffTE.c:341-470) on the
   padded (N_PX+2)x(N_PY+2) layout, OFFSET = 0 */
static void ntff_frequency_te_padded(const orc_sim *s, const dcx *Ex, const dcx *Ey, const dcx *Hz, dcx *res)
{
  const int SPX = s->N_PX + 2, SPY = s->N_PY + 2, NPY = SPY;
  double cx = s->N_PX / 2, cy = s->N_PY / 2, K = s->K_s;
  dcx coef = csqrt(I * K / (8 * M_PI * s->R0)) * cexp(I * K * s->R0);
  int tp = s->top, bm = s->bottom, rt = s->right, lt = s->left;
  for (int ang = 0; ang < 360; ang++) {
    double rad = ang * M_PI / 180.0;
    double rx = cos(rad), ry = sin(rad);
    dcx Nx = 0, Ny = 0, Lz = 0;
    if (0 < bm && bm < SPY - 1) {
      const int L = IMAX(1, lt), Rr = IMIN(SPX - 1, rt);
      for (int i = L; i < Rr; i++) {
        const double r2x = i - cx + 0.5, r2y = bm - cy;
        int k = IDX(i, bm);
        dcx C_HZ = 0.5 * (Hz[k] + Hz[k - 1]);
        dcx C_EX = Ex[k];
        double ip = rx * r2x + ry * r2y;
        Nx -= C_HZ * cexp(I * K * ip);
        Lz -= C_EX * cexp(I * K * ip);
      }
    }
    if (0 < rt && rt < SPX - 1) {
      int T = IMIN(SPY - 1, tp), B = IMAX(1, bm);
      for (int j = B; j < T; j++) {
        double r2x = rt - cx, r2y = j - cy + 0.5;
        int k = IDX(rt, j);
        dcx C_HZ = 0.5 * (Hz[k] + Hz[k - SPY]);
        dcx C_EY = Ey[k];
        double ip = rx * r2x + ry * r2y;
        Ny -= C_HZ * cexp(I * K * ip);
        Lz -= C_EY * cexp(I * K * ip);
      }
    }
    if (0 < tp && tp < SPY - 1) {
      int L = IMAX(1, lt), Rr = IMIN(SPX - 1, rt);
      for (int i = L; i < Rr; i++) {
        const double r2x = i - cx + 0.5, r2y = tp - cy;
        int k = IDX(i, tp);
        dcx C_HZ = 0.5 * (Hz[k] + Hz[k - 1]);
        dcx C_EX = Ex[k];
        double ip = rx * r2x + ry * r2y;
        Nx += C_HZ * cexp(I * K * ip);
        Lz += C_EX * cexp(I * K * ip);
      }
    }
    if (0 < lt && lt < SPX - 1) {
      int T = IMIN(SPY - 1, tp), B = IMAX(1, bm);
      for (int j = B; j < T; j++) {
        const double r2x = lt - cx, r2y = j - cy + 0.5;
        int k = IDX(lt, j);
        dcx C_HZ = 0.5 * (Hz[k] + Hz[k - SPY]);
        dcx C_EY = Ey[k];
        double ip = rx * r2x + ry * r2y;
        Ny += C_HZ * cexp(I * K * ip);
        Lz += C_EY * cexp(I * K * ip);
      }
    }
    dcx Nphi = -Nx * sin(rad) + Ny * cos(rad);
    res[ang] = coef * (s->Z0 * Nphi - Lz) * sqrt(s->cfg.h_u_nm);
  }
}

void orc_ntff_frequency(const orc_sim *s, double *out360)
{
  dcx *res = (dcx *)out360;
  size_t np = (size_t)s->N_PX * s->N_PY;
  if (s->tm) {
    if (!s->mpi) {
      ntff_frequency_tm_plain(s, s->F[TM_HX], s->F[TM_HY], s->F[TM_EZ], res);
      return;
    }
    /* repack the padded arrays as mpiTM_UPML.c:542-544 does */
    dcx *P[3];
    int src[3] = {TM_HX, TM_HY, TM_EZ};
    for (int a = 0; a < 3; a++) {
      P[a] = (dcx *)malloc(np * sizeof(dcx));
      for (int x = 0; x < s->N_PX; x++)
        for (int y = 0; y < s->N_PY; y++)
          P[a][x * s->N_PY + y] = s->F[src[a]][(x + 1) * s->NPYa + (y + 1)];
    }
    ntff_frequency_tm_plain(s, P[0], P[1], P[2], res);
    for (int a = 0; a < 3; a++) free(P[a]);
  } else {
    if (s->mpi) {
      ntff_frequency_te_padded(s, s->F[TE_EX], s->F[TE_EY], s->F[TE_HZ], res);
      return;
    }
    /* plain TE: embed into the padded layout (ghost ring = 0) */
    size_t npp = (size_t)(s->N_PX + 2) * (s->N_PY + 2);
    dcx *P[3];
    int src[3] = {TE_EX, TE_EY, TE_HZ};
    for (int a = 0; a < 3; a++) {
      P[a] = (dcx *)calloc(npp, sizeof(dcx));
      for (int x = 0; x < s->N_PX; x++)
        for (int y = 0; y < s->N_PY; y++)
          P[a][(x + 1) * (s->N_PY + 2) + (y + 1)] = s->F[src[a]][x * s->NPYa + y];
    }
    ntff_frequency_te_padded(s, P[0], P[1], P[2], res);
    for (int a = 0; a < 3; a++) free(P[a]);
  }
}

/* ------------------------------------------------------------------ FFT + spectra */
/* fft1x with flag 0 (cfft.c:104-179): radix-2 decimation in frequency,
   twiddle cexp(+i*pi*k/xp2), 1-based bit-reversal, unnormalised */
void orc_cfft(double *data, int n)
{
  dcx *a = (dcx *)data;
  int iter = 0, i = n;
  while (i >>= 1) iter++;
  if (n < 2 || n != (1 << iter)) return;
  double sign = -1.;
  int xp2 = n;
  for (int it = 0; it < iter; it++) {
    int xp = xp2;
    xp2 = xp / 2;
    double w = -M_PI / xp2;
    for (int k = 0; k < xp2; k++) {
      dcx ww = cexp(I * sign * w * k);
      /* butterflies (k + m*xp, k + m*xp + xp2) for m = 0 .. n/xp-1 */
      for (int base = k; base < n; base += xp) {
        dcx *p = a + base, *q = a + base + xp2;
        dcx t = (*p) - (*q);
        *p = (*p) + (*q);
        *q = t * (ww);
      }
    }
  }
  int j1 = n / 2, j2 = n - 1, j = 1;
  for (i = 1; i <= j2; i++) {
    if (i < j) { dcx t = a[i - 1]; a[i - 1] = a[j - 1]; a[j - 1] = t; }
    int k = j1;
    while (k < j) { j -= k; k /= 2; }
    j += k;
  }
}

/* ntffTM_TimeTranslate (ntffTM.c:161-194) and ntffTE_TimeTranslate (ntffTE.c:20-55)
   for the plain solvers; for the MPI solvers the same transform as their
   ntffOutput (mpiTM_UPML.c:799-838, mpiTE_UPML.c:840-879), whose coefficient
   omits 1/(4 pi c) because the per-step accumulation already applied 1/(4 pi c R) */
void orc_time_translate(const orc_sim *s, double *eth_out, double *eph_out)
{
  dcx *Eth = (dcx *)eth_out, *Eph = (dcx *)eph_out;
  const double w_s = s->Omega_s;
  const dcx coef = s->mpi ? csqrt(2 * M_PI * s->C0 / (I * w_s))
                          : 1.0 / (4 * M_PI * s->C0) * csqrt(2 * M_PI * s->C0 / (I * w_s));
  const int maxTime = (int)s->maxTime;
  double theta = 0;
  const double ToRad = M_PI / 180.0;
  for (int ang = 0, k = 0; ang < 360; ang++, k += s->arraySize) {
    /* ntffTM_TimeTranslate forms the angle as ang*ToRad (ntffTM.c:169-173);
       ntffTE_TimeTranslate and the MPI ntffOutput as ang*M_PI/180.0 */
    double phi = (s->tm && !s->mpi) ? ang * ToRad : ang * M_PI / 180.0;
    double sx = cos(theta) * cos(phi), sy = cos(theta) * sin(phi), sz = -cos(theta);
    double px = -sin(phi), py = cos(phi);
    for (int i = 0; i < maxTime; i++) {
      dcx WTH, WPH, UTH, UPH;
      if (s->tm) {
        const dcx *Ux = s->NT[0], *Uy = s->NT[1], *Wz = s->NT[2];
        WTH = 0 + 0 + Wz[k + i] * sz;
        WPH = 0 + 0;
        UTH = Ux[k + i] * sx + Uy[k + i] * sy + 0;
        UPH = Ux[k + i] * px + Uy[k + i] * py;
      } else {
        const dcx *Wx = s->NT[0], *Wy = s->NT[1], *Uz = s->NT[2];
        WTH = Wx[k + i] * sx + Wy[k + i] * sy + 0;
        WPH = Wx[k + i] * px + Wy[k + i] * py;
        UTH = 0 + 0 + Uz[k + i] * sz;
        UPH = 0 + 0;
      }
      Eth[k + i] = coef * (-s->Z0 * WTH - UPH);
      Eph[k + i] = coef * (-s->Z0 * WPH + UTH);
    }
  }
}

/* ntffTM_TimeOutput / ntffTE_TimeOutput spectra (ntffTM.c:197-233, ntffTE.c:160-196).
   TM transforms Eth, TE transforms Eph.  Samples beyond NTFF_NUM are not copied
   (the reference overruns its 8192 buffer there, Q10). */
void orc_time_spectra(const orc_sim *s, double *out)
{
  size_t n = (size_t)360 * s->arraySize;
  dcx *Eth = (dcx *)calloc(n, sizeof(dcx)), *Eph = (dcx *)calloc(n, sizeof(dcx));
  orc_time_translate(s, (double *)Eth, (double *)Eph);
  const dcx *E = s->tm ? Eth : Eph;
  dcx *buf = (dcx *)malloc(NTFF_NUM * sizeof(dcx));
  int maxTime = (int)s->maxTime;
  int ncopy = maxTime < NTFF_NUM ? maxTime : NTFF_NUM;
  for (int ang = 0; ang < 360; ang++) {
    memset(buf, 0, NTFF_NUM * sizeof(dcx));
    memcpy(buf, &E[(size_t)ang * s->arraySize], ncopy * sizeof(dcx));
    orc_cfft((double *)buf, NTFF_NUM);
    for (int lambda_nm = LAMBDA_ST_NM; lambda_nm <= LAMBDA_EN_NM; lambda_nm++) {
      double p = s->C0 * s->cfg.h_u_nm * NTFF_NUM / lambda_nm;
      int index = floor(p);
      p = p - index;
      dcx e0 = buf[index], e1 = buf[index + 1];
      double n0 = creal(e0) * creal(e0) + cimag(e0) * cimag(e0);  /* cnorm (myComplex.c:34-39) */
      double n1 = creal(e1) * creal(e1) + cimag(e1) * cimag(e1);
      out[(lambda_nm - LAMBDA_ST_NM) * 360 + ang] = ((1 - p) * n0 + p * n1) / NTFF_NUM;
    }
  }
  free(buf);
  free(Eth);
  free(Eph);
}

/* eps map over the N_PX x N_PY physical cells: out[x*N_PY+y] = models_eps(x+dx, y+dy, mode) */
void orc_model_eps_map(const orc_sim *s, int mode, double dx, double dy, double *out)
{
  for (int x = 0; x < s->N_PX; x++)
    for (int y = 0; y < s->N_PY; y++)
      out[(size_t)x * s->N_PY + y] = models_eps(s, x + dx, y + dy, mode);
}
