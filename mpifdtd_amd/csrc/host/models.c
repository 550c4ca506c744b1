/*
 * models.c -- dielectric model interface (host side).
 *
 * Model selection, size requirements, parameter sweeps and the per-model
 * derived state follow models.c:126-194 and the five model files
 * (noModel.c, circleModel.c, multiLayerModel.c, zigzagModel.c,
 * morphoScaleModel.c).  The eps callbacks themselves are the shared
 * host/device functions of fdtd_model_eps.h; the solvers fill their eps maps
 * on the GPU from models_getState().
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>
#include "mpifdtd_field.h"
#include "mpifdtd_models.h"
#include "fdtd_model_eps.h"

#ifndef M_PI
#define M_PI 3.1415926535897932384626433832795
#endif

static enum MODEL model_ = NO_MODEL;
static fdtd_model_params p_;
static int p_init_ = 0;
static fdtd_model_state st_;

void models_defaultParams(fdtd_model_params *p)
{
  memset(p, 0, sizeof *p);
  p->circle_radius_nm = 500; p->circle_n = 1.6;                          /* circleModel.c:6-10 */
  p->ml_n0 = 1.0; p->ml_n1 = 1.56; p->ml_n0x = 1.0; p->ml_n1x = 1.1;      /* multiLayerModel.c:7-15 */
  p->ml_width_nm0 = 300; p->ml_width_nm1 = 300;
  p->ml_thick_nm0 = 100; p->ml_thick_nm1 = 90;
  p->ml_layer_num = 4;
  p->zz_width_nm = 300; p->zz_thick_nm = 80; p->zz_layer_num = 5;       /* zigzagModel.c:7-24 */
  p->zz_degree = 10; p->zz_n = 1.56;
  p->mo_width_nm = 300; p->mo_thick_nm0 = 90; p->mo_thick_nm1 = 90;     /* morphoScaleModel.c:18-60 */
  p->mo_layer_num = 6; p->mo_n = 1.56; p->mo_nx = 1.1;
}

static void ensure_params(void)
{
  if (!p_init_) {
    models_defaultParams(&p_);
    p_init_ = 1;
  }
}

void models_setParams(const fdtd_model_params *p) { p_ = *p; p_init_ = 1; }
void models_getParams(fdtd_model_params *p) { ensure_params(); *p = p_; }
enum MODEL models_getModel(void) { return model_; }
const fdtd_model_state *models_getState(void) { return &st_; }

void models_setModel(enum MODEL model)
{
  ensure_params();
  switch (model) {
  case NO_MODEL: case MIE_CYLINDER: case LAYER: case MORPHO_SCALE: case ZIGZAG:
    model_ = model;
    break;
  case CONCENTRIC_CIRCLE: /* models.c:116-124 */
    printf("not implemented concentricCircle Model");
    exit(2);
  default:
    printf("model %d is not supported by the UPML solvers of this library\n", (int)model);
    exit(2);
  }
}

double models_eps(double x, double y, enum MODE mode)
{
  switch (mode) {
  case D_X: return fdtd_model_eps(&st_, x, y, 1, 0);
  case D_Y: return fdtd_model_eps(&st_, x, y, 0, 1);
  default: return fdtd_model_eps(&st_, x, y, 1, 1);
  }
}

/* sweep steps: circleModel.c:105-109, multiLayerModel.c:235-269,
   zigzagModel.c:80-104, morphoScaleModel.c:256-285 */
int models_isFinish(void)
{
  ensure_params();
  fdtd_model_params *p = &p_;
  switch (model_) {
  case NO_MODEL:
    return 1;
  case MIE_CYLINDER:
    p->circle_radius_nm += 100;
    return p->circle_radius_nm > 500;
  case LAYER: {
    p->ml_left_gap_nm += 0;
    if (p->ml_left_gap_nm >= p->ml_thick_nm0 + p->ml_thick_nm1 || !p->ml_use_gap) {
      p->ml_left_gap_nm = 0;
      p->ml_thick_nm0 += 10;
      if (p->ml_thick_nm0 > p->ml_thick_nm1 + 50) {
        p->ml_thick_nm1 += 10;
        p->ml_thick_nm0 = p->ml_thick_nm1 + 10;
        if (p->ml_thick_nm1 > 150) {
          p->ml_thick_nm1 = 90;
          p->ml_edge_rate += 0.5;
          if (p->ml_edge_rate > 1.0) {
            p->ml_edge_rate = 0.0;
            p->ml_branch_nm += 50;
            if (p->ml_branch_nm > 50) {
              p->ml_branch_nm = 0;
              p->ml_layer_num += 4;
              if (p->ml_layer_num > 10) {
                printf("there are no models which hasn't been simulated yet\n");
                return 1;
              }
            }
          }
        }
      }
    }
    return 0;
  }
  case ZIGZAG:
    p->zz_degree += 10;
    if (p->zz_degree > 80) {
      p->zz_degree = 10;
      p->zz_thick_nm += 10;
      if (p->zz_thick_nm > 150) {
        p->zz_thick_nm = 80;
        p->zz_width_nm += 10;
        if (p->zz_width_nm > 300) {
          p->zz_width_nm = 300;
          p->zz_layer_num += 2;
          if (p->zz_layer_num > 11) {
            printf("there are no models which hasn't been simulated yet\n");
            return 1;
          }
        }
      }
    }
    return 0;
  case MORPHO_SCALE:
    p->mo_left_gap_nm = 0;
    p->mo_thick_nm0 += 10;
    p->mo_thick_nm1 += 10;
    if (p->mo_thick_nm1 > 160) {
      p->mo_thick_nm0 = 90;
      p->mo_thick_nm1 = 90;
      p->mo_edge_rate += 0.5;
      if (p->mo_edge_rate > 1.0) {
        p->mo_edge_rate = 0.0;
        p->mo_branch_nm += 25;
        if (p->mo_branch_nm > 50) {
          p->mo_layer_num += 2;
          p->mo_branch_nm = 0;
          if (p->mo_layer_num > 10) {
            printf("there are no models which hasn't been simulated yet\n");
            return 1;
          }
        }
      }
    }
    return 0;
  default:
    return 1;
  }
}

/* noModel.c:36-40, circleModel.c:128-133, multiLayerModel.c:277-283,
   zigzagModel.c:137-144, morphoScaleModel.c:435-440 */
void models_needSize(int *x_nm, int *y_nm)
{
  ensure_params();
  const fdtd_model_params *p = &p_;
  switch (model_) {
  case MIE_CYLINDER:
    *x_nm = (int)(2.3 * p->circle_radius_nm * 2);
    *y_nm = (int)(2.3 * p->circle_radius_nm * 2);
    break;
  case LAYER:
    *x_nm = (p->ml_width_nm0 > p->ml_width_nm1 ? p->ml_width_nm0 : p->ml_width_nm1) + p->ml_branch_nm;
    *y_nm = (p->ml_thick_nm0 + p->ml_thick_nm1) * p->ml_layer_num + (p->ml_use_gap ? p->ml_thick_nm0 + p->ml_thick_nm0 : 0);
    break;
  case ZIGZAG: {
    double max_rad = 80 * M_PI / 180.0;
    *x_nm = (int)(cos(max_rad) * p->zz_width_nm + p->zz_thick_nm * sin(max_rad));
    *y_nm = (int)(sin(max_rad) * p->zz_width_nm * p->zz_layer_num + (1 + p->zz_layer_num) * p->zz_thick_nm * cos(max_rad));
    break;
  }
  case MORPHO_SCALE:
    *x_nm = p->mo_width_nm + p->mo_branch_nm;
    *y_nm = (p->mo_thick_nm0 + p->mo_thick_nm1) * p->mo_layer_num;
    break;
  default:
    *x_nm = 1000;
    *y_nm = 1000;
  }
}

/* circleModel.c:119-126, multiLayerModel.c:331-348, zigzagModel.c:146-160,
   morphoScaleModel.c:379-422: derived state in cell units */
void models_initModel(void)
{
  ensure_params();
  const fdtd_model_params *p = &p_;
  FieldInfo_S fs = field_getFieldInfo_S();
  const double eps0 = 1.0;
  memset(&st_, 0, sizeof st_);
  st_.model = (int)model_;
  st_.N_PX = fs.N_PX; st_.N_PY = fs.N_PY; st_.N_X = fs.N_X; st_.N_Y = fs.N_Y; st_.N_PML = fs.N_PML;
  st_.eps0 = eps0;
  switch (model_) {
  case MIE_CYLINDER:
    st_.ci_radius_s = field_toCellUnit(p->circle_radius_nm);
    st_.ci_posx = fs.N_PX / 2;
    st_.ci_posy = fs.N_PY / 2;
    st_.ci_eps_s = p->circle_n * p->circle_n * eps0;
    break;
  case LAYER:
    st_.ml_layer_num = p->ml_layer_num;
    st_.ml_uniaxial = p->ml_uniaxial; st_.ml_asymmetry = p->ml_asymmetry; st_.ml_use_gap = p->ml_use_gap;
    st_.ml_width_s[0] = field_toCellUnit(p->ml_width_nm0);
    st_.ml_width_s[1] = field_toCellUnit(p->ml_width_nm1);
    st_.ml_thick_s[0] = field_toCellUnit(p->ml_thick_nm0);
    st_.ml_thick_s[1] = field_toCellUnit(p->ml_thick_nm1);
    st_.ml_gap_s = field_toCellUnit(p->ml_left_gap_nm);
    st_.ml_ep[0] = p->ml_n0 * p->ml_n0 * eps0;
    st_.ml_ep[1] = p->ml_n1 * p->ml_n1 * eps0;
    st_.ml_epx[0] = p->ml_n0x * p->ml_n0x * eps0;
    st_.ml_epx[1] = p->ml_n1x * p->ml_n1x * eps0;
    st_.ml_branch_s = field_toCellUnit(p->ml_branch_nm);
    st_.ml_c0 = -4 * st_.ml_width_s[0] * p->ml_curve / st_.ml_thick_s[0] / st_.ml_thick_s[0];
    st_.ml_c1 = -4 * st_.ml_width_s[1] * p->ml_curve / st_.ml_thick_s[1] / st_.ml_thick_s[1];
    st_.ml_edge_rate = p->ml_edge_rate;
    break;
  case ZIGZAG:
    st_.zz_layer_num = p->zz_layer_num;
    st_.zz_rad = p->zz_degree * M_PI / 180.0;
    st_.zz_width_s = field_toCellUnit(p->zz_width_nm);
    st_.zz_thick_s = field_toCellUnit(p->zz_thick_nm);
    st_.zz_height_s = (st_.zz_width_s * sin(st_.zz_rad)) * p->zz_layer_num + p->zz_layer_num * st_.zz_thick_s / 2 * cos(st_.zz_rad);
    st_.zz_exw = st_.zz_thick_s / 2.0 * sin(st_.zz_rad);
    st_.zz_exh = st_.zz_thick_s / 2.0 * cos(st_.zz_rad);
    st_.zz_ox = fs.N_PX / 2.0;
    st_.zz_oy = (fs.N_PY - st_.zz_height_s) / 2.0;
    st_.zz_ep = p->zz_n * p->zz_n * eps0;
    break;
  case MORPHO_SCALE: {
    int L = p->mo_layer_num;
    if (L > FDTD_MAX_LAYERS) {
      fprintf(stderr, "morpho scale: %d layers exceed the supported %d\n", L, FDTD_MAX_LAYERS);
      exit(2);
    }
    st_.mo_layer_num = L;
    st_.mo_uniaxial = p->mo_uniaxial;
    st_.mo_width_s = field_toCellUnit(p->mo_width_nm);
    st_.mo_thick_s[0] = field_toCellUnit(p->mo_thick_nm0);
    st_.mo_thick_s[1] = field_toCellUnit(p->mo_thick_nm1);
    st_.mo_branch_s = field_toCellUnit(p->mo_branch_nm);
    st_.mo_gap_s = field_toCellUnit(p->mo_left_gap_nm);
    st_.mo_height_s = (st_.mo_thick_s[0] + st_.mo_thick_s[1]) * L + (st_.mo_thick_s[0] + st_.mo_thick_s[1]);
    st_.mo_ep = p->mo_n * p->mo_n * eps0;
    st_.mo_epx = p->mo_nx * p->mo_nx * eps0;
    st_.mo_oy = fs.N_PY / 2 - st_.mo_height_s / 2;
    st_.mo_ox = fs.N_PX / 2;
    st_.mo_edge_rate = p->mo_edge_rate;
    st_.mo_curve = p->mo_curve;
    srand(p->mo_seed);
    double TO_RAD = M_PI / 180.0;
    int R = p->mo_randomness;
    for (int i = 0; i < L; i++) {
      st_.mo_rand[1][i] = (rand() % (R + 1) - R / 2) * TO_RAD;
      st_.mo_rand[0][i] = (rand() % (R + 1) - R / 2) * TO_RAD;
    }
    double thick = st_.mo_thick_s[0] + st_.mo_thick_s[1];
    double half_thick = 0.5 * st_.mo_thick_s[0];
    double off_l = p->mo_asymmetry ? st_.mo_thick_s[1] + half_thick : half_thick;
    double off_r = half_thick;
    for (int i = 0; i < L; i++) {
      st_.mo_lam_a[0][i] = tan(st_.mo_rand[0][i]);
      st_.mo_lam_b1[0][i] = thick * i + off_l - half_thick / cos(st_.mo_rand[0][i]);
      st_.mo_lam_b2[0][i] = thick * i + off_l + half_thick / cos(st_.mo_rand[0][i]);
      st_.mo_lam_a[1][i] = tan(st_.mo_rand[1][i]);
      st_.mo_lam_b1[1][i] = thick * i + off_r - half_thick / cos(st_.mo_rand[1][i]);
      st_.mo_lam_b2[1][i] = thick * i + off_r + half_thick / cos(st_.mo_rand[1][i]);
    }
    break;
  }
  default:
    break;
  }
}

static void make_and_move(const char *name)
{
  mkdir(name, 0775);
  if (chdir(name) != 0) {
    printf("cannot move to %s\n", name);
    exit(2);
  }
}

/* models.c:132-137 + the per-model moveDirectory functions */
void models_moveDirectory(void)
{
  ensure_params();
  const fdtd_model_params *p = &p_;
  char buf[512];
  switch (model_) {
  case NO_MODEL:
    make_and_move("NoModel");
    make_and_move("NoModel");
    break;
  case MIE_CYLINDER:
    make_and_move("MieCylinderModel");
    snprintf(buf, sizeof buf, "radius_%dnm", (int)p->circle_radius_nm);
    make_and_move(buf);
    break;
  case LAYER:
    make_and_move("MultiLayerModel");
    make_and_move(p->ml_asymmetry ? "asymmetry" : "symmetry");
    if (p->ml_uniaxial) snprintf(buf, sizeof buf, "uniaxial_n0y%.2lf_n0x%.2lf_n1x%.2lf_n1x%.2lf", p->ml_n0, p->ml_n0x, p->ml_n1, p->ml_n1x);
    else snprintf(buf, sizeof buf, "n_%.2lf_%.2lf", p->ml_n0, p->ml_n1);
    make_and_move(buf);
    snprintf(buf, sizeof buf, "curve_%.2lf", p->ml_curve); make_and_move(buf);
    snprintf(buf, sizeof buf, "width%d_%d", p->ml_width_nm0, p->ml_width_nm0); make_and_move(buf);
    snprintf(buf, sizeof buf, "thick%d_%d", p->ml_thick_nm0, p->ml_thick_nm1); make_and_move(buf);
    snprintf(buf, sizeof buf, "gap%d", p->ml_left_gap_nm); make_and_move(buf);
    snprintf(buf, sizeof buf, "layer%d", p->ml_layer_num); make_and_move(buf);
    snprintf(buf, sizeof buf, "edge%.1lf", p->ml_edge_rate); make_and_move(buf);
    snprintf(buf, sizeof buf, "branch%d", p->ml_branch_nm); make_and_move(buf);
    break;
  case ZIGZAG:
    make_and_move("ZigZagModel");
    snprintf(buf, sizeof buf, "n_%.2lf", p->zz_n); make_and_move(buf);
    snprintf(buf, sizeof buf, "width%d_thick%d_layer%d", p->zz_width_nm, p->zz_thick_nm, p->zz_layer_num); make_and_move(buf);
    snprintf(buf, sizeof buf, "degree_%d", p->zz_degree); make_and_move(buf);
    break;
  case MORPHO_SCALE:
    make_and_move("MorphoScaleModel");
    make_and_move(p->mo_asymmetry ? "asymmetry" : "symmetry");
    if (p->mo_uniaxial) snprintf(buf, sizeof buf, "uniaxial_ny%.2lf_nx%.2lf", p->mo_n, p->mo_nx);
    else snprintf(buf, sizeof buf, "n_%.2lf", p->mo_n);
    make_and_move(buf);
    snprintf(buf, sizeof buf, "curve_%.2lf", p->mo_curve); make_and_move(buf);
    snprintf(buf, sizeof buf, "width%d", p->mo_width_nm); make_and_move(buf);
    snprintf(buf, sizeof buf, "thick%d_%d", p->mo_thick_nm0, p->mo_thick_nm1); make_and_move(buf);
    snprintf(buf, sizeof buf, "gap%d", p->mo_left_gap_nm); make_and_move(buf);
    snprintf(buf, sizeof buf, "layer%d", p->mo_layer_num); make_and_move(buf);
    snprintf(buf, sizeof buf, "edge%.1lf", p->mo_edge_rate); make_and_move(buf);
    snprintf(buf, sizeof buf, "branch%d", p->mo_branch_nm); make_and_move(buf);
    snprintf(buf, sizeof buf, "random%d_%d", p->mo_randomness, p->mo_seed); make_and_move(buf);
    break;
  default:
    break;
  }
}
