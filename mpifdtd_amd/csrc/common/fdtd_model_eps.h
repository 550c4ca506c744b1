/*
 * fdtd_model_eps.h -- permittivity callbacks of the five in-scope models.
 *
 * Written in the common subset of C99 and HIP C++ so that one source serves
 * both the host-side models_eps() (compiled by gcc) and the device eps-fill
 * kernel (compiled by hipcc for gfx950).  Arithmetic follows the reference's
 * operation order so the maps agree with it bit for bit; pow(v, 2) is written
 * v*v (glibc returns the correctly rounded square for exponent 2).
 *
 *   noModel      noModel.c:5-8
 *   circle       circleModel.c:72-103
 *   multilayer   multiLayerModel.c:87-185
 *   zigzag       zigzagModel.c:42-78
 *   morphoScale  morphoScaleModel.c:101-253
 * col/row select the sub-sampling axis: models_eps maps D_X -> (1,0),
 * D_Y -> (0,1), D_XY -> (1,1) (models.c:166-184).
 */
#ifndef FDTD_MODEL_EPS_H
#define FDTD_MODEL_EPS_H

#include <math.h>
#include "fdtd_hip.h"

#if defined(__HIPCC__)
#define FDTD_HD __host__ __device__ static inline
#else
#define FDTD_HD static inline
#endif

FDTD_HD double fdtd_sq(double v) { return v * v; }

FDTD_HD double fdtd_eps_circle(const fdtd_model_state *m, double x, double y, int col, int row)
{
  if (x < m->N_PML || y < m->N_PML || x > m->N_X + m->N_PML || y > m->N_Y + m->N_PML) return m->eps0;
  double dx = x - m->ci_posx, dy = y - m->ci_posy;
  double len = dx * dx + dy * dy;
  double r = m->ci_radius_s;
  if (len >= (r + 1) * (r + 1)) return m->eps0;
  if (len <= (r - 1) * (r - 1)) return m->ci_eps_s;
  double split = 10, half = split / 2, sum = 0;
  for (double i = -half + 0.5; i < half; i += 1)
    for (double j = -half + 0.5; j < half; j += 1)
      if (fdtd_sq(dx + col * i / split) + fdtd_sq(dy + row * j / split) <= r * r) sum += 1;
  sum /= split * split;
  return m->ci_eps_s * sum + m->eps0 * (1 - sum);
}

FDTD_HD double fdtd_ml_width(const fdtd_model_state *m, double sx, double sy, double wid, double hei, double modY, int k)
{
  double p = 1 - sy / hei;
  double new_wid = (wid + m->ml_branch_s) * (p + (1 - p) * m->ml_edge_rate);
  double dh = k == 0 ? modY : modY - m->ml_thick_s[0];
  double c = k == 0 ? m->ml_c0 : m->ml_c1;
  if (m->ml_asymmetry && sx < 0) dh = (k == 1 ? modY : modY - m->ml_thick_s[1]);
  return c * fdtd_sq(dh - m->ml_thick_s[k] / 2) + new_wid;
}

FDTD_HD double fdtd_eps_multilayer(const fdtd_model_state *m, double x, double y, int col, int row)
{
  double width = m->ml_width_s[0] > m->ml_width_s[1] ? m->ml_width_s[0] : m->ml_width_s[1];
  double thick = m->ml_thick_s[0] + m->ml_thick_s[1];
  double height = thick * m->ml_layer_num + m->ml_gap_s;
  int oy = (int)(m->N_PY / 2 - height / 2);
  int ox = m->N_PX / 2;
  double _x = x - ox, _y = y - oy;
  if (fabs(_x) > (width / 2 + 0.5) || _y < -0.5 || _y > height + 0.5) return m->eps0;
  height = thick * m->ml_layer_num;
  double s0 = 0, s1 = 0;
  double split = 10, half = split / 2;
  for (double i = -half + 0.5; i < half; i += 1) {
    for (double j = -half + 0.5; j < half; j += 1) {
      double sx = _x + col * i / split;
      double sy = _y + row * j / split;
      if (sx < 0 && m->ml_use_gap) sy -= m->ml_gap_s;
      if (sy < 0 || sy > height) continue;
      double p = 1 - sy / height;
      if (fabs(sx) < m->ml_branch_s * (p + (1 - p) * m->ml_edge_rate)) { s1 += 1; continue; }
      double modY = sy - floor(sy / thick) * thick;
      if (modY == m->ml_thick_s[0]) {
        s0 += 0.5 * (fabs(sx) < m->ml_width_s[0] / 2);
        s1 += 0.5 * (fabs(sx) < m->ml_width_s[1] / 2);
        continue;
      }
      int k;
      if (sx < 0 && m->ml_asymmetry) k = (modY < m->ml_thick_s[1]);
      else k = (modY > m->ml_thick_s[0]);
      double wid = fdtd_ml_width(m, sx, sy, m->ml_width_s[k], height, modY, k);
      if (fabs(sx) < wid / 2) {
        if (k) s1 += 1; else s0 += 1;
      }
    }
  }
  s0 /= split * split;
  s1 /= split * split;
  if (m->ml_uniaxial && col == 0 && row == 1)
    return m->eps0 * (1 - s0 - s1) + m->ml_epx[0] * s0 + m->ml_epx[1] * s1;
  return m->eps0 * (1 - s0 - s1) + m->ml_ep[0] * s0 + m->ml_ep[1] * s1;
}

FDTD_HD double fdtd_eps_zigzag(const fdtd_model_state *m, double _x, double _y, int col, int row)
{
  double x = _x - m->zz_ox, y = _y - m->zz_oy;
  double rad = m->zz_rad, ws = m->zz_width_s, ts = m->zz_thick_s, hs = m->zz_height_s;
  if (fabs(x) > ws / 2.0 + m->zz_exw + 0.5 || y < -m->zz_exh - 0.5 || y > hs + m->zz_exh + 0.5) return m->eps0;
  double sm = 0, split = 10, half = split / 2;
  for (double i = -half + 0.5; i < half; i += 1) {
    for (double j = -half + 0.5; j < half; j += 1) {
      double sx = x + col * i / split;
      double sy = y + row * j / split;
      if (fabs(sx) > ws / 2 + m->zz_exw || sy < -m->zz_exh || sy > hs + m->zz_exh) continue;
      /* the reference's min/max macros evaluate in double before truncating to int */
      double q = sy / (ws * sin(rad) + cos(rad) * ts / 2);
      double mx = (0 > q ? 0 : q);
      int k = (int)((m->zz_layer_num - 1) < mx ? (double)(m->zz_layer_num - 1) : mx);
      double px = -ws * cos(rad) / 2.0;
      double py = floor((double)((k + 1) / 2)) * 2 * (ws * sin(rad)) + k * cos(rad) * ts / 2;
      double vx = (1 - ((k & 1) << 1)) * cos(rad);
      double vy = sin(rad);
      double sqrDist = fdtd_sq(sx - px) + fdtd_sq(sy - py) - fdtd_sq((sx - px) * vx + (sy - py) * vy);
      if (sqrDist <= ts * ts) sm += 1;
    }
  }
  sm /= split * split;
  return m->eps0 * (1 - sm) + m->zz_ep * sm;
}

FDTD_HD int fdtd_in_lamela(double a, double b1, double b2, double x, double y)
{
  return (a * x + b1 <= y && y <= a * x + b2);
}

FDTD_HD double fdtd_mo_width(const fdtd_model_state *m, double sx, double sy, int side, int id)
{
  double a = m->mo_lam_a[side][id], b1 = m->mo_lam_b1[side][id], b2 = m->mo_lam_b2[side][id];
  double b = (sy - sx * a);
  int signX = sx < 0 ? 0 : 1;
  double rad = m->mo_rand[signX][id];
  double offset = (signX == 0 ? -1 : 1) * (0.5 * (b2 + b1) - b) * sin(rad);
  double p = 1 - b / m->mo_height_s;
  double width_p = (p + (1 - p) * m->mo_edge_rate);
  double yy = (b - b1) / (b2 - b1) - 0.5;
  double half_width = m->mo_width_s / 2.0;
  double c0 = 4.0 * half_width * m->mo_curve;
  return width_p * (half_width * (1 - m->mo_curve)) + c0 * (0.25 - yy * yy) + offset;
}

/* calcLamela (morphoScaleModel.c:128-145): index of the first lamela whose band
 * holds a corner of the cell square, or -1 */
FDTD_HD int fdtd_mo_find(const fdtd_model_state *m, double lft, double rht, double btm, double top)
{
  int side = lft < 0 ? 0 : 1;
  for (int i = 0; i < m->mo_layer_num; i++) {
    double a = m->mo_lam_a[side][i], b1 = m->mo_lam_b1[side][i], b2 = m->mo_lam_b2[side][i];
    if (fdtd_in_lamela(a, b1, b2, lft, btm) || fdtd_in_lamela(a, b1, b2, rht, btm) ||
        fdtd_in_lamela(a, b1, b2, lft, top) || fdtd_in_lamela(a, b1, b2, rht, top))
      return i;
  }
  return -1;
}

FDTD_HD double fdtd_eps_morpho(const fdtd_model_state *m, double x, double y, int col, int row)
{
  double _x = x - m->mo_ox, _y = y - m->mo_oy;
  if (fabs(_x) > 0.5 * sqrt(fdtd_sq(m->mo_width_s) + fdtd_sq(m->mo_thick_s[0])) + 0.5) return m->eps0;
  int lid[2] = {-1, -1};
  double gap = m->mo_gap_s;
  if (fabs(_x) > 0.5) {
    if (_x < 0) lid[0] = fdtd_mo_find(m, _x - 0.5, _x + 0.5, _y - gap - 0.5, _y - gap + 0.5);
    else lid[1] = fdtd_mo_find(m, _x - 0.5, _x + 0.5, _y - 0.5, _y + 0.5);
  } else {
    lid[0] = fdtd_mo_find(m, _x - 0.5, 0, _y - gap - 0.5, _y - gap + 0.5);
    lid[1] = fdtd_mo_find(m, 0, _x + 0.5, _y - 0.5, _y + 0.5);
  }
  double sm = 0, split = 10, half = split / 2;
  for (double i = -half + 0.5; i < half; i += 1) {
    for (double j = -half + 0.5; j < half; j += 1) {
      double sx = _x + col * i / split;
      double sy = _y + row * j / split;
      if (sx < 0) sy -= gap;
      if (sy >= 0 && sy <= m->mo_height_s) {
        double p = 1 - sy / m->mo_height_s;
        double width_p = (p + (1 - p) * m->mo_edge_rate);
        if (fabs(sx) < m->mo_branch_s / 2.0 * width_p) { sm += 1; continue; }
      }
      int side = sx < 0 ? 0 : 1;
      int id = lid[side];
      if (id < 0) continue;
      double a = m->mo_lam_a[side][id];
      if (fdtd_in_lamela(a, m->mo_lam_b1[side][id], m->mo_lam_b2[side][id], sx, sy)) {
        double sqrWidth = fdtd_sq(sx) * (1 + fdtd_sq(a));
        double newWidth = fdtd_mo_width(m, sx, sy, side, id);
        if (sqrWidth < fdtd_sq(newWidth)) { sm += 1; continue; }
      }
      for (int k = -1; k < 2; k += 2) {
        int id2 = id + k;
        if (id2 < 0 || id2 >= m->mo_layer_num) continue;
        double a2 = m->mo_lam_a[side][id2];
        if (fdtd_in_lamela(a2, m->mo_lam_b1[side][id2], m->mo_lam_b2[side][id2], sx, sy)) {
          double sqrWidth = fdtd_sq(sx) * (1 + fdtd_sq(a2));
          double newWidth = fdtd_mo_width(m, sx, sy, side, id2);
          if (sqrWidth < fdtd_sq(newWidth)) { sm += 1; break; }
        }
      }
    }
  }
  sm /= split * split;
  if (m->mo_uniaxial && col == 0 && row == 1) return m->eps0 * (1 - sm) + m->mo_epx * sm;
  return m->eps0 * (1 - sm) + m->mo_ep * sm;
}

FDTD_HD double fdtd_model_eps(const fdtd_model_state *m, double x, double y, int col, int row)
{
  switch (m->model) {
  case FDTD_MIE_CYLINDER: return fdtd_eps_circle(m, x, y, col, row);
  case FDTD_LAYER: return fdtd_eps_multilayer(m, x, y, col, row);
  case FDTD_ZIGZAG: return fdtd_eps_zigzag(m, x, y, col, row);
  case FDTD_MORPHO_SCALE: return fdtd_eps_morpho(m, x, y, col, row);
  default: return m->eps0;
  }
}

#endif
