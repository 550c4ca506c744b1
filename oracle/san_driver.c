/*
 * san_driver.c -- runs every function of the CPU oracle (fdtd_oracle.c) once on
 * small inputs, for the AddressSanitizer / UndefinedBehaviorSanitizer build of
 * `make -C oracle sanitize` (oracle/_san/san_driver).  TEST INFRASTRUCTURE
 * ONLY: the oracle is the parity checker, and this driver only looks for out of
 * bounds accesses, leaks, overflows and other undefined behaviour in it (the
 * sanitizers abort on the first finding); it also checks that every output is
 * finite.  Exit status 0 when every case ran clean.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "fdtd_oracle.h"

static int bad = 0;

static void finite_or_fail(const char *what, const double *v, size_t n)
{
  for (size_t i = 0; i < n; i++)
    if (!isfinite(v[i])) {
      fprintf(stderr, "%s: non-finite value at %zu\n", what, i);
      bad = 1;
      return;
    }
}

static void run_case(const char *name, int solver, int model, int w, int h, int hu, int pml, int angle, int steps,
                     int ntff_time, int point_source)
{
  orc_config c;
  orc_default_config(&c);
  c.width_nm = w;
  c.height_nm = h;
  c.h_u_nm = hu;
  c.pml = pml;
  c.lambda_nm = 500;
  c.angle_deg = angle;
  c.stepNum = steps;
  c.solver = solver;
  c.model = model;
  c.ntff_time = ntff_time;
  c.point_source = point_source;
  orc_sim *s = orc_create(&c);
  orc_run(s, steps);
  int d[8];
  orc_dims(s, d);
  const size_t n = (size_t)d[2] * d[3];
  double *f = malloc(2 * n * sizeof(double));
  for (int q = 0; q < 9; q++) {
    orc_get_field(s, q, f);
    finite_or_fail(name, f, 2 * n);
    if (q == 0) orc_set_field(s, q, f);
  }
  double *e = malloc(n * sizeof(double));
  for (int q = 0; q < 3; q++) orc_get_eps(s, q, e);
  const int upml = solver == ORC_TM_UPML || solver == ORC_TE_UPML || solver == ORC_MPI_TM_UPML ||
                   solver == ORC_MPI_TE_UPML;
  if (upml) {
    double ff[720];
    orc_ntff_frequency(s, ff);
    finite_or_fail(name, ff, 720);
  }
  if (upml && ntff_time) {
    const size_t na = (size_t)360 * d[6];
    double *u = malloc(2 * na * sizeof(double)), *v = malloc(2 * na * sizeof(double));
    for (int q = 0; q < 3; q++) {
      orc_get_ntff_time(s, q, u);
      finite_or_fail(name, u, 2 * na);
    }
    orc_time_translate(s, u, v);
    double *sp = malloc(321 * 360 * sizeof(double));
    orc_time_spectra(s, sp);
    finite_or_fail(name, sp, 321 * 360);
    free(sp);
    free(u);
    free(v);
  }
  if (!upml) {
    double el[181];
    orc_elliptic(s, el);
  }
  double *m = malloc((size_t)d[0] * d[1] * sizeof(double));
  orc_model_eps_map(s, ORC_D_XY, 0.0, 0.0, m);
  orc_model_eps_map(s, ORC_D_X, 0.0, 0.5, m);
  orc_model_eps_map(s, ORC_D_Y, 0.5, 0.0, m);
  finite_or_fail(name, m, (size_t)d[0] * d[1]);
  (void)orc_sigma_x(s, 3.5, 0);
  (void)orc_sigma_y(s, 0, 3.5);
  orc_set_angle(s, angle + 30);
  orc_reset(s);
  orc_run(s, 3);
  free(m);
  free(e);
  free(f);
  orc_destroy(s);
  fprintf(stderr, "%s ok\n", name);
}

int main(void)
{
  /* every solver of enum SOLVER; square and non-square grids (the reference's
     NTFF bins overflow on strongly non-square ones, which the oracle guards) */
  run_case("mpi_tm_mie", ORC_MPI_TM_UPML, ORC_MIE_CYLINDER, 2000, 2000, 50, 8, 0, 60, 1, 0);
  run_case("mpi_tm_point", ORC_MPI_TM_UPML, ORC_NO_MODEL, 1500, 1500, 50, 8, 0, 40, 1, 1);
  run_case("mpi_te_zigzag", ORC_MPI_TE_UPML, ORC_ZIGZAG, 2400, 1800, 50, 8, 30, 60, 1, 0);
  run_case("tm_pulse_morpho", ORC_TM_UPML, ORC_MORPHO_SCALE, 2400, 2000, 50, 8, 90, 60, 1, 0);
  run_case("te_pulse_layer", ORC_TE_UPML, ORC_LAYER, 2000, 2400, 50, 8, 45, 60, 1, 0);
  run_case("mpi_tm_rect", ORC_MPI_TM_UPML, ORC_MIE_CYLINDER, 3000, 1500, 50, 8, 60, 50, 1, 0);
  run_case("split_tm", ORC_TM_2D, ORC_MIE_CYLINDER, 2000, 2000, 50, 8, 0, 40, 0, 0);
  run_case("split_te", ORC_TE_2D, ORC_MIE_CYLINDER, 2000, 2000, 50, 8, 0, 40, 0, 0);
  run_case("ns_tm", ORC_NS_TM_2D, ORC_ZIGZAG, 2000, 2000, 50, 8, 30, 40, 0, 0);
  run_case("ns_te", ORC_NS_TE_2D, ORC_LAYER, 2000, 2000, 50, 8, 45, 40, 0, 0);
  double buf[2 * 64];
  for (int i = 0; i < 128; i++) buf[i] = sin(0.37 * i);
  orc_cfft(buf, 64);
  finite_or_fail("cfft", buf, 128);
  if (bad) return 1;
  fprintf(stderr, "san_driver: every case ran clean\n");
  return 0;
}
