/*
 * ref_harness.c -- driver that runs the reference's own UPML solvers (compiled
 * from the sources under /root/reference, see oracle/Makefile target `ref`) on a
 * given configuration and dumps raw outputs.  It replaces the reference's
 * main.c, whose model/solver choice is hard-coded (main.c:155-156).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py to write the
 * golden fixtures that pin oracle/fdtd_oracle.c, and timed (mode=time) as the
 * CPU baseline of bench.py.  Never shipped or called by the product.
 *
 * usage: ref_harness key=value ...
 *   mode=run|eps   solver=<enum SOLVER>  model=<enum MODEL>  sweep=<models_isFinish calls>
 *   width=<nm> height=<nm> h=<nm> pml=<cells> lambda=<nm> angle=<deg> steps=<n>
 *   out=<directory>
 * mode=run : simulator_init, `steps` x simulator_calc, then dumps (solvers TM_2D, TE_2D, NS_TM/TE_2D,
 *            TM/TE_UPML_2D, MPI_TM/TE_UPML_2D)
 *            f0/f1/f2.bin (TM: Ez,Hx,Hy  TE: Ex,Ey,Hz; solver's own layout),
 *            eps.bin (getEps), meta.txt, ff.bin (TM: ntffTM_Frequency on the
 *            fields repacked to N_PX x N_PY, after ntffTM_init), then
 *            plain solvers: simulator_reset() -> spectra file in `out`;
 *            MPI TE: finish() -> MPI_TE_UPML/Eth_*.txt, Eph_*.txt in `out`.
 * mode=eps : field_init + models_initModel, dumps eps_xy.bin (D_XY at (x,y)),
 *            eps_dy.bin (D_Y at (x+.5,y)), eps_dx.bin (D_X at (x,y+.5)) for all
 *            N_PX x N_PY cells.
 * mode=time: simulator_init, `warm` x simulator_calc, then `steps` x
 *            simulator_calc timed with CLOCK_MONOTONIC; prints one JSON line
 *            {"N_PX", "N_PY", "steps", "seconds", "ntff_seconds"} and writes
 *            nothing.  ntff=1 (plain TM): then `steps` x ntffTM_TimeCalc alone on
 *            the same fields (ntff_seconds; -1 otherwise).  The CPU baseline of
 *            bench.py (the reference timed beside the GPU run).
 * mode=util: field_init, `steps` x field_nextStep, then field.h / myComplex.h
 *            helpers on fixed inputs: eps.bin, base.bin (the inputs), src.bin
 *            (field_scatteredWave, _NotUPML, field_nsScatteredWaveNotUPML and
 *            field_scatteredPulse each added to a copy of base), scal.bin
 *            (pmlCoef, _LXY, ns_beta, ns_beta_aster on a 3 x 3 set), bil.bin /
 *            nrm.bin (cbilinear, cnorm), meta.txt (subdomain getters, ind), and
 *            ell.txt, all_c.txt, all_d.txt from the three writers.
 * mode=config: path=<file>: the 11 values main.c's readConfig (main.c:319-366,
 *            commented out in the reference) takes from a config.txt, through
 *            the reference's own parser_nextLine (parser.c:3-20) and the same
 *            atoi / strtod conversions, printed as one JSON list; [] when the
 *            file holds fewer than 11 value lines.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#include <time.h>
#include <unistd.h>
#include <sys/stat.h>
#include <mpi.h>
#include "simulator.h"
#include "field.h"
#include "models.h"
#include "ntffTM.h"
#include "mpiTM_UPML.h"
#include "mpiTE_UPML.h"
#include "fdtdTM_upml.h"
#include "fdtdTE_upml.h"
#include "fdtdTM.h"
#include "fdtdTE.h"
#include "nsFdtdTM.h"
#include "nsFdtdTE.h"
#include "parser.h"

static const char *arg(int argc, char **argv, const char *key, const char *def)
{
  size_t n = strlen(key);
  for (int i = 1; i < argc; i++)
    if (strncmp(argv[i], key, n) == 0 && argv[i][n] == '=') return argv[i] + n + 1;
  return def;
}

static void dump(const char *dir, const char *name, const void *p, size_t bytes)
{
  char path[1024];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *fp = fopen(path, "wb");
  if (!fp) { perror(path); exit(3); }
  fwrite(p, 1, bytes, fp);
  fclose(fp);
}

int main(int argc, char **argv)
{
  const char *mode = arg(argc, argv, "mode", "run");
  int solver = atoi(arg(argc, argv, "solver", "4"));
  int model = atoi(arg(argc, argv, "model", "1"));
  int sweep = atoi(arg(argc, argv, "sweep", "0"));
  FieldInfo fi;
  fi.width_nm = atoi(arg(argc, argv, "width", "4000"));
  fi.height_nm = atoi(arg(argc, argv, "height", "4000"));
  fi.h_u_nm = atoi(arg(argc, argv, "h", "50"));
  fi.pml = atoi(arg(argc, argv, "pml", "10"));
  fi.lambda_nm = atoi(arg(argc, argv, "lambda", "500"));
  fi.angle_deg = atoi(arg(argc, argv, "angle", "0"));
  fi.stepNum = atoi(arg(argc, argv, "steps", "100"));
  const char *out = arg(argc, argv, "out", ".");
  mkdir(out, 0755);

  if (strcmp(mode, "config") == 0) {
    FILE *fp = fopen(arg(argc, argv, "path", "config.txt"), "r");
    if (!fp) { perror("config"); return 3; }
    char buf[1024];
    int v[11];
    for (int i = 0; i < 11; i++) {
      if (!parser_nextLine(fp, buf)) { printf("[]\n"); fclose(fp); return 0; }
      v[i] = i == 4 ? (int)strtod(buf, NULL) : atoi(buf); /* lambda: strtod, main.c:340 */
    }
    fclose(fp);
    printf("[");
    for (int i = 0; i < 11; i++) printf("%d%s", v[i], i < 10 ? ", " : "]\n");
    return 0;
  }

  MPI_Init(0, 0);
  models_setModel((enum MODEL)model);
  for (int k = 0; k < sweep; k++) models_isFinish();

  if (strcmp(mode, "util") == 0) {
    field_init(fi);
    for (int t = 0; t < fi.stepNum; t++) field_nextStep();
    FieldInfo_S s = field_getFieldInfo_S();
    size_t n = (size_t)s.N_PX * s.N_PY;
    double *eps = malloc(n * sizeof(double));
    double complex *base = malloc(n * sizeof(double complex)), *src = malloc(4 * n * sizeof(double complex));
    for (int x = 0; x < s.N_PX; x++)
      for (int y = 0; y < s.N_PY; y++) {
        size_t k = (size_t)x * s.N_PY + y;
        eps[k] = ((x * 7 + y * 3) % 5 == 0) ? 1.0 : 1.0 + ((x * 13 + y * 7) % 11) / 4.0;
        base[k] = (x - 0.5 * y) + I * (0.25 * x * y / s.N_PX);
      }
    for (int q = 0; q < 4; q++) memcpy(src + q * n, base, n * sizeof(double complex));
    field_scatteredWave(src, eps, 0.5, 0);
    field_scatteredWaveNotUPML(src + n, eps, 0, 0.5);
    field_nsScatteredWaveNotUPML(src + 2 * n, eps, 0, 0.5, 0.6);
    field_scatteredPulse(src + 3 * n, eps, 0.5, 0, 0.8);
    double scal[36];
    const double av[3] = {0, 0.3, 1.7}, bv[3] = {0.2, 1.0, 2.5};
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double *o = scal + 4 * (3 * a + b);
        o[0] = field_pmlCoef(1 + av[a], bv[b]);
        o[1] = field_pmlCoef_LXY(1 + av[a], bv[b]);
        o[2] = field_ns_beta(av[a], bv[b]);
        o[3] = field_ns_beta_aster(av[a], bv[b]);
      }
    const double px[5] = {1.0, 2.5, 10.25, s.N_PX / 2.0 + 0.3, s.N_PX - 2.5};
    const double py[5] = {1.0, 3.75, 20.5, s.N_PY / 2.0 - 0.7, s.N_PY - 2.25};
    double complex bil[5];
    double nrm[5];
    for (int q = 0; q < 5; q++) {
      bil[q] = cbilinear(base, px[q], py[q], s.N_PX, s.N_PY);
      nrm[q] = cnorm(bil[q]);
    }
    dump(out, "eps.bin", eps, n * sizeof(double));
    dump(out, "base.bin", base, n * sizeof(double complex));
    dump(out, "src.bin", src, 4 * n * sizeof(double complex));
    dump(out, "scal.bin", scal, sizeof scal);
    dump(out, "bil.bin", bil, sizeof bil);
    dump(out, "nrm.bin", nrm, sizeof nrm);
    char meta[256];
    int len = snprintf(meta, sizeof meta, "%d %d %d %d %d %d %d %d %d %d\n", s.N_PX, s.N_PY, field_getOffsetX(),
                       field_getOffsetY(), field_getSubNx(), field_getSubNy(), field_getSubNpx(), field_getSubNpy(),
                       field_getSubNcell(), ind(3, 4));
    dump(out, "meta.txt", meta, len);
    if (chdir(out) != 0) { perror(out); return 3; }
    field_outputElliptic("ell.txt", src);
    field_outputAllDataComplex("all_c.txt", src + n);
    field_outputAllDataDouble("all_d.txt", eps);
    MPI_Finalize();
    return 0;
  }

  if (strcmp(mode, "eps") == 0) {
    field_init(fi);
    models_initModel();
    FieldInfo_S s = field_getFieldInfo_S();
    size_t n = (size_t)s.N_PX * s.N_PY;
    double *a = malloc(n * sizeof(double)), *b = malloc(n * sizeof(double)), *c = malloc(n * sizeof(double));
    for (int x = 0; x < s.N_PX; x++)
      for (int y = 0; y < s.N_PY; y++) {
        size_t k = (size_t)x * s.N_PY + y;
        a[k] = models_eps(x, y, D_XY);
        b[k] = models_eps(x + 0.5, y, D_Y);
        c[k] = models_eps(x, y + 0.5, D_X);
      }
    dump(out, "eps_xy.bin", a, n * sizeof(double));
    dump(out, "eps_dy.bin", b, n * sizeof(double));
    dump(out, "eps_dx.bin", c, n * sizeof(double));
    char meta[256];
    int len = snprintf(meta, sizeof meta, "%d %d\n", s.N_PX, s.N_PY);
    dump(out, "meta.txt", meta, len);
    MPI_Finalize();
    return 0;
  }

  simulator_setSolver((enum SOLVER)solver);
  simulator_init(fi);
  if (strcmp(mode, "time") == 0) {
    const int warm = atoi(arg(argc, argv, "warm", "2"));
    for (int n = 0; n < warm; n++) simulator_calc();
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int n = 0; n < fi.stepNum; n++) simulator_calc();
    clock_gettime(CLOCK_MONOTONIC, &t1);
    FieldInfo_S st = field_getFieldInfo_S();
    /* ntff=1 (plain TM): the per-step time-domain NTFF the plain TM update calls
     * (fdtdTM_upml.c:65), timed alone for the same number of calls on the same
     * fields into accumulators of the reference's size, so that the step's rate
     * without it follows by difference (the update's other parts are file-static) */
    double ntff_s = -1.0;
    if (atoi(arg(argc, argv, "ntff", "0")) && solver == TM_UPML_2D) {
      const size_t na = (size_t)360 * field_getNTFFInfo().arraySize;
      double complex *U0 = calloc(na, sizeof *U0), *U1 = calloc(na, sizeof *U1), *U2 = calloc(na, sizeof *U2);
      if (!U0 || !U1 || !U2) { fprintf(stderr, "out of memory\n"); _exit(1); }
      struct timespec n0, n1;
      clock_gettime(CLOCK_MONOTONIC, &n0);
      for (int n = 0; n < fi.stepNum; n++)
        ntffTM_TimeCalc(fdtdTM_upml_getHx(), fdtdTM_upml_getHy(), fdtdTM_upml_getEz(), U0, U1, U2);
      clock_gettime(CLOCK_MONOTONIC, &n1);
      ntff_s = (double)(n1.tv_sec - n0.tv_sec) + 1e-9 * (double)(n1.tv_nsec - n0.tv_nsec);
    }
    printf("{\"N_PX\": %d, \"N_PY\": %d, \"steps\": %d, \"seconds\": %.6f, \"ntff_seconds\": %.6f}\n", st.N_PX,
           st.N_PY, fi.stepNum, (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec), ntff_s);
    fflush(stdout);
    _exit(0);  /* no finish(): it would write the reference's output files */
  }
  for (int n = 0; n < fi.stepNum; n++) simulator_calc();

  FieldInfo_S s = field_getFieldInfo_S();
  int mpi = (solver == MPI_TM_UPML_2D || solver == MPI_TE_UPML_2D);
  int tm = (solver == MPI_TM_UPML_2D || solver == TM_UPML_2D || solver == TM_2D || solver == NS_TM_2D);
  int npx = mpi ? s.N_PX + 2 : s.N_PX, npy = mpi ? s.N_PY + 2 : s.N_PY;
  size_t n = (size_t)npx * npy;
  double complex *f[3];
  double *eps;
  if (solver == MPI_TM_UPML_2D) {
    f[0] = mpi_fdtdTM_upml_getEz(); f[1] = mpi_fdtdTM_upml_getHx(); f[2] = mpi_fdtdTM_upml_getHy();
    eps = mpi_fdtdTM_upml_getEps();
  } else if (solver == MPI_TE_UPML_2D) {
    f[0] = mpi_fdtdTE_upml_getEx(); f[1] = mpi_fdtdTE_upml_getEy(); f[2] = mpi_fdtdTE_upml_getHz();
    eps = mpi_fdtdTE_upml_getEps();
  } else if (solver == TM_UPML_2D) {
    f[0] = fdtdTM_upml_getEz(); f[1] = fdtdTM_upml_getHx(); f[2] = fdtdTM_upml_getHy();
    eps = fdtdTM_upml_getEps();
  } else if (solver == TM_2D) {
    f[0] = fdtdTM_getEz(); f[1] = fdtdTM_getHx(); f[2] = fdtdTM_getHy();
    eps = fdtdTM_getEps();
  } else if (solver == TE_2D) {
    f[0] = fdtdTE_getEx(); f[1] = fdtdTE_getEy(); f[2] = fdtdTE_getHz();
    eps = fdtdTE_getEps();
  } else if (solver == NS_TE_2D) {
    f[0] = nsFdtdTE_getEx(); f[1] = nsFdtdTE_getEy(); f[2] = nsFdtdTE_getHz();
    eps = nsFdtdTE_getEps();
  } else if (solver == NS_TM_2D) {
    f[0] = nsFdtdTM_getEz(); f[1] = nsFdtdTM_getHx(); f[2] = nsFdtdTM_getHy();
    eps = nsFdtdTM_getEps();
  } else {
    f[0] = fdtdTE_upml_getEx(); f[1] = fdtdTE_upml_getEy(); f[2] = fdtdTE_upml_getHz();
    eps = fdtdTE_upml_getEps();
  }
  dump(out, "f0.bin", f[0], n * sizeof(double complex));
  dump(out, "f1.bin", f[1], n * sizeof(double complex));
  dump(out, "f2.bin", f[2], n * sizeof(double complex));
  /* the MPI ghost ring of eps is uninitialised in the reference: zero it */
  double *e = malloc(n * sizeof(double));
  memcpy(e, eps, n * sizeof(double));
  if (mpi)
    for (int i = 0; i < npx; i++)
      for (int j = 0; j < npy; j++)
        if (i == 0 || j == 0 || i == npx - 1 || j == npy - 1) e[i * npy + j] = 0;
  dump(out, "eps.bin", e, n * sizeof(double));
  NTFFInfo ni = field_getNTFFInfo();
  char meta[512];
  int len = snprintf(meta, sizeof meta, "%d %d %d %d %d %d %d %d %d %d %.17g %d\n", s.N_PX, s.N_PY, npx, npy,
                     ni.top, ni.bottom, ni.left, ni.right, ni.cx, ni.cy, ni.RFperC, ni.arraySize);
  dump(out, "meta.txt", meta, len);

  if (tm) {
    /* repack to N_PX x N_PY as mpiTM_UPML.c:542-544 does, then the plain-layout far field */
    size_t np = (size_t)s.N_PX * s.N_PY;
    double complex *P[3];
    int order[3] = {1, 2, 0}; /* Hx, Hy, Ez */
    for (int a = 0; a < 3; a++) {
      P[a] = malloc(np * sizeof(double complex));
      for (int x = 0; x < s.N_PX; x++)
        for (int y = 0; y < s.N_PY; y++)
          P[a][(size_t)x * s.N_PY + y] = mpi ? f[order[a]][(size_t)(x + 1) * npy + (y + 1)] : f[order[a]][(size_t)x * npy + y];
    }
    double complex res[360];
    ntffTM_init();
    ntffTM_Frequency(P[0], P[1], P[2], res);
    dump(out, "ff.bin", res, sizeof res);
  }

  if (chdir(out) != 0) { perror(out); return 3; }
  if (!mpi) {
    simulator_reset(); /* plain UPML solvers write the spectra files here, TM_2D / TE_2D /
                          NS_TM_2D / NS_TE_2D tm_, te_, ns_tm_, ns_te_<h>nm.txt */
  } else if (!tm) {
    mkdir("MPI_TE_UPML", 0755);
    void (*fin)(void) = mpi_fdtdTE_upml_getFinish();
    fin(); /* ntffOutput -> Eth/Eph text files; also calls MPI_Finalize */
    return 0;
  }
  MPI_Finalize();
  return 0;
}
