/*
 * ntff_out.c -- spectra writers and the config.txt reader.
 *
 * ntff_outputEnormTxt / ntff_outputEnormBin keep the reference's file formats
 * (ntff.c:6-32), including the text writer's "%lf.20" conversion, which
 * prints six decimals followed by a literal ".20".  The binary file holds
 * 321 x 360 doubles, wavelength-major.
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mpifdtd_simulator.h"

static FILE *open_or_die(const char *name, const char *mode)
{
  FILE *fp = fopen(name, mode);
  if (!fp) {
    printf("cannot open file %s \n", name);
    exit(2);
  }
  return fp;
}

void ntff_outputEnormTxt(double **e_norm, const char *file_name)
{
  FILE *fp = open_or_die(file_name, "w");
  for (int lambda_nm = LAMBDA_ST_NM; lambda_nm <= LAMBDA_EN_NM; lambda_nm++) {
    fprintf(fp, "%d ", lambda_nm);
    for (int ang = 0; ang < 360; ang++) fprintf(fp, "%lf.20 ", e_norm[lambda_nm - LAMBDA_ST_NM][ang]);
    fprintf(fp, "\n");
  }
  fclose(fp);
}

void ntff_outputEnormBin(double **e_norm, const char *file_name)
{
  FILE *fp = open_or_die(file_name, "wb");
  for (int l = 0; l <= LAMBDA_EN_NM - LAMBDA_ST_NM; l++) {
    if (fwrite(&e_norm[l][0], sizeof(double), 360, fp) < 360) {
      printf("error in write binary %s\n", file_name);
      exit(2);
    }
  }
  fclose(fp);
}

/* next value line (parser.c:3-20): lines whose first character is '#', NUL or
 * '\n' are skipped, any other line is a value line (a line of blanks, or the
 * "\r\n" of a CRLF file, reads as 0, as in the reference); the value stops at a
 * trailing "# note" because the conversion stops at the first non-digit */
static int next_line(FILE *fp, char *buf, size_t n)
{
  while (fgets(buf, (int)n, fp)) {
    if (buf[0] == '#' || buf[0] == '\0' || buf[0] == '\n') continue;
    return 1;
  }
  return 0;
}

/* readConfig field order (main.c:319-366) */
int fdtd_readConfig(const char *path, fdtd_run_config *out)
{
  FILE *fp = fopen(path, "r");
  if (!fp) {
    printf("cannot find %s\n", path);
    return -1;
  }
  char buf[1024];
  int v[11];
  for (int i = 0; i < 11; i++) {
    if (!next_line(fp, buf, sizeof buf)) {
      fclose(fp);
      printf("parse error in %s: expected 11 values, got %d\n", path, i);
      return -2;
    }
    /* lambda with strtod into the int field, the others with atoi (main.c:330-358) */
    v[i] = i == 4 ? (int)strtod(buf, NULL) : atoi(buf);
  }
  fclose(fp);
  memset(out, 0, sizeof *out);
  out->field_info.width_nm = v[0];
  out->field_info.height_nm = v[1];
  out->field_info.h_u_nm = v[2];
  out->field_info.pml = v[3];
  out->field_info.lambda_nm = v[4];
  out->field_info.stepNum = v[5];
  out->startAngle = v[6];
  out->endAngle = v[7];
  out->deltaAngle = v[8];
  out->field_info.angle_deg = v[6];
  out->model = v[9];
  out->solver = v[10];
  return 0;
}
