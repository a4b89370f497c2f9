/*
 * save.c -- restart directories and statistics output in the reference's
 * format (src/save.c:13-122, src/statistics/output.c:9-77,
 * src/fileio.c:109-365), and their readers (src/fluid/init.c:33-85,
 * src/suspensions/init.c:24-86).
 *
 * A directory holds one .npy file per dataset: 0-d int32 step, itot,
 * jtot, n_particles (and num for statistics); 0-d float64 time, lx, ly,
 * Re; 1-d float64 xf, xc, yf, yc and the nine particle arrays; 2-d
 * float64 fields without their y-halo rows, shape (jtot, itot+1) for
 * ux-like and (jtot, itot+2) for uy- and p-like arrays.
 */
#define _POSIX_C_SOURCE 200809L
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include "eif_host.h"

#define NREC 30
/* particle_t member offsets of the saved quantities
 * (include/suspensions.h:8-31) */
enum { R_DEN = 0, R_A = 1, R_B = 2, R_X = 3, R_Y = 4, R_AZ = 5, R_UX = 9, R_UY = 10, R_VZ = 11 };

static const char *const particle_names[9] = {
    "particle_dens", "particle_as", "particle_bs", "particle_xs", "particle_ys",
    "particle_azs", "particle_uxs", "particle_uys", "particle_vzs"};
static const int particle_members[9] = {R_DEN, R_A, R_B, R_X, R_Y, R_AZ, R_UX, R_UY, R_VZ};

static void npy_path(char *buf, size_t cap, const char *dir, const char *name) {
  snprintf(buf, cap, "%s/%s.npy", dir, name);
}

int eifh_w_0d_int(const char *dir, const char *name, int v) {
  char path[4200];
  npy_path(path, sizeof(path), dir, name);
  const int v32 = v;
  return eifh_npy_write(path, EIFH_NPY_INT, 0, NULL, &v32, sizeof(v32));
}

int eifh_w_0d_double(const char *dir, const char *name, double v) {
  char path[4200];
  npy_path(path, sizeof(path), dir, name);
  return eifh_npy_write(path, EIFH_NPY_DOUBLE, 0, NULL, &v, sizeof(v));
}

int eifh_w_1d(const char *dir, const char *name, size_t n, const double *v) {
  char path[4200];
  npy_path(path, sizeof(path), dir, name);
  return eifh_npy_write(path, EIFH_NPY_DOUBLE, 1, &n, v, n * sizeof(double));
}

int eifh_w_field(const char *dir, const char *name, int jtot, int ncols,
                 const double *halo_padded) {
  char path[4200];
  npy_path(path, sizeof(path), dir, name);
  const size_t shape[2] = {(size_t)jtot, (size_t)ncols};
  /* the first row is the y-halo (j = 0) and is not stored */
  return eifh_npy_write(path, EIFH_NPY_DOUBLE, 2, shape, halo_padded + ncols,
                        shape[0] * shape[1] * sizeof(double));
}

int eifh_r_field(const char *dir, const char *name, int jtot, int ncols,
                 double *halo_padded) {
  char path[4200];
  npy_path(path, sizeof(path), dir, name);
  const size_t shape[2] = {(size_t)jtot, (size_t)ncols};
  return eifh_npy_read_doubles(path, 2, shape, halo_padded + ncols);
}

void eifh_save_dirname(char *buf, size_t cap, const char *root, int step) {
  snprintf(buf, cap, "%s/step%010d", root, step);
}

int eifh_mkdirs(const char *path) {
  char tmp[4200];
  snprintf(tmp, sizeof(tmp), "%s", path);
  const size_t len = strlen(tmp);
  for (size_t k = 1; k <= len; k++) {
    if (tmp[k] != '/' && tmp[k] != '\0') continue;
    const char c = tmp[k];
    tmp[k] = '\0';
    if (mkdir(tmp, 0777) != 0 && errno != EEXIST) {
      perror(tmp);
      return -1;
    }
    tmp[k] = c;
  }
  return 0;
}

/* src/save.c:13-39 and src/statistics/output.c:9-35: scalars and global
 * coordinates; yf[j] = j dy, yc[j] = (yf[j] + yf[j+1]) / 2 */
static int save_param(const char *dir, const eifh_param *p, const double *xf,
                      const double *xc) {
  int rc = 0;
  rc |= eifh_w_0d_int(dir, "step", p->step);
  rc |= eifh_w_0d_int(dir, "itot", p->itot);
  rc |= eifh_w_0d_int(dir, "jtot", p->jtot);
  rc |= eifh_w_0d_double(dir, "time", p->time);
  rc |= eifh_w_0d_double(dir, "lx", p->lx);
  rc |= eifh_w_0d_double(dir, "ly", p->ly);
  rc |= eifh_w_0d_double(dir, "Re", p->Re);
  rc |= eifh_w_1d(dir, "xf", (size_t)p->itot + 1, xf);
  rc |= eifh_w_1d(dir, "xc", (size_t)p->itot + 2, xc);
  const int jtot = p->jtot;
  const double dy = p->ly / jtot;
  double *yf = calloc((size_t)jtot + 1, sizeof(double));
  double *yc = calloc((size_t)jtot, sizeof(double));
  if (!yf || !yc) {
    free(yf);
    free(yc);
    return -1;
  }
  for (int j = 0; j < jtot + 1; j++) yf[j] = 1. * j * dy;
  for (int j = 0; j < jtot; j++) yc[j] = 0.5 * (yf[j] + yf[j + 1]);
  rc |= eifh_w_1d(dir, "yf", (size_t)jtot + 1, yf);
  rc |= eifh_w_1d(dir, "yc", (size_t)jtot, yc);
  free(yf);
  free(yc);
  return rc ? -1 : 0;
}

int eifh_save(const char *dirname, const eifh_param *p, const double *xf,
              const double *xc, const double *ux, const double *uy,
              const double *pr, int n, const double *records) {
  if (eifh_mkdirs(dirname)) return -1;
  int rc = save_param(dirname, p, xf, xc);
  rc |= eifh_w_field(dirname, "ux", p->jtot, p->itot + 1, ux);
  rc |= eifh_w_field(dirname, "uy", p->jtot, p->itot + 2, uy);
  rc |= eifh_w_field(dirname, "p", p->jtot, p->itot + 2, pr);
  rc |= eifh_w_0d_int(dirname, "n_particles", n);
  double *buf = calloc(n > 0 ? (size_t)n : 1, sizeof(double));
  if (!buf) return -1;
  for (int m = 0; m < 9; m++) {
    for (int k = 0; k < n; k++) buf[k] = records[(size_t)k * NREC + particle_members[m]];
    rc |= eifh_w_1d(dirname, particle_names[m], (size_t)n, buf);
  }
  free(buf);
  return rc ? -1 : 0;
}

int eifh_load_fields(const char *dir, int itot, int jtot, double *ux, double *uy,
                     double *pr) {
  if (eifh_r_field(dir, "ux", jtot, itot + 1, ux)) return -1;
  if (eifh_r_field(dir, "uy", jtot, itot + 2, uy)) return -1;
  if (eifh_r_field(dir, "p", jtot, itot + 2, pr)) return -1;
  return 0;
}

int eifh_load_n_particles(const char *dir, int *n) {
  char path[4200];
  npy_path(path, sizeof(path), dir, "n_particles");
  if (eifh_npy_read_int(path, n)) return -1;
  if (*n < 0) {
    fprintf(stderr, "%s: negative particle count %d\n", path, *n);
    return -1;
  }
  return 0;
}

/* src/suspensions/init.c:37-72: the first n values of each array */
int eifh_load_particles(const char *dir, int n, double *soa) {
  for (int m = 0; m < 9; m++) {
    char path[4200];
    char dtype[64];
    int nd, fo;
    size_t sh[8];
    npy_path(path, sizeof(path), dir, particle_names[m]);
    if (n == 0) continue;
    if (!eifh_npy_read_header(path, dtype, sizeof(dtype), &nd, sh, &fo)) return -1;
    if (nd != 1 || sh[0] < (size_t)n) {
      fprintf(stderr, "%s: expected at least %d values\n", path, n);
      return -1;
    }
    double *all = malloc(sh[0] * sizeof(double));
    if (!all) return -1;
    const size_t want[1] = {sh[0]};
    const int rc = eifh_npy_read_doubles(path, 1, want, all);
    if (!rc) memcpy(soa + (size_t)m * n, all, (size_t)n * sizeof(double));
    free(all);
    if (rc) return -1;
  }
  return 0;
}

int eifh_stat_output(const char *dirname, const eifh_param *p, const double *xf,
                     const double *xc, int num, const double *ux1,
                     const double *ux2, const double *uy1, const double *uy2,
                     const double *phi) {
  if (eifh_mkdirs(dirname)) return -1;
  int rc = eifh_w_0d_int(dirname, "num", num);
  rc |= save_param(dirname, p, xf, xc);
  rc |= eifh_w_field(dirname, "ux1", p->jtot, p->itot + 1, ux1);
  rc |= eifh_w_field(dirname, "uy1", p->jtot, p->itot + 2, uy1);
  rc |= eifh_w_field(dirname, "ux2", p->jtot, p->itot + 1, ux2);
  rc |= eifh_w_field(dirname, "uy2", p->jtot, p->itot + 2, uy2);
  rc |= eifh_w_field(dirname, "phi", p->jtot, p->itot + 2, phi);
  return rc ? -1 : 0;
}
