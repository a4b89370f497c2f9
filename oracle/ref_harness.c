/*
 * ref_harness.c -- thin C entry points over the reference's OWN compiled
 * sources (NaokiHori/EllipsesInFlows src/ellipse.c, src/suspensions/
 * {auxiliary,collision,reset,init}.c, src/param/init.c, src/fileio.c,
 * src/simple_npyio.c, src/parallel/{others,init,halo}.c, src/common.c),
 * built from where they lie by oracle/Makefile into oracle/_ref/.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests to pin the oracle restatement.
 * The reference's fluid, Poisson, inertia, exchange, increment and update
 * translation units include <fftw3.h>, which this image does not have, so
 * they are not built here.
 *
 * The reference calls MPI inside these functions; every entry point runs
 * them as a single rank (MPI_Init on first use).
 */
#define _POSIX_C_SOURCE 200112L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mpi.h>
#include "param.h"
#include "parallel.h"
#include "suspensions.h"
#include "ellipse.h"
#include "fileio.h"

#define NREC 30

static void ensure_mpi(void) {
  int flag = 0;
  MPI_Initialized(&flag);
  if (!flag) MPI_Init(NULL, NULL);
}

int ref_record_size(void) { return (int)(sizeof(particle_t) / sizeof(double)); }

double ref_find_normal_t(double a, double b, double x, double y) {
  return find_normal_t(a, b, x, y);
}
double ref_compute_ex(double a, double b, double t) { return compute_ex(a, b, t); }
double ref_compute_ey(double a, double b, double t) { return compute_ey(a, b, t); }
double ref_curvature(double a, double b, double t) { return compute_curvature(a, b, t); }

double ref_s_weight(double g, double pa, double pb, double px, double py,
                    double paz, double x, double y) {
  return suspensions_s_weight(g, pa, pb, px, py, paz, x, y);
}
double ref_v_weight(double g, double pa, double pb, double px, double py,
                    double paz, double x, double y) {
  return suspensions_v_weight(g, pa, pb, px, py, paz, x, y);
}
int ref_decide_loop_size(int lb, int ub, double h, double r, double c,
                         int *mn, int *mx) {
  return suspensions_decide_loop_size(lb, ub, h, r, c, mn, mx);
}
double ref_mass(double den, double a, double b) {
  return suspensions_compute_mass(den, a, b);
}
double ref_moment_of_inertia(double den, double a, double b) {
  return suspensions_compute_moment_of_inertia(den, a, b);
}

static suspensions_t *wrap_records(int n, double *records) {
  suspensions_t *s = calloc(1, sizeof(*s));
  s->n_particles = n;
  s->particles = calloc(n > 0 ? n : 1, sizeof(particle_t *));
  for (int k = 0; k < n; k++) {
    s->particles[k] = calloc(1, sizeof(particle_t));
    memcpy(s->particles[k], records + (size_t)k * NREC, sizeof(particle_t));
  }
  s->buf = calloc(3 * (size_t)(n > 0 ? n : 1), sizeof(double));
  return s;
}

static void unwrap_records(suspensions_t *s, double *records) {
  for (int k = 0; k < s->n_particles; k++) {
    memcpy(records + (size_t)k * NREC, s->particles[k], sizeof(particle_t));
    free(s->particles[k]);
  }
  free(s->particles);
  free(s->buf);
  free(s);
}

/* suspensions_compute_collision_force on n records (in/out) */
int ref_collision(int n, double lx, double ly, int cnstep, double *records) {
  ensure_mpi();
  param_t param;
  memset(&param, 0, sizeof(param));
  param.lx = lx;
  param.ly = ly;
  parallel_t parallel = {1, 0, 0, 0};
  suspensions_t *s = wrap_records(n, records);
  suspensions_compute_collision_force(&param, &parallel, cnstep, s);
  unwrap_records(s, records);
  return 0;
}

int ref_reset(int n, double *records) {
  suspensions_t *s = wrap_records(n, records);
  suspensions_reset_particle_increments(s);
  unwrap_records(s, records);
  return 0;
}

/* param_init with the given environment; returns grid vectors and RK
 * coefficients (alpha[3], beta[3], gamma[3]) */
int ref_param_grid(int itot, int jtot, double ly, double *xf, double *xc,
                   double *dxf, double *dxc, double *yf, double *yc,
                   double *rk, double *dxdy) {
  ensure_mpi();
  char buf[64];
  unsetenv("dirname_restart");
  snprintf(buf, sizeof(buf), "%d", itot);
  setenv("itot", buf, 1);
  snprintf(buf, sizeof(buf), "%d", jtot);
  setenv("jtot", buf, 1);
  snprintf(buf, sizeof(buf), "%.17g", ly);
  setenv("ly", buf, 1);
  FILE *saved = stdout;
  stdout = fopen("/dev/null", "w");
  param_t *p = param_init();
  fclose(stdout);
  stdout = saved;
  memcpy(xf, p->xf, sizeof(double) * (itot + 1));
  memcpy(xc, p->xc, sizeof(double) * (itot + 2));
  memcpy(dxf, p->dxf, sizeof(double) * itot);
  memcpy(dxc, p->dxc, sizeof(double) * (itot + 1));
  memcpy(yf, p->yf, sizeof(double) * (jtot + 1));
  memcpy(yc, p->yc, sizeof(double) * jtot);
  for (int k = 0; k < 3; k++) {
    rk[k] = p->rkcoefs[k].alpha;
    rk[3 + k] = p->rkcoefs[k].beta;
    rk[6 + k] = p->rkcoefs[k].gamma;
  }
  dxdy[0] = p->dx;
  dxdy[1] = p->dy;
  param_finalise(p);
  return 0;
}

/* npy writers/readers of src/fileio.c */
int ref_w_1d(const char *dir, const char *name, int n, const double *data) {
  return fileio_w_1d_serial(dir, name, NPYIO_DOUBLE, sizeof(double), n, data);
}
int ref_w_0d_int(const char *dir, const char *name, int v) {
  return fileio_w_0d_serial(dir, name, NPYIO_INT, sizeof(int), &v);
}
int ref_w_0d_double(const char *dir, const char *name, double v) {
  return fileio_w_0d_serial(dir, name, NPYIO_DOUBLE, sizeof(double), &v);
}
int ref_r_1d(const char *dir, const char *name, int n, double *data) {
  return fileio_r_1d_serial(dir, name, sizeof(double), n, data);
}
int ref_r_0d_int(const char *dir, const char *name, int *v) {
  return fileio_r_0d_serial(dir, name, sizeof(int), v);
}

/* src/suspensions/init.c via a restart directory: returns the particle
 * count, fills up to cap records */
int ref_load_particles(const char *dir, int cap, double *records) {
  ensure_mpi();
  param_t param;
  memset(&param, 0, sizeof(param));
  param.load_flow_field = true;
  param.dirname_restart = (char *)dir;
  param.itot = 4;
  param.jtot = 4;
  parallel_t parallel = {1, 0, 0, 0};
  suspensions_t *s = suspensions_init(&param, &parallel);
  const int n = s->n_particles;
  for (int k = 0; k < n && k < cap; k++)
    memcpy(records + (size_t)k * NREC, s->particles[k], sizeof(particle_t));
  suspensions_finalise(s);
  return n;
}

/* param_init (src/param/init.c:218-239) with the current environment:
 * out = {timemax, wtimemax, log.rate, log.after, log.next, save.rate,
 *        save.after, save.next, stat.rate, stat.after, stat.next, itot, jtot,
 *        lx, ly, sf[0], sf[1], sf[2], Re, Fr, extfrcy, step, time,
 *        load_flow_field} (24 values) */
int ref_param_values(double *out) {
  ensure_mpi();
  FILE *saved = stdout;
  stdout = fopen("/dev/null", "w");
  param_t *p = param_init();
  fclose(stdout);
  stdout = saved;
  const double v[24] = {p->timemax, p->wtimemax, p->log.rate, p->log.after, p->log.next,
                        p->save.rate, p->save.after, p->save.next, p->stat.rate,
                        p->stat.after, p->stat.next, p->itot, p->jtot, p->lx, p->ly,
                        p->safefactors[0], p->safefactors[1], p->safefactors[2], p->Re,
                        p->Fr, p->extfrcy, p->step, p->time, p->load_flow_field};
  memcpy(out, v, sizeof(v));
  param_finalise(p);
  return 0;
}

/* src/fileio.c:283-365: 2-d fields through the reference's MPI-IO writer
 * and reader for one rank; kind 0 ux-like, 1 uy-like, 2 p-like; the data
 * are halo-padded arrays of the matching layout */
static int field_io(int write, const char *dir, const char *name, int kind, int itot,
                    int jtot, double *data) {
  ensure_mpi();
  param_t param;
  memset(&param, 0, sizeof(param));
  param.itot = itot;
  param.jtot = jtot;
  parallel_t parallel = {1, 0, 0, 0};
  if (write) {
    if (kind == 0) return fileio_w_ux_like_parallel(dir, name, &param, &parallel, data);
    if (kind == 1) return fileio_w_uy_like_parallel(dir, name, &param, &parallel, data);
    return fileio_w_p_like_parallel(dir, name, &param, &parallel, data);
  }
  if (kind == 0) return fileio_r_ux_like_parallel(dir, name, &param, &parallel, data);
  if (kind == 1) return fileio_r_uy_like_parallel(dir, name, &param, &parallel, data);
  return fileio_r_p_like_parallel(dir, name, &param, &parallel, data);
}
int ref_w_field(const char *dir, const char *name, int kind, int itot, int jtot,
                double *data) {
  return field_io(1, dir, name, kind, itot, jtot, data);
}
int ref_r_field(const char *dir, const char *name, int kind, int itot, int jtot,
                double *data) {
  return field_io(0, dir, name, kind, itot, jtot, data);
}
