/*
 * main.c -- eif_gpu_main, the driver of the GPU time-step path: a drop-in
 * for the reference's a.out (src/main.c:57-118).
 *
 * Same environment variables, same inputs (the initparticles npy files or the
 * restart directory named by dirname_restart), same outputs (output/log,
 * output/save/stepNNNNNNNNNN, output/stat/stepNNNNNNNNNN), same main loop:
 * decide dt and integrate (one eif_gpu_step), advance step and time, then
 * log, save and collect statistics when their schedules fall due, until
 * time > timemax or the wall time exceeds wtimemax; a final save and the
 * statistics output close the run.  One process drives one GPU.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "eif_gpu.h"
#include "eif_host.h"

#define NREC EIF_PARTICLE_NDOUBLES

#define GPU(call)                                                              \
  do {                                                                         \
    int _rc = (call);                                                          \
    if (_rc != EIF_OK) {                                                       \
      fprintf(stderr, "%s failed (%d): %s\n", #call, _rc, eif_gpu_last_error()); \
      exit(1);                                                                 \
    }                                                                          \
  } while (0)

static double wtime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "out of host memory\n");
    exit(1);
  }
  return p;
}

typedef struct {
  eifh_param P;
  eif_ctx *ctx;
  double *xf, *xc;
  double *ux, *uy, *p;  /* host staging in the reference layouts */
  double *rec;          /* particle records */
  int np;
} run_t;

static void save(run_t *r) {
  char dir[4200];
  eifh_save_dirname(dir, sizeof(dir), EIFH_DIR_SAVE, r->P.step);
  GPU(fluid_gpu_download(r->ctx, r->ux, r->uy, r->p));
  if (r->np) GPU(suspensions_gpu_download(r->ctx, r->rec));
  if (eifh_save(dir, &r->P, r->xf, r->xc, r->ux, r->uy, r->p, r->np, r->rec))
    fprintf(stderr, "save to %s failed\n", dir);
}

static void logging(run_t *r) {
  double divmax, divsum, moms[2], energy[2];
  GPU(fluid_gpu_divergence(r->ctx, &divmax, &divsum));
  GPU(fluid_gpu_momentum(r->ctx, moms));
  GPU(fluid_gpu_energy(r->ctx, energy));
  if (r->np) GPU(suspensions_gpu_download(r->ctx, r->rec));
  eifh_logging(EIFH_DIR_LOG, &r->P, divmax, divsum, moms, energy, r->np, r->rec);
}

static void statistics_output(run_t *r) {
  const int nx = r->P.itot, ny = r->P.jtot;
  const size_t nu = (size_t)(nx + 1) * (ny + 2), npad = (size_t)(nx + 2) * (ny + 2);
  double *ux1 = xcalloc(nu, sizeof(double)), *ux2 = xcalloc(nu, sizeof(double));
  double *uy1 = xcalloc(npad, sizeof(double)), *uy2 = xcalloc(npad, sizeof(double));
  double *phi = xcalloc(npad, sizeof(double));
  int num = 0;
  GPU(statistics_gpu_download(r->ctx, ux1, ux2, uy1, uy2, phi, &num));
  char dir[4200];
  eifh_save_dirname(dir, sizeof(dir), EIFH_DIR_STAT, r->P.step);
  if (eifh_stat_output(dir, &r->P, r->xf, r->xc, num, ux1, ux2, uy1, uy2, phi))
    fprintf(stderr, "statistics output to %s failed\n", dir);
  free(ux1);
  free(ux2);
  free(uy1);
  free(uy2);
  free(phi);
}

int main(void) {
  run_t r;
  memset(&r, 0, sizeof(r));
  const double t0 = wtime();
  if (eifh_param_load(&r.P, 1)) return 1;
  if (eifh_mkdirs(EIFH_DIR_SAVE) || eifh_mkdirs(EIFH_DIR_LOG) || eifh_mkdirs(EIFH_DIR_STAT))
    return 1;
  eifh_param *P = &r.P;
  const int nx = P->itot, ny = P->jtot;
  eif_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.itot = nx;
  cfg.jtot = ny;
  cfg.ly = P->ly;
  cfg.Re = P->Re;
  cfg.Fr = P->Fr;
  cfg.extfrcy = P->extfrcy;
  cfg.uy_wall_xm = P->uy_wall_xm;
  cfg.uy_wall_xp = P->uy_wall_xp;
  for (int k = 0; k < 3; k++) cfg.safefactors[k] = P->safefactors[k];
  cfg.poisson_mode = P->poisson_ysolve ? EIF_POISSON_PARTITION : EIF_POISSON_THOMAS;
  GPU(fluid_gpu_create(&cfg, &r.ctx));
  r.xf = xcalloc((size_t)nx + 1, sizeof(double));
  r.xc = xcalloc((size_t)nx + 2, sizeof(double));
  GPU(fluid_gpu_get_grid(r.ctx, r.xf, r.xc, NULL, NULL, NULL, NULL));
  r.ux = xcalloc((size_t)(nx + 1) * (ny + 2), sizeof(double));
  r.uy = xcalloc((size_t)(nx + 2) * (ny + 2), sizeof(double));
  r.p = xcalloc((size_t)(nx + 2) * (ny + 2), sizeof(double));
  /* fluid: zero with boundary values, or loaded (src/fluid/init.c:33-85) */
  if (P->load_flow_field) {
    if (eifh_load_fields(P->dirname_restart, nx, ny, r.ux, r.uy, r.p)) return 1;
    GPU(fluid_gpu_upload(r.ctx, r.ux, r.uy, r.p));
    GPU(fluid_gpu_update_boundaries_ux(r.ctx));
    GPU(fluid_gpu_update_boundaries_uy(r.ctx));
    GPU(fluid_gpu_update_boundaries_p(r.ctx));
  }
  /* particles (src/suspensions/init.c:24-86) */
  {
    const char *dir = P->load_flow_field ? P->dirname_restart : EIFH_DIR_INIT;
    if (eifh_load_n_particles(dir, &r.np)) return 1;
    double *soa = xcalloc((size_t)9 * r.np, sizeof(double));
    if (eifh_load_particles(dir, r.np, soa)) return 1;
    const size_t n = (size_t)r.np;
    GPU(suspensions_gpu_create(r.ctx, r.np, soa, soa + n, soa + 2 * n, soa + 3 * n,
                               soa + 4 * n, soa + 5 * n, soa + 6 * n, soa + 7 * n,
                               soa + 8 * n));
    free(soa);
    r.rec = xcalloc(n * NREC, sizeof(double));
  }
  double t1 = wtime();
  for (;;) {
    GPU(eif_gpu_step(r.ctx, &P->dt));
    P->step += 1;
    P->time += P->dt;
    if (P->log.next < P->time) {
      logging(&r);
      P->log.next += P->log.rate;
    }
    if (P->save.next < P->time) {
      save(&r);
      P->save.next += P->save.rate;
    }
    if (P->stat.next < P->time) {
      GPU(statistics_gpu_collect(r.ctx));
      P->stat.next += P->stat.rate;
    }
    t1 = wtime();
    if (P->time > P->timemax) break;
    if (t1 - t0 > P->wtimemax) break;
  }
  printf("elapsed: %.2f [s]\n", t1 - t0);
  save(&r);
  statistics_output(&r);
  GPU(fluid_gpu_destroy(r.ctx));
  free(r.xf);
  free(r.xc);
  free(r.ux);
  free(r.uy);
  free(r.p);
  free(r.rec);
  return 0;
}
