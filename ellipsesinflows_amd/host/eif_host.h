/*
 * eif_host.h -- host side of the EllipsesInFlows driver: run parameters
 * from the environment, the npy save/restart format, logging files and
 * statistics output.  Plain C99, no GPU calls: the driver (main.c) moves
 * data between these functions and the device through include/eif_gpu.h.
 *
 * Mirrors the reference's host modules
 *   src/param/init.c      (environment variables, schedules)
 *   src/fileio.c, src/simple_npyio.c (npy v1.0 files)
 *   src/save.c            (output/save/stepNNNNNNNNNN)
 *   src/logging.c         (output/log/NAME.dat)
 *   src/statistics/output.c (output/stat/stepNNNNNNNNNN)
 * with the same names, defaults, file names and formats, for one process
 * owning the whole domain.
 */
#ifndef EIF_HOST_H
#define EIF_HOST_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* directory names (include/fileio.h:10-12) */
#define EIFH_DIR_SAVE "output/save"
#define EIFH_DIR_LOG "output/log"
#define EIFH_DIR_STAT "output/stat"
#define EIFH_DIR_INIT "initparticles"

/* npy dtype strings as the reference writes them (include/fileio.h:21-22) */
#define EIFH_NPY_INT "'int32'"
#define EIFH_NPY_DOUBLE "'float64'"

typedef struct {
  double rate, after, next;
} eifh_schedule;

/* the members of param_t (include/param.h:29-53) that the driver needs,
 * plus the GPU path's options */
typedef struct {
  int load_flow_field;
  char dirname_restart[4096];
  int itot, jtot;
  double lx, ly;
  double safefactors[3];
  double Re, Fr, extfrcy;
  double time, dt;
  int step;
  double timemax, wtimemax;
  eifh_schedule log, save, stat;
  /* options of this implementation (not in the reference) */
  double uy_wall_xm, uy_wall_xp; /* wall values of uy (reference: 0, 0) */
  int poisson_ysolve;            /* 0 bordering Thomas, 1 partitioned  */
} eifh_param;

/* --- parameters (src/param/init.c:46-103, :193-216, :218-239) ------- */
/* Loads every variable from the environment with the reference's names
 * and defaults, printing them as the reference does when verbose != 0;
 * with dirname_restart set, reads step and time from that directory.
 * Returns 0, or -1 if the restart directory cannot be read. */
int eifh_param_load(eifh_param *p, int verbose);
double eifh_next_schedule(double rate, double after, double time);

/* --- npy v1.0 (src/simple_npyio.c) ----------------------------------- */
/* Writes a header in the reference's byte format followed by nbytes of
 * data.  Returns 0 or -1. */
int eifh_npy_write(const char *path, const char *dtype, int ndim,
                   const size_t *shape, const void *data, size_t nbytes);
/* Parses a header (versions 1-3).  dtype receives the descr value as
 * written (quotes included), shape up to 8 dimensions.  Returns the data
 * offset in bytes, or 0 on error. */
size_t eifh_npy_read_header(const char *path, char *dtype, size_t dtype_cap,
                            int *ndim, size_t shape[8], int *fortran_order);
/* Reads an array of doubles: the file must hold float64 data in C order
 * with exactly the given shape.  Returns 0 or -1. */
int eifh_npy_read_doubles(const char *path, int ndim, const size_t *shape,
                          double *out);
/* Reads a 0-d or one-element integer (int32 or int64).  Returns 0 or -1. */
int eifh_npy_read_int(const char *path, int *out);
/* Reads a 0-d or one-element double.  Returns 0 or -1. */
int eifh_npy_read_double(const char *path, double *out);

/* --- datasets in a directory (src/fileio.c:109-365) -------------------- */
int eifh_w_0d_int(const char *dir, const char *name, int v);
int eifh_w_0d_double(const char *dir, const char *name, double v);
int eifh_w_1d(const char *dir, const char *name, size_t n, const double *v);
/* 2-d fields without their y-halo rows: the file holds rows j = 1..jtot
 * of a halo-padded array with ncols values per row, shape (jtot, ncols);
 * ncols is itot+1 for ux-like and itot+2 for uy-/p-like arrays. */
int eifh_w_field(const char *dir, const char *name, int jtot, int ncols,
                 const double *halo_padded);
int eifh_r_field(const char *dir, const char *name, int jtot, int ncols,
                 double *halo_padded);

/* --- save / restart (src/save.c:13-122, src/fluid/init.c:33-85,
 *     src/suspensions/init.c:24-86) ------------------------------------ */
/* "output/save/step%010d" */
void eifh_save_dirname(char *buf, size_t cap, const char *root, int step);
/* mkdir -p; returns 0 or -1 */
int eifh_mkdirs(const char *path);
/* Writes a complete restart directory.  xf[itot+1], xc[itot+2]; ux, uy,
 * p in the halo-padded layouts; records are n particle_t records of 30
 * doubles. */
int eifh_save(const char *dirname, const eifh_param *p, const double *xf,
              const double *xc, const double *ux, const double *uy,
              const double *pr, int n, const double *records);
/* Reads ux, uy, p of a restart directory into halo-padded arrays (rows
 * 1..jtot; halo rows are left as they are). */
int eifh_load_fields(const char *dir, int itot, int jtot, double *ux,
                     double *uy, double *pr);
/* Reads n_particles of a particle directory (initparticles/ or a restart
 * directory). */
int eifh_load_n_particles(const char *dir, int *n);
/* Reads the nine particle arrays (den, a, b, x, y, az, ux, uy, vz) of n
 * particles; soa is 9 x n doubles. */
int eifh_load_particles(const char *dir, int n, double *soa);

/* --- statistics output (src/statistics/output.c:59-77) ----------------- */
int eifh_stat_output(const char *dirname, const eifh_param *p, const double *xf,
                     const double *xc, int num, const double *ux1,
                     const double *ux2, const double *uy1, const double *uy2,
                     const double *phi);

/* --- logging (src/logging.c:12-207) --------------------------------------
 * Appends one line to output/log/{progress,divergence,momentum,energy}.dat
 * and to output/log/particleNNNNNNNNNN.dat for each particle (files are
 * truncated instead when step is 0), and prints the progress line. */
int eifh_logging(const char *logdir, const eifh_param *p, double divmax,
                 double divsum, const double moms[2], const double energy[2],
                 int n, const double *records);
int eifh_log_digits(double rate);

#ifdef __cplusplus
}
#endif
#endif
