/*
 * logging.c -- the log files of src/logging.c:12-207 with their file
 * names and line formats:
 *   progress.dat     "step %8d, time %*.*f, dt %.2e"   (also on stdout)
 *   divergence.dat   "%*.*f % .1e % .1e"               max |div|, sum div
 *   momentum.dat     "%*.*f % 18.15e % 18.15e"          x and y momentum
 *   energy.dat       "%*.*f % 18.15e % 18.15e"          x and y kinetic energy
 *   particleNNNNNNNNNN.dat "%*.*f" then x, y, az, ux, uy, vz as "% .7e"
 * where the time field has width ndigits+3 and ndigits decimals, ndigits
 * being the number of decimal digits of (int)(1/log_rate).
 * The reductions themselves are computed on the device by the caller.
 */
#include <stdio.h>
#include "eif_host.h"

#define NREC 30

/* number of decimal digits of (int)(1/rate), src/logging.c:12-19 */
int eifh_log_digits(double rate) {
  int d = 1;
  int inv = (int)(1. / rate);
  while (inv /= 10) d++;
  return d;
}

static FILE *open_log(const char *dir, const char *name, int step) {
  char path[4200];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *fp = fopen(path, step == 0 ? "w" : "a");
  if (!fp) perror(path);
  return fp;
}

int eifh_logging(const char *logdir, const eifh_param *p, double divmax,
                 double divsum, const double moms[2], const double energy[2],
                 int n, const double *records) {
  const int nd = eifh_log_digits(p->log.rate);
  int rc = 0;
  FILE *fp = open_log(logdir, "progress.dat", p->step);
  if (fp) {
    fprintf(fp, "step %8d, time %*.*f, dt %.2e\n", p->step, nd + 3, nd, p->time, p->dt);
    fclose(fp);
  } else {
    rc = -1;
  }
  fprintf(stdout, "step %8d, time %*.*f, dt %.2e\n", p->step, nd + 3, nd, p->time, p->dt);
  if ((fp = open_log(logdir, "divergence.dat", p->step))) {
    fprintf(fp, "%*.*f % .1e % .1e\n", nd + 3, nd, p->time, divmax, divsum);
    fclose(fp);
  } else {
    rc = -1;
  }
  if ((fp = open_log(logdir, "momentum.dat", p->step))) {
    fprintf(fp, "%*.*f % 18.15e % 18.15e\n", nd + 3, nd, p->time, moms[0], moms[1]);
    fclose(fp);
  } else {
    rc = -1;
  }
  if ((fp = open_log(logdir, "energy.dat", p->step))) {
    fprintf(fp, "%*.*f % 18.15e % 18.15e\n", nd + 3, nd, p->time, energy[0], energy[1]);
    fclose(fp);
  } else {
    rc = -1;
  }
  for (int k = 0; k < n; k++) {
    char name[64];
    snprintf(name, sizeof(name), "particle%010d.dat", k);
    if (!(fp = open_log(logdir, name, p->step))) {
      rc = -1;
      continue;
    }
    const double *q = records + (size_t)k * NREC;
    /* x, y, az, ux, uy, vz (include/suspensions.h:8-31 offsets 3,4,5,9,10,11) */
    fprintf(fp, "%*.*f % .7e % .7e % .7e % .7e % .7e % .7e\n", nd + 3, nd, p->time, q[3], q[4],
            q[5], q[9], q[10], q[11]);
    fclose(fp);
  }
  return rc;
}
