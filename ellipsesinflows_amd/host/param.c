/*
 * param.c -- run parameters from environment variables, with the names,
 * defaults and console lines of src/param/init.c:46-103, and the output
 * schedules of src/param/init.c:193-216.
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "eif_host.h"

/* src/param/init.c:37-51: strtol, clamped to the int range */
static int load_int(const char *name, int def, int verbose) {
  const char *s = getenv(name);
  int v = def;
  if (s) {
    long t = strtol(s, NULL, 10);
    t = t < INT_MIN ? INT_MIN : t;
    t = t > INT_MAX ? INT_MAX : t;
    v = (int)t;
  }
  if (verbose) printf("  %s: %d%s\n", name, v, s ? "" : " (default)");
  return v;
}

/* src/param/init.c:53-64 */
static double load_double(const char *name, double def, int verbose) {
  const char *s = getenv(name);
  const double v = s ? strtod(s, NULL) : def;
  if (verbose) printf("  %s: %.3e%s\n", name, v, s ? "" : " (default)");
  return v;
}

/* The next output time strictly after `after` (before it is reached) or
 * after the current time, on the grid of multiples of rate; the 1e-8
 * shift rounds an exact multiple up, src/param/init.c:193-216. */
double eifh_next_schedule(double rate, double after, double time) {
  const double small = 1.e-8;
  if (time < after) return rate * ceil((after + small) / rate);
  return rate * ceil((time + small) / rate);
}

int eifh_param_load(eifh_param *p, int verbose) {
  memset(p, 0, sizeof(*p));
  if (verbose) printf("------- parameters are loaded -------\n");
  const char *restart = getenv("dirname_restart");
  if (restart) {
    p->load_flow_field = 1;
    snprintf(p->dirname_restart, sizeof(p->dirname_restart), "%s", restart);
    if (verbose) printf("  flow fields are loaded: %s\n", p->dirname_restart);
  } else {
    p->load_flow_field = 0;
    if (verbose) printf("  flow fields are initialised\n");
  }
  p->timemax = load_double("timemax", 1.0e+3, verbose);
  p->wtimemax = load_double("wtimemax", 6.0e+2, verbose);
  p->log.rate = load_double("log_rate", 1.0e+0, verbose);
  p->log.after = load_double("log_after", 0.0e+0, verbose);
  p->save.rate = load_double("save_rate", 1.0e+3, verbose);
  p->save.after = load_double("save_after", 0.0e+0, verbose);
  p->stat.rate = load_double("stat_rate", 1.0e-1, verbose);
  p->stat.after = load_double("stat_after", 2.0e+3, verbose);
  p->itot = load_int("itot", 32, verbose);
  p->jtot = load_int("jtot", 32, verbose);
  p->lx = 1.;
  p->ly = load_double("ly", 1.0e+0, verbose);
  p->safefactors[0] = load_double("safefactor_adv", 7.5e-1, verbose);
  p->safefactors[1] = load_double("safefactor_dif", 7.5e-1, verbose);
  p->safefactors[2] = load_double("safefactor_par", 9.5e-1, verbose);
  p->Re = load_double("Re", 2.334e+3, verbose);
  p->Fr = load_double("Fr", DBL_MAX, verbose);
  p->extfrcy = load_double("extfrcy", 2.337e-4, verbose);
  /* options of this implementation */
  p->uy_wall_xm = load_double("uy_wall_xm", 0., verbose);
  p->uy_wall_xp = load_double("uy_wall_xp", 0., verbose);
  {
    const char *s = getenv("poisson_ysolve");
    p->poisson_ysolve = !(s && strcmp(s, "thomas") == 0);
    if (verbose)
      printf("  poisson_ysolve: %s%s\n", p->poisson_ysolve ? "partition" : "thomas",
             s ? "" : " (default)");
  }
  if (verbose) printf("-------------------------------------\n");
  if (p->load_flow_field) {
    char path[4200];
    snprintf(path, sizeof(path), "%s/step.npy", p->dirname_restart);
    if (eifh_npy_read_int(path, &p->step)) return -1;
    snprintf(path, sizeof(path), "%s/time.npy", p->dirname_restart);
    if (eifh_npy_read_double(path, &p->time)) return -1;
  } else {
    p->step = 0;
    p->time = 0.;
  }
  p->log.next = eifh_next_schedule(p->log.rate, p->log.after, p->time);
  p->save.next = eifh_next_schedule(p->save.rate, p->save.after, p->time);
  p->stat.next = eifh_next_schedule(p->stat.rate, p->stat.after, p->time);
  return 0;
}
