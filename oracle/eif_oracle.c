/*
 * eif_oracle.c -- CPU restatement of the EllipsesInFlows time step
 * (NaokiHori/EllipsesInFlows, src/main.c:13-55 and everything it calls),
 * for one process owning the whole y-range.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * Arithmetic follows the reference expression by expression (same operand
 * order, same libm calls) and is compiled like the reference (-O3 -std=c99,
 * i.e. no floating-point contraction), so that functions which also build
 * from the reference's own sources agree with them bit for bit.
 *
 * How each piece is pinned (tests/test_oracle_*.py):
 *   find_normal_t, s/v_weight, decide_loop_size, mass/inertia  -> bit-exact
 *       against the reference's own ellipse.c / auxiliary.c (oracle/_ref)
 *   collisions (pair + wall)                                    -> bit-exact
 *       against the reference's collision.c (oracle/_ref)
 *   grid (xf, xc, dxf, dxc, yf, yc)                             -> bit-exact
 *       against the reference's param_init (oracle/_ref)
 *   DCT-II/III                 -> O(n^2) long-double sums, 2n round trip
 *   periodic tridiagonal       -> dense LAPACK solve, circulant eigenvalues
 *   Poisson + projection       -> discrete eigenfunctions, zero divergence
 *   predictor stencils         -> manufactured fields (Couette/Poiseuille
 *                                 steady states, polynomial exactness,
 *                                 discrete energy conservation of advection)
 *   inertia / exchange sums    -> reference weights (oracle/_ref) summed in
 *                                 numpy on small boxes: the force and torque
 *                                 of a translating and rotating body in
 *                                 fluid at rest, spin-up by the enclosed
 *                                 fluid, momentum conservation of exchange
 *                                 + increment
 *   increment / update         -> free fall under buoyancy plus a constant
 *                                 contact force integrated exactly over a
 *                                 step; periodic wrap; face-average split
 *                                 and conservation of the field update
 *   update_pressure            -> p_new = p_old + psi; gauge invariance of
 *                                 a whole step
 *   integrate (src/main.c)     -> restated call for call in the tests and
 *                                 compared bit for bit (contacts, residuals
 *                                 crossing 1e-8, the 102-sub-iteration cap);
 *                                 the reference's validation cases (Feng
 *                                 1994 migration, Aidun 1998 rotation)
 *                                 against the literature data it ships
 * FFTW and the reference's fluid sources are not buildable here.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "eif_oracle.h"

/* the reference's value of pi (include/common.h:7) */
#define ORC_PI 3.1415926535897932

/* ---------------------------------------------------------------------
 * layouts (include/arrays/fluid.h, include/arrays/param.h,
 * include/arrays/suspensions.h), nx = itot, ny = jtot (one process)
 * ------------------------------------------------------------------- */
#define IX_UX(i, j)  ((long)(j) * (nx + 1) + ((i) - 1))
#define IX_P(i, j)   ((long)(j) * (nx + 2) + (i))
#define IX_SX(i, j)  ((long)((j) - 1) * (nx - 1) + ((i) - 2))
#define IX_SY(i, j)  ((long)((j) - 1) * nx + ((i) - 1))
#define XF(i)  (s->xf[(i) - 1])
#define XC(i)  (s->xc[(i)])
#define DXF(i) (s->dxf[(i) - 1])
#define DXC(i) (s->dxc[(i) - 1])
#define YF(j)  (s->yf[(j) - 1])
#define YC(j)  (s->yc[(j) - 1])

/* particle record member offsets (include/suspensions.h:8-31) */
enum {
  P_DEN = 0, P_A, P_B, P_X, P_Y, P_AZ, P_DX, P_DY, P_DAZ,
  P_UX, P_UY, P_VZ, P_DUX, P_DUY, P_DVZ, P_FUX, P_FUY, P_TVZ,
  P_IUX = 18, P_IUY = 20, P_IVZ = 22, P_CFX = 24, P_CFY = 26, P_CTZ = 28
};
#define NREC ORC_PARTICLE_NDOUBLES

struct orc {
  int itot, jtot;
  double lx, ly, dx, dy;
  double *xf, *xc, *dxf, *dxc, *yf, *yc;
  double Re, Fr, extfrcy, uy_xm, uy_xp;
  double alpha[3], beta[3], gamma[3];
  double sf[3];
  double dt, time;
  int step;
  double *f[ORC_NFIELDS];
  long flen[ORC_NFIELDS];
  double *qx, *qy;
  int np;
  double *prt;
  int substeps[3];
  int *pairs;
  int npairs, cap_pairs;
};

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "eif_oracle: out of memory\n");
    abort();
  }
  return p;
}

int orc_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  return omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

/* grid construction, src/param/init.c:105-167 (one process, yoffset 0);
 * RK coefficients src/param/init.c:169-191 */
orc_t *orc_create(int itot, int jtot, double ly, double Re, double Fr,
                  double extfrcy, double uy_wall_xm, double uy_wall_xp,
                  const double safefactors[3]) {
  if (itot < 2 || jtot < 2) return NULL;
  orc_t *s = xcalloc(1, sizeof(*s));
  const int nx = itot, ny = jtot;
  s->itot = itot;
  s->jtot = jtot;
  s->lx = 1.;
  s->ly = ly;
  s->Re = Re;
  s->Fr = Fr;
  s->extfrcy = extfrcy;
  s->uy_xm = uy_wall_xm;
  s->uy_xp = uy_wall_xp;
  for (int k = 0; k < 3; k++) s->sf[k] = safefactors[k];
  s->xf = xcalloc(nx + 1, sizeof(double));
  s->xc = xcalloc(nx + 2, sizeof(double));
  s->dxf = xcalloc(nx, sizeof(double));
  s->dxc = xcalloc(nx + 1, sizeof(double));
  s->yf = xcalloc(ny + 1, sizeof(double));
  s->yc = xcalloc(ny, sizeof(double));
  {
    const double lx = s->lx;
    const double dx = lx / itot;
    s->dx = dx;
    for (int i = 1; i <= nx + 1; i++) XF(i) = 1. * (i - 1) * dx;
    XF(1) = 0.;
    XF(nx + 1) = lx;
    for (int i = 1; i <= nx; i++) DXF(i) = XF(i + 1) - XF(i);
    XC(0) = XF(1);
    for (int i = 1; i <= nx; i++) XC(i) = 0.5 * (XF(i) + XF(i + 1));
    XC(nx + 1) = XF(nx + 1);
    for (int i = 1; i <= nx + 1; i++) DXC(i) = XC(i) - XC(i - 1);
    const double dy = ly / jtot;
    s->dy = dy;
    const double yoffset = dy * 0;
    for (int j = 1; j <= ny + 1; j++) YF(j) = yoffset + 1. * (j - 1) * dy;
    for (int j = 1; j <= ny; j++) YC(j) = yoffset + 0.5 * (2 * j - 1) * dy;
  }
  s->alpha[0] = 32. / 60.;
  s->beta[0] = 0. / 60.;
  s->alpha[1] = 25. / 60.;
  s->beta[1] = -17. / 60.;
  s->alpha[2] = 45. / 60.;
  s->beta[2] = -25. / 60.;
  for (int k = 0; k < 3; k++) s->gamma[k] = s->alpha[k] + s->beta[k];
  const long lux = (long)(nx + 1) * (ny + 2);
  const long lp = (long)(nx + 2) * (ny + 2);
  const long lsx = (long)(nx - 1) * ny;
  const long lsy = (long)nx * ny;
  const long lens[ORC_NFIELDS] = {lux, lp, lp, lp, lsx, lsx, lsx,
                                  lsy, lsy, lsy, lp, lp};
  for (int k = 0; k < ORC_NFIELDS; k++) {
    s->flen[k] = lens[k];
    s->f[k] = xcalloc(lens[k], sizeof(double));
  }
  s->qx = xcalloc((size_t)nx * ny, sizeof(double));
  s->qy = xcalloc((size_t)nx * ny, sizeof(double));
  return s;
}

void orc_destroy(orc_t *s) {
  if (!s) return;
  free(s->xf); free(s->xc); free(s->dxf); free(s->dxc); free(s->yf);
  free(s->yc);
  for (int k = 0; k < ORC_NFIELDS; k++) free(s->f[k]);
  free(s->qx); free(s->qy); free(s->prt); free(s->pairs);
  free(s);
}

double *orc_field(orc_t *s, int which) {
  return (which >= 0 && which < ORC_NFIELDS) ? s->f[which] : NULL;
}
long orc_field_len(orc_t *s, int which) {
  return (which >= 0 && which < ORC_NFIELDS) ? s->flen[which] : -1;
}
double *orc_grid(orc_t *s, int which) {
  switch (which) {
    case ORC_XF: return s->xf;
    case ORC_XC: return s->xc;
    case ORC_DXF: return s->dxf;
    case ORC_DXC: return s->dxc;
    case ORC_YF: return s->yf;
    case ORC_YC: return s->yc;
  }
  return NULL;
}
double orc_get_dt(const orc_t *s) { return s->dt; }
void orc_set_dt(orc_t *s, double dt) { s->dt = dt; }
double orc_get_time(const orc_t *s) { return s->time; }
int orc_get_step(const orc_t *s) { return s->step; }

/* particle input as loaded by src/suspensions/init.c:24-86 */
int orc_set_particles(orc_t *s, int n, const double *den, const double *a,
                      const double *b, const double *x, const double *y,
                      const double *az, const double *ux, const double *uy,
                      const double *vz) {
  free(s->prt);
  s->np = n;
  s->prt = xcalloc((size_t)(n > 0 ? n : 1) * NREC, sizeof(double));
  for (int k = 0; k < n; k++) {
    double *p = s->prt + (size_t)k * NREC;
    p[P_DEN] = den[k]; p[P_A] = a[k]; p[P_B] = b[k];
    p[P_X] = x[k]; p[P_Y] = y[k]; p[P_AZ] = az[k];
    p[P_UX] = ux[k]; p[P_UY] = uy[k]; p[P_VZ] = vz[k];
  }
  return 0;
}
int orc_n_particles(const orc_t *s) { return s->np; }
int orc_get_particle_records(const orc_t *s, double *out) {
  memcpy(out, s->prt, sizeof(double) * NREC * (size_t)s->np);
  return 0;
}
int orc_set_particle_records(orc_t *s, const double *in) {
  memcpy(s->prt, in, sizeof(double) * NREC * (size_t)s->np);
  return 0;
}

/* ---------------------------------------------------------------------
 * boundary conditions, src/fluid/boundary_conditions.c:7-62; with one
 * process the y-halo exchange is the periodic copy of rows ny -> 0 and
 * 1 -> ny+1 (src/parallel/halo.c:7-31 with ym = yp = self)
 * ------------------------------------------------------------------- */
int orc_update_boundaries_ux(orc_t *s, double *ux) {
  const int nx = s->itot, ny = s->jtot;
  memcpy(&ux[IX_UX(2, 0)], &ux[IX_UX(2, ny)], sizeof(double) * (nx - 1));
  memcpy(&ux[IX_UX(2, ny + 1)], &ux[IX_UX(2, 1)], sizeof(double) * (nx - 1));
  for (int j = 0; j <= ny + 1; j++) {
    ux[IX_UX(1, j)] = 0.;
    ux[IX_UX(nx + 1, j)] = 0.;
  }
  return 0;
}

int orc_update_boundaries_uy(orc_t *s, double *uy) {
  const int nx = s->itot, ny = s->jtot;
  memcpy(&uy[IX_P(1, 0)], &uy[IX_P(1, ny)], sizeof(double) * nx);
  memcpy(&uy[IX_P(1, ny + 1)], &uy[IX_P(1, 1)], sizeof(double) * nx);
  for (int j = 0; j <= ny + 1; j++) {
    uy[IX_P(0, j)] = s->uy_xm;
    uy[IX_P(nx + 1, j)] = s->uy_xp;
  }
  return 0;
}

int orc_update_boundaries_p(orc_t *s, double *p) {
  const int nx = s->itot, ny = s->jtot;
  memcpy(&p[IX_P(1, 0)], &p[IX_P(1, ny)], sizeof(double) * nx);
  memcpy(&p[IX_P(1, ny + 1)], &p[IX_P(1, 1)], sizeof(double) * nx);
  for (int j = 0; j <= ny + 1; j++) {
    p[IX_P(0, j)] = p[IX_P(1, j)];
    p[IX_P(nx + 1, j)] = p[IX_P(nx, j)];
  }
  return 0;
}

/* ---------------------------------------------------------------------
 * time-step size, src/param/decide_dt.c:11-81
 * ------------------------------------------------------------------- */
int orc_decide_dt(orc_t *s) {
  const int nx = s->itot, ny = s->jtot;
  const double small = 1.e-8;
  const double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY];
  const double dy = s->dy;
  double dt_adv = 5.e-2;
  for (int j = 1; j <= ny; j++)
    for (int i = 2; i <= nx; i++) {
      double vel = fmax(fabs(ux[IX_UX(i, j)]), small);
      dt_adv = fmin(dt_adv, DXC(i) / vel);
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      double vel = fmax(fabs(uy[IX_P(i, j)]), small);
      dt_adv = fmin(dt_adv, dy / vel);
    }
  dt_adv *= s->sf[0];
  double dxmin = s->lx;
  for (int i = 2; i <= nx; i++) dxmin = fmin(dxmin, DXC(i));
  double dt_dif_x = 0.25 * s->Re * pow(dxmin, 2.);
  double dt_dif_y = 0.25 * s->Re * pow(dy, 2.);
  dt_dif_x *= s->sf[1];
  dt_dif_y *= s->sf[1];
  const double delta = fmin(s->dx, s->dy);
  double dt_par = DBL_MAX;
  for (int n = 0; n < s->np; n++) {
    const double *p = s->prt + (size_t)n * NREC;
    double pux = fmax(fabs(p[P_UX]) + fabs(fmax(p[P_A], p[P_B]) * p[P_VZ]), small);
    double puy = fmax(fabs(p[P_UY]) + fabs(fmax(p[P_A], p[P_B]) * p[P_VZ]), small);
    double val = fmax(pux, puy);
    dt_par = fmin(dt_par, delta / val);
  }
  dt_par *= s->sf[2];
  s->dt = fmin(fmin(dt_adv, fmin(dt_dif_x, dt_dif_y)), dt_par);
  return 0;
}

/* ---------------------------------------------------------------------
 * predictor, src/fluid/update_velocity.c:9-237
 * ------------------------------------------------------------------- */
static void src_ux(orc_t *s) {
  const int nx = s->itot, ny = s->jtot;
  const double dy = s->dy, Re = s->Re;
  const double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY], *p = s->f[ORC_P];
  double *sa = s->f[ORC_SRCUXA], *sb = s->f[ORC_SRCUXB], *sg = s->f[ORC_SRCUXG];
  memcpy(sb, sa, sizeof(double) * s->flen[ORC_SRCUXA]);
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++) {
    for (int i = 2; i <= nx; i++) {
      const double u_c = ux[IX_UX(i, j)];
      const double u_w = ux[IX_UX(i - 1, j)];
      const double u_e = ux[IX_UX(i + 1, j)];
      const double u_s = ux[IX_UX(i, j - 1)];
      const double u_n = ux[IX_UX(i, j + 1)];
      double g_w = (-u_w + u_c) / DXF(i - 1);
      double g_e = (-u_c + u_e) / DXF(i);
      double g_s = (-u_s + u_c) / dy;
      double g_n = (-u_c + u_n) / dy;
      double cw = DXF(i - 1) / (2. * DXC(i));
      double ce = DXF(i) / (2. * DXC(i));
      double aw = 0.5 * u_w + 0.5 * u_c;
      double ae = 0.5 * u_c + 0.5 * u_e;
      double adv1 = -cw * aw * g_w - ce * ae * g_e;
      double vs = cw * uy[IX_P(i - 1, j)] + ce * uy[IX_P(i, j)];
      double vn = cw * uy[IX_P(i - 1, j + 1)] + ce * uy[IX_P(i, j + 1)];
      double adv2 = -0.5 * vs * g_s - 0.5 * vn * g_n;
      double dif1 = 1. / Re * (g_e - g_w) / DXC(i);
      double dif2 = 1. / Re * (g_n - g_s) / dy;
      double pre = -(p[IX_P(i, j)] - p[IX_P(i - 1, j)]) / DXC(i);
      sa[IX_SX(i, j)] = (adv1 + adv2) + (dif1 + dif2);
      sg[IX_SX(i, j)] = pre;
    }
  }
}

static void src_uy(orc_t *s) {
  const int nx = s->itot, ny = s->jtot;
  const double dy = s->dy, Re = s->Re, ext = s->extfrcy;
  const double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY], *p = s->f[ORC_P];
  double *sa = s->f[ORC_SRCUYA], *sb = s->f[ORC_SRCUYB], *sg = s->f[ORC_SRCUYG];
  memcpy(sb, sa, sizeof(double) * s->flen[ORC_SRCUYA]);
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++) {
    for (int i = 1; i <= nx; i++) {
      const double v_c = uy[IX_P(i, j)];
      const double v_w = uy[IX_P(i - 1, j)];
      const double v_e = uy[IX_P(i + 1, j)];
      const double v_s = uy[IX_P(i, j - 1)];
      const double v_n = uy[IX_P(i, j + 1)];
      double g_w = (-v_w + v_c) / DXC(i);
      double g_e = (-v_c + v_e) / DXC(i + 1);
      double g_s = (-v_s + v_c) / dy;
      double g_n = (-v_c + v_n) / dy;
      double cw = DXC(i) / (2. * DXF(i));
      double ce = DXC(i + 1) / (2. * DXF(i));
      double uw = 0.5 * ux[IX_UX(i, j - 1)] + 0.5 * ux[IX_UX(i, j)];
      double ue = 0.5 * ux[IX_UX(i + 1, j - 1)] + 0.5 * ux[IX_UX(i + 1, j)];
      double adv1 = -cw * uw * g_w - ce * ue * g_e;
      double as = 0.5 * v_s + 0.5 * v_c;
      double an = 0.5 * v_c + 0.5 * v_n;
      double adv2 = -0.5 * as * g_s - 0.5 * an * g_n;
      double dif1 = 1. / Re * (g_e - g_w) / DXF(i);
      double dif2 = 1. / Re * (g_n - g_s) / dy;
      double pre = -(p[IX_P(i, j)] - p[IX_P(i, j - 1)]) / dy;
      sa[IX_SY(i, j)] = (adv1 + adv2) + (dif1 + dif2) + ext;
      sg[IX_SY(i, j)] = pre;
    }
  }
}

int orc_update_velocity(orc_t *s, int rk) {
  const int nx = s->itot, ny = s->jtot;
  const double al = s->alpha[rk], be = s->beta[rk], ga = s->gamma[rk];
  const double dt = s->dt;
  src_ux(s);
  src_uy(s);
  {
    double *ux = s->f[ORC_UX];
    const double *sa = s->f[ORC_SRCUXA], *sb = s->f[ORC_SRCUXB], *sg = s->f[ORC_SRCUXG];
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 2; i <= nx; i++)
        ux[IX_UX(i, j)] += al * dt * sa[IX_SX(i, j)] + be * dt * sb[IX_SX(i, j)]
                           + ga * dt * sg[IX_SX(i, j)];
    orc_update_boundaries_ux(s, ux);
  }
  {
    double *uy = s->f[ORC_UY];
    const double *sa = s->f[ORC_SRCUYA], *sb = s->f[ORC_SRCUYB], *sg = s->f[ORC_SRCUYG];
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        uy[IX_P(i, j)] += al * dt * sa[IX_SY(i, j)] + be * dt * sb[IX_SY(i, j)]
                          + ga * dt * sg[IX_SY(i, j)];
    orc_update_boundaries_uy(s, uy);
  }
  return 0;
}

/* ---------------------------------------------------------------------
 * DCT (the FFTW REDFT10 / REDFT01 definitions):
 *   REDFT10: Y_k = 2 sum_{j<n} X_j cos(pi (j+1/2) k / n)
 *   REDFT01: Y_k = X_0 + 2 sum_{1<=j<n} X_j cos(pi j (k+1/2) / n)
 * Powers of two use Makhoul's n-point complex-FFT algorithm (iterative
 * radix-2 decimation in time), other lengths up to DCT_MIXED_MAX the same
 * with a Stockham mixed-radix FFT, longer ones the direct sum.  Every
 * trigonometric factor is read from one table C[m] = cos(pi m / (2n)),
 * m = 0..2n, using cos(pi m/2n) = C[4n-m] and sin(pi k/2n) = C[|n-k|].
 * ------------------------------------------------------------------- */
static double *cos_table(int n) {
  double *C = xcalloc(2 * (size_t)n + 1, sizeof(double));
  for (int m = 0; m <= 2 * n; m++) C[m] = cos(ORC_PI * m / (2. * n));
  return C;
}

/* cos(pi m / (2n)) for any m >= 0 */
static double cos_lookup(const double *C, int n, long m) {
  m %= 4 * (long)n;
  return m <= 2 * n ? C[m] : C[4 * (long)n - m];
}

/* in-place radix-2 DIT FFT, sign -1 forward / +1 backward (unnormalised) */
static void fft_pow2(int n, const double *C, double *re, double *im, int sign) {
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1;
    for (int k = 0; k < half; k++) {
      const int m = k * (n / len);
      const double wr = C[4 * m];
      const double wi = sign * C[n >= 4 * m ? n - 4 * m : 4 * m - n];
      for (int b = 0; b < n; b += len) {
        const int i0 = b + k, i1 = b + k + half;
        const double tr = wr * re[i1] - wi * im[i1];
        const double ti = wr * im[i1] + wi * re[i1];
        re[i1] = re[i0] - tr; im[i1] = im[i0] - ti;
        re[i0] = re[i0] + tr; im[i0] = im[i0] + ti;
      }
    }
  }
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* lengths that are not powers of two: the Stockham FFT up to this many
 * values (the device's LDS bound), the direct sum beyond */
#define DCT_MIXED_MAX 5120

/* exp(sign 2 pi i e / n) from the table: cos(2 pi e/n) = C(4e),
 * sin(2 pi e/n) = C(|n - 4e|) */
static void unit_root(const double *C, int n, long e, double sign, double *wr, double *wi) {
  const long d = (long)n - 4 * e;
  *wr = cos_lookup(C, n, 4 * e);
  *wi = sign * cos_lookup(C, n, d < 0 ? -d : d);
}

/* Stockham autosort FFT of any length n (sign -1 forward, +1 backward,
 * unnormalised); stage radices are n's prime factors in increasing order.
 * In the stage of radix R after radices of product Ns, butterfly j
 * (k = j mod Ns) takes v_r = x[j + r n/R] w^(k r), w = e^(sign 2 pi i/(Ns R)),
 * and stores sum_r v_r e^(sign 2 pi i q r / R) at (j/Ns) Ns R + k + q Ns.
 * Factors with exponent 0 are not multiplied.  The result is left in
 * (re, im); (tr, ti) is the second buffer. */
static void fft_mixed(int n, const double *C, double *re, double *im, double *tr, double *ti,
                      int sign) {
  double *xr = re, *xi = im, *yr = tr, *yi = ti;
  double *vr = xcalloc(n, sizeof(double)), *vi = xcalloc(n, sizeof(double));
  int Ns = 1, m = n;
  while (m > 1) {
    int R = 2;
    while (m % R) R++;
    m /= R;
    const int nR = n / R, tw = n / (Ns * R);
    for (int j = 0; j < nR; j++) {
      const int k = j % Ns;
      const int o = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; r++) {
        vr[r] = xr[j + r * nR];
        vi[r] = xi[j + r * nR];
        const long e = (long)k * r * tw;
        if (e != 0) {
          double wr, wi;
          unit_root(C, n, e, sign, &wr, &wi);
          const double a = vr[r] * wr - vi[r] * wi, b = vr[r] * wi + vi[r] * wr;
          vr[r] = a;
          vi[r] = b;
        }
      }
      for (int q = 0; q < R; q++) {
        double sr = vr[0], si = vi[0];
        for (int r = 1; r < R; r++) {
          const long e2 = (long)((q * r) % R) * nR;
          if (e2 == 0) {
            sr = sr + vr[r];
            si = si + vi[r];
          } else {
            double cr, ci;
            unit_root(C, n, e2, sign, &cr, &ci);
            sr = sr + (vr[r] * cr - vi[r] * ci);
            si = si + (vr[r] * ci + vi[r] * cr);
          }
        }
        yr[o + q * Ns] = sr;
        yi[o + q * Ns] = si;
      }
    }
    double *t = xr; xr = yr; yr = t;
    t = xi; xi = yi; yi = t;
    Ns *= R;
  }
  if (xr != re) {
    memcpy(re, xr, sizeof(double) * n);
    memcpy(im, xi, sizeof(double) * n);
  }
  free(vr);
  free(vi);
}

static void dct_direct(int n, const double *C, double *x, int type3) {
  double *y = xcalloc(n, sizeof(double));
  for (int k = 0; k < n; k++) {
    double acc = type3 ? x[0] : 0.;
    for (int j = type3 ? 1 : 0; j < n; j++) {
      const long m = type3 ? (long)j * (2 * k + 1) : (long)(2 * j + 1) * k;
      acc = acc + 2. * x[j] * cos_lookup(C, n, m);
    }
    y[k] = acc;
  }
  memcpy(x, y, sizeof(double) * n);
  free(y);
}

static void dct2_tab(int n, const double *C, double *x, double *re, double *im) {
  if (!is_pow2(n) && n > DCT_MIXED_MAX) {
    dct_direct(n, C, x, 0);
    return;
  }
  if (!is_pow2(n)) {  /* Makhoul order, Stockham FFT */
    double *tr = xcalloc(n, sizeof(double)), *ti = xcalloc(n, sizeof(double));
    for (int i = 0; i < n; i++) {
      const int m = (i & 1) ? n - 1 - (i >> 1) : (i >> 1);
      re[m] = x[i];
      im[m] = 0.;
    }
    fft_mixed(n, C, re, im, tr, ti, -1);
    for (int k = 0; k < n; k++) x[k] = 2. * (C[k] * re[k] + C[n - k] * im[k]);
    free(tr);
    free(ti);
    return;
  }
  for (int j = 0; j < n / 2; j++) {
    re[j] = x[2 * j];
    re[n - 1 - j] = x[2 * j + 1];
  }
  for (int j = 0; j < n; j++) im[j] = 0.;
  fft_pow2(n, C, re, im, -1);
  for (int k = 0; k < n; k++) x[k] = 2. * (C[k] * re[k] + C[n - k] * im[k]);
}

static void dct3_tab(int n, const double *C, double *x, double *re, double *im) {
  if (!is_pow2(n) && n > DCT_MIXED_MAX) {
    dct_direct(n, C, x, 1);
    return;
  }
  if (!is_pow2(n)) {  /* pre-twiddled spectrum, Stockham FFT, Makhoul order */
    double *tr = xcalloc(n, sizeof(double)), *ti = xcalloc(n, sizeof(double));
    for (int k = 0; k < n; k++) {
      const double xr = x[k];
      const double xi = k == 0 ? 0. : -x[n - k];
      re[k] = C[k] * xr - C[n - k] * xi;
      im[k] = C[k] * xi + C[n - k] * xr;
    }
    fft_mixed(n, C, re, im, tr, ti, +1);
    for (int i = 0; i < n; i++) x[i] = (i & 1) ? re[n - 1 - (i >> 1)] : re[i >> 1];
    free(tr);
    free(ti);
    return;
  }
  for (int k = 0; k < n; k++) {
    const double xr = x[k];
    const double xi = k == 0 ? 0. : -x[n - k];
    re[k] = C[k] * xr - C[n - k] * xi;
    im[k] = C[k] * xi + C[n - k] * xr;
  }
  fft_pow2(n, C, re, im, +1);
  for (int j = 0; j < n / 2; j++) {
    x[2 * j] = re[j];
    x[2 * j + 1] = re[n - 1 - j];
  }
}

/* Two real rows through one complex FFT (n a power of two): row a is the
 * real and row b the imaginary part of the Makhoul-ordered input.  With Z
 * the transform and kk = (n - k) mod n, row a's spectrum is
 * (Z[k] + conj Z[kk]) / 2 and row b's (Z[k] - conj Z[kk]) / 2i; the halves
 * cancel against REDFT10's factor 2. */
static void dct2_pair_tab(int n, const double *C, double *xa, double *xb, double *re,
                          double *im) {
  for (int j = 0; j < n / 2; j++) {
    re[j] = xa[2 * j];
    re[n - 1 - j] = xa[2 * j + 1];
    im[j] = xb[2 * j];
    im[n - 1 - j] = xb[2 * j + 1];
  }
  fft_pow2(n, C, re, im, -1);
  for (int k = 0; k < n; k++) {
    const int kk = (n - k) & (n - 1);
    xa[k] = C[k] * (re[k] + re[kk]) + C[n - k] * (im[k] - im[kk]);
    xb[k] = C[k] * (im[k] + im[kk]) + C[n - k] * (re[kk] - re[k]);
  }
}

/* The inverse pair: the pre-twiddled spectra Va (row a) and Vb (row b) go
 * in as Va + i Vb, and the two real rows come out as the real and the
 * imaginary part. */
static void dct3_pair_tab(int n, const double *C, double *xa, double *xb, double *re,
                          double *im) {
  for (int k = 0; k < n; k++) {
    const double ar = xa[k], ai = k == 0 ? 0. : -xa[n - k];
    const double br = xb[k], bi = k == 0 ? 0. : -xb[n - k];
    const double var = C[k] * ar - C[n - k] * ai, vai = C[k] * ai + C[n - k] * ar;
    const double vbr = C[k] * br - C[n - k] * bi, vbi = C[k] * bi + C[n - k] * br;
    re[k] = var - vbi;
    im[k] = vai + vbr;
  }
  fft_pow2(n, C, re, im, +1);
  for (int j = 0; j < n / 2; j++) {
    xa[2 * j] = re[j];
    xa[2 * j + 1] = re[n - 1 - j];
    xb[2 * j] = im[j];
    xb[2 * j + 1] = im[n - 1 - j];
  }
}

/* row lengths whose Poisson solve pairs rows (the device's register-FFT
 * sizes) */
static int pairs_rows(int n) { return is_pow2(n) && n >= 64 && n <= 8192; }

int orc_dct2_pair(int n, double *xa, double *xb) {
  if (!is_pow2(n)) return -1;
  double *C = cos_table(n);
  double *re = xcalloc(n, sizeof(double)), *im = xcalloc(n, sizeof(double));
  dct2_pair_tab(n, C, xa, xb, re, im);
  free(C); free(re); free(im);
  return 0;
}

int orc_dct3_pair(int n, double *xa, double *xb) {
  if (!is_pow2(n)) return -1;
  double *C = cos_table(n);
  double *re = xcalloc(n, sizeof(double)), *im = xcalloc(n, sizeof(double));
  dct3_pair_tab(n, C, xa, xb, re, im);
  free(C); free(re); free(im);
  return 0;
}

int orc_dct2(int n, double *x) {
  double *C = cos_table(n);
  double *re = xcalloc(n, sizeof(double)), *im = xcalloc(n, sizeof(double));
  dct2_tab(n, C, x, re, im);
  free(C); free(re); free(im);
  return 0;
}

int orc_dct3(int n, double *x) {
  double *C = cos_table(n);
  double *re = xcalloc(n, sizeof(double)), *im = xcalloc(n, sizeof(double));
  dct3_tab(n, C, x, re, im);
  free(C); free(re); free(im);
  return 0;
}

/* ---------------------------------------------------------------------
 * tridiagonal solvers, src/tdm.c:37-111 (double path only)
 * ------------------------------------------------------------------- */
static void gtsv_thomas(int n, double *l, double *c, double *u, double *q) {
  u[0] = u[0] / c[0];
  q[0] = q[0] / c[0];
  for (int i = 1; i <= n - 1; i++) {
    double val = c[i] - l[i] * u[i - 1];
    if (fabs(val) < DBL_EPSILON) {
      u[i] = 0.;
      q[i] = 0.;
    } else {
      val = 1. / val;
      u[i] = val * u[i];
      q[i] = val * (q[i] - l[i] * q[i - 1]);
    }
  }
  for (int i = n - 2; i >= 0; i--) q[i] -= u[i] * q[i + 1];
}

/* work: 7*(n-1) doubles */
static void gtsv_periodic(int n, const double *l, const double *c,
                          const double *u, double *q, double *work) {
  double *l0 = work, *c0 = l0 + (n - 1), *u0 = c0 + (n - 1);
  double *l1 = u0 + (n - 1), *c1 = l1 + (n - 1), *u1 = c1 + (n - 1);
  double *q1 = u1 + (n - 1);
  for (int i = 0; i < n - 1; i++) {
    l0[i] = l[i]; c0[i] = c[i]; u0[i] = u[i];
    l1[i] = l[i]; c1[i] = c[i]; u1[i] = u[i];
    q1[i] = i == 0 ? -l[i] : i == n - 2 ? -u[i] : 0.;
  }
  gtsv_thomas(n - 1, l0, c0, u0, q);
  gtsv_thomas(n - 1, l1, c1, u1, q1);
  double num = q[n - 1] - u[n - 1] * q[0] - l[n - 1] * q[n - 2];
  double den = c[n - 1] + u[n - 1] * q1[0] + l[n - 1] * q1[n - 2];
  q[n - 1] = fabs(den) < DBL_EPSILON ? 0. : num / den;
  for (int i = 0; i <= n - 2; i++) q[i] = q[i] + q[n - 1] * q1[i];
}

int orc_tdm_periodic(int n, const double *l, const double *c, const double *u,
                     double *q) {
  double *work = xcalloc(7 * (size_t)(n - 1), sizeof(double));
  gtsv_periodic(n, l, c, u, q, work);
  free(work);
  return 0;
}

/* ---------------------------------------------------------------------
 * Poisson solve, src/fluid/compute_potential.c:48-138 (one process:
 * the transposes are plain index transposes)
 * ------------------------------------------------------------------- */
int orc_compute_potential(orc_t *s, int rk) {
  const int nx = s->itot, ny = s->jtot;
  const double dx = s->dx, dy = s->dy;
  const double gamma = s->gamma[rk], dt = s->dt;
  const double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY];
  double *qx = s->qx, *qy = s->qy, *psi = s->f[ORC_PSI];
  double *C = cos_table(nx);
  const int paired = pairs_rows(nx);
#pragma omp parallel
  {
  double *re = xcalloc(nx, sizeof(double)), *im = xcalloc(nx, sizeof(double));
  double *zero = xcalloc(nx, sizeof(double));
#pragma omp for schedule(static)
  for (int j = 1; j <= ny; j++) {
    for (int i = 1; i <= nx; i++) {
      double ux_xm = ux[IX_UX(i, j)];
      double ux_xp = ux[IX_UX(i + 1, j)];
      double uy_ym = uy[IX_P(i, j)];
      double uy_yp = uy[IX_P(i, j + 1)];
      qx[(long)(j - 1) * nx + (i - 1)] =
          1. / (gamma * dt) * (+(ux_xp - ux_xm) / dx + (uy_yp - uy_ym) / dy);
    }
  }
  /* rows 2p and 2p+1 (from 0) share one FFT; an odd last row pairs with a
   * zero row */
#pragma omp for schedule(static)
  for (int jp = 0; jp < (paired ? (ny + 1) / 2 : ny); jp++) {
    if (!paired) {
      dct2_tab(nx, C, &qx[(long)jp * nx], re, im);
      continue;
    }
    double *ra = &qx[(long)(2 * jp) * nx];
    double *rb = 2 * jp + 1 < ny ? &qx[(long)(2 * jp + 1) * nx] : zero;
    if (rb == zero) memset(zero, 0, sizeof(double) * nx);
    dct2_pair_tab(nx, C, ra, rb, re, im);
  }
  free(re); free(im); free(zero);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) qy[(long)i * ny + j] = qx[(long)j * nx + i];
#pragma omp parallel
  {
    double *l = xcalloc(ny, sizeof(double));
    double *c = xcalloc(ny, sizeof(double));
    double *u = xcalloc(ny, sizeof(double));
    double *work = xcalloc(7 * (size_t)(ny - 1), sizeof(double));
#pragma omp for schedule(static)
    for (int i = 1; i <= nx; i++) {
      double eigenvalue = -4. / pow(dx, 2.) * pow(sin(ORC_PI * (i - 1) / (2. * nx)), 2.);
      for (int j = 0; j < ny; j++) {
        l[j] = 1. / dy / dy;
        u[j] = 1. / dy / dy;
        c[j] = -l[j] - u[j] + eigenvalue;
      }
      gtsv_periodic(ny, l, c, u, &qy[(long)(i - 1) * ny], work);
    }
    free(l); free(c); free(u); free(work);
  }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) qx[(long)j * nx + i] = qy[(long)i * ny + j];
#pragma omp parallel
  {
  double *re = xcalloc(nx, sizeof(double)), *im = xcalloc(nx, sizeof(double));
  double *zero = xcalloc(nx, sizeof(double));
#pragma omp for schedule(static)
  for (int jp = 0; jp < (paired ? (ny + 1) / 2 : ny); jp++) {
    if (!paired) {
      dct3_tab(nx, C, &qx[(long)jp * nx], re, im);
      continue;
    }
    double *ra = &qx[(long)(2 * jp) * nx];
    double *rb = 2 * jp + 1 < ny ? &qx[(long)(2 * jp + 1) * nx] : zero;
    if (rb == zero) memset(zero, 0, sizeof(double) * nx);
    dct3_pair_tab(nx, C, ra, rb, re, im);
  }
  free(re); free(im); free(zero);
  }
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) psi[IX_P(i, j)] = qx[(long)(j - 1) * nx + i - 1] / (2. * nx);
  free(C);
  orc_update_boundaries_p(s, psi);
  return 0;
}

/* projection, src/fluid/correct_velocity.c:6-42 */
int orc_correct_velocity(orc_t *s, int rk) {
  const int nx = s->itot, ny = s->jtot;
  const double dy = s->dy, gamma = s->gamma[rk], dt = s->dt;
  const double *psi = s->f[ORC_PSI];
  double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY];
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 2; i <= nx; i++) {
      double dxi = 1. / DXC(i);
      ux[IX_UX(i, j)] -= gamma * dt * (psi[IX_P(i, j)] - psi[IX_P(i - 1, j)]) * dxi;
    }
  orc_update_boundaries_ux(s, ux);
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++)
      uy[IX_P(i, j)] -= gamma * dt * (psi[IX_P(i, j)] - psi[IX_P(i, j - 1)]) / dy;
  orc_update_boundaries_uy(s, uy);
  return 0;
}

/* pressure, src/fluid/update_pressure.c:7-24 */
int orc_update_pressure(orc_t *s) {
  const int nx = s->itot, ny = s->jtot;
  const double *psi = s->f[ORC_PSI];
  double *p = s->f[ORC_P];
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) p[IX_P(i, j)] += psi[IX_P(i, j)];
  orc_update_boundaries_p(s, p);
  return 0;
}

/* ---------------------------------------------------------------------
 * ellipse geometry, src/ellipse.c:7-65
 * ------------------------------------------------------------------- */
static double evolute_x(double a, double b, double t) {
  return a * (1. - pow(b / a, 2.)) * pow(cos(t), 3.);
}
static double evolute_y(double a, double b, double t) {
  return b * (1. - pow(a / b, 2.)) * pow(sin(t), 3.);
}
static double curvature(double a, double b, double t) {
  const double ct = cos(t), st = sin(t);
  const double ct2 = pow(ct, 2.), st2 = pow(st, 2.);
  const double a2 = pow(a, 2.), b2 = pow(b, 2.);
  const double num = a * b * (ct2 + st2);
  const double den = pow(+a2 * st2 + b2 * ct2, 1.5);
  return num / den;
}

double orc_find_normal_t(double a, double b, double x_, double y_) {
  const double px = fabs(x_), py = fabs(y_);
  double t = 0.25 * ORC_PI;
  for (int n = 0;; n++) {
    const double x = a * cos(t), y = b * sin(t);
    const double ex = evolute_x(a, b, t), ey = evolute_y(a, b, t);
    const double rx = x - ex, ry = y - ey;
    const double qx = px - ex, qy = py - ey;
    const double r = sqrt(pow(ry, 2.) + pow(rx, 2.));
    const double q = sqrt(pow(qy, 2.) + pow(qx, 2.));
    const double dc = r * asin((rx * qy - ry * qx) / (r * q));
    const double dt = dc / sqrt(a * a + b * b - x * x - y * y);
    t += dt;
    t = fmin(0.5 * ORC_PI, t);
    t = fmax(0., t);
    if (fabs(dt) < 1.e-8) break;
    if (n > 100) break;
  }
  if (x_ < 0.) t = ORC_PI - t;
  if (y_ < 0.) t = -t;
  return t;
}

/* ---------------------------------------------------------------------
 * IBM helpers, src/suspensions/auxiliary.c:11-75
 * ------------------------------------------------------------------- */
#define NEXTRA 3
int orc_decide_loop_size(int lbound, int ubound, double grid_size,
                         double radius, double grav_center, int *min, int *max) {
  *min = (int)(floor((grav_center - radius) / grid_size) - NEXTRA);
  *max = (int)(ceil((grav_center + radius) / grid_size) + NEXTRA);
  *min = *min < lbound ? lbound : *min;
  *max = *max > ubound ? ubound : *max;
  return 0;
}

static double signed_dist(double grid_size, double pa, double pb, double px,
                          double py, double paz, double x, double y) {
  double x_, y_;
  {
    const double dx = x - px, dy = y - py;
    const double c = cos(-paz), sn = sin(-paz);
    x_ = c * dx - sn * dy;
    y_ = sn * dx + c * dy;
  }
  const double f0 = 1. - (x_ / pa) * (x_ / pa) - (y_ / pb) * (y_ / pb);
  const double sign = f0 > 0. ? +1. : -1.;
  const double t = orc_find_normal_t(pa, pb, x_, y_);
  const double ex = pa * cos(t) - x_, ey = pb * sin(t) - y_;
  const double dist = sign * sqrt(+ex * ex + ey * ey);
  return dist / grid_size;
}

#define BETA 2.
double orc_s_weight(double grid_size, double pa, double pb, double px,
                    double py, double paz, double x, double y) {
  const double d = signed_dist(grid_size, pa, pb, px, py, paz, x, y);
  const double th = tanh(BETA * d);
  return 0.5 * BETA * (1. - th * th);
}
double orc_v_weight(double grid_size, double pa, double pb, double px,
                    double py, double paz, double x, double y) {
  const double d = signed_dist(grid_size, pa, pb, px, py, paz, x, y);
  return 0.5 * (1. + tanh(BETA * d));
}

static double p_mass(const double *p) {
  const double vol = ORC_PI * p[P_A] * p[P_B];
  return p[P_DEN] * vol;
}
static double p_moi(const double *p) {
  const double m = p_mass(p);
  return 0.25 * m * (p[P_A] * p[P_A] + p[P_B] * p[P_B]);
}

/* ---------------------------------------------------------------------
 * suspensions
 * ------------------------------------------------------------------- */
/* src/suspensions/reset.c:7-18 */
int orc_reset_particle_increments(orc_t *s) {
  for (int n = 0; n < s->np; n++) {
    double *p = s->prt + (size_t)n * NREC;
    p[P_DUX] = 0.; p[P_DUY] = 0.; p[P_DVZ] = 0.;
    p[P_DX] = 0.; p[P_DY] = 0.; p[P_DAZ] = 0.;
  }
  return 0;
}

/* src/suspensions/inertia.c:9-70 (+ the identity Allreduce :72-94) */
int orc_compute_inertia(orc_t *s, int cnstep) {
  const int nx = s->itot, ny = s->jtot;
  const double ly = s->ly, dx = s->dx, dy = s->dy;
  const double grid_size = fmin(dx, dy);
  const double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY];
#pragma omp parallel for schedule(dynamic, 1)
  for (int n = 0; n < s->np; n++) {
    double *p = s->prt + (size_t)n * NREC;
    const double pa = p[P_A], pb = p[P_B];
    const double px = p[P_X] + p[P_DX];
    const double py = p[P_Y] + p[P_DY];
    const double paz = p[P_AZ] + p[P_DAZ];
    const double pm = p_mass(p), pim = p_moi(p);
    double iux = 0., iuy = 0., ivz = 0.;
    for (int periodic = -1; periodic <= 1; periodic++) {
      const double py_ = py + ly * periodic;
      int imin, imax, jmin, jmax;
      orc_decide_loop_size(1, nx, dx, fmax(pa, pb), px, &imin, &imax);
      orc_decide_loop_size(1, ny, dy, fmax(pa, pb), py_ - YF(1), &jmin, &jmax);
      for (int j = jmin; j <= jmax; j++) {
        const double y = YC(j);
        for (int i = imin; i <= imax; i++) {
          const double x = XC(i);
          const double w = orc_v_weight(grid_size, pa, pb, px, py_, paz, x, y);
          const double valx = w * 0.5 * (ux[IX_UX(i, j)] + ux[IX_UX(i + 1, j)]) * (dx * dy);
          const double valy = w * 0.5 * (uy[IX_P(i, j)] + uy[IX_P(i, j + 1)]) * (dx * dy);
          iux += valx / pm;
          ivz += -(y - py_) * valx / pim;
          iuy += valy / pm;
          ivz += +(x - px) * valy / pim;
        }
      }
    }
    p[P_IUX + cnstep] = iux;
    p[P_IUY + cnstep] = iuy;
    p[P_IVZ + cnstep] = ivz;
  }
  return 0;
}

/* src/suspensions/exchange.c:13-88 (+ identity Allreduce :90-112).
 * Each particle's per-cell forcing is computed in parallel into its own
 * buffer; the field accumulation then runs in the reference's particle
 * order, so the sums are the same as the serial loop.  The reference
 * refreshes the halo of dux/duy after every particle (:80-81); the halo
 * is a function of the interior only, so one refresh at the end gives the
 * same arrays. */
int orc_exchange_momentum(orc_t *s) {
  const int nx = s->itot, ny = s->jtot;
  const double dt = s->dt, ly = s->ly, dx = s->dx, dy = s->dy;
  const double grid_size = fmin(dx, dy);
  const double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY];
  double *dux = s->f[ORC_DUX], *duy = s->f[ORC_DUY];
  memset(dux, 0, sizeof(double) * s->flen[ORC_DUX]);
  memset(duy, 0, sizeof(double) * s->flen[ORC_DUY]);
  const int np = s->np;
  int (*box)[3][4] = xcalloc(np > 0 ? np : 1, sizeof(*box));
  long *off = xcalloc((size_t)np + 1, sizeof(long));
  for (int n = 0; n < np; n++) {
    const double *p = s->prt + (size_t)n * NREC;
    long cnt = 0;
    for (int k = 0; k < 3; k++) {
      const double py_ = p[P_Y] + ly * (k - 1);
      int *b = box[n][k];
      orc_decide_loop_size(1, nx, dx, fmax(p[P_A], p[P_B]), p[P_X], &b[0], &b[1]);
      orc_decide_loop_size(1, ny, dy, fmax(p[P_A], p[P_B]), py_ - YF(1), &b[2], &b[3]);
      if (b[1] >= b[0] && b[3] >= b[2]) cnt += (long)(b[1] - b[0] + 1) * (b[3] - b[2] + 1);
    }
    off[n + 1] = off[n] + cnt;
  }
  double *cx = xcalloc((size_t)off[np] + 1, sizeof(double));
  double *cy = xcalloc((size_t)off[np] + 1, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
  for (int n = 0; n < np; n++) {
    double *p = s->prt + (size_t)n * NREC;
    const double pa = p[P_A], pb = p[P_B];
    const double pm = p_mass(p), pim = p_moi(p);
    const double px = p[P_X], py = p[P_Y], paz = p[P_AZ];
    const double pux = p[P_UX], puy = p[P_UY], pvz = p[P_VZ];
    double fux = 0., fuy = 0., tvz = 0.;
    long c = off[n];
    for (int k = 0; k < 3; k++) {
      const double py_ = py + ly * (k - 1);
      const int *b = box[n][k];
      for (int j = b[2]; j <= b[3]; j++) {
        const double y = YC(j);
        for (int i = b[0]; i <= b[1]; i++) {
          const double x = XC(i);
          const double w = orc_s_weight(grid_size, pa, pb, px, py_, paz, x, y);
          const double ux_p = pux - pvz * (y - py_);
          const double uy_p = puy + pvz * (x - px);
          const double ux_f = 0.5 * (ux[IX_UX(i, j)] + ux[IX_UX(i + 1, j)]);
          const double uy_f = 0.5 * (uy[IX_P(i, j)] + uy[IX_P(i, j + 1)]);
          const double fx = w * (ux_p - ux_f) / dt;
          const double fy = w * (uy_p - uy_f) / dt;
          cx[c] = fx * dt;
          cy[c] = fy * dt;
          c++;
          fux -= +(fx * dx * dy) / pm;
          tvz -= -(y - py_) * (fx * dx * dy) / pim;
          fuy -= +(fy * dx * dy) / pm;
          tvz -= +(x - px) * (fy * dx * dy) / pim;
        }
      }
    }
    p[P_FUX] = fux;
    p[P_FUY] = fuy;
    p[P_TVZ] = tvz;
  }
  for (int n = 0; n < np; n++) {
    long c = off[n];
    for (int k = 0; k < 3; k++) {
      const int *b = box[n][k];
      for (int j = b[2]; j <= b[3]; j++)
        for (int i = b[0]; i <= b[1]; i++) {
          dux[IX_P(i, j)] += cx[c];
          duy[IX_P(i, j)] += cy[c];
          c++;
        }
    }
  }
  orc_update_boundaries_p(s, dux);
  orc_update_boundaries_p(s, duy);
  free(cx); free(cy); free(off); free(box);
  return 0;
}

/* src/suspensions/update.c:36-59 */
int orc_update_momentum_field(orc_t *s) {
  const int nx = s->itot, ny = s->jtot;
  double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY];
  const double *dux = s->f[ORC_DUX], *duy = s->f[ORC_DUY];
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 2; i <= nx; i++)
      ux[IX_UX(i, j)] += 0.5 * (dux[IX_P(i - 1, j)] + dux[IX_P(i, j)]);
  orc_update_boundaries_ux(s, ux);
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++)
      uy[IX_P(i, j)] += 0.5 * (duy[IX_P(i, j - 1)] + duy[IX_P(i, j)]);
  orc_update_boundaries_uy(s, uy);
  return 0;
}

/* ---------------------------------------------------------------------
 * collisions, src/suspensions/collision.c:41-367
 * ------------------------------------------------------------------- */
typedef struct { double a, b, x, y, angle; } ell_t;
typedef struct { double x, y, r; } circ_t;

static int pair_left(int np, int n0) { return (2 * np - n0 - 1) * (n0) / 2; }
static int pair_right(int np, int n0) { return (2 * np - n0 - 2) * (n0 + 1) / 2 - 1; }
static void pair_indices(int np, int n, int *n0, int *n1) {
  for (*n0 = 0;; (*n0)++)
    if (pair_left(np, *n0) <= n && n <= pair_right(np, *n0)) break;
  *n1 = n - pair_left(np, *n0) + (*n0) + 1;
}

static double harmonic(double v0, double v1) {
  const double r = 0.5 * (1. / v0 + 1. / v1);
  return 1. / r;
}

static void evolute_fit(const ell_t *e, double x, double y, double *ex,
                        double *ey, double *ix, double *iy, double *r) {
  double x_, y_;
  {
    const double x__ = x - e->x, y__ = y - e->y;
    x_ = cos(-e->angle) * x__ - sin(-e->angle) * y__;
    y_ = sin(-e->angle) * x__ + cos(-e->angle) * y__;
  }
  const double t = orc_find_normal_t(e->a, e->b, x_, y_);
  *r = 1. / fmax(curvature(e->a, e->b, t), DBL_EPSILON);
  x_ = evolute_x(e->a, e->b, t);
  y_ = evolute_y(e->a, e->b, t);
  *ex = cos(+e->angle) * x_ - sin(+e->angle) * y_;
  *ey = sin(+e->angle) * x_ + cos(+e->angle) * y_;
  *ex = *ex + e->x;
  *ey = *ey + e->y;
  x_ = e->a * cos(t);
  y_ = e->b * sin(t);
  *ix = cos(+e->angle) * x_ - sin(+e->angle) * y_;
  *iy = sin(+e->angle) * x_ + cos(+e->angle) * y_;
}

static void equivalent_circles(const ell_t *e0, const ell_t *e1, circ_t *c0,
                               circ_t *c1, double *ix0, double *iy0,
                               double *ix1, double *iy1) {
  const double small = 1.e-8;
  c0->x = e0->x; c0->y = e0->y;
  c1->x = e1->x; c1->y = e1->y;
  for (;;) {
    double c0x, c0y, c1x, c1y;
    evolute_fit(e0, c1->x, c1->y, &c0x, &c0y, ix0, iy0, &c0->r);
    evolute_fit(e1, c0->x, c0->y, &c1x, &c1y, ix1, iy1, &c1->r);
    const double d0x = c0x - c0->x, d0y = c0y - c0->y;
    const double d1x = c1x - c1->x, d1y = c1y - c1->y;
    c0->x += d0x; c0->y += d0y;
    c1->x += d1x; c1->y += d1y;
    if (fmax(fmax(fabs(d0x), fabs(d0y)), fmax(fabs(d1x), fabs(d1y))) < small) break;
  }
}

static void equivalent_circle_wall(const ell_t *e, double wallx, circ_t *c,
                                   double *ix, double *iy) {
  const double small = 1.e-8;
  c->x = e->x;
  c->y = e->y;
  for (;;) {
    double cx, cy;
    const double mirrorx = 2. * wallx - c->x;
    evolute_fit(e, mirrorx, c->y, &cx, &cy, ix, iy, &c->r);
    const double ddx = cx - c->x, ddy = cy - c->y;
    c->x += ddx;
    c->y += ddy;
    if (fmax(fabs(ddx), fabs(ddy)) < small) break;
  }
}

/* returns 1 when the pair overlaps (force applied), 0 otherwise */
int orc_collide_pair(double ly, int cnstep, double *p0, double *p1) {
  double yoffset = 0.;
  {
    const double p0y = p0[P_Y] + p0[P_DY];
    const double p1y = p1[P_Y] + p1[P_DY];
    double minval = DBL_MAX;
    for (int periodic = -1; periodic <= 1; periodic++) {
      const double val = fabs((p1y + ly * periodic) - p0y);
      if (val < minval) {
        minval = val;
        yoffset = ly * periodic;
      }
    }
  }
  {
    const double x0 = p0[P_X] + p0[P_DX], y0 = p0[P_Y] + p0[P_DY];
    const double r0 = fmax(p0[P_A], p0[P_B]);
    const double x1 = p1[P_X] + p1[P_DX], y1 = p1[P_Y] + p1[P_DY] + yoffset;
    const double r1 = fmax(p1[P_A], p1[P_B]);
    const double norm = fmax(hypot(x1 - x0, y1 - y0), DBL_EPSILON);
    if (r0 + r1 - norm < 0.) return 0;
  }
  ell_t e0 = {p0[P_A], p0[P_B], p0[P_X] + p0[P_DX], p0[P_Y] + p0[P_DY], p0[P_AZ]};
  ell_t e1 = {p1[P_A], p1[P_B], p1[P_X] + p1[P_DX], p1[P_Y] + p1[P_DY] + yoffset,
              p1[P_AZ]};
  circ_t c0, c1;
  double ix0, iy0, ix1, iy1;
  equivalent_circles(&e0, &e1, &c0, &c1, &ix0, &iy0, &ix1, &iy1);
  const double m0 = p_mass(p0), m1 = p_mass(p1);
  const double i0 = p_moi(p0), i1 = p_moi(p1);
  double k;
  {
    const double mass = harmonic(m0, m1);
    const double r0 = sqrt(p0[P_A] * p0[P_B]);
    const double r1 = sqrt(p1[P_A] * p1[P_B]);
    const double reft = pow(harmonic(r0, r1), 1.5);
    k = mass * pow(ORC_PI, 2.) / pow(reft, 2.);
  }
  double nxv = c1.x - c0.x, nyv = c1.y - c0.y;
  const double norm = fmax(hypot(nxv, nyv), DBL_EPSILON);
  nxv /= norm;
  nyv /= norm;
  const double overlap = c0.r + c1.r - norm;
  if (overlap > 0.) {
    const double cfx = k * overlap * nxv;
    const double cfy = k * overlap * nyv;
    p0[P_CFX + cnstep] += 1. / m0 * (-cfx);
    p1[P_CFX + cnstep] += 1. / m1 * (+cfx);
    p0[P_CFY + cnstep] += 1. / m0 * (-cfy);
    p1[P_CFY + cnstep] += 1. / m1 * (+cfy);
    p0[P_CTZ + cnstep] += 1. / i0 * (ix0 * (-cfy) - iy0 * (-cfx));
    p1[P_CTZ + cnstep] += 1. / i1 * (ix1 * (+cfy) - iy1 * (+cfx));
    return 1;
  }
  return 0;
}

int orc_collide_wall(double wallx, int cnstep, double *p) {
  {
    const double x = p[P_X] + p[P_DX];
    const double r = fmax(p[P_A], p[P_B]);
    const double norm = fabs(wallx - x);
    if (r - norm < 0.) return 0;
  }
  ell_t e = {p[P_A], p[P_B], p[P_X] + p[P_DX], p[P_Y] + p[P_DY], p[P_AZ]};
  circ_t c;
  double ix, iy;
  equivalent_circle_wall(&e, wallx, &c, &ix, &iy);
  const double m = p_mass(p), im = p_moi(p);
  double k;
  {
    const double mass = m;
    const double r = sqrt(p[P_A] * p[P_B]);
    const double reft = pow(r, 1.5);
    k = mass * pow(ORC_PI, 2.) / pow(reft, 2.);
  }
  double nxv = wallx - c.x;
  const double norm = fabs(nxv);
  nxv /= norm;
  const double overlap = c.r - norm;
  if (overlap > 0.) {
    const double cfx = k * overlap * nxv;
    const double cfy = 0.;
    p[P_CFX + cnstep] += 1. / m * (-cfx);
    p[P_CFY + cnstep] += 1. / m * (-cfy);
    p[P_CTZ + cnstep] += 1. / im * (ix * (-cfy) - iy * (-cfx));
    return 1;
  }
  return 0;
}

static void record_pair(orc_t *s, int n0, int n1) {
  if (s->npairs == s->cap_pairs) {
    s->cap_pairs = s->cap_pairs ? 2 * s->cap_pairs : 64;
    int *np_ = realloc(s->pairs, sizeof(int) * 2 * (size_t)s->cap_pairs);
    if (!np_) abort();
    s->pairs = np_;
  }
  s->pairs[2 * s->npairs] = n0;
  s->pairs[2 * s->npairs + 1] = n1;
  s->npairs++;
}

int orc_compute_collision_force(orc_t *s, int cnstep) {
  const int np = s->np;
  s->npairs = 0;
  for (int n = 0; n < np; n++) {
    double *p = s->prt + (size_t)n * NREC;
    p[P_CFX + cnstep] = 0.;
    p[P_CFY + cnstep] = 0.;
    p[P_CTZ + cnstep] = 0.;
  }
  const int n_total = np * (np - 1) / 2;
  for (int n = 0; n < n_total; n++) {
    int n0, n1;
    pair_indices(np, n, &n0, &n1);
    if (orc_collide_pair(s->ly, cnstep, s->prt + (size_t)n0 * NREC,
                         s->prt + (size_t)n1 * NREC))
      record_pair(s, n0, n1);
  }
  const double walls[2] = {0., s->lx};
  for (int w = 0; w < 2; w++)
    for (int n = 0; n < np; n++)
      if (orc_collide_wall(walls[w], cnstep, s->prt + (size_t)n * NREC))
        record_pair(s, n, -1 - w);
  return 0;
}

int orc_n_colliding(const orc_t *s) { return s->npairs; }
int orc_get_colliding(const orc_t *s, int *pairs) {
  memcpy(pairs, s->pairs, sizeof(int) * 2 * (size_t)s->npairs);
  return s->npairs;
}

/* src/suspensions/increment.c:10-72 */
int orc_increment_particles(orc_t *s, int rk, double *residual) {
  const double Fr = s->Fr, dt = s->dt, gamma = s->gamma[rk];
  *residual = 0.;
  for (int n = 0; n < s->np; n++) {
    double *p = s->prt + (size_t)n * NREC;
    const double den = p[P_DEN];
    const double dux_prev = p[P_DUX], duy_prev = p[P_DUY], dvz_prev = p[P_DVZ];
    p[P_DUX] = +0.5 * gamma * dt * (p[P_CFX] + p[P_CFX + 1]) + dt * p[P_FUX]
               + (p[P_IUX + 1] - p[P_IUX]) + 1.0 * gamma * dt * (1. - den) / pow(Fr, 2.);
    p[P_DUY] = +0.5 * gamma * dt * (p[P_CFY] + p[P_CFY + 1]) + dt * p[P_FUY]
               + (p[P_IUY + 1] - p[P_IUY]);
    p[P_DVZ] = +0.5 * gamma * dt * (p[P_CTZ] + p[P_CTZ + 1]) + dt * p[P_TVZ]
               + (p[P_IVZ + 1] - p[P_IVZ]);
    *residual = fmax(*residual, fabs(p[P_DUX] - dux_prev));
    *residual = fmax(*residual, fabs(p[P_DUY] - duy_prev));
    *residual = fmax(*residual, fabs(p[P_DVZ] - dvz_prev));
  }
  for (int n = 0; n < s->np; n++) {
    double *p = s->prt + (size_t)n * NREC;
    const double ux0 = p[P_UX], uy0 = p[P_UY], vz0 = p[P_VZ];
    const double ux1 = p[P_UX] + p[P_DUX];
    const double uy1 = p[P_UY] + p[P_DUY];
    const double vz1 = p[P_VZ] + p[P_DVZ];
    p[P_DX] = 0.5 * gamma * dt * (ux0 + ux1);
    p[P_DY] = 0.5 * gamma * dt * (uy0 + uy1);
    p[P_DAZ] = 0.5 * gamma * dt * (vz0 + vz1);
  }
  return 0;
}

/* src/suspensions/update.c:8-34,61-65 */
int orc_update_particles(orc_t *s) {
  const double ly = s->ly;
  for (int n = 0; n < s->np; n++) {
    double *p = s->prt + (size_t)n * NREC;
    p[P_UX] += p[P_DUX];
    p[P_UY] += p[P_DUY];
    p[P_VZ] += p[P_DVZ];
  }
  for (int n = 0; n < s->np; n++) {
    double *p = s->prt + (size_t)n * NREC;
    p[P_X] += p[P_DX];
    p[P_Y] += p[P_DY];
    p[P_AZ] += p[P_DAZ];
    if (p[P_Y] < 0.) p[P_Y] += ly;
    if (p[P_Y] > ly) p[P_Y] -= ly;
    if (p[P_AZ] < 0.) p[P_AZ] += 2. * ORC_PI;
    if (p[P_AZ] > 2. * ORC_PI) p[P_AZ] -= 2. * ORC_PI;
  }
  return 0;
}

/* src/main.c:13-55 */
int orc_integrate(orc_t *s) {
  for (int rk = 0; rk < 3; rk++) {
    orc_reset_particle_increments(s);
    orc_update_boundaries_ux(s, s->f[ORC_UX]);
    orc_update_boundaries_uy(s, s->f[ORC_UY]);
    orc_update_boundaries_p(s, s->f[ORC_P]);
    orc_compute_inertia(s, 0);
    orc_update_velocity(s, rk);
    orc_exchange_momentum(s);
    orc_update_momentum_field(s);
    orc_compute_potential(s, rk);
    orc_correct_velocity(s, rk);
    orc_update_pressure(s);
    orc_compute_collision_force(s, 0);
    int substep;
    for (substep = 0;; substep++) {
      orc_compute_inertia(s, 1);
      orc_compute_collision_force(s, 1);
      double residual;
      orc_increment_particles(s, rk, &residual);
      if (residual < 1.e-8) break;
      if (substep > 100) break;
    }
    s->substeps[rk] = substep + 1;
    orc_update_particles(s);
  }
  return 0;
}

/* src/main.c:71-76 */
int orc_step(orc_t *s) {
  orc_decide_dt(s);
  orc_integrate(s);
  s->step += 1;
  s->time += s->dt;
  return 0;
}

int orc_last_substeps(const orc_t *s, int out[3]) {
  for (int k = 0; k < 3; k++) out[k] = s->substeps[k];
  return 0;
}

/* ---------------------------------------------------------------------
 * diagnostics, src/logging.c:40-172 (single process sums)
 * ------------------------------------------------------------------- */
int orc_divergence(const orc_t *s, double *divmax, double *divsum) {
  const int nx = s->itot, ny = s->jtot;
  const double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY];
  double mx = 0., sm = 0.;
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double div = (ux[IX_UX(i + 1, j)] - ux[IX_UX(i, j)]) / DXF(i)
                         + (uy[IX_P(i, j + 1)] - uy[IX_P(i, j)]) / s->dy;
      mx = fmax(mx, fabs(div));
      sm += div;
    }
  *divmax = mx;
  *divsum = sm;
  return 0;
}

int orc_momentum(const orc_t *s, double moms[2]) {
  const int nx = s->itot, ny = s->jtot;
  const double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY];
  moms[0] = 0.;
  moms[1] = 0.;
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx + 1; i++) moms[0] += ux[IX_UX(i, j)] * (DXC(i) * s->dy);
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) moms[1] += uy[IX_P(i, j)] * (DXF(i) * s->dy);
  return 0;
}

/* src/statistics/collect.c:10-95, whole-grid loops as in the reference;
 * pow(v, 1.) is v and pow(v, 2.) is v*v (GCC folds both in the reference
 * build) */
int orc_statistics_collect(const orc_t *s, double *ux1, double *ux2, double *uy1,
                           double *uy2, double *phi) {
  const int nx = s->itot, ny = s->jtot;
  const double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY];
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx + 1; i++) {
      const double v = ux[IX_UX(i, j)];
      ux1[IX_UX(i, j)] += v;
      ux2[IX_UX(i, j)] += v * v;
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 0; i <= nx + 1; i++) {
      const double v = uy[IX_P(i, j)];
      uy1[IX_P(i, j)] += v;
      uy2[IX_P(i, j)] += v * v;
    }
  for (int n = 0; n < s->np; n++) {
    const double *p = s->prt + (size_t)n * NREC;
    const double pa = p[P_A], pb = p[P_B], px = p[P_X], py = p[P_Y], paz = p[P_AZ];
    for (int periodic = -1; periodic <= 1; periodic++) {
      const double py_ = py + s->ly * periodic;
      for (int j = 1; j <= ny; j++) {
        const double y = YC(j);
        for (int i = 1; i <= nx; i++) {
          const double x = XC(i);
          const double dx = x - px;
          const double dy = y - py_;
          const double c = cos(-paz);
          const double sn = sin(-paz);
          const double x_ = c * dx - sn * dy;
          const double y_ = sn * dx + c * dy;
          const double ex = x_ / pa, ey = y_ / pb;
          phi[IX_P(i, j)] += 1. - ex * ex - ey * ey > 0. ? 1. : 0.;
        }
      }
    }
  }
  return 0;
}

int orc_energy(const orc_t *s, double q[2]) {
  const int nx = s->itot, ny = s->jtot;
  const double *ux = s->f[ORC_UX], *uy = s->f[ORC_UY];
  q[0] = 0.;
  q[1] = 0.;
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx + 1; i++)
      q[0] += 0.5 * pow(ux[IX_UX(i, j)], 2.) * (DXC(i) * s->dy);
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++)
      q[1] += 0.5 * pow(uy[IX_P(i, j)], 2.) * (DXF(i) * s->dy);
  return 0;
}
