// eif_suspensions.hip -- ellipses coupled to the fluid on gfx950:
// Eulerian IBM weights, inertia integrals, momentum exchange and its
// box-local deposit, soft-sphere collisions, Crank-Nicolson increments.
//
// Per-element arithmetic follows NaokiHori/EllipsesInFlows:
//   ellipse geometry      src/ellipse.c:7-65
//   IBM weights / boxes   src/suspensions/auxiliary.c:11-75
//   inertia               src/suspensions/inertia.c:9-70
//   momentum exchange     src/suspensions/exchange.c:13-88
//   field update          src/suspensions/update.c:36-59
//   collisions            src/suspensions/collision.c:75-367
//   increments / update   src/suspensions/increment.c:10-72, update.c:8-34
//
// Work decomposition (gfx950):
//   * one workgroup per particle for the IBM integrals; its three periodic
//     images' boxes are flattened and strided over the workgroup, and the
//     per-particle sums are reduced in a fixed tree;
//   * the signed distance of every box cell is cached per particle with the
//     position it belongs to; the reference evaluates it three times at
//     identical positions in every RK stage (inertia(0), exchange and the
//     first inertia(1), src/main.c:21-38), the cache serves the repeats with
//     the same bits;
//   * the exchange writes each box cell's forcing to box-local scratch and
//     registers the box with the 64x4 tiles it touches; a gather over the
//     tiles then adds the contributions cell by cell in the reference's
//     particle/image order and applies them to ux/uy (no atomics, no
//     full-field pass);
//   * collisions: one thread per particle pair finds the colliding pairs,
//     one thread per particle gathers its partners' forces in the
//     reference's pair order.
#include "eif_internal.hpp"
#include <hip/hip_ext.h>
#include <algorithm>
#include <cfloat>
#include <cmath>

namespace eif {

#define TRY_S(x)          \
  do {                    \
    const int _r = (x);   \
    if (_r) return _r;    \
  } while (0)

#define NEXTRA 3
#define BETA 2.

// x^3 to nearly correct rounding (error-free square, then product), the
// value glibc's pow(x, 3.) returns.
__device__ inline double cube(double x) {
  const double p = x * x;
  const double e = fma(x, x, -p);
  const double h = p * x;
  const double eh = fma(p, x, -h);
  return h + (eh + e * x);
}

// src/suspensions/auxiliary.c:11-17
// ih = 1/h, used when h is a power of two (pow2): then x * ih == x / h
__device__ inline void loop_size(int lb, int ub, double h, double ih, bool pow2, double r,
                                 double c, int &mn, int &mx) {
  mn = (int)(floor(pow2 ? (c - r) * ih : (c - r) / h) - NEXTRA);
  mx = (int)(ceil(pow2 ? (c + r) * ih : (c + r) / h) + NEXTRA);
  mn = mn < lb ? lb : mn;
  mx = mx > ub ? ub : mx;
}

struct EllipseShape {
  double a, b;
  double kex, key;  // a(1-(b/a)^2), b(1-(a/b)^2)   (ellipse.c:7-13)
  double a2b2;      // a*a + b*b                     (ellipse.c:49)
  double ia, ib;    // 1/a, 1/b
};

__device__ inline EllipseShape make_shape(double a, double b) {
  EllipseShape s;
  s.a = a;
  s.b = b;
  const double ba = b / a, ab = a / b;
  s.kex = a * (1. - ba * ba);
  s.key = b * (1. - ab * ab);
  s.a2b2 = a * a + b * b;
  s.ia = 1. / a;
  s.ib = 1. / b;
  return s;
}

// src/ellipse.c:30-65
__device__ double find_normal_t(const EllipseShape &e, double x_, double y_, int &iters) {
  const double px = fabs(x_), py = fabs(y_);
  double t = 0.25 * EIF_PI;
  int n;
  for (n = 0;; n++) {
    double s, c;
    sincos(t, &s, &c);
    const double x = e.a * c, y = e.b * s;
    const double ex = e.kex * cube(c), ey = e.key * cube(s);
    const double rx = x - ex, ry = y - ey;
    const double qx = px - ex, qy = py - ey;
    const double r = sqrt(ry * ry + rx * rx);
    const double q = sqrt(qy * qy + qx * qx);
    const double dc = r * asin((rx * qy - ry * qx) / (r * q));
    const double dt = dc / sqrt(e.a2b2 - x * x - y * y);
    t += dt;
    t = fmin(0.5 * EIF_PI, t);
    t = fmax(0., t);
    if (fabs(dt) < 1.e-8) break;
    if (n > 100) break;
  }
  iters += n + 1;
  if (x_ < 0.) t = EIF_PI - t;
  if (y_ < 0.) t = -t;
  return t;
}

// compute_signed_dist, src/suspensions/auxiliary.c:19-47; (cr, sr) are
// cos(-paz), sin(-paz)
__device__ inline double signed_dist(const EllipseShape &e, double grid_size,
                                     double px, double py, double cr, double sr,
                                     double x, double y, int &iters) {
  const double ddx = x - px, ddy = y - py;
  const double x_ = cr * ddx - sr * ddy;
  const double y_ = sr * ddx + cr * ddy;
  const double xa = x_ / e.a, yb = y_ / e.b;
  const double f0 = 1. - xa * xa - yb * yb;
  const double sign = f0 > 0. ? +1. : -1.;
  const double t = find_normal_t(e, x_, y_, iters);
  double st, ct;
  sincos(t, &st, &ct);
  const double ex = e.a * ct - x_, ey = e.b * st - y_;
  const double dist = sign * sqrt(+ex * ex + ey * ey);
  return dist / grid_size;
}

// ---------------------------------------------------------------------
// Foot of the normal for the IBM weights.
//
// src/ellipse.c:30-65 iterates t <- clamp(t + dt) from t = pi/4 until
// |dt| < 1e-8, every iteration in FP64 with sin, cos, asin and three
// square roots.  The iteration converges quadratically to the unique foot
// on the first-quadrant arc, so here it starts in FP32 (hardware sin/cos,
// stopped at |dt| < 1e-3, where the FP32 rounding noise takes over) and
// finishes with the reference's FP64 iteration and stopping rule, which
// then needs one or two iterations instead of about four.  After the first
// FP64 step every update of t is below 2^-10, so cos and sin of the new t
// follow from the old ones by the addition theorem with the Taylor series
// of cos(delta) and sin(delta) (truncation below 1e-21), and only one FP64
// sincos is evaluated per cell.  The converged (cos t, sin t) agree with
// the reference's to a few units in the last place; the distance to the
// foot is stationary in t, so it agrees to rounding.
// ---------------------------------------------------------------------
struct EllipseShape32 {
  float a, b, kex, key, a2b2;
};

__device__ inline EllipseShape32 shape32(const EllipseShape &e) {
  return {(float)e.a, (float)e.b, (float)e.kex, (float)e.key, (float)e.a2b2};
}

// sin and cos on [0, pi/2] (slightly beyond is fine): Taylor polynomials
// on |u| <= pi/4 (truncation below 1e-20) after u = pi/2 - t for t > pi/4,
// with pi/2 split into two doubles so that the reduction is exact to
// rounding.  The general libm routines carry large-argument reduction
// that is never needed here and costs registers in the cell loop.
__device__ inline double sin_poly(double u) {
  const double u2 = u * u;
  double p = 1. / 355687428096000.;
  p = fma(p, u2, -1. / 1307674368000.);
  p = fma(p, u2, 1. / 6227020800.);
  p = fma(p, u2, -1. / 39916800.);
  p = fma(p, u2, 1. / 362880.);
  p = fma(p, u2, -1. / 5040.);
  p = fma(p, u2, 1. / 120.);
  p = fma(p, u2, -1. / 6.);
  return fma(u * u2, p, u);
}
__device__ inline double cos_poly(double u) {
  const double u2 = u * u;
  double p = -1. / 6402373705728000.;
  p = fma(p, u2, 1. / 20922789888000.);
  p = fma(p, u2, -1. / 87178291200.);
  p = fma(p, u2, 1. / 479001600.);
  p = fma(p, u2, -1. / 3628800.);
  p = fma(p, u2, 1. / 40320.);
  p = fma(p, u2, -1. / 720.);
  p = fma(p, u2, 1. / 24.);
  p = fma(p, u2, -0.5);
  return fma(u2, p, 1.);
}
__device__ inline void sincos_q1(double t, double &s, double &c) {
  constexpr double PIO2_HI = 1.5707963267948966;      // nearest double to pi/2
  constexpr double PIO2_LO = 6.123233995736766e-17;   // pi/2 - PIO2_HI
  // one pair of polynomials for both halves (no divergent branch): the
  // argument and the roles of the results are selected instead
  const bool low = t <= 0.78539816339744831;
  const double u = low ? t : (PIO2_HI - t) + PIO2_LO;
  const double ps = sin_poly(u), pc = cos_poly(u);
  s = low ? ps : pc;
  c = low ? pc : ps;
}

// 1/sqrt(x) from the hardware estimate and one Newton step (relative error
// near 1e-16); used where only the size of a step depends on it
__device__ inline double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-(x * y), y, 1.);
  return fma(0.5 * y, e, y);
}

// asin by its Maclaurin series, exact to rounding for |x| <= 2^-6; for
// larger |x| a monotone under-estimate, which only shortens a step of the
// fixed-point iteration (its fixed point is x = 0)
__device__ inline double asin_series(double x) {
  const double x2 = x * x;
  double p = 0.022372159090909091;  // 63/2816
  p = fma(p, x2, 0.030381944444444444);  // 35/1152
  p = fma(p, x2, 0.044642857142857144);  // 5/112
  p = fma(p, x2, 0.075);                 // 3/40
  p = fma(p, x2, 1. / 6.);
  return fma(x * x2, p, x);
}

// cos and sin of t + delta from c = cos t, s = sin t, |delta| <= 2^-10
__device__ inline void rotate_small(double &c, double &s, double delta) {
  const double d2 = delta * delta;
  const double hc = d2 * (0.5 - d2 * (1. / 24.));           // 1 - cos(delta)
  const double sd = delta * (1. - d2 * (1. / 6. - d2 * (1. / 120.)));  // sin(delta)
  const double cn = c - (s * sd + c * hc);
  const double sn = s + (c * sd - s * hc);
  c = cn;
  s = sn;
}

// FP32 part: from t = pi/4 until |dt| < 1e-3 (at most 8 iterations)
__device__ inline float foot_fp32(const EllipseShape32 &f, double px, double py, int &it32) {
  float tf = (float)(0.25 * EIF_PI);
  const float fx = (float)px, fy = (float)py;
  for (int n = 0; n < 8; n++) {
    float sn, cs;
    __sincosf(tf, &sn, &cs);
    const float x = f.a * cs, y = f.b * sn;
    const float ex = f.kex * (cs * cs * cs), ey = f.key * (sn * sn * sn);
    const float rx = x - ex, ry = y - ey;
    const float qx = fx - ex, qy = fy - ey;
    const float rr = rx * rx + ry * ry, qq = qx * qx + qy * qy;
    float arg = (rx * qy - ry * qx) * __builtin_amdgcn_rsqf(rr * qq);
    arg = fminf(fmaxf(arg, -1.f), 1.f);
    const float den = fmaxf(f.a2b2 - x * x - y * y, 1e-30f);
    const float a2 = arg * arg;
    const float as = arg + arg * a2 * (0.16666667f + a2 * (0.075f + a2 * 0.044642857f));
    const float dt = __builtin_amdgcn_sqrtf(rr) * as * __builtin_amdgcn_rsqf(den);
    it32++;
    if (!(fabsf(dt) < 1e30f)) break;  // no usable step: hand over to FP64
    tf = fminf(fmaxf(tf + dt, 0.f), 1.5707963f);
    if (fabsf(dt) < 1e-3f) break;
  }
  return tf;
}

// FP64 part: the reference's iteration and stopping rule from t (updated
// in place); (c, s) = (cos t, sin t) of the foot on the first-quadrant arc
// for the folded point (px, py) >= 0
__device__ inline void foot_fp64(const EllipseShape &e, double px, double py, double &t,
                                 double &c, double &s, int &it64) {
  constexpr double HALF_PI = 0.5 * EIF_PI;
  sincos_q1(t, s, c);
  for (int n = 0;; n++) {
    const double x = e.a * c, y = e.b * s;
    const double ex = e.kex * cube(c), ey = e.key * cube(s);
    const double rx = x - ex, ry = y - ey;
    const double qx = px - ex, qy = py - ey;
    // dt = r asin(cross / (r q)) / sqrt(a^2 + b^2 - x^2 - y^2): the
    // residual (cross) is formed exactly as in ellipse.c:41-46; the scale
    // factors only set the step length, so reciprocal square roots do
    const double rr = ry * ry + rx * rx, qq = qy * qy + qx * qx;
    const double cross = rx * qy - ry * qx;
    const double arg = cross * rsqrt_nr(rr * qq);
    const double dt = asin_series(arg) * (rr * rsqrt_nr(rr * (e.a2b2 - x * x - y * y)));
    double tn = t + dt;
    tn = fmin(HALF_PI, tn);
    tn = fmax(0., tn);
    const double delta = tn - t;
    if (fabs(delta) <= 0x1p-10 && tn != HALF_PI && tn != 0.)
      rotate_small(c, s, delta);
    else
      sincos_q1(tn, s, c);
    t = tn;
    it64++;
    if (fabs(dt) < 1.e-8) break;
    if (n > 100) break;
  }
}

// compute_signed_dist (src/suspensions/auxiliary.c:19-47) with the foot
// from foot_fp32 / foot_fp64; the unfolding of ellipse.c:62-63 maps
// (cos t, sin t) of the folded foot to (-cos t, sin t) for x_ < 0 and
// (cos t, -sin t) for y_ < 0.  The foot iteration starts from the parameter this cell's foot had at the
// particle's previous evaluation in the same RK stage (warm), which the
// particle has barely moved from, or else from the FP32 estimate; the root
// on the first-quadrant arc is unique, so the start changes only the number
// of iterations.  The new parameter is stored back.
__device__ inline double signed_dist_warm(const EllipseShape &e, const EllipseShape32 &f,
                                          double inv_grid, double px, double py, double cr,
                                          double sr, double x, double y, bool warm,
                                          double &tslot, int &it32, int &it64) {
  const double ddx = x - px, ddy = y - py;
  const double x_ = cr * ddx - sr * ddy;
  const double y_ = sr * ddx + cr * ddy;
  const double xa = x_ * e.ia, yb = y_ * e.ib;
  const double f0 = 1. - xa * xa - yb * yb;
  const double sign = f0 > 0. ? +1. : -1.;
  const double ax = fabs(x_), ay = fabs(y_);
  double t = warm ? tslot : (double)foot_fp32(f, ax, ay, it32);
  double ct, st;
  foot_fp64(e, ax, ay, t, ct, st, it64);
  tslot = t;
  const double fx = x_ < 0. ? -(e.a * ct) : e.a * ct;
  const double fy = y_ < 0. ? -(e.b * st) : e.b * st;
  const double ex = fx - x_, ey = fy - y_;
  const double dist = sign * sqrt(+ex * ex + ey * ey);
  return dist * inv_grid;
}

// src/suspensions/auxiliary.c:49-59.  tanh(x) as 1 - 2 / (exp(2|x|) + 1)
// with the sign of x: like the libm routine it rounds to exactly +-1 once
// 2 / (exp(2|x|) + 1) drops below half an ulp of 1, so the weights are
// exactly 0 or 1 on the same cells.
__device__ inline double tanh_w(double x) {
  const double th = 1. - 2. / (exp(2. * fabs(x)) + 1.);
  return copysign(th, x);
}
// Both weights are functions of th = tanh(BETA d); the per-cell cache keeps
// th, so a cached cell needs no exp.
__device__ inline double weight_th(double d) { return tanh_w(BETA * d); }
__device__ inline double v_weight_th(double th) { return 0.5 * (1. + th); }
__device__ inline double s_weight_th(double th) { return 0.5 * BETA * (1. - th * th); }

// mass and moment of inertia, src/suspensions/auxiliary.c:63-75
__device__ inline double p_mass(const double *q) {
  const double vol = EIF_PI * q[R_A] * q[R_B];
  return q[R_DEN] * vol;
}
__device__ inline double p_moi(const double *q) {
  const double m = p_mass(q);
  return 0.25 * m * (q[R_A] * q[R_A] + q[R_B] * q[R_B]);
}

// ---------------------------------------------------------------------
// IBM integrals
// ---------------------------------------------------------------------
constexpr int IBM_THREADS = 256;
constexpr int IBM_SLICE = 256;  // box cells per wave in the inertia / exchange kernels
constexpr int SLICE_VALS = 5;   // per-slice partials: 3 sums, FP64 and FP32 iterations

// Sub-iteration `sub` of a stage's speculatively queued Crank-Nicolson
// loop (src/main.c:37-50) runs unless the previous one converged (residual
// below 1e-8) or the stage has already run 101 of them.  res[s] is the
// residual of sub-iteration s, cleared at the stage start, so a sub-iteration
// that did not run leaves 0 and closes every later one.  res == nullptr:
// not gated.
__device__ inline bool sub_open(const double *res, int sub) {
  return !res || sub == 0 || (!(res[sub - 1] < 1.e-8) && sub <= 101);
}

struct SliceHead;

struct IbmArgs {
  Layout L;
  int nx, ny;
  double ly, dx, dy, grid_size;
  double idx, idy;  // 1/dx, 1/dy
  int pow2;         // dx, dy powers of two: divisions by them are products with idx, idy
  const double *xc, *yc;
  const double *ux, *uy;
  double *prt;
  const double *dscal;
  Box *boxes;
  const long *box_off;
  double *sc_fx, *sc_fy;
  double *sc_dist;  // per box cell: tanh(BETA d) of the signed distance d
  double *sc_pos;
  double *sc_t;   // per box cell: foot parameter of the last evaluation
  int *sc_key;    // per particle: boxes of that evaluation (8 ints)
  int *tile_cnt, *tile_list;
  int ntx, nty;
  int *flags;
  unsigned long long *stats;  // per particle: cells, cells evaluated, FP64 and FP32 iterations
  double *cpos;               // np x 4: the broad phase's packed positions and radii
  int np;
  const int *slice_first;  // np+1: first slice of each particle
  const int *slice_owner;  // particle of each slice
  long *slice_info;        // np x 2: box cells, cells evaluated (this call)
  const double *res;       // sub-iteration residuals (null: not gated)
  int sub;                 // sub-iteration index
  // k_exchange only: with us_x, also inertia(0) of the stage-start
  // velocity (us_x, us_y), its slice sums into slice_part0
  const double *us_x, *us_y;
  double *slice_part0;
  double *slice_part;
  const int *halt;   // graph-step stop flag
  SliceHead *heads;  // per particle: the pass's position, shape, boxes, cache state      // per-slice partial sums (3 each)
};

// x of lane i + n of the same row of 16 lanes (DPP row_shl:n, CTRL = 0x100 + n);
// lanes whose source leaves the row get 0
template <int CTRL>
__device__ inline double row_shl_f64(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// fixed-order tree over the 64 lanes of a wave (lane 0 receives the sum):
// v[i] += v[i + off] for off = 32, 16, 8, 4, 2, 1; the last four levels stay
// inside lanes 0-15 and move through DPP instead of LDS permutes
__device__ inline double wave_sum(double v) {
  v += __shfl_down(v, 32, 64);
  v += __shfl_down(v, 16, 64);
  v += row_shl_f64<0x108>(v);
  v += row_shl_f64<0x104>(v);
  v += row_shl_f64<0x102>(v);
  v += row_shl_f64<0x101>(v);
  return v;
}

// lane l's x, l the same on every lane of the wave
__device__ inline double lane_value(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// A slice's partial sums (lane 0); k_ibm_finalize adds them per particle in
// slice order after the kernel has finished, so no fence or atomic is needed.
__device__ inline void slice_store(const IbmArgs &a, int slice, const double v[SLICE_VALS]) {
  if (threadIdx.x == 0)
    for (int k = 0; k < SLICE_VALS; k++) a.slice_part[SLICE_VALS * (long)slice + k] = v[k];
}

// box of one periodic image (auxiliary.c:11-17); three separate members
// rather than an array, so that selecting an image stays in registers
struct Img {
  int imin, imax, jmin, jmax;
  int cnt;
  float inv_w;  // 1 / box width, for the quotient in cell_of
};
struct BoxSet {
  Img b0, b1, b2;
  int total;
};

__device__ inline Img make_img(const IbmArgs &a, double px, double py_, double r) {
  Img m;
  loop_size(1, a.nx, a.dx, a.idx, a.pow2, r, px, m.imin, m.imax);
  loop_size(1, a.ny, a.dy, a.idy, a.pow2, r, py_ - 0., m.jmin, m.jmax);
  const int w = m.imax - m.imin + 1, h = m.jmax - m.jmin + 1;
  m.cnt = (w > 0 && h > 0) ? w * h : 0;
  m.inv_w = w > 0 ? 1.f / (float)w : 0.f;
  return m;
}

__device__ inline void make_boxes(const IbmArgs &a, double px, double py,
                                  double r, BoxSet &bs) {
  bs.b0 = make_img(a, px, py + a.ly * (0 - 1), r);
  bs.b1 = make_img(a, px, py + a.ly * (1 - 1), r);
  bs.b2 = make_img(a, px, py + a.ly * (2 - 1), r);
  bs.total = bs.b0.cnt + bs.b1.cnt + bs.b2.cnt;
}

// boxes of the last distance evaluation (cells map to the same flattened
// index only when every image's box is unchanged)
__device__ inline bool same_boxes(const int *key, const BoxSet &bs) {
  return key[7] != 0 && key[0] == bs.b0.imin && key[1] == bs.b0.jmin &&
         key[2] == bs.b1.imin && key[3] == bs.b1.jmin && key[4] == bs.b2.imin &&
         key[5] == bs.b2.jmin && key[6] == bs.total;
}
__device__ inline void store_boxes(int *key, const BoxSet &bs) {
  key[0] = bs.b0.imin;
  key[1] = bs.b0.jmin;
  key[2] = bs.b1.imin;
  key[3] = bs.b1.jmin;
  key[4] = bs.b2.imin;
  key[5] = bs.b2.jmin;
  key[6] = bs.total;
  key[7] = 1;
}

__device__ inline Img pick(const BoxSet &bs, int k) {
  Img m = bs.b0;
  if (k == 1) m = bs.b1;
  if (k == 2) m = bs.b2;
  return m;
}

// image k and cell (i, j) of flattened box cell t.  The image's fields are
// blended arithmetically (a select between struct members can be turned
// into a load from a stack copy), and the quotient by the box width comes
// from a float reciprocal, corrected once (exact for t < 2^24).
__device__ inline void cell_of(const BoxSet &bs, int t, int &k, int &i, int &j) {
  const int c0 = bs.b0.cnt, c01 = c0 + bs.b1.cnt;
  const int g1 = t >= c0, g2 = t >= c01;
  k = g1 + g2;
  const int tk = t - g1 * c0 - g2 * bs.b1.cnt;
  const int imin = bs.b0.imin + g1 * (bs.b1.imin - bs.b0.imin) + g2 * (bs.b2.imin - bs.b1.imin);
  const int imax = bs.b0.imax + g1 * (bs.b1.imax - bs.b0.imax) + g2 * (bs.b2.imax - bs.b1.imax);
  const int jmin = bs.b0.jmin + g1 * (bs.b1.jmin - bs.b0.jmin) + g2 * (bs.b2.jmin - bs.b1.jmin);
  const int w = imax - imin + 1;
  int q = (int)((float)tk * __builtin_amdgcn_rcpf((float)w));
  int r = tk - q * w;
  if (r < 0) {
    q--;
    r += w;
  } else if (r >= w) {
    q++;
    r -= w;
  }
  j = jmin + q;
  i = imin + r;
}

// What every slice wave of one IBM pass needs of its particle: computed
// once per particle and pass, by the kernel that last moved the particle
// or updated its cache (k_reset, the finalise passes, k_increment, or
// k_heads), instead of by each of the particle's waves.  The arithmetic is
// that the waves would do themselves.
struct SliceHead {
  EllipseShape e;
  double px, py, paz;  // position of the pass
  double cr, sr;       // cos, sin of -paz
  double inv_pm, inv_pim;
  long off, cap;       // the particle's per-cell scratch range
  BoxSet bs;
  EllipseShape32 f;
  int cached;          // the distance cache holds this position
  int warm;            // the foot parameters hold these boxes
};

// same_boxes against the key store_boxes(k) would leave
__device__ inline bool same_boxset(const BoxSet &k, const BoxSet &bs) {
  return k.b0.imin == bs.b0.imin && k.b0.jmin == bs.b0.jmin && k.b1.imin == bs.b1.imin &&
         k.b1.jmin == bs.b1.jmin && k.b2.imin == bs.b2.imin && k.b2.jmin == bs.b2.jmin &&
         k.total == bs.total;
}

// full: also the parts that depend only on the particle's shape and mass
// (written at every stage start, so records changed between calls are
// picked up); otherwise only the position-dependent parts
// prev: the header whose pass the cache records now describe (known to the
// caller), else the records are read
// rad: max(a, b) of the particle, when the caller has it (< 0: read here)
__device__ inline void write_head(const IbmArgs &a, int n, double px, double py, double paz,
                                  bool full, const SliceHead *prev = nullptr,
                                  double rad = -1.) {
  const double *q = a.prt + (long)n * R_N;
  SliceHead *h = a.heads + n;
  if (full) {
    h->e = make_shape(q[R_A], q[R_B]);
    h->f = shape32(h->e);
    h->inv_pm = 1. / p_mass(q);
    h->inv_pim = 1. / p_moi(q);
    h->off = a.box_off[n];
    h->cap = a.box_off[n + 1] - h->off;
  }
  h->px = px;
  h->py = py;
  h->paz = paz;
  double sr, cr;
  sincos(-paz, &sr, &cr);
  h->cr = cr;
  h->sr = sr;
  BoxSet bs;
  make_boxes(a, px, py, rad >= 0. ? rad : fmax(q[R_A], q[R_B]), bs);
  h->bs = bs;
  if (prev) {  // the cache records were just made from prev
    h->cached = prev->px == px && prev->py == py && prev->paz == paz;
    h->warm = same_boxset(prev->bs, bs);
  } else {
    const double *cp = a.sc_pos + 4 * (long)n;
    h->cached = cp[3] != 0. && cp[0] == px && cp[1] == py && cp[2] == paz;
    h->warm = same_boxes(a.sc_key + 8 * (long)n, bs);
  }
}

// heads at the particles' positions, with the increments when with_dx
struct HeadsB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, IbmArgs a, int with_dx) {
  if (halted(a.halt)) return;
  const int n = vb.bx * blockDim.x + threadIdx.x;
  if (n >= a.np) return;
  const double *q = a.prt + (long)n * R_N;
  if (with_dx)
    write_head(a, n, q[R_X] + q[R_DX], q[R_Y] + q[R_DY], q[R_AZ] + q[R_DAZ], true);
  else
    write_head(a, n, q[R_X], q[R_Y], q[R_AZ], true);
}
};
__global__ void k_heads(IbmArgs a, int with_dx) {
  HeadsB::run(VB::here(), a, with_dx);
}

// src/suspensions/inertia.c:9-70.  One wave per slice of IBM_SLICE box
// cells of one particle (its three images' boxes flattened); the last
// slice of a particle to finish adds the slices' sums in slice order.
// open: the sub-iteration gate, evaluated by the caller; tested after this
// slice's own loads are issued, so the gate's load is not waited for alone
__device__ __forceinline__ void inertia_slice(const IbmArgs &a, int cnstep, int slice,
                                              bool open = true) {
  const int n = a.slice_owner[slice];
  const int s0 = a.slice_first[n], nsl = a.slice_first[n + 1] - s0;
  const SliceHead &hd = a.heads[n];
  const BoxSet bs = hd.bs;
  const int total = bs.total;
  if (!open) return;
  if (total > hd.cap || total > nsl * IBM_SLICE) {
    if (threadIdx.x == 0) atomicOr(&a.flags[0], 1);
    return;
  }
  const EllipseShape e = hd.e;
  const EllipseShape32 f = hd.f;
  const double px = hd.px, py = hd.py;
  const double inv_pm = hd.inv_pm, inv_pim = hd.inv_pim;
  const bool cached = hd.cached, warm = hd.warm;
  double *dist = a.sc_dist + hd.off;
  double *tcell = a.sc_t + hd.off;
  const double cr = hd.cr, sr = hd.sr;
  const double dxdy = a.dx * a.dy;
  double acc[3] = {0., 0., 0.};  // iux, iuy, ivz
  int it64 = 0, it32 = 0;
  const int t0 = (slice - s0) * IBM_SLICE;
  const int t1 = t0 + IBM_SLICE < total ? t0 + IBM_SLICE : total;
  if (cached) {
    // all loads of the lane's cells first (a cell past t1 loads the last
    // cell instead and is not added), then the sums in cell order
    if (t0 < t1) {
      constexpr int U = IBM_SLICE / 64;
      double th_[U], x_[U], y_[U], vx0[U], vx1[U], vy0[U], vy1[U];
      int k_[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int t = min(t0 + (int)threadIdx.x + 64 * u, t1 - 1);
        int k, i, j;
        cell_of(bs, t, k, i, j);
        k_[u] = k;
        x_[u] = a.xc[i];
        y_[u] = a.yc[j - 1];
        th_[u] = dist[t];
        vx0[u] = a.ux[a.L.ux(i, j)];
        vx1[u] = a.ux[a.L.ux(i + 1, j)];
        vy0[u] = a.uy[a.L.p(i, j)];
        vy1[u] = a.uy[a.L.p(i, j + 1)];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (t0 + (int)threadIdx.x + 64 * u >= t1) continue;
        const double py_ = py + a.ly * (k_[u] - 1);
        const double w = v_weight_th(th_[u]);
        const double valx = w * 0.5 * (vx0[u] + vx1[u]) * dxdy;
        const double valy = w * 0.5 * (vy0[u] + vy1[u]) * dxdy;
        acc[0] += valx * inv_pm;
        acc[2] += -(y_[u] - py_) * valx * inv_pim;
        acc[1] += valy * inv_pm;
        acc[2] += +(x_[u] - px) * valy * inv_pim;
      }
    }
  } else
  for (int t = t0 + (int)threadIdx.x; t < t1; t += 64) {
    int k, i, j;
    cell_of(bs, t, k, i, j);
    const double py_ = py + a.ly * (k - 1);
    const double x = a.xc[i], y = a.yc[j - 1];
    // the velocities do not depend on the distance: loaded first
    const double vx0 = a.ux[a.L.ux(i, j)], vx1 = a.ux[a.L.ux(i + 1, j)];
    const double vy0 = a.uy[a.L.p(i, j)], vy1 = a.uy[a.L.p(i, j + 1)];
    const double th = weight_th(signed_dist_warm(e, f, 1. / a.grid_size, px, py_, cr, sr, x,
                                                 y, warm, tcell[t], it32, it64));
    dist[t] = th;
    const double w = v_weight_th(th);
    const double valx = w * 0.5 * (vx0 + vx1) * dxdy;
    const double valy = w * 0.5 * (vy0 + vy1) * dxdy;
    acc[0] += valx * inv_pm;
    acc[2] += -(y - py_) * valx * inv_pim;
    acc[1] += valy * inv_pm;
    acc[2] += +(x - px) * valy * inv_pim;
  }
  double part[SLICE_VALS] = {wave_sum(acc[0]), wave_sum(acc[1]), wave_sum(acc[2]),
                             wave_sum((double)it64), wave_sum((double)it32)};
  slice_store(a, slice, part);
  if (slice == s0 && threadIdx.x == 0) {
    a.slice_info[2 * (long)n] = total;
    a.slice_info[2 * (long)n + 1] = cached ? 0 : total;
  }
}

struct InertiaB {
  static constexpr int T = 64, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, IbmArgs a, int cnstep) {
  if (halted(a.halt)) return;
  inertia_slice(a, cnstep, vb.bx, sub_open(a.res, a.sub));
}
};
__global__ __launch_bounds__(64) void k_inertia(IbmArgs a, int cnstep) {
  InertiaB::run(VB::here(), a, cnstep);
}

// Per-particle totals of the sliced inertia / exchange kernels, by the 64
// lanes of one wave: the lanes load the slices' partial sums together, and
// they are added in slice order through lane shuffles.  Lane 0 also
// records the position whose distances are now cached and the work
// counters.  which: 0..1 inertia (cnstep), 2 exchange.
// the particle's slice sums in slice order (every lane gets them)
__device__ inline void slice_sums(const IbmArgs &a, int s0, int s1, double tot[SLICE_VALS]) {
  const int lane = threadIdx.x & 63;
  for (int k = 0; k < SLICE_VALS; k++) tot[k] = 0.;
  for (int base = s0; base < s1; base += 64) {
    double v[SLICE_VALS] = {0., 0., 0., 0., 0.};
    if (base + lane < s1)
      for (int k = 0; k < SLICE_VALS; k++)
        v[k] = a.slice_part[SLICE_VALS * (long)(base + lane) + k];
    const int cnt = min(64, s1 - base);
    // the slice index is the same on every lane: lane l's values are read
    // as scalars (no LDS round trip), added in the same order
    for (int l = 0; l < cnt; l++)
      for (int k = 0; k < SLICE_VALS; k++) tot[k] += lane_value(v[k], l);
  }
}

// the cache records of the pass with header hd, and the work counters
__device__ inline void record_pass(const IbmArgs &a, int n, const SliceHead &hd,
                                   const double tot[SLICE_VALS]) {
  double *cw = a.sc_pos + 4 * (long)n;
  cw[0] = hd.px;
  cw[1] = hd.py;
  cw[2] = hd.paz;
  cw[3] = 1.;
  store_boxes(a.sc_key + 8 * (long)n, hd.bs);
  if (n == 0) a.stats[4 * (long)a.np] += 1ull;  // IBM calls
  unsigned long long *c = a.stats + 4 * (long)n;
  c[0] += (unsigned long long)a.slice_info[2 * (long)n];
  c[1] += (unsigned long long)a.slice_info[2 * (long)n + 1];
  c[2] += (unsigned long long)tot[3];
  c[3] += (unsigned long long)tot[4];
}

__device__ void finalize_particle(const IbmArgs &a, int n, int which, bool record = true) {
  double tot[SLICE_VALS];
  slice_sums(a, a.slice_first[n], a.slice_first[n + 1], tot);
  if ((threadIdx.x & 63) != 0) return;
  double *q = a.prt + (long)n * R_N;
  if (which < 2) {
    q[R_IUX + which] = tot[0];
    q[R_IUY + which] = tot[1];
    q[R_IVZ + which] = tot[2];
  } else {
    q[R_FUX] = tot[0];
    q[R_FUY] = tot[1];
    q[R_TVZ] = tot[2];
  }
  // the cache records and counters (unless another call's); the pass's
  // header holds its position and boxes
  if (record) record_pass(a, n, a.heads[n], tot);
}

// The exchange's sums of particle n and those of the inertia(0) computed
// beside it, by one wave; the exchange's call records the distance cache
// and the counters.
__device__ inline void finalize_x0_particle(const IbmArgs &a, int n) {
  finalize_particle(a, n, 2);
  IbmArgs a0 = a;
  a0.slice_part = a.slice_part0;
  finalize_particle(a0, n, 0, false);
  // the first sub-iteration's pass, at the cached stage-start positions
  if ((threadIdx.x & 63) == 0) {
    const double *q = a.prt + (long)n * R_N;
    write_head(a, n, q[R_X] + q[R_DX], q[R_Y] + q[R_DY], q[R_AZ] + q[R_DAZ], false);
  }
}

// one wave per particle, four per workgroup
struct IbmFinalizeB {
  static constexpr int T = 256, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, IbmArgs a, int which) {
  if (halted(a.halt)) return;
  if (!sub_open(a.res, a.sub)) return;
  const int n = vb.bx * 4 + (int)(threadIdx.x >> 6);
  if (n < a.np) finalize_particle(a, n, which);
}
};
__global__ __launch_bounds__(256) void k_ibm_finalize(IbmArgs a, int which) {
  IbmFinalizeB::run(VB::here(), a, which);
}

// Measurement only: the reference's FP64 iteration (find_normal_t,
// src/ellipse.c:30-65) at every cell of the particles' current inertia
// boxes, counting cells and iterations into out[0], out[1].
__global__ __launch_bounds__(IBM_THREADS) void k_ref_iterations(IbmArgs a,
                                                                unsigned long long *out) {
  __shared__ BoxSet bs;
  __shared__ unsigned long long sum;
  const int n = blockIdx.x;
  const double *q = a.prt + (long)n * R_N;
  const double pa = q[R_A], pb = q[R_B];
  const double px = q[R_X] + q[R_DX];
  const double py = q[R_Y] + q[R_DY];
  const double paz = q[R_AZ] + q[R_DAZ];
  const EllipseShape e = make_shape(pa, pb);
  if (threadIdx.x == 0) {
    make_boxes(a, px, py, fmax(pa, pb), bs);
    sum = 0;
  }
  __syncthreads();
  double cr, sr;
  sincos(-paz, &sr, &cr);
  int iters = 0;
  for (long t = threadIdx.x; t < bs.total; t += IBM_THREADS) {
    int k, i, j;
    cell_of(bs, t, k, i, j);
    const double py_ = py + a.ly * (k - 1);
    (void)signed_dist(e, a.grid_size, px, py_, cr, sr, a.xc[i], a.yc[j - 1], iters);
  }
  if (iters) atomicAdd(&sum, (unsigned long long)iters);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], (unsigned long long)bs.total);
    atomicAdd(&out[1], sum);
  }
}

// Tiles reached by a box's forcing in the field update (update.c:42-57:
// UX(i) reads DUX(i-1), UY(j) reads DUY(j-1), and DUY(i,0) is the periodic
// copy of DUY(i,jtot)).  Up to two rectangles of tiles per box.
struct TileRange {
  int tx0, tx1, ty0, ty1;
  __device__ int count() const { return tx1 < tx0 || ty1 < ty0 ? 0 : (tx1 - tx0 + 1) * (ty1 - ty0 + 1); }
};

__device__ inline void box_tile_ranges(const IbmArgs &a, int imin, int imax, int jmin,
                                       int jmax, TileRange &r0, TileRange &r1) {
  const int i_hi = imax + 1 > a.nx ? a.nx : imax + 1;
  const int j_hi = jmax + 1 > a.ny ? a.ny : jmax + 1;
  r0 = {(imin - 1) / TILE_X, (i_hi - 1) / TILE_X, (jmin - 1) / TILE_Y, (j_hi - 1) / TILE_Y};
  if (jmax == a.ny && jmin > 1)
    r1 = {(imin - 1) / TILE_X, (imax - 1) / TILE_X, 0, 0};
  else
    r1 = {0, -1, 0, -1};
}

__device__ inline void register_tile(const IbmArgs &a, int tile, int box_id) {
  const int slot = atomicAdd(&a.tile_cnt[tile], 1);
  if (slot < TILE_CAP)
    a.tile_list[(long)tile * TILE_CAP + slot] = box_id;
  else
    atomicOr(&a.flags[0], 2);
}

// Signed distances of the box cells at the particles' positions (with the
// increments dx, dy, daz when with_dx) where the per-particle cache does not
// already hold them; sliced like k_inertia.  Used before k_exchange when
// the cache may be stale.
struct IbmDistB {
  static constexpr int T = 64, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, IbmArgs a) {
  if (halted(a.halt)) return;
  if (!sub_open(a.res, a.sub)) return;
  const int slice = vb.bx;
  const int n = a.slice_owner[slice];
  const int s0 = a.slice_first[n], nsl = a.slice_first[n + 1] - s0;
  const SliceHead &hd = a.heads[n];
  const BoxSet bs = hd.bs;
  const int total = bs.total;
  if (total > hd.cap || total > nsl * IBM_SLICE) {
    if (threadIdx.x == 0) atomicOr(&a.flags[0], 1);
    return;
  }
  const bool cached = hd.cached;
  const double px = hd.px, py = hd.py;
  int it64 = 0, it32 = 0;
  if (!cached) {
    const EllipseShape e = hd.e;
    const EllipseShape32 f = hd.f;
    const bool warm = hd.warm;
    double *dist = a.sc_dist + hd.off;
    double *tcell = a.sc_t + hd.off;
    const double cr = hd.cr, sr = hd.sr;
    const int t0 = (slice - s0) * IBM_SLICE;
    const int t1 = t0 + IBM_SLICE < total ? t0 + IBM_SLICE : total;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += 64) {
      int k, i, j;
      cell_of(bs, t, k, i, j);
      const double py_ = py + a.ly * (k - 1);
      dist[t] = weight_th(signed_dist_warm(e, f, 1. / a.grid_size, px, py_, cr, sr, a.xc[i],
                                           a.yc[j - 1], warm, tcell[t], it32, it64));
    }
  }
  const double c64 = wave_sum((double)it64), c32 = wave_sum((double)it32);
  if (threadIdx.x == 0) {
    a.slice_part[SLICE_VALS * (long)slice + 3] = c64;
    a.slice_part[SLICE_VALS * (long)slice + 4] = c32;
    if (slice == s0) a.slice_info[2 * (long)n + 1] = cached ? 0 : total;
  }
}
};
__global__ __launch_bounds__(64) void k_ibm_dist(IbmArgs a) {
  IbmDistB::run(VB::here(), a);
}

// src/suspensions/exchange.c:13-88, sliced like k_inertia, on distances
// already in the per-particle cache (the position is the one of the
// preceding inertia(0) or k_ibm_dist).  The first slice of a particle also
// publishes its boxes and registers them with the tiles they reach.
// own_counts: no k_ibm_dist ran before, so this kernel reports no
// distance work.
// The heads hold the particles' positions without increments.
struct ExchangeB {
  static constexpr int T = 64, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, IbmArgs a, int own_counts) {
  if (halted(a.halt)) return;
  const int slice = vb.bx;
  const int n = a.slice_owner[slice];
  const int s0 = a.slice_first[n], nsl = a.slice_first[n + 1] - s0;
  const double *q = a.prt + (long)n * R_N;
  const SliceHead &hd = a.heads[n];
  const double dt = a.dscal[S_DT];
  const double inv_dt = 1. / dt;
  const double inv_pm = hd.inv_pm, inv_pim = hd.inv_pim;
  const double px = hd.px, py = hd.py;
  const double pux = q[R_UX], puy = q[R_UY], pvz = q[R_VZ];
  const BoxSet bs = hd.bs;
  const int total = bs.total;
  if (total > hd.cap || total > nsl * IBM_SLICE) {
    if (threadIdx.x == 0) atomicOr(&a.flags[0], 1);
    return;
  }
  if (slice == s0) {
    if (threadIdx.x == 0) {
      long off = hd.off;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const Img m = pick(bs, k);
        Box b;
        b.imin = m.imin;
        b.imax = m.imax;
        b.jmin = m.jmin;
        b.jmax = m.jmax;
        b.off = off;
        a.boxes[(long)n * 3 + k] = b;
        off += m.cnt;
      }
    }
    // register the boxes with their tiles, one tile per lane: the first
    // tile of each of the six ranges is claimed by atomics issued together,
    // then the list entries stored (the registrations of one range do not
    // wait for those of the range before); a range of more than 64 tiles
    // registers the rest one by one
    int tile_[6], slot_[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
      tile_[r] = -1;
      const Img m = pick(bs, r / 2);
      if (m.cnt <= 0) continue;
      TileRange tr[2];
      box_tile_ranges(a, m.imin, m.imax, m.jmin, m.jmax, tr[0], tr[1]);
      const TileRange &rg = tr[r & 1];
      const int cnt = rg.count();
      const int w = rg.tx1 - rg.tx0 + 1;
      if ((int)threadIdx.x < cnt) tile_[r] = (rg.ty0 + (int)threadIdx.x / w) * a.ntx + rg.tx0 + (int)threadIdx.x % w;
      for (int e_ = threadIdx.x + 64; e_ < cnt; e_ += 64)
        register_tile(a, (rg.ty0 + e_ / w) * a.ntx + rg.tx0 + e_ % w, n * 3 + r / 2);
    }
#pragma unroll
    for (int r = 0; r < 6; r++) slot_[r] = tile_[r] >= 0 ? atomicAdd(&a.tile_cnt[tile_[r]], 1) : -1;
#pragma unroll
    for (int r = 0; r < 6; r++) {
      if (tile_[r] < 0) continue;
      if (slot_[r] < TILE_CAP)
        a.tile_list[(long)tile_[r] * TILE_CAP + slot_[r]] = n * 3 + r / 2;
      else
        atomicOr(&a.flags[0], 2);
    }
  }
  const double *dist = a.sc_dist + hd.off;
  double *fxo = a.sc_fx + hd.off;
  double *fyo = a.sc_fy + hd.off;
  const double dx = a.dx, dy = a.dy;
  const double dxdy = a.dx * a.dy;
  double acc[3] = {0., 0., 0.};  // fux, fuy, tvz
  double ai[3] = {0., 0., 0.};   // inertia(0): iux, iuy, ivz
  const int t0 = (slice - s0) * IBM_SLICE;
  const int t1 = t0 + IBM_SLICE < total ? t0 + IBM_SLICE : total;
  // all loads of the lane's cells first (a cell past t1 loads the last cell
  // instead and is skipped), then the forcing and the sums in cell order
  if (t0 < t1) {
    constexpr int U = IBM_SLICE / 64;
    double th_[U], x_[U], y_[U], fx0[U], fx1[U], fy0[U], fy1[U], sx0[U], sx1[U], sy0[U],
        sy1[U];
    int k_[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = min(t0 + (int)threadIdx.x + 64 * u, t1 - 1);
      int k, i, j;
      cell_of(bs, t, k, i, j);
      k_[u] = k;
      x_[u] = a.xc[i];
      y_[u] = a.yc[j - 1];
      th_[u] = dist[t];
      fx0[u] = a.ux[a.L.ux(i, j)];
      fx1[u] = a.ux[a.L.ux(i + 1, j)];
      fy0[u] = a.uy[a.L.p(i, j)];
      fy1[u] = a.uy[a.L.p(i, j + 1)];
      if (a.us_x) {
        sx0[u] = a.us_x[a.L.ux(i, j)];
        sx1[u] = a.us_x[a.L.ux(i + 1, j)];
        sy0[u] = a.us_y[a.L.p(i, j)];
        sy1[u] = a.us_y[a.L.p(i, j + 1)];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = t0 + (int)threadIdx.x + 64 * u;
      if (t >= t1) continue;
      const double py_ = py + a.ly * (k_[u] - 1);
      const double x = x_[u], y = y_[u];
      const double th = th_[u];
      if (a.us_x) {  // inertia(0), src/suspensions/inertia.c:55-60, at dx = 0
        const double wv = v_weight_th(th);
        const double valx = wv * 0.5 * (sx0[u] + sx1[u]) * dxdy;
        const double valy = wv * 0.5 * (sy0[u] + sy1[u]) * dxdy;
        ai[0] += valx * inv_pm;
        ai[2] += -(y - py_) * valx * inv_pim;
        ai[1] += valy * inv_pm;
        ai[2] += +(x - px) * valy * inv_pim;
      }
      const double w = s_weight_th(th);
      const double ux_p = pux - pvz * (y - py_);
      const double uy_p = puy + pvz * (x - px);
      const double ux_f = 0.5 * (fx0[u] + fx1[u]);
      const double uy_f = 0.5 * (fy0[u] + fy1[u]);
      const double fx = w * (ux_p - ux_f) * inv_dt;
      const double fy = w * (uy_p - uy_f) * inv_dt;
      fxo[t] = fx * dt;
      fyo[t] = fy * dt;
      acc[0] -= +(fx * dx * dy) * inv_pm;
      acc[2] -= -(y - py_) * (fx * dx * dy) * inv_pim;
      acc[1] -= +(fy * dx * dy) * inv_pm;
      acc[2] -= +(x - px) * (fy * dx * dy) * inv_pim;
    }
  }
  const double s0v = wave_sum(acc[0]), s1v = wave_sum(acc[1]), s2v = wave_sum(acc[2]);
  if (a.us_x) {
    const double i0 = wave_sum(ai[0]), i1 = wave_sum(ai[1]), i2 = wave_sum(ai[2]);
    if (threadIdx.x == 0) {
      double *pp = a.slice_part0 + SLICE_VALS * (long)slice;
      pp[0] = i0;
      pp[1] = i1;
      pp[2] = i2;
      pp[3] = 0.;
      pp[4] = 0.;
    }
  }
  if (threadIdx.x == 0) {
    double *pp = a.slice_part + SLICE_VALS * (long)slice;
    pp[0] = s0v;
    pp[1] = s1v;
    pp[2] = s2v;
    if (own_counts) {
      pp[3] = 0.;
      pp[4] = 0.;
    }
    if (slice == s0) {
      a.slice_info[2 * (long)n] = total;
      if (own_counts) a.slice_info[2 * (long)n + 1] = 0;
    }
  }
}
};
__global__ __launch_bounds__(64) void k_exchange(IbmArgs a, int own_counts) {
  ExchangeB::run(VB::here(), a, own_counts);
}

// ---------------------------------------------------------------------
// deposit + apply over active tiles (update.c:36-59)
// ---------------------------------------------------------------------
struct TileArgs {
  Layout L;
  int nx, ny, ntx;
  const Box *boxes;
  const double *sc_fx, *sc_fy;
  const int *tile_cnt, *tile_list;
  double *ux, *uy;
  double *dux, *duy;  // materialisation target (or null)
  int apply;
  const int *halt;
};

__device__ inline double box_value(const Box &b, const double *sc, int i, int j) {
  if (i < b.imin || i > b.imax || j < b.jmin || j > b.jmax) return 0.;
  const long w = b.imax - b.imin + 1;
  return sc[b.off + (long)(j - b.jmin) * w + (i - b.imin)];
}

// One 64-lane wave per active tile (lane = tile column, all TILE_Y rows),
// so that many tiles are in flight per CU.  The tile's box registrations
// (in registration order) are ranked in parallel into the reference's order
// (particle index, then image; a box registered twice counts once), the
// boxes are staged in LDS, and every cell sums its boxes' values in that
// order.
constexpr int TILE_THREADS = 64;
static_assert(TILE_X == TILE_THREADS, "one lane per tile column");
static_assert(TILE_CAP <= TILE_THREADS, "one lane per registration");

__device__ __forceinline__ void tile_body(const TileArgs &a, int tile) {
  __shared__ int raw[TILE_CAP];
  __shared__ bool rep[TILE_CAP];
  __shared__ Box sbox[TILE_CAP];
  __shared__ int nlist;
  // tiles in grid order, so neighbouring workgroups touch neighbouring
  // memory; tiles no box reaches return at once
  const int cnt = min(a.tile_cnt[tile], TILE_CAP);
  if (cnt == 0) return;
  const int lane = threadIdx.x;
  const Layout L = a.L;
  const int tx = tile % a.ntx, ty = tile / a.ntx;
  const int i0 = tx * TILE_X + 1, j0 = ty * TILE_Y + 1;
  if (lane < cnt) raw[lane] = a.tile_list[(long)tile * TILE_CAP + lane];
  if (lane == 0) nlist = 0;
  __syncthreads();
  int v = 0;
  if (lane < cnt) {
    v = raw[lane];
    bool dup = false;
    for (int s = 0; s < lane; s++) dup |= raw[s] == v;
    rep[lane] = dup;
  }
  __syncthreads();
  if (lane < cnt && !rep[lane]) {
    int pos = 0;
    for (int s = 0; s < cnt; s++) pos += raw[s] < v && !rep[s];
    sbox[pos] = a.boxes[v];
    atomicMax(&nlist, pos + 1);
  }
  __syncthreads();
  const int nl = nlist;
  // column i0-1+lane of DUX and column i0+lane of DUY on rows j0-1 .. j0+3
  // (DUY's row j0-1 = 0 is row jtot), then column i0+lane of the tile.
  // Boxes outermost: the 13 loads of one box are in flight together, and
  // every cell still adds its boxes in list order.
  const int i = i0 + lane;
  double dxw[TILE_Y], dxc[TILE_Y], dy[TILE_Y + 1];
  int jy[TILE_Y + 1];
#pragma unroll
  for (int r = 0; r < TILE_Y; r++) dxw[r] = dxc[r] = 0.;
#pragma unroll
  for (int r = 0; r <= TILE_Y; r++) {
    dy[r] = 0.;
    jy[r] = j0 - 1 + r == 0 ? a.ny : j0 - 1 + r;
  }
  const bool iw = i - 1 >= 1 && i - 1 <= a.nx, ic = i <= a.nx;
  for (int s = 0; s < nl; s++) {
    const Box b = sbox[s];
    double tw[TILE_Y], tc[TILE_Y], ty_[TILE_Y + 1];
#pragma unroll
    for (int r = 0; r < TILE_Y; r++) {
      const bool rin = j0 + r <= a.ny;
      tw[r] = rin && iw ? box_value(b, a.sc_fx, i - 1, j0 + r) : 0.;
      tc[r] = rin && ic ? box_value(b, a.sc_fx, i, j0 + r) : 0.;
    }
#pragma unroll
    for (int r = 0; r <= TILE_Y; r++)
      ty_[r] = ic && jy[r] <= a.ny ? box_value(b, a.sc_fy, i, jy[r]) : 0.;
#pragma unroll
    for (int r = 0; r < TILE_Y; r++) {
      dxw[r] += tw[r];
      dxc[r] += tc[r];
    }
#pragma unroll
    for (int r = 0; r <= TILE_Y; r++) dy[r] += ty_[r];
  }
  if (i > a.nx) return;
#pragma unroll
  for (int r = 0; r < TILE_Y; r++) {
    const int j = j0 + r;
    if (j > a.ny) break;
    const double dux_w = dxw[r], dux_c = dxc[r], duy_s = dy[r], duy_c = dy[r + 1];
    if (a.apply) {
      // cells no box reaches get an increment of exactly zero: adding it
      // changes no value (a -0.0 could become +0.0), so they are skipped
      if (i >= 2 && (dux_w != 0. || dux_c != 0.)) {
        const long k = L.ux(i, j);
        const double u = a.ux[k] + 0.5 * (dux_w + dux_c);
        a.ux[k] = u;
        if (j == a.ny) a.ux[L.ux(i, 0)] = u;
        if (j == 1) a.ux[L.ux(i, a.ny + 1)] = u;
      }
      if (duy_s != 0. || duy_c != 0.) {
        const long k = L.p(i, j);
        const double u = a.uy[k] + 0.5 * (duy_s + duy_c);
        a.uy[k] = u;
        if (j == a.ny) a.uy[L.p(i, 0)] = u;
        if (j == 1) a.uy[L.p(i, a.ny + 1)] = u;
      }
    } else {
      a.dux[L.p(i, j)] = dux_c;
      a.duy[L.p(i, j)] = duy_c;
    }
  }
}

struct TilesB {
  static constexpr int T = TILE_THREADS, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, TileArgs a) {
  if (halted(a.halt)) return;
  tile_body(a, vb.bx);
}
};
__global__ __launch_bounds__(TILE_THREADS) void k_tiles(TileArgs a) {
  TilesB::run(VB::here(), a);
}

// The stage start's deposit with the exchange's per-particle sums (and
// inertia(0)'s) in the same launch: the first np workgroups finalise one
// particle each, the rest are the tiles.  Neither reads what the other
// writes.
struct TilesX0B {
  static constexpr int T = TILE_THREADS, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, TileArgs t, IbmArgs a) {
  if (halted(a.halt)) return;
  if ((int)vb.bx < a.np)
    finalize_x0_particle(a, vb.bx);
  else
    tile_body(t, vb.bx - a.np);
}
};
__global__ __launch_bounds__(TILE_THREADS) void k_tiles_x0(TileArgs t, IbmArgs a) {
  TilesX0B::run(VB::here(), t, a);
}

// ---------------------------------------------------------------------
// collisions, src/suspensions/collision.c:75-367
// ---------------------------------------------------------------------
struct EllD {
  double a, b, x, y, angle;
};

// compute_evolute (collision.c:75-109) with the foot from foot_fp64,
// started at tw, the folded parameter of the previous fit in the same
// fixed-point iteration (or the FP32 estimate when tw < 0); the root on
// the first-quadrant arc is unique, so the start only changes the number
// of iterations.  (cm, sm) = cos/sin(-angle), (cp, sp) = cos/sin(+angle).
__device__ void evolute_fit_fast(const EllD &e, const EllipseShape &s,
                                 const EllipseShape32 &f, double cm, double sm, double cp,
                                 double sp, double x, double y, double &ex, double &ey,
                                 double &ix, double &iy, double &r, double &tw) {
  double x_, y_;
  {
    const double x__ = x - e.x, y__ = y - e.y;
    x_ = cm * x__ - sm * y__;
    y_ = sm * x__ + cm * y__;
  }
  const double px = fabs(x_), py = fabs(y_);
  int it32 = 0, it64 = 0;
  if (tw < 0.) tw = (double)foot_fp32(f, px, py, it32);
  double c, sn;
  foot_fp64(s, px, py, tw, c, sn, it64);
  // unfolded foot (ellipse.c:62-63)
  const double ct = x_ < 0. ? -c : c;
  const double st = y_ < 0. ? -sn : sn;
  {
    // compute_curvature, ellipse.c:15-25
    const double ct2 = ct * ct, st2 = st * st;
    const double a2 = e.a * e.a, b2 = e.b * e.b;
    const double num = e.a * e.b * (ct2 + st2);
    const double den = +a2 * st2 + b2 * ct2;
    r = 1. / fmax(num / (den * sqrt(den)), DBL_EPSILON);
  }
  x_ = s.kex * cube(ct);
  y_ = s.key * cube(st);
  ex = cp * x_ - sp * y_;
  ey = sp * x_ + cp * y_;
  ex = ex + e.x;
  ey = ey + e.y;
  x_ = e.a * ct;
  y_ = e.b * st;
  ix = cp * x_ - sp * y_;
  iy = sp * x_ + cp * y_;
}

struct CollArgs {
  const double *res;  // sub-iteration residuals (null: not gated)
  int sub;            // sub-iteration index
  int np;
  double lx, ly;
  double *prt;
  int cnstep;
  int *rows;       // np x CONTACT_CAP: colliding partners of each particle, ascending
  int *row_cnt;    // np
  int *wall_hit;   // np x 2
  int *flags;
  const int *halt;
  // the cell list (use_cells): partners are sought in a particle's cell and
  // its eight neighbours (periodic in y) instead of among all particles
  int use_cells;
  int ncx, ncy;
  double hx, hy;
  int *cl_start, *cl_items, *cl_cell, *cl_slot;
  double *cpos;  // np x 4: x + dx, y + dy, max(a, b) (written by k_coll_pack and k_increment)
};

__device__ inline double harmonic(double v0, double v1) {
  const double r = 0.5 * (1. / v0 + 1. / v1);
  return 1. / r;
}

__device__ inline double min_image_offset(double p0y, double p1y, double ly) {
  double yoffset = 0., minval = DBL_MAX;
  for (int periodic = -1; periodic <= 1; periodic++) {
    const double val = fabs((p1y + ly * periodic) - p0y);
    if (val < minval) {
      minval = val;
      yoffset = ly * periodic;
    }
  }
  return yoffset;
}

// The early exit of collision.c:156-189 for the pair n0 < n1: the bounding
// circles (radius max(a, b)) of the minimum image overlap
// (the positions x + dx, y + dy and radii max(a, b) from the packed cpos,
// the same values as the records give)
__device__ inline bool pair_candidate(const CollArgs &a, int n0, int n1) {
  const double *p0 = a.cpos + 4 * (long)n0, *p1 = a.cpos + 4 * (long)n1;
  const double x0 = p0[0], y0 = p0[1];
  const double x1 = p1[0], y1r = p1[1];
  const double y1 = y1r + min_image_offset(y0, y1r, a.ly);
  const double norm = fmax(hypot(x1 - x0, y1 - y0), DBL_EPSILON);
  return !(p0[2] + p1[2] - norm < 0.);
}

// the packed positions and radii of the broad phase, from the records
struct CollPackB {
  static constexpr int T = 256, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, CollArgs a) {
    if (halted(a.halt)) return;
    if (!sub_open(a.res, a.sub)) return;
    const int n = vb.bx * blockDim.x + threadIdx.x;
    if (n >= a.np) return;
    const double *q = a.prt + (long)n * R_N;
    double *o = a.cpos + 4 * (long)n;
    o[0] = q[R_X] + q[R_DX];
    o[1] = q[R_Y] + q[R_DY];
    o[2] = fmax(q[R_A], q[R_B]);
    o[3] = 0.;
  }
};
__global__ __launch_bounds__(256) void k_coll_pack(CollArgs a) {
  CollPackB::run(VB::here(), a);
}

// One contact of collision.c:156-304, by the two lanes l and l^1.
// A pair n0 < n1 (wall < 0): the lane of side 0 fits ellipse n0 and the
// lane of side 1 ellipse n1, each towards the other's current equivalent-
// circle centre, exchanged every iteration (the reference evaluates both
// fits from the previous centres too).  A wall (wall = 0: x = 0, 1: x = lx)
// of particle n0: both lanes fit n0 towards its mirror image.  Returns
// whether the circles overlap, with the contributions to n0 in v[0..2] and,
// for a pair, to n1 in v[3..5]; both lanes return the same values.
__device__ bool contact_force(const CollArgs &a, int n0, int n1, int wall, int side,
                              double v[6]) {
  const double wallx = wall == 0 ? 0. : a.lx;
  const double *p0 = a.prt + (long)n0 * R_N;
  const double *p1 = wall < 0 ? a.prt + (long)n1 * R_N : p0;
  const double *pm = side && wall < 0 ? p1 : p0;
  double yoffset = 0.;
  if (wall < 0) {
    yoffset = min_image_offset(p0[R_Y] + p0[R_DY], p1[R_Y] + p1[R_DY], a.ly);
  } else {  // wall early exit, collision.c:259-266
    const double x = p0[R_X] + p0[R_DX];
    const double r = fmax(p0[R_A], p0[R_B]);
    const double norm = fabs(wallx - x);
    if (r - norm < 0.) return false;
  }
  const double ym = pm[R_Y] + pm[R_DY];
  const EllD e = {pm[R_A], pm[R_B], pm[R_X] + pm[R_DX],
                  side && wall < 0 ? ym + yoffset : ym, pm[R_AZ]};
  // find_equivalent_circles, collision.c:111-135 (guarded iteration count)
  const EllipseShape sh = make_shape(e.a, e.b);
  const EllipseShape32 f = shape32(sh);
  double cm, sm, cp, sp;
  sincos(-e.angle, &sm, &cm);
  sincos(+e.angle, &sp, &cp);
  double tw = -1.;
  double cx = e.x, cy = e.y, cr = 0., ix = 0., iy = 0.;
  for (int it = 0;; it++) {
    const double ox = __shfl_xor(cx, 1), oy = __shfl_xor(cy, 1);
    const double tx = wall < 0 ? ox : 2. * wallx - cx, ty = wall < 0 ? oy : cy;
    double nx, ny;
    evolute_fit_fast(e, sh, f, cm, sm, cp, sp, tx, ty, nx, ny, ix, iy, cr, tw);
    const double dx = nx - cx, dy = ny - cy;
    cx += dx;
    cy += dy;
    const double dm = fmax(fabs(dx), fabs(dy));
    const double dmo = __shfl_xor(dm, 1);
    const double dmax = wall >= 0 ? dm : side ? fmax(dmo, dm) : fmax(dm, dmo);
    if (dmax < 1.e-8) break;
    if (it >= EQC_ITER_CAP) {
      if (side == 0) atomicOr(&a.flags[0], 8);
      break;
    }
  }
  const double ocx = __shfl_xor(cx, 1), ocy = __shfl_xor(cy, 1), ocr = __shfl_xor(cr, 1);
  const double oix = __shfl_xor(ix, 1), oiy = __shfl_xor(iy, 1);
  if (wall >= 0) {  // collision.c:268-304
    const double m = p_mass(p0), im = p_moi(p0);
    double k;
    {
      const double mass = m;
      const double r = sqrt(p0[R_A] * p0[R_B]);
      const double reft = pow(r, 1.5);
      k = mass * (EIF_PI * EIF_PI) / (reft * reft);
    }
    double nxv = wallx - cx;
    const double norm = fabs(nxv);
    nxv /= norm;
    const double overlap = cr - norm;
    if (!(overlap > 0.)) return false;
    const double cfx = k * overlap * nxv;
    const double cfy = 0.;
    v[0] = 1. / m * (-cfx);
    v[1] = 1. / m * (-cfy);
    v[2] = 1. / im * (ix * (-cfy) - iy * (-cfx));
    return true;
  }
  const double c0x = side ? ocx : cx, c0y = side ? ocy : cy, c0r = side ? ocr : cr;
  const double c1x = side ? cx : ocx, c1y = side ? cy : ocy, c1r = side ? cr : ocr;
  const double ix0 = side ? oix : ix, iy0 = side ? oiy : iy;
  const double ix1 = side ? ix : oix, iy1 = side ? iy : oiy;
  const double m0 = p_mass(p0), m1 = p_mass(p1);
  const double i0 = p_moi(p0), i1 = p_moi(p1);
  double k;
  {
    const double mass = harmonic(m0, m1);
    const double r0 = sqrt(p0[R_A] * p0[R_B]);
    const double r1 = sqrt(p1[R_A] * p1[R_B]);
    const double reft = pow(harmonic(r0, r1), 1.5);
    k = mass * (EIF_PI * EIF_PI) / (reft * reft);
  }
  double nxv = c1x - c0x, nyv = c1y - c0y;
  const double norm = fmax(hypot(nxv, nyv), DBL_EPSILON);
  nxv /= norm;
  nyv /= norm;
  const double overlap = c0r + c1r - norm;
  if (!(overlap > 0.)) return false;
  const double cfx = k * overlap * nxv;
  const double cfy = k * overlap * nyv;
  v[0] = 1. / m0 * (-cfx);
  v[1] = 1. / m0 * (-cfy);
  v[2] = 1. / i0 * (ix0 * (-cfy) - iy0 * (-cfx));
  v[3] = 1. / m1 * (+cfx);
  v[4] = 1. / m1 * (+cfy);
  v[5] = 1. / i1 * (ix1 * (+cfy) - iy1 * (+cfx));
  return true;
}

// The collision force of particle n (collision.c:306-367), by one wave:
// the lanes test all partners, then fit the candidate pairs and the two
// walls in parallel, two lanes per contact (each pair is fitted by both of
// its particles' waves, with the same operands), and the contributions are
// added in the order of the reference's loops: partners in ascending
// index, then wall x = 0, then wall x = lx.
constexpr int COLL_CAND_CAP = 256;  // candidate partners per particle

// the candidates of particle n among the particles of cell c (in the cell
// list's order), appended to s_cand
__device__ inline void cell_candidates(const CollArgs &a, int n, int c, int *s_cand, int &cnt) {
  const int lane = threadIdx.x;
  const int s0 = a.cl_start[c], s1 = a.cl_start[c + 1];
  for (int base = s0; base < s1; base += 64) {
    const int k = base + lane;
    const int m = k < s1 ? a.cl_items[k] : n;
    const bool hit = m != n && pair_candidate(a, min(n, m), max(n, m));
    const unsigned long long b = __ballot(hit);
    if (hit) {
      const int pos = cnt + __popcll(b & ((1ull << lane) - 1ull));
      if (pos < COLL_CAND_CAP) s_cand[pos] = m;
    }
    cnt += __popcll(b);
  }
}

__device__ void collide_particle(const CollArgs &a, int n) {
  __shared__ int s_cand[COLL_CAND_CAP];
  __shared__ int s_sorted[COLL_CAND_CAP];
  const int lane = threadIdx.x;
  int cnt = 0;
  const int *cand = s_cand;
  if (a.use_cells) {
    // the cell and its neighbours (ncy >= 3: the three rows are distinct);
    // two particles within reach are in the same or adjacent cells
    const int c0 = a.cl_cell[n], cx0 = c0 % a.ncx, cy0 = c0 / a.ncx;
    for (int dyc = -1; dyc <= 1; dyc++) {
      const int cy = (cy0 + dyc + a.ncy) % a.ncy;
      for (int dxc = -1; dxc <= 1; dxc++) {
        const int cx = cx0 + dxc;
        if (cx < 0 || cx >= a.ncx) continue;
        cell_candidates(a, n, cy * a.ncx + cx, s_cand, cnt);
      }
    }
    if (cnt > COLL_CAND_CAP) {
      if (lane == 0) atomicOr(&a.flags[0], 4);
      cnt = COLL_CAND_CAP;
    }
    __syncthreads();
    // ascending partner order, as the scan over all particles finds them:
    // each candidate's rank is the number of smaller ones (all distinct)
    for (int k = lane; k < cnt; k += 64) {
      const int v = s_cand[k];
      int r = 0;
      for (int j = 0; j < cnt; j++) r += s_cand[j] < v;
      s_sorted[r] = v;
    }
    cand = s_sorted;
  } else {
    for (int base = 0; base < a.np; base += 64) {
      const int m = base + lane;
      const bool c = m < a.np && m != n && pair_candidate(a, min(n, m), max(n, m));
      const unsigned long long b = __ballot(c);
      if (c) {
        const int pos = cnt + __popcll(b & ((1ull << lane) - 1ull));
        if (pos < COLL_CAND_CAP) s_cand[pos] = m;
      }
      cnt += __popcll(b);
    }
    if (cnt > COLL_CAND_CAP) {
      if (lane == 0) atomicOr(&a.flags[0], 4);
      cnt = COLL_CAND_CAP;
    }
  }
  __syncthreads();
  double fx = 0., fy = 0., tz = 0.;
  int ncont = 0;
  int *row = a.rows + (long)n * CONTACT_CAP;
  const int njobs = cnt + 2;  // the candidates, then the two walls
  for (int r = 0; r < njobs; r += 32) {
    const int k = r + (lane >> 1), side = lane & 1;
    int m = -1;
    bool hit = false;
    double w0 = 0., w1 = 0., w2 = 0.;
    if (k < njobs) {
      const int wall = k < cnt ? -1 : k - cnt;
      if (wall < 0) m = cand[k];
      double v[6] = {0., 0., 0., 0., 0., 0.};
      hit = contact_force(a, wall < 0 ? min(n, m) : n, wall < 0 ? max(n, m) : n, wall, side, v);
      const bool first = wall >= 0 || n < m;  // n's own contributions
      w0 = first ? v[0] : v[3];
      w1 = first ? v[1] : v[4];
      w2 = first ? v[2] : v[5];
      if (wall >= 0 && side == 0) a.wall_hit[2 * (long)n + wall] = hit ? 1 : 0;
    }
    unsigned long long hb = __ballot(hit && side == 0);
    while (hb) {
      const int l = __ffsll((long long)hb) - 1;
      hb &= hb - 1ull;
      const double u0 = lane_value(w0, l), u1 = lane_value(w1, l), u2 = lane_value(w2, l);
      const int mm = __builtin_amdgcn_readlane(m, l);
      if (mm >= 0) {  // a partner
        if (ncont >= CONTACT_CAP) {
          if (lane == 0) atomicOr(&a.flags[0], 16);
          continue;
        }
        if (lane == 0) row[ncont] = mm;
        ncont++;
      }
      fx += u0;
      fy += u1;
      tz += u2;
    }
  }
  if (lane == 0) {
    a.row_cnt[n] = ncont;
    double *q = a.prt + (long)n * R_N;
    q[R_CFX + a.cnstep] = fx;
    q[R_CFY + a.cnstep] = fy;
    q[R_CTZ + a.cnstep] = tz;
  }
}

// one 64-lane workgroup per particle
struct CollideB {
  static constexpr int T = 64, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, CollArgs a) {
  if (halted(a.halt)) return;
  if (!sub_open(a.res, a.sub)) return;
  collide_particle(a, vb.bx);
}
};
__global__ __launch_bounds__(64) void k_collide(CollArgs a) {
  CollideB::run(VB::here(), a);
}

// The cell list of the collision broad phase at the particles' current
// positions (x + dx, y + dy), in three launches:
//  - count: each particle's cell, and its place in the cell from a device
//    atomic on the cell's count (the order within a cell follows the
//    atomics; the broad phase sorts each particle's candidates);
//  - scan: one workgroup turns the counts into offsets and clears them, so
//    that every build starts from zero counts;
//  - place: each particle's index stored at its cell's offset plus its place.
// Count and place are gated like the collisions they serve; the scan only
// by the sub-iteration (not by the halt flag, which another stream may
// raise between the launches), so it runs whenever the count ran.

// the cell of a particle at (x, y): reciprocals may move a cell boundary by
// rounding, which the 1e-6 margin of the cell size covers
__device__ inline int cell_of(const CollArgs &a, double x, double y) {
  const double yw = y - floor(y * (1. / a.ly)) * a.ly;  // into [0, ly), for the cell only
  const int cx = min(max((int)floor(x * (1. / a.hx)), 0), a.ncx - 1);
  const int cy = min(max((int)floor(yw * (1. / a.hy)), 0), a.ncy - 1);
  return cy * a.ncx + cx;
}

struct CellCountB {
  static constexpr int T = 256, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, CollArgs a, int *counts) {
    if (halted(a.halt)) return;
    if (!sub_open(a.res, a.sub)) return;
    const int n = vb.bx * blockDim.x + threadIdx.x;
    if (n >= a.np) return;
    const double *q = a.prt + (long)n * R_N;
    const int c = cell_of(a, q[R_X] + q[R_DX], q[R_Y] + q[R_DY]);
    a.cl_cell[n] = c;
    a.cl_slot[n] = atomicAdd(&counts[c], 1);
  }
};
__global__ __launch_bounds__(256) void k_cell_count(CollArgs a, int *counts) {
  CellCountB::run(VB::here(), a, counts);
}

struct CellScanB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, CollArgs a, int *counts) {
    if (!sub_open(a.res, a.sub)) return;
    extern __shared__ int cl_lds[];
    const int nc = a.ncx * a.ncy;
    int *cnt = cl_lds;          // nc counts
    int *part = cl_lds + nc;    // blockDim.x partial sums
    for (int k = threadIdx.x; k < nc; k += blockDim.x) {
      cnt[k] = counts[k];
      counts[k] = 0;
    }
    __syncthreads();
    // exclusive scan: each thread sums a contiguous run; the runs' sums
    // are scanned within each wave by lane shuffles, the waves' totals by
    // the first wave; each run is offset in LDS, then stored coalesced
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (nc + blockDim.x - 1) / blockDim.x;
    const int k0 = min(t * per, nc), k1 = min(k0 + per, nc);
    int sum = 0;
    for (int k = k0; k < k1; k++) sum += cnt[k];
    int inc = sum;  // inclusive scan of the wave's run sums
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(inc, d, 64);
      if (lane >= d) inc += v;
    }
    if (lane == 63) part[w] = inc;
    __syncthreads();
    if (w == 0) {
      const int nw = blockDim.x >> 6;
      int tot = lane < nw ? part[lane] : 0;
      for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(tot, d, 64);
        if (lane >= d) tot += v;
      }
      if (lane < nw) part[32 + lane] = tot;  // inclusive totals of waves
    }
    __syncthreads();
    int run = (w > 0 ? part[32 + w - 1] : 0) + inc - sum;  // exclusive offset of this run
    for (int k = k0; k < k1; k++) {
      const int v = cnt[k];
      cnt[k] = run;
      run += v;
    }
    __syncthreads();
    for (int k = t; k < nc; k += blockDim.x) a.cl_start[k] = cnt[k];
    if (t == blockDim.x - 1) a.cl_start[nc] = part[32 + (blockDim.x >> 6) - 1];
  }
};
__global__ __launch_bounds__(1024) void k_cell_scan(CollArgs a, int *counts) {
  CellScanB::run(VB::here(), a, counts);
}

struct CellPlaceB {
  static constexpr int T = 256, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, CollArgs a) {
    if (halted(a.halt)) return;
    if (!sub_open(a.res, a.sub)) return;
    const int n = vb.bx * blockDim.x + threadIdx.x;
    if (n >= a.np) return;
    a.cl_items[a.cl_start[a.cl_cell[n]] + a.cl_slot[n]] = n;
  }
};
__global__ __launch_bounds__(256) void k_cell_place(CollArgs a) {
  CellPlaceB::run(VB::here(), a);
}

// One launch for a sub-iteration's inertia(1) and collision(1), which read
// the same positions and write disjoint particle members: the first np
// workgroups are the collisions (long, started first), the rest the IBM
// slices.  Bounded to 4 waves per SIMD (128 VGPRs): the collision path
// alone needs more and spills a little, which costs less than the
// occupancy its registers would take from the slices.
struct InertiaCollideB {
  static constexpr int T = 64, M = 4;
  static __device__ __forceinline__ void run(const VB &vb, IbmArgs a, CollArgs ca, int cnstep) {
  if (halted(a.halt)) return;
  const bool open = sub_open(a.res, a.sub);
  if ((int)vb.bx < ca.np) {
    if (open) collide_particle(ca, vb.bx);
  } else {
    inertia_slice(a, cnstep, vb.bx - ca.np, open);
  }
}
};
__global__ __launch_bounds__(64, 4) void k_inertia_collide(IbmArgs a, CollArgs ca, int cnstep) {
  InertiaCollideB::run(VB::here(), a, ca, cnstep);
}

struct ZeroIntB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, int *p, int n) {
  const int t = vb.bx * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0;
}
};
__global__ void k_zero_int(int *p, int n) {
  ZeroIntB::run(VB::here(), p, n);
}


// ---------------------------------------------------------------------
// particle increments / update / reset
// ---------------------------------------------------------------------
__device__ inline void atomic_max_nonneg(double *addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long *>(addr),
            (unsigned long long)__double_as_longlong(v));
}

struct ZeroResidualB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, double *dscal) { dscal[S_RESIDUAL] = 0.; }
};
__global__ void k_zero_residual(double *dscal) {
  ZeroResidualB::run(VB::here(), dscal);
}

// src/suspensions/increment.c:10-64
// fin: also finalises the preceding inertia(1) call (its slice sums and
// cache records) for this particle first, saving a launch per sub-iteration
// one wave per particle (its lanes finalise the inertia sums), four per
// workgroup; lane 0 computes the increment
struct IncrementB {
  static constexpr int T = 256, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, double *prt, int np, const double *dscal,
                                                   double ga, double Fr, double *res_out,
                                                   IbmArgs ia, int fin) {
  if (halted(ia.halt)) return;
  const bool open = sub_open(ia.res, ia.sub);
  __shared__ double red[4];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = vb.bx * 4 + w;
  // the slice range and lane 0's record inputs are loaded before the gate
  // is tested (all threads of the workgroup see the same gate), so that
  // their latencies overlap
  int s0 = 0, s1 = 0;
  if (n < np && fin) {
    s0 = ia.slice_first[n];
    s1 = ia.slice_first[n + 1];
  }
  double dt = 0., den = 0., dux_prev = 0., duy_prev = 0., dvz_prev = 0.;
  double cfx0 = 0., cfy0 = 0., ctz0 = 0., cfx1 = 0., cfy1 = 0., ctz1 = 0.;
  double fux = 0., fuy = 0., tvz = 0., iux0 = 0., iuy0 = 0., ivz0 = 0.;
  double ux0 = 0., uy0 = 0., vz0 = 0., x0 = 0., y0 = 0., az0 = 0., rad = 0.;
  double iux1 = 0., iuy1 = 0., ivz1 = 0.;
  SliceHead hd{};
  double *p = prt + (long)n * R_N;
  if (n < np && lane == 0) {
    dt = dscal[S_DT];
    den = p[R_DEN];
    dux_prev = p[R_DUX]; duy_prev = p[R_DUY]; dvz_prev = p[R_DVZ];
    cfx0 = p[R_CFX]; cfy0 = p[R_CFY]; ctz0 = p[R_CTZ];
    cfx1 = p[R_CFX + 1]; cfy1 = p[R_CFY + 1]; ctz1 = p[R_CTZ + 1];
    fux = p[R_FUX]; fuy = p[R_FUY]; tvz = p[R_TVZ];
    iux0 = p[R_IUX]; iuy0 = p[R_IUY]; ivz0 = p[R_IVZ];
    ux0 = p[R_UX]; uy0 = p[R_UY]; vz0 = p[R_VZ];
    x0 = p[R_X]; y0 = p[R_Y]; az0 = p[R_AZ];
    rad = fmax(p[R_A], p[R_B]);
    if (fin) {
      hd = ia.heads[n];
    } else {
      iux1 = p[R_IUX + 1];
      iuy1 = p[R_IUY + 1];
      ivz1 = p[R_IVZ + 1];
    }
  }
  if (!open) return;
  double res = 0.;
  double tot[SLICE_VALS];
  if (n < np && fin) slice_sums(ia, s0, s1, tot);
  if (n < np && lane == 0) {
    if (fin) {  // inertia(1) of this sub-iteration: its sums and cache records
      iux1 = tot[0];
      iuy1 = tot[1];
      ivz1 = tot[2];
      p[R_IUX + 1] = iux1;
      p[R_IUY + 1] = iuy1;
      p[R_IVZ + 1] = ivz1;
      record_pass(ia, n, hd, tot);
    }
    if (ia.res && ia.sub == 0) {  // collision(1) at the stage-start positions
      cfx1 = cfx0;
      cfy1 = cfy0;
      ctz1 = ctz0;
      p[R_CFX + 1] = cfx1;
      p[R_CFY + 1] = cfy1;
      p[R_CTZ + 1] = ctz1;
    }
    const double dux = +0.5 * ga * dt * (cfx0 + cfx1) + dt * fux + (iux1 - iux0)
                       + 1.0 * ga * dt * (1. - den) / (Fr * Fr);
    const double duy = +0.5 * ga * dt * (cfy0 + cfy1) + dt * fuy + (iuy1 - iuy0);
    const double dvz = +0.5 * ga * dt * (ctz0 + ctz1) + dt * tvz + (ivz1 - ivz0);
    p[R_DUX] = dux;
    p[R_DUY] = duy;
    p[R_DVZ] = dvz;
    res = fmax(res, fabs(dux - dux_prev));
    res = fmax(res, fabs(duy - duy_prev));
    res = fmax(res, fabs(dvz - dvz_prev));
    const double ux1 = ux0 + dux, uy1 = uy0 + duy, vz1 = vz0 + dvz;
    const double dxn = 0.5 * ga * dt * (ux0 + ux1);
    const double dyn = 0.5 * ga * dt * (uy0 + uy1);
    const double dazn = 0.5 * ga * dt * (vz0 + vz1);
    p[R_DX] = dxn;
    p[R_DY] = dyn;
    p[R_DAZ] = dazn;
    // the broad phase's packed position for the next sub-iteration's collisions
    double *cp = ia.cpos + 4 * (long)n;
    cp[0] = x0 + dxn;
    cp[1] = y0 + dyn;
    cp[2] = rad;
    // the next sub-iteration's inertia pass
    write_head(ia, n, x0 + dxn, y0 + dyn, az0 + dazn, false, fin ? &hd : nullptr, rad);
  }
  if (lane == 0) red[w] = res;
  __syncthreads();
  if (threadIdx.x == 0)
    atomic_max_nonneg(res_out, fmax(fmax(red[0], red[1]), fmax(red[2], red[3])));
}
};
__global__ __launch_bounds__(256) void k_increment(double *prt, int np, const double *dscal,
                                                   double ga, double Fr, double *res_out,
                                                   IbmArgs ia, int fin) {
  IncrementB::run(VB::here(), prt, np, dscal, ga, Fr, res_out, ia, fin);
}

// src/suspensions/update.c:8-34
struct UpdateParticlesB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, double *prt, int np, double ly, const int *halt) {
  if (halted(halt)) return;
  const int n = vb.bx * blockDim.x + threadIdx.x;
  if (n >= np) return;
  double *p = prt + (long)n * R_N;
  p[R_UX] += p[R_DUX];
  p[R_UY] += p[R_DUY];
  p[R_VZ] += p[R_DVZ];
  p[R_X] += p[R_DX];
  p[R_Y] += p[R_DY];
  p[R_AZ] += p[R_DAZ];
  if (p[R_Y] < 0.) p[R_Y] += ly;
  if (p[R_Y] > ly) p[R_Y] -= ly;
  if (p[R_AZ] < 0.) p[R_AZ] += 2. * EIF_PI;
  if (p[R_AZ] > 2. * EIF_PI) p[R_AZ] -= 2. * EIF_PI;
}
};
__global__ void k_update_particles(double *prt, int np, double ly, const int *halt) {
  UpdateParticlesB::run(VB::here(), prt, np, ly, halt);
}

// src/suspensions/reset.c:7-18
// res (RES_SLOTS sub-iteration residuals) and tile_cnt (ntiles): cleared
// when given, for the stage's sub-iterations and the next exchange's tile
// registrations
// Also the heads of the stage-start pass (positions without increments).
struct ResetB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, IbmArgs a, double *res, int *tile_cnt, int ntiles) {
  if (halted(a.halt)) return;
  const int n = vb.bx * blockDim.x + threadIdx.x;
  if (res && n < RES_SLOTS) res[n] = 0.;
  if (tile_cnt && n < ntiles) tile_cnt[n] = 0;
  if (n >= a.np) return;
  // foot-point warm starts reach back no further than the stage start, so
  // that a run restarted from a save repeats the same operations
  a.sc_key[8 * (long)n + 7] = 0;
  double *p = a.prt + (long)n * R_N;
  p[R_DUX] = 0.;
  p[R_DUY] = 0.;
  p[R_DVZ] = 0.;
  p[R_DX] = 0.;
  p[R_DY] = 0.;
  p[R_DAZ] = 0.;
  write_head(a, n, p[R_X], p[R_Y], p[R_AZ], true);
}
};
__global__ void k_reset(IbmArgs a, double *res, int *tile_cnt, int ntiles) {
  ResetB::run(VB::here(), a, res, tile_cnt, ntiles);
}

// ---------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------
static IbmArgs ibm_args(Ctx *c) {
  IbmArgs a;
  a.L = c->L;
  a.nx = c->nx;
  a.ny = c->ny;
  a.ly = c->ly;
  a.dx = c->dx;
  a.dy = c->dy;
  a.grid_size = fmin(c->dx, c->dy);
  a.idx = 1. / c->dx;
  a.idy = 1. / c->dy;
  a.pow2 = c->pow2_grid ? 1 : 0;
  a.xc = c->xc;
  a.yc = c->yc;
  a.ux = c->ux;
  a.uy = c->uy;
  a.prt = c->prt;
  a.dscal = c->dscal;
  a.boxes = c->boxes;
  a.box_off = c->box_off;
  a.sc_fx = c->sc_fx;
  a.sc_fy = c->sc_fy;
  a.sc_dist = c->sc_dist;
  a.sc_pos = c->sc_pos;
  a.sc_t = c->sc_t;
  a.sc_key = c->sc_key;
  a.tile_cnt = c->tile_cnt;
  a.tile_list = c->tile_list;
  a.ntx = c->ntx;
  a.nty = c->nty;
  a.flags = c->flags;
  a.stats = c->ibm_pp;
  a.cpos = c->cpos;
  a.np = c->np;
  a.slice_first = c->slice_first;
  a.slice_owner = c->slice_owner;
  a.slice_info = c->slice_info;
  a.res = nullptr;
  a.sub = 0;
  a.us_x = a.us_y = nullptr;
  a.slice_part0 = c->slice_part0;
  a.slice_part = c->slice_part;
  a.heads = static_cast<SliceHead *>(c->heads);
  a.halt = c->halt;
  return a;
}

int launch_ref_iterations(Ctx *c, unsigned long long *dout) {
  if (c->np == 0) return EIF_OK;
  IbmArgs a = ibm_args(c);
  hipLaunchKernelGGL(k_ref_iterations, dim3(c->np), dim3(IBM_THREADS), 0, c->stream, a, dout);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

int susp_init(Ctx *c) {
  // per-particle scratch capacity: three images of a box no wider than
  // floor(2r/h)+10 cells in each direction
  std::vector<double> rec((size_t)c->np * R_N);
  EIF_HIP(hipMemcpy(rec.data(), c->prt, rec.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<long> off(c->np + 1, 0);
  // slices of IBM_SLICE cells per particle: the boxes of at most two images
  // meet the domain when ly > 2r + 8 dy, and two clipped images together
  // are at most two rows taller than one box; otherwise allow three images
  std::vector<int> first(c->np + 1, 0);
  for (int n = 0; n < c->np; n++) {
    const double r = fmax(rec[(size_t)n * R_N + R_A], rec[(size_t)n * R_N + R_B]);
    const long cx = (long)(2. * r / c->dx) + 10, cy = (long)(2. * r / c->dy) + 10;
    const long wx = cx < c->nx ? cx : c->nx, wy = cy < c->ny ? cy : c->ny;
    off[n + 1] = off[n] + 3 * wx * wy;
    const long cells = c->ly > 2. * r + 16. * c->dy ? wx * (wy + 4) : 3 * wx * wy;
    first[n + 1] = first[n] + (int)((cells + IBM_SLICE - 1) / IBM_SLICE);
  }
  c->scratch_cells = off[c->np];
  c->nslices = first[c->np];
  TRY_S(coll_cells_setup(c, rec.data()));
  EIF_HIP(hipMalloc(&c->slice_first, (c->np + 1) * sizeof(int)));
  EIF_HIP(hipMemcpy(c->slice_first, first.data(), (c->np + 1) * sizeof(int), hipMemcpyHostToDevice));
  {
    std::vector<int> owner(c->nslices > 0 ? c->nslices : 1, 0);
    for (int n = 0; n < c->np; n++)
      for (int sl = first[n]; sl < first[n + 1]; sl++) owner[sl] = n;
    EIF_HIP(hipMalloc(&c->slice_owner, owner.size() * sizeof(int)));
    EIF_HIP(hipMemcpy(c->slice_owner, owner.data(), owner.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  EIF_HIP(hipMalloc(&c->slice_info, 2 * (size_t)(c->np > 0 ? c->np : 1) * sizeof(long)));
  EIF_HIP(hipMemset(c->slice_info, 0, 2 * (size_t)(c->np > 0 ? c->np : 1) * sizeof(long)));
  EIF_HIP(hipMalloc(&c->slice_part, SLICE_VALS * (size_t)(c->nslices > 0 ? c->nslices : 1) * sizeof(double)));
  EIF_HIP(hipMalloc(&c->slice_part0, SLICE_VALS * (size_t)(c->nslices > 0 ? c->nslices : 1) * sizeof(double)));
  EIF_HIP(hipMalloc(&c->heads, sizeof(SliceHead) * (size_t)(c->np > 0 ? c->np : 1)));
  EIF_HIP(hipMalloc(&c->ibm_pp, (4 * (size_t)c->np + 1) * sizeof(unsigned long long)));
  EIF_HIP(hipMemset(c->ibm_pp, 0, (4 * (size_t)c->np + 1) * sizeof(unsigned long long)));
  EIF_HIP(hipMalloc(&c->box_off, (c->np + 1) * sizeof(long)));
  EIF_HIP(hipMemcpy(c->box_off, off.data(), (c->np + 1) * sizeof(long), hipMemcpyHostToDevice));
  const size_t sc = (size_t)(c->scratch_cells > 0 ? c->scratch_cells : 1) * sizeof(double);
  EIF_HIP(hipMalloc(&c->sc_fx, sc));
  EIF_HIP(hipMalloc(&c->sc_fy, sc));
  EIF_HIP(hipMalloc(&c->sc_dist, sc));
  EIF_HIP(hipMalloc(&c->sc_t, sc));
  EIF_HIP(hipMalloc(&c->sc_key, 8 * (size_t)(c->np > 0 ? c->np : 1) * sizeof(int)));
  EIF_HIP(hipMemset(c->sc_key, 0, 8 * (size_t)(c->np > 0 ? c->np : 1) * sizeof(int)));
  EIF_HIP(hipMalloc(&c->sc_pos, 4 * (size_t)(c->np > 0 ? c->np : 1) * sizeof(double)));
  EIF_HIP(hipMemset(c->sc_pos, 0, 4 * (size_t)(c->np > 0 ? c->np : 1) * sizeof(double)));
  EIF_HIP(hipMalloc(&c->boxes, 3 * (size_t)(c->np > 0 ? c->np : 1) * sizeof(Box)));
  EIF_HIP(hipMemset(c->boxes, 0, 3 * (size_t)(c->np > 0 ? c->np : 1) * sizeof(Box)));
  c->ntx = (c->nx + TILE_X - 1) / TILE_X;
  c->nty = (c->ny + TILE_Y - 1) / TILE_Y;
  const size_t ntiles = (size_t)c->ntx * c->nty;
  EIF_HIP(hipMalloc(&c->tile_cnt, ntiles * sizeof(int)));
  EIF_HIP(hipMemset(c->tile_cnt, 0, ntiles * sizeof(int)));
  EIF_HIP(hipMalloc(&c->tile_list, ntiles * TILE_CAP * sizeof(int)));
  // collisions
  {
    const size_t npp = (size_t)(c->np > 0 ? c->np : 1);
    EIF_HIP(hipMalloc(&c->rows, npp * CONTACT_CAP * sizeof(int)));
    EIF_HIP(hipMalloc(&c->cpos, 4 * npp * sizeof(double)));
    EIF_HIP(hipMemset(c->cpos, 0, 4 * npp * sizeof(double)));
    EIF_HIP(hipMalloc(&c->row_cnt, npp * sizeof(int)));
    EIF_HIP(hipMemset(c->row_cnt, 0, npp * sizeof(int)));
  }
  EIF_HIP(hipMalloc(&c->wall_hit, 2 * (size_t)(c->np > 0 ? c->np : 1) * sizeof(int)));
  EIF_HIP(hipMemset(c->wall_hit, 0, 2 * (size_t)(c->np > 0 ? c->np : 1) * sizeof(int)));
  return EIF_OK;
}

int launch_reset(Ctx *c, hipStream_t st, bool clear_tiles) {
  if (!st) st = c->stream;
  // also clears the sub-iteration residuals, so it runs without particles too
  const int ntiles = c->np > 0 && clear_tiles ? c->ntx * c->nty : 0;
  const int n = std::max(std::max(c->np, ntiles), RES_SLOTS);
  go<ResetB>(k_reset, dim3((n + 255) / 256), dim3(256), 0, st, ibm_args(c),
                     c->res_slots, ntiles > 0 ? c->tile_cnt : nullptr, ntiles);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

int launch_inertia(Ctx *c, int cnstep, hipStream_t st, bool defer_finalize) {
  if (c->np == 0) return EIF_OK;
  if (!st) st = c->stream;
  ProfScope ps(c, "ibm_inertia", st);
  IbmArgs a = ibm_args(c);
  go<HeadsB>(k_heads, dim3((c->np + 255) / 256), dim3(256), 0, st, a, 1);
  go<InertiaB>(k_inertia, dim3(c->nslices), dim3(64), 0, st, a, cnstep);
  if (!defer_finalize)
    go<IbmFinalizeB>(k_ibm_finalize, dim3((c->np + 3) / 4), dim3(256), 0, st, a, cnstep);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

int launch_exchange(Ctx *c, bool cached_known, bool tiles_clear, bool defer_finalize) {
  if (c->np == 0) return EIF_OK;
  ProfScope ps(c, "ibm_exchange");
  const size_t ntiles = (size_t)c->ntx * c->nty;
  if (!tiles_clear) {  // clear the previous call's tile registrations
    go<ZeroIntB>(k_zero_int, dim3((unsigned)((ntiles + 255) / 256)), dim3(256), 0,
                       c->stream, c->tile_cnt, (int)ntiles);
  }
  IbmArgs a = ibm_args(c);
  go<HeadsB>(k_heads, dim3((c->np + 255) / 256), dim3(256), 0, c->stream, a, 0);
  if (!cached_known)
    go<IbmDistB>(k_ibm_dist, dim3(c->nslices), dim3(64), 0, c->stream, a);
  go<ExchangeB>(k_exchange, dim3(c->nslices), dim3(64), 0, c->stream, a,
                     cached_known ? 1 : 0);
  EIF_HIP(hipGetLastError());
  if (!defer_finalize) return launch_exchange_finalize(c, c->stream);
  return EIF_OK;
}

// Signed distances (as tanh(BETA d)) of the box cells at the particles'
// positions without increments, into the per-cell cache
int launch_ibm_dist(Ctx *c, hipStream_t st) {
  if (c->np == 0) return EIF_OK;
  go<IbmDistB>(k_ibm_dist, dim3(c->nslices), dim3(64), 0, st ? st : c->stream,
                     ibm_args(c));
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// The exchange at the stage start (distances cached by launch_ibm_dist, tile
// registrations cleared by the reset) with inertia(0) of the stage-start
// velocity (us_x, us_y) in the same pass over the cells; the sums are added
// by the deposit launch (launch_update_momentum_x0)
int launch_exchange_inertia0(Ctx *c, const double *us_x, const double *us_y) {
  if (c->np == 0) return EIF_OK;
  ProfScope ps(c, "ibm_exchange");
  IbmArgs a = ibm_args(c);
  a.us_x = us_x;
  a.us_y = us_y;
  go<ExchangeB>(k_exchange, dim3(c->nslices), dim3(64), 0, c->stream, a, 0);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// the exchange's per-particle sums (slice order); nothing on the fluid side
// reads them, so this may run on another stream after the exchange
int launch_exchange_finalize(Ctx *c, hipStream_t st) {
  if (c->np == 0) return EIF_OK;
  go<IbmFinalizeB>(k_ibm_finalize, dim3((c->np + 3) / 4), dim3(256), 0, st, ibm_args(c), 2);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

static TileArgs tile_args(Ctx *c, int apply) {
  TileArgs t;
  t.L = c->L;
  t.nx = c->nx;
  t.ny = c->ny;
  t.ntx = c->ntx;
  t.boxes = c->boxes;
  t.sc_fx = c->sc_fx;
  t.sc_fy = c->sc_fy;
  t.tile_cnt = c->tile_cnt;
  t.tile_list = c->tile_list;
  t.ux = c->ux;
  t.uy = c->uy;
  t.dux = c->dux;
  t.duy = c->duy;
  t.apply = apply;
  t.halt = c->halt;
  return t;
}

static int tile_grid(Ctx *c) {
  const long ntiles = (long)c->ntx * c->nty;
  return (int)ntiles;  // one workgroup per tile; idle tiles return at once
}

int launch_update_momentum(Ctx *c) {
  if (c->np == 0) return EIF_OK;
  ProfScope ps(c, "ibm_apply");
  TileArgs t = tile_args(c, 1);
  go<TilesB>(k_tiles, dim3(tile_grid(c)), dim3(TILE_THREADS), 0, c->stream, t);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// the stage start's deposit and the exchange's particle sums, one launch
int launch_update_momentum_x0(Ctx *c) {
  if (c->np == 0) return EIF_OK;
  ProfScope ps(c, "ibm_apply");
  TileArgs t = tile_args(c, 1);
  go<TilesX0B>(k_tiles_x0, dim3(c->np + tile_grid(c)), dim3(TILE_THREADS), 0, c->stream,
                     t, ibm_args(c));
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

int launch_materialise_du(Ctx *c) {
  const size_t bytes = (size_t)c->L.p_len() * sizeof(double);
  EIF_HIP(hipMemsetAsync(c->dux, 0, bytes, c->stream));
  EIF_HIP(hipMemsetAsync(c->duy, 0, bytes, c->stream));
  if (c->np > 0) {
    TileArgs t = tile_args(c, 0);
    go<TilesB>(k_tiles, dim3(tile_grid(c)), dim3(TILE_THREADS), 0, c->stream,
                       t);
    EIF_HIP(hipGetLastError());
  }
  int r = launch_bc_p(c, c->dux);
  if (r) return r;
  return launch_bc_p(c, c->duy);
}

// whether the collision broad phase uses the cell list
static bool coll_use_cells(const Ctx *c) {
  if (c->cl_ncy < 3 || c->coll_mode == 1) return false;
  return c->coll_mode == 2 || c->np >= COLL_CELLS_MIN;
}

int coll_cells_setup(Ctx *c, const double *rec) {
  double rmax = 0.;
  for (int n = 0; n < c->np; n++)
    rmax = fmax(rmax, fmax(rec[(size_t)n * R_N + R_A], rec[(size_t)n * R_N + R_B]));
  int ncx = 0, ncy = 0;
  if (rmax > 0.) {
    // cells at least the widest reach (with a margin above rounding), and
    // at most COLL_CELLS_CAP of them
    double h = fmax(2. * rmax * (1. + 1e-6), sqrt(c->lx * c->ly / COLL_CELLS_CAP));
    for (;;) {
      const double fx = fmax(floor(c->lx / h), 1.), fy = fmax(floor(c->ly / h), 1.);
      if (fx * fy <= COLL_CELLS_CAP) {
        ncx = (int)fx;
        ncy = (int)fy;
        break;
      }
      h *= 1.05;
    }
  }
  // sized for the most cells once: graphs keep the addresses.  The counts
  // start at zero and every build leaves them at zero; cells and places
  // start at zero, so that a build whose count was halted still places
  // within bounds
  if (!c->cl_start) {
    EIF_HIP(hipMalloc(&c->cl_start, (COLL_CELLS_CAP + 1) * sizeof(int)));
    EIF_HIP(hipMalloc(&c->cl_count, COLL_CELLS_CAP * sizeof(int)));
    const size_t n = (size_t)(c->np > 0 ? c->np : 1) * sizeof(int);
    EIF_HIP(hipMalloc(&c->cl_items, n));
    EIF_HIP(hipMalloc(&c->cl_cell, n));
    EIF_HIP(hipMalloc(&c->cl_slot, n));
    EIF_HIP(hipMemset(c->cl_start, 0, (COLL_CELLS_CAP + 1) * sizeof(int)));
    EIF_HIP(hipMemset(c->cl_count, 0, COLL_CELLS_CAP * sizeof(int)));
    EIF_HIP(hipMemset(c->cl_cell, 0, n));
    EIF_HIP(hipMemset(c->cl_slot, 0, n));
  }
  c->cl_ncx = ncx;
  c->cl_ncy = ncy;
  c->cl_hx = ncx > 0 ? c->lx / ncx : 0.;
  c->cl_hy = ncy > 0 ? c->ly / ncy : 0.;
  return EIF_OK;
}

int set_collision_search(Ctx *c, int mode) {
  if (mode < 0 || mode > 2) {
    set_error("collision search must be 0 (by particle count), 1 (all pairs) or 2 (cell list)");
    return EIF_ERR_ARG;
  }
  c->coll_mode = mode;
  return EIF_OK;
}

static CollArgs coll_args(Ctx *c, int cnstep) {
  CollArgs a;
  a.res = nullptr;
  a.sub = 0;
  a.np = c->np;
  a.lx = c->lx;
  a.ly = c->ly;
  a.prt = c->prt;
  a.cnstep = cnstep;
  a.rows = c->rows;
  a.row_cnt = c->row_cnt;
  a.wall_hit = c->wall_hit;
  a.flags = c->flags;
  a.halt = c->halt;
  a.use_cells = coll_use_cells(c) ? 1 : 0;
  a.ncx = c->cl_ncx;
  a.ncy = c->cl_ncy;
  a.hx = c->cl_hx;
  a.hy = c->cl_hy;
  a.cl_start = c->cl_start;
  a.cl_items = c->cl_items;
  a.cl_cell = c->cl_cell;
  a.cl_slot = c->cl_slot;
  a.cpos = c->cpos;
  return a;
}

// the cell list for a collision launch with these (gated) arguments
static void launch_cell_list(Ctx *c, const CollArgs &a, hipStream_t st) {
  if (!a.use_cells) return;
  const size_t lds = ((size_t)c->cl_ncx * c->cl_ncy + 1024) * sizeof(int);
  const dim3 g((c->np + 255) / 256);
  go<CellCountB>(k_cell_count, g, dim3(256), 0, st, a, c->cl_count);
  go<CellScanB>(k_cell_scan, dim3(1), dim3(1024), lds, st, a, c->cl_count);
  go<CellPlaceB>(k_cell_place, g, dim3(256), 0, st, a);
}

int launch_collision(Ctx *c, int cnstep, hipStream_t st) {
  if (c->np == 0) return EIF_OK;
  if (!st) st = c->stream;
  ProfScope ps(c, "collision", st);
  const CollArgs a = coll_args(c, cnstep);
  go<CollPackB>(k_coll_pack, dim3((c->np + 255) / 256), dim3(256), 0, st, a);
  launch_cell_list(c, a, st);
  go<CollideB>(k_collide, dim3(c->np), dim3(64), 0, st, a);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

int launch_substep_forces(Ctx *c, int sub) {
  if (c->np == 0) return EIF_OK;
  IbmArgs a = ibm_args(c);
  a.res = c->res_slots;
  a.sub = sub;
  if (sub == 0) {
    // The first sub-iteration's positions are the stage start's (dx = 0
    // after the reset): collision(1) would repeat collision(0) operand for
    // operand, so the increment copies collision(0)'s forces instead
    ProfScope ps(c, "ibm_inertia");
    go<InertiaB>(k_inertia, dim3(c->nslices), dim3(64), 0, c->stream, a, 1);
    EIF_HIP(hipGetLastError());
    return EIF_OK;
  }
  CollArgs ca = coll_args(c, 1);
  ca.res = a.res;
  ca.sub = sub;
  launch_cell_list(c, ca, c->stream);
  hipEvent_t e0, e1;
  if (prof_kernel_events(c, "ibm_substep", &e0, &e1))  // a sampled launch stamps its events
    hipExtLaunchKernelGGL(k_inertia_collide, dim3(c->np + c->nslices), dim3(64), 0, c->stream,
                          e0, e1, 0, a, ca, 1);
  else
    go<InertiaCollideB>(k_inertia_collide, dim3(c->np + c->nslices), dim3(64), 0, c->stream, a,
                       ca, 1);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

int launch_increment(Ctx *c, int rk, int sub) {
  IbmArgs ia = ibm_args(c);
  double *res_out = c->dscal + S_RESIDUAL;
  if (sub < 0) {  // not gated: the residual starts from zero
    go<ZeroResidualB>(k_zero_residual, dim3(1), dim3(1), 0, c->stream, c->dscal);
  } else {  // the residual slot was cleared at the stage start
    ia.res = c->res_slots;
    ia.sub = sub;
    res_out = c->res_slots + sub;
  }
  if (c->np > 0)
    go<IncrementB>(k_increment, dim3((c->np + 3) / 4), dim3(256), 0, c->stream,
                       c->prt, c->np, c->dscal, c->gamma[rk], c->cfg.Fr, res_out, ia,
                       sub >= 0 ? 1 : 0);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// one thread copies the state of the first `queued` sub-iterations into
// pinned host memory: how many ran, and the residual of the last one.  The
// record number is stored last with system-scope release, so a host that
// sees it sees the rest.
__global__ void k_status(const double *res, int queued, const int *flags,
                         const double *dscal, HostStatus *h, unsigned seq) {
  int done = 0;
  double r = 0.;
  for (int s = 0; s < queued && sub_open(res, s); s++) {
    done = s + 1;
    r = res[s];
  }
  h->residual = r;
  h->dt = dscal[S_DT];
  h->substeps = done;
  h->flags = flags[0];
  __hip_atomic_store(&h->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

int launch_status(Ctx *c, int queued) {
  hipLaunchKernelGGL(k_status, dim3(1), dim3(1), 0, c->stream, c->res_slots, queued,
                     c->flags, c->dscal, c->hstat, ++c->hstat_seq);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// Graph steps: the end of a stage's captured sub-iterations.  Records how
// many of the first K ran and the last residual; if the loop would go on
// (sub-iteration K is open), stops the rest of the step (halt = stage + 1)
// for the host to resume.
struct StageGateB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, const double *res, int K, int stage, int *counts, double *resid,
                             int *halt) {
  if (*halt) return;
  int done = 0;
  double r = 0.;
  for (int s = 0; s < K && sub_open(res, s); s++) {
    done = s + 1;
    r = res[s];
  }
  counts[stage] = done;
  resid[stage] = r;
  if (sub_open(res, K)) *halt = stage + 1;
}
};
__global__ void k_stage_gate(const double *res, int K, int stage, int *counts, double *resid,
                             int *halt) {
  StageGateB::run(VB::here(), res, K, stage, counts, resid, halt);
}

int launch_stage_gate(Ctx *c, int stage, int K) {
  go<StageGateB>(k_stage_gate, dim3(1), dim3(1), 0, c->stream, c->res_slots, K, stage,
                     c->g_counts, c->g_resid, c->halt);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// Graph steps: the step's record for the host (two slots, by the parity of
// the device's step counter), the record number stored last with
// system-scope release.  Not gated: a stopped step reports its stop.
struct GraphStatusB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, const int *halt, const int *counts, const double *resid,
                               const double *dscal, const int *flags, unsigned *seq,
                               GraphStatus *h) {
  const unsigned s = *seq + 1;
  *seq = s;
  GraphStatus *o = h + (s & 1);
  o->dt = dscal[S_DT];
  for (int k = 0; k < 3; k++) {
    o->counts[k] = counts[k];
    o->resid[k] = resid[k];
  }
  o->halt = *halt;
  o->flags = flags[0];
  __hip_atomic_store(&o->seq, s, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
};
__global__ void k_graph_status(const int *halt, const int *counts, const double *resid,
                               const double *dscal, const int *flags, unsigned *seq,
                               GraphStatus *h) {
  GraphStatusB::run(VB::here(), halt, counts, resid, dscal, flags, seq, h);
}

int launch_graph_status(Ctx *c) {
  go<GraphStatusB>(k_graph_status, dim3(1), dim3(1), 0, c->stream, c->halt, c->g_counts,
                     c->g_resid, c->dscal, c->flags, c->g_seq, c->gstat);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

int launch_update_particles(Ctx *c, hipStream_t st) {
  if (c->np == 0) return EIF_OK;
  go<UpdateParticlesB>(k_update_particles, dim3((c->np + 255) / 256), dim3(256), 0,
                     st ? st : c->stream, c->prt, c->np, c->ly, c->halt);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

}  // namespace eif
