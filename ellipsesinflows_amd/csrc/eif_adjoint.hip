// eif_adjoint.hip -- hand-written adjoint of one RK stage of integrate()
// (src/main.c:15-53) for contact-free suspensions, with dt held at the
// value the forward used.
//
// Forward stage (the reference's call order, src/main.c:15-53):
//   I0  = inertia(X0, u^k)                          inertia.c:9-70
//   u*  = u^k + a dt A(u^k) + b dt A(u^{k-1}) + g dt G(p)   update_velocity.c
//   F, DU = exchange(X0, U0, u*)                    exchange.c:13-88
//   u** = u* + faceavg(DU)                          update.c:36-59
//   psi = L^-1 div(u**) / (g dt);  u^{k+1} = u** - g dt grad(psi);
//   p^{k+1} = p + psi                                compute_potential.c,
//                                                    correct_velocity.c,
//                                                    update_pressure.c
//   loop s: X1_s = X0 + dX_{s-1};  dU_s = dt F + inertia(X1_s, u^{k+1}) - I0;
//           dX_s = g dt (U0 + dU_s / 2)             increment.c:10-64
//   X' = X0 + dX, U' = U0 + dU                      update.c:8-34
// Its adjoint runs these backwards.  The stencils' local partial
// derivatives are written out by hand (the transposed advection and
// diffusion stencils) and scattered with atomic adds; the projection
// u -> u - G L^-1 D u is self-adjoint on the uniform staggered grid, so its
// adjoint is the forward projection applied to the cotangent, with the
// pressure cotangent added to the right-hand side; the weights' derivatives
// use d(signed distance)/d(point) = the unit normal through the foot point
// (the foot is a stationary point of the distance, src/ellipse.c:30-65), and
// d v_weight / d dist = s_weight (auxiliary.c:49-59).
//
// What is recorded per stage: u^k, u*, u^{k+1} and the particle records at
// the stage start and before each sub-iteration.  Collisions are
// forward-only: a recorded step reports an error if any contact occurs.
#include "eif_internal.hpp"
#include <cmath>
#include <cstring>
#include <vector>

namespace eif {

#define BETA_W 2.
#define TRY(x)         \
  do {                 \
    int _r = (x);      \
    if (_r) return _r; \
  } while (0)

// ---------------------------------------------------------------------
// cotangent scatter into the reference layouts: halo rows fold onto the
// periodic rows they copy; wall values are boundary data, not unknowns
// ---------------------------------------------------------------------
__device__ inline int fold_j(int j, int ny) { return j == 0 ? ny : (j == ny + 1 ? 1 : j); }

__device__ inline void add_ux(const Layout &L, double *a, int i, int j, double v) {
  if (i < 2 || i > L.nx || v == 0.) return;
  atomicAdd(a + L.ux(i, fold_j(j, L.ny)), v);
}
__device__ inline void add_p(const Layout &L, double *a, int i, int j, double v) {
  if (i < 1 || i > L.nx || v == 0.) return;
  atomicAdd(a + L.p(i, fold_j(j, L.ny)), v);
}

// ---------------------------------------------------------------------
// predictor: transposed stencils of update_velocity.c:9-237
// ---------------------------------------------------------------------
struct PredAdj {
  Layout L;
  const double *sx, *sy;            // seeds: cotangent of u*
  const double *ux, *uy;            // u^k
  const double *uxo, *uyo;          // u^{k-1} (null at stage 0)
  double *ox, *oy;                  // cotangent of u^k (accumulated)
  double *oxo, *oyo;                // cotangent of u^{k-1} (accumulated)
  double *op;                       // cotangent of p (accumulated)
  double *ore;                      // cotangent of 1/Re (accumulated)
  const double *dxf, *dxc;
  double dy, rre, aldt, bedt, gadt;
};

// reverse of src_x at x-face (i, j) with seed S, from the window of (ux, uy)
__device__ inline double src_x_rev(const PredAdj &a, const double *ux, const double *uy,
                                   double *ox, double *oy, int i, int j, double S) {
  const Layout &L = a.L;
  const double dxf_im = a.dxf[i - 2], dxf_i = a.dxf[i - 1], dxc_i = a.dxc[i - 1], dy = a.dy;
  const double uxc_w = ux[L.ux(i - 1, j)], uxc_c = ux[L.ux(i, j)], uxc_e = ux[L.ux(i + 1, j)];
  const double uxm_c = ux[L.ux(i, j - 1)], uxp_c = ux[L.ux(i, j + 1)];
  const double uyc_w = uy[L.p(i - 1, j)], uyc_c = uy[L.p(i, j)];
  const double uyp_w = uy[L.p(i - 1, j + 1)], uyp_c = uy[L.p(i, j + 1)];
  const double g_w = (uxc_c - uxc_w) / dxf_im, g_e = (uxc_e - uxc_c) / dxf_i;
  const double g_s = (uxc_c - uxm_c) / dy, g_n = (uxp_c - uxc_c) / dy;
  const double cw = dxf_im / (2. * dxc_i), ce = dxf_i / (2. * dxc_i);
  const double aw = 0.5 * (uxc_w + uxc_c), ae = 0.5 * (uxc_c + uxc_e);
  const double vs = cw * uyc_w + ce * uyc_c, vn = cw * uyp_w + ce * uyp_c;
  const double rre = a.rre;
  const double gwb = S * (-cw * aw - rre / dxc_i), geb = S * (-ce * ae + rre / dxc_i);
  const double gsb = S * (-0.5 * vs - rre / dy), gnb = S * (-0.5 * vn + rre / dy);
  const double awb = S * (-cw * g_w), aeb = S * (-ce * g_e);
  const double vsb = S * (-0.5 * g_s), vnb = S * (-0.5 * g_n);
  add_ux(L, ox, i - 1, j, -gwb / dxf_im + 0.5 * awb);
  add_ux(L, ox, i, j, gwb / dxf_im - geb / dxf_i + gsb / dy - gnb / dy + 0.5 * awb + 0.5 * aeb);
  add_ux(L, ox, i + 1, j, geb / dxf_i + 0.5 * aeb);
  add_ux(L, ox, i, j - 1, -gsb / dy);
  add_ux(L, ox, i, j + 1, gnb / dy);
  add_p(L, oy, i - 1, j, cw * vsb);
  add_p(L, oy, i, j, ce * vsb);
  add_p(L, oy, i - 1, j + 1, cw * vnb);
  add_p(L, oy, i, j + 1, ce * vnb);
  return S * ((g_e - g_w) / dxc_i + (g_n - g_s) / dy);  // d/d(1/Re)
}

// reverse of src_y at y-face (i, j)
__device__ inline double src_y_rev(const PredAdj &a, const double *ux, const double *uy,
                                   double *ox, double *oy, int i, int j, double S) {
  const Layout &L = a.L;
  const double dxf_i = a.dxf[i - 1], dxc_i = a.dxc[i - 1], dxc_ip = a.dxc[i], dy = a.dy;
  const double uyc_w = uy[L.p(i - 1, j)], uyc_c = uy[L.p(i, j)], uyc_e = uy[L.p(i + 1, j)];
  const double uym_c = uy[L.p(i, j - 1)], uyp_c = uy[L.p(i, j + 1)];
  const double uxm_c = ux[L.ux(i, j - 1)], uxc_c = ux[L.ux(i, j)];
  const double uxm_e = ux[L.ux(i + 1, j - 1)], uxc_e = ux[L.ux(i + 1, j)];
  const double g_w = (uyc_c - uyc_w) / dxc_i, g_e = (uyc_e - uyc_c) / dxc_ip;
  const double g_s = (uyc_c - uym_c) / dy, g_n = (uyp_c - uyc_c) / dy;
  const double cw = dxc_i / (2. * dxf_i), ce = dxc_ip / (2. * dxf_i);
  const double uw = 0.5 * (uxm_c + uxc_c), ue = 0.5 * (uxm_e + uxc_e);
  const double as = 0.5 * (uym_c + uyc_c), an = 0.5 * (uyc_c + uyp_c);
  const double rre = a.rre;
  const double gwb = S * (-cw * uw - rre / dxf_i), geb = S * (-ce * ue + rre / dxf_i);
  const double gsb = S * (-0.5 * as - rre / dy), gnb = S * (-0.5 * an + rre / dy);
  const double uwb = S * (-cw * g_w), ueb = S * (-ce * g_e);
  const double asb = S * (-0.5 * g_s), anb = S * (-0.5 * g_n);
  add_p(L, oy, i - 1, j, -gwb / dxc_i);
  add_p(L, oy, i, j, gwb / dxc_i - geb / dxc_ip + gsb / dy - gnb / dy + 0.5 * asb + 0.5 * anb);
  add_p(L, oy, i + 1, j, geb / dxc_ip);
  add_p(L, oy, i, j - 1, -gsb / dy + 0.5 * asb);
  add_p(L, oy, i, j + 1, gnb / dy + 0.5 * anb);
  add_ux(L, ox, i, j - 1, 0.5 * uwb);
  add_ux(L, ox, i, j, 0.5 * uwb);
  add_ux(L, ox, i + 1, j - 1, 0.5 * ueb);
  add_ux(L, ox, i + 1, j, 0.5 * ueb);
  return S * ((g_e - g_w) / dxf_i + (g_n - g_s) / dy);
}

__global__ __launch_bounds__(256) void k_adj_predictor(PredAdj a) {
  const Layout L = a.L;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1;
  double rre_b = 0.;
  if (i <= L.nx && j <= L.ny) {
    if (i >= 2) {
      const double S = a.sx[L.ux(i, j)];
      if (S != 0.) {
        rre_b += src_x_rev(a, a.ux, a.uy, a.ox, a.oy, i, j, a.aldt * S);
        if (a.uxo) rre_b += src_x_rev(a, a.uxo, a.uyo, a.oxo, a.oyo, i, j, a.bedt * S);
        // pressure gradient -(P(i) - P(i-1)) / DXC(i)
        const double pb = a.gadt * S / a.dxc[i - 1];
        add_p(L, a.op, i, j, -pb);
        add_p(L, a.op, i - 1, j, pb);
      }
    }
    const double S = a.sy[L.p(i, j)];
    if (S != 0.) {
      rre_b += src_y_rev(a, a.ux, a.uy, a.ox, a.oy, i, j, a.aldt * S);
      if (a.uxo) rre_b += src_y_rev(a, a.uxo, a.uyo, a.oxo, a.oyo, i, j, a.bedt * S);
      const double pb = a.gadt * S / a.dy;
      add_p(L, a.op, i, j, -pb);
      add_p(L, a.op, i, j - 1, pb);
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = rre_b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != 0.) atomicAdd(a.ore, red[0]);
}

// out = a + b (whole arrays), and b cleared
__global__ void k_adj_sum_clear(double *out, const double *a, double *b, long n) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < n) {
    out[t] = a[t] + b[t];
    b[t] = 0.;
  }
}

// cotangent fields in the layout the forward kernels read: the periodic
// halo rows copy the interior, the wall values are zero
__global__ void k_adj_halo(Layout L, double *ux, double *uy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = L.nx, ny = L.ny;
  if (i <= nx + 1) {
    const bool wall_x = i < 2 || i > nx;
    for (int j = 0; j <= ny + 1; j++)
      if (wall_x && i >= 1) ux[L.ux(i, j)] = 0.;
    if (i >= 1) {
      ux[L.ux(i, 0)] = wall_x ? 0. : ux[L.ux(i, ny)];
      ux[L.ux(i, ny + 1)] = wall_x ? 0. : ux[L.ux(i, 1)];
    }
    const bool wall_y = i < 1 || i > nx;
    for (int j = 0; j <= ny + 1; j++)
      if (wall_y) uy[L.p(i, j)] = 0.;
    if (!wall_y) {
      uy[L.p(i, 0)] = uy[L.p(i, ny)];
      uy[L.p(i, ny + 1)] = uy[L.p(i, 1)];
    }
  }
}

// halo rows and wall values of cotangent fields cleared
__global__ void k_adj_clean(Layout L, double *ux, double *uy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = L.nx, ny = L.ny;
  if (i > nx + 1) return;
  if (i >= 1) {
    const bool wall_x = i < 2 || i > nx;
    ux[L.ux(i, 0)] = 0.;
    ux[L.ux(i, ny + 1)] = 0.;
    if (wall_x)
      for (int j = 1; j <= ny; j++) ux[L.ux(i, j)] = 0.;
  }
  uy[L.p(i, 0)] = 0.;
  uy[L.p(i, ny + 1)] = 0.;
  if (i < 1 || i > nx)
    for (int j = 1; j <= ny; j++) uy[L.p(i, j)] = 0.;
}

// ---------------------------------------------------------------------
// IBM: inertia (inertia.c:9-70) and exchange (exchange.c:13-88) adjoints,
// one workgroup per particle over the cells of its three image boxes
// ---------------------------------------------------------------------
__device__ double adj_foot(double a, double b, double x_, double y_) {
  // the reference's iteration (src/ellipse.c:30-65) in FP64
  const double px = fabs(x_), py = fabs(y_);
  double t = 0.25 * EIF_PI;
  for (int n = 0;; n++) {
    const double ct = cos(t), st = sin(t);
    const double x = a * ct, y = b * st;
    const double ex = a * (1. - (b * b) / (a * a)) * (ct * ct * ct);
    const double ey = b * (1. - (a * a) / (b * b)) * (st * st * st);
    const double rx = x - ex, ry = y - ey, qx = px - ex, qy = py - ey;
    const double r = sqrt(ry * ry + rx * rx), q = sqrt(qy * qy + qx * qx);
    if (!(r * q > 0.)) break;
    const double arg = fmin(1., fmax(-1., (rx * qy - ry * qx) / (r * q)));
    const double dc = r * asin(arg);
    const double dt = dc / sqrt(a * a + b * b - x * x - y * y);
    t += dt;
    t = fmin(0.5 * EIF_PI, t);
    t = fmax(0., t);
    if (fabs(dt) < 1.e-8 || n > 100) break;
  }
  if (x_ < 0.) t = EIF_PI - t;
  if (y_ < 0.) t = -t;
  return t;
}

struct IbmAdj {
  Layout L;
  int nx, ny, np;
  double ly, dx, dy, grid_size, dt;
  const double *xc, *yc;
  const double *prt;       // particle records (positions: x + dx, y + dy, az + daz)
  int with_dx;             // inertia(1): positions with the increments
  const double *ux, *uy;   // velocity field the forward read
  const double *seed;      // per particle 3 values: I bar or F bar
  const double *gx, *gy;   // exchange: cotangent of u** (gathered)
  double *sx, *sy;         // cotangent scatter target
  double *out;             // per particle 6 values: x, y, az, U, V, Omega bars (added)
  int exchange;            // 0: inertia, 1: exchange
};

// cells of one workgroup of k_adj_ibm (the three image boxes of a particle
// are cut into chunks of this many cells)
constexpr int ADJ_CHUNK = 256;

__device__ inline void adj_box(int lb, int ub, double h, double r, double c, int &lo, int &hi) {
  lo = (int)(floor((c - r) / h) - 3);
  hi = (int)(ceil((c + r) / h) + 3);
  lo = lo < lb ? lb : lo;
  hi = hi > ub ? ub : hi;
}

__global__ __launch_bounds__(256) void k_adj_ibm(IbmAdj a) {
  const int n = blockIdx.y;
  const long c0 = (long)blockIdx.x * ADJ_CHUNK;
  const Layout L = a.L;
  const double *q = a.prt + (long)n * R_N;
  const double pa = q[R_A], pb = q[R_B];
  const double px = q[R_X] + (a.with_dx ? q[R_DX] : 0.);
  const double py0 = q[R_Y] + (a.with_dx ? q[R_DY] : 0.);
  const double paz = q[R_AZ] + (a.with_dx ? q[R_DAZ] : 0.);
  const double m = q[R_DEN] * (EIF_PI * pa * pb);
  const double I = 0.25 * m * (pa * pa + pb * pb);
  const double U = q[R_UX], V = q[R_UY], W = q[R_VZ];
  const double s0 = a.seed[3 * n], s1 = a.seed[3 * n + 1], s2 = a.seed[3 * n + 2];
  const double dxdy = a.dx * a.dy, h = a.grid_size;
  const double cth = cos(-paz), sth = sin(-paz);
  double acc[6] = {0., 0., 0., 0., 0., 0.};
  const double r = fmax(pa, pb);
  int i0, i1;
  adj_box(1, a.nx, a.dx, r, px, i0, i1);
  const int ni = i1 - i0 + 1;
  long base = 0;  // first flattened cell of this image's box
  for (int per = -1; per <= 1; per++) {
    const double py = py0 + a.ly * per;
    int j0, j1;
    adj_box(1, a.ny, a.dy, r, py - 0., j0, j1);
    const int nj = j1 - j0 + 1;
    if (ni <= 0 || nj <= 0) continue;
    const long cnt = (long)ni * nj;
    const long lo = c0 > base ? c0 - base : 0, hi = c0 + ADJ_CHUNK - base;
    const long t1 = hi < cnt ? hi : cnt;
    base += cnt;
    for (long t = lo + threadIdx.x; t < t1; t += blockDim.x) {
      const int i = i0 + (int)(t % ni), j = j0 + (int)(t / ni);
      const double x = a.xc[i], y = a.yc[j - 1];
      // signed distance and its gradient with respect to the local point
      const double ddx = x - px, ddy = y - py;
      const double xl = cth * ddx - sth * ddy, yl = sth * ddx + cth * ddy;
      const double f0 = 1. - (xl / pa) * (xl / pa) - (yl / pb) * (yl / pb);
      const double sign = f0 > 0. ? 1. : -1.;
      const double tt = adj_foot(pa, pb, xl, yl);
      const double fx = pa * cos(tt), fy = pb * sin(tt);
      const double vx = xl - fx, vy = yl - fy;
      const double dr = sqrt(vx * vx + vy * vy);
      const double dist = sign * dr / h;
      const double th = tanh(BETA_W * dist);
      double gxl = 0., gyl = 0.;
      if (dr > 0.) {
        gxl = sign * vx / (dr * h);
        gyl = sign * vy / (dr * h);
      }
      double wb = 0.;  // cotangent of the weight
      double dwdd;     // d weight / d dist
      double w;
      if (!a.exchange) {
        // v_weight; iux += valx/m, iuy += valy/m, ivz += (-(y-py) valx + (x-px) valy)/I
        w = 0.5 * (1. + th);
        dwdd = 0.5 * BETA_W * (1. - th * th);
        const double ufx = 0.5 * (a.ux[L.ux(i, j)] + a.ux[L.ux(i + 1, j)]);
        const double ufy = 0.5 * (a.uy[L.p(i, j)] + a.uy[L.p(i, j + 1)]);
        const double vxb = s0 / m - s2 * (y - py) / I;
        const double vyb = s1 / m + s2 * (x - px) / I;
        const double valx = w * ufx * dxdy, valy = w * ufy * dxdy;
        acc[1] += s2 * valx / I;   // explicit lever arm d/dpy
        acc[0] += -s2 * valy / I;  // d/dpx
        wb = (vxb * ufx + vyb * ufy) * dxdy;
        const double fxb = vxb * w * dxdy * 0.5, fyb = vyb * w * dxdy * 0.5;
        add_ux(L, a.sx, i, j, fxb);
        add_ux(L, a.sx, i + 1, j, fxb);
        add_p(L, a.sy, i, j, fyb);
        add_p(L, a.sy, i, j + 1, fyb);
      } else {
        // s_weight; e = w (u_p - u_f) deposited (dux = fx dt), and
        // fux = -sum e_x dxdy/(m dt), fuy likewise,
        // tvz = -sum (-(y-py) e_x + (x-px) e_y) dxdy/(I dt)
        w = 0.5 * BETA_W * (1. - th * th);
        dwdd = -BETA_W * BETA_W * th * (1. - th * th);
        const double upx = U - W * (y - py), upy = V + W * (x - px);
        const double ufx = 0.5 * (a.ux[L.ux(i, j)] + a.ux[L.ux(i + 1, j)]);
        const double ufy = 0.5 * (a.uy[L.p(i, j)] + a.uy[L.p(i, j + 1)]);
        const double ex = w * (upx - ufx), ey = w * (upy - ufy);
        // cotangents of the deposited values: the face average of update.c:36-59
        double dgx = 0., dgy = 0.;
        if (i >= 2) dgx += 0.5 * a.gx[L.ux(i, j)];
        if (i + 1 <= a.nx) dgx += 0.5 * a.gx[L.ux(i + 1, j)];
        dgy = 0.5 * (a.gy[L.p(i, j)] + a.gy[L.p(i, j < a.ny ? j + 1 : 1)]);
        const double kf = dxdy / (m * a.dt), kt = dxdy / (I * a.dt);
        const double exb = dgx - s0 * kf + s2 * (y - py) * kt;
        const double eyb = dgy - s1 * kf - s2 * (x - px) * kt;
        acc[1] += -s2 * ex * kt;  // explicit lever arms of tvz
        acc[0] += s2 * ey * kt;
        wb = exb * (upx - ufx) + eyb * (upy - ufy);
        const double upxb = exb * w, upyb = eyb * w;
        acc[3] += upxb;
        acc[4] += upyb;
        acc[5] += -(y - py) * upxb + (x - px) * upyb;
        acc[1] += W * upxb;
        acc[0] += -W * upyb;
        const double fxb = -exb * w * 0.5, fyb = -eyb * w * 0.5;
        add_ux(L, a.sx, i, j, fxb);
        add_ux(L, a.sx, i + 1, j, fxb);
        add_p(L, a.sy, i, j, fyb);
        add_p(L, a.sy, i, j + 1, fyb);
      }
      (void)w;
      const double db = wb * dwdd;
      // x_ = c dx - s dy, y_ = s dx + c dy with c = cos(-az), s = sin(-az)
      acc[0] += db * (gxl * (-cth) + gyl * (-sth));
      acc[1] += db * (gxl * sth + gyl * (-cth));
      acc[2] += db * (gxl * yl - gyl * xl);
    }
  }
  __shared__ double red[6][256];
  for (int k = 0; k < 6; k++) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 6; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 6 && red[threadIdx.x][0] != 0.) atomicAdd(a.out + 6 * n + threadIdx.x, red[threadIdx.x][0]);
}

// ---------------------------------------------------------------------
// the particle chain of the sub-iteration loop, on the device (no host
// round trip per pass): per particle 3 values each
// ---------------------------------------------------------------------
struct Chain {
  int np;
  double gdt, dt;
  double *prt_b;       // np x 6: cotangents of x, y, az, ux, uy, vz (stage end -> start)
  double *force_b;     // np x 3: force seeds (consumed)
  double *X0b, *U0b, *dUb, *Fb, *I0b, *seed, *pout;
};

enum { CH_START, CH_PRE, CH_POST, CH_POST_LAST, CH_SEED_F, CH_ADD_F, CH_SEED_I0, CH_ADD_I0, CH_END };

__global__ void k_adj_chain(Chain c, int mode) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * c.np) return;
  const int n = t / 3, k = t % 3;
  switch (mode) {
    case CH_START: {  // update.c:8-34 and the last increment (increment.c:46-63)
      const double xb = c.prt_b[6 * n + k], ub = c.prt_b[6 * n + 3 + k];
      c.X0b[t] = xb;
      c.U0b[t] = ub + c.gdt * xb;
      c.dUb[t] = ub + 0.5 * c.gdt * xb;
      c.Fb[t] = c.force_b[t];
      c.force_b[t] = 0.;
      c.I0b[t] = 0.;
      break;
    }
    case CH_PRE:  // dU_s = dt F + I1_s - I0 (increment.c:23-38)
      c.Fb[t] += c.dt * c.dUb[t];
      c.I0b[t] -= c.dUb[t];
      c.seed[t] = c.dUb[t];
      c.pout[6 * n + k] = 0.;
      c.pout[6 * n + 3 + k] = 0.;
      break;
    case CH_POST:  // X1_s = X0 + dX_{s-1}, dX = g dt (U0 + dU / 2)
    case CH_POST_LAST: {
      const double x1b = c.pout[6 * n + k];
      c.X0b[t] += x1b;
      if (mode == CH_POST) {
        c.dUb[t] = 0.5 * c.gdt * x1b;
        c.U0b[t] += c.gdt * x1b;
      }
      break;
    }
    case CH_SEED_F:
    case CH_SEED_I0:
      c.seed[t] = mode == CH_SEED_F ? c.Fb[t] : c.I0b[t];
      c.pout[6 * n + k] = 0.;
      c.pout[6 * n + 3 + k] = 0.;
      break;
    case CH_ADD_F:
      c.X0b[t] += c.pout[6 * n + k];
      c.U0b[t] += c.pout[6 * n + 3 + k];
      break;
    case CH_ADD_I0:
      c.X0b[t] += c.pout[6 * n + k];
      break;
    case CH_END:
      c.prt_b[6 * n + k] = c.X0b[t];
      c.prt_b[6 * n + 3 + k] = c.U0b[t];
      break;
  }
}

// ---------------------------------------------------------------------
// tape and host orchestration
// ---------------------------------------------------------------------
struct StageRec {
  int rk = 0, S = 0;
  double dt = 0.;
  double *uk_x = nullptr, *uk_y = nullptr, *us_x = nullptr, *us_y = nullptr;
  double *u1_x = nullptr, *u1_y = nullptr;
  double *p0 = nullptr;         // particle records at the stage start
  std::vector<double *> ps;     // records before each sub-iteration's inertia(1)
};

struct AdjTape {
  std::vector<StageRec> recs;
  size_t used = 0;
  double *a_ux = nullptr, *a_uy = nullptr, *a_p = nullptr;   // cotangents of u, p
  double *pend_x = nullptr, *pend_y = nullptr;               // of the previous stage's u^k
  double *w_ux = nullptr, *w_uy = nullptr;                   // work
  double *a_re = nullptr;                                    // of 1/Re
  double *seed = nullptr, *pout = nullptr;                   // per particle 3 / 6
  double *prt_b = nullptr, *force_b = nullptr;               // particle cotangents, force seeds
  double *X0b = nullptr, *U0b = nullptr, *dUb = nullptr, *Fb = nullptr, *I0b = nullptr;
  int chunks = 1;  // cell chunks of the largest particle's three boxes
  int *contact = nullptr;  // set when a recorded stage met a contact
};

static int dmalloc0(double **p, size_t n, hipStream_t st) {
  EIF_HIP(hipMalloc(p, n * sizeof(double)));
  EIF_HIP(hipMemsetAsync(*p, 0, n * sizeof(double), st));
  return EIF_OK;
}

void tape_free(Ctx *c) {
  AdjTape *t = c->tape;
  if (!t) return;
  for (auto &r : t->recs) {
    double *bufs[] = {r.uk_x, r.uk_y, r.us_x, r.us_y, r.u1_x, r.u1_y, r.p0};
    for (auto b : bufs)
      if (b) (void)hipFree(b);
    for (auto b : r.ps)
      if (b) (void)hipFree(b);
  }
  double *bufs[] = {t->a_ux, t->a_uy, t->a_p, t->pend_x, t->pend_y, t->w_ux, t->w_uy,
                    t->a_re, t->seed, t->pout, t->prt_b, t->force_b, t->X0b, t->U0b,
                    t->dUb, t->Fb, t->I0b};
  for (auto b : bufs)
    if (b) (void)hipFree(b);
  if (t->contact) (void)hipFree(t->contact);
  delete t;
  c->tape = nullptr;
}

static int tape_zero(Ctx *c) {
  AdjTape *t = c->tape;
  const size_t nu = c->L.ux_len() * sizeof(double), npb = c->L.p_len() * sizeof(double);
  EIF_HIP(hipMemsetAsync(t->a_ux, 0, nu, c->stream));
  EIF_HIP(hipMemsetAsync(t->pend_x, 0, nu, c->stream));
  EIF_HIP(hipMemsetAsync(t->a_uy, 0, npb, c->stream));
  EIF_HIP(hipMemsetAsync(t->pend_y, 0, npb, c->stream));
  EIF_HIP(hipMemsetAsync(t->a_p, 0, npb, c->stream));
  EIF_HIP(hipMemsetAsync(t->a_re, 0, sizeof(double), c->stream));
  const size_t npp = (size_t)(c->np > 0 ? c->np : 1);
  EIF_HIP(hipMemsetAsync(t->prt_b, 0, 6 * npp * sizeof(double), c->stream));
  EIF_HIP(hipMemsetAsync(t->force_b, 0, 3 * npp * sizeof(double), c->stream));
  EIF_HIP(hipMemsetAsync(t->contact, 0, sizeof(int), c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  return EIF_OK;
}

int tape_begin(Ctx *c) {
  if (!c->tape) {
    AdjTape *t = new AdjTape();
    c->tape = t;
    const size_t nu = c->L.ux_len(), npl = c->L.p_len();
    const size_t npp = (size_t)(c->np > 0 ? c->np : 1);
    int rc;
    if ((rc = dmalloc0(&t->a_ux, nu, c->stream)) || (rc = dmalloc0(&t->a_uy, npl, c->stream)) ||
        (rc = dmalloc0(&t->a_p, npl, c->stream)) || (rc = dmalloc0(&t->pend_x, nu, c->stream)) ||
        (rc = dmalloc0(&t->pend_y, npl, c->stream)) || (rc = dmalloc0(&t->w_ux, nu, c->stream)) ||
        (rc = dmalloc0(&t->w_uy, npl, c->stream)) || (rc = dmalloc0(&t->a_re, 1, c->stream)) ||
        (rc = dmalloc0(&t->seed, 3 * npp, c->stream)) || (rc = dmalloc0(&t->pout, 6 * npp, c->stream)) ||
        (rc = dmalloc0(&t->prt_b, 6 * npp, c->stream)) ||
        (rc = dmalloc0(&t->force_b, 3 * npp, c->stream)) ||
        (rc = dmalloc0(&t->X0b, 3 * npp, c->stream)) || (rc = dmalloc0(&t->U0b, 3 * npp, c->stream)) ||
        (rc = dmalloc0(&t->dUb, 3 * npp, c->stream)) || (rc = dmalloc0(&t->Fb, 3 * npp, c->stream)) ||
        (rc = dmalloc0(&t->I0b, 3 * npp, c->stream)))
      return rc;
    EIF_HIP(hipMalloc(&t->contact, sizeof(int)));
    // the largest particle's boxes, in chunks of ADJ_CHUNK cells
    if (c->np > 0) {
      std::vector<double> rec((size_t)c->np * R_N);
      EIF_HIP(hipMemcpy(rec.data(), c->prt, rec.size() * sizeof(double), hipMemcpyDeviceToHost));
      long most = 1;
      for (int n = 0; n < c->np; n++) {
        const double r = fmax(rec[(size_t)n * R_N + R_A], rec[(size_t)n * R_N + R_B]);
        const long w = (long)ceil(2. * r / c->dx) + 9, hgt = (long)ceil(2. * r / c->dy) + 9;
        most = std::max(most, 3 * w * hgt);
      }
      t->chunks = (int)((most + ADJ_CHUNK - 1) / ADJ_CHUNK);
    }
  }
  c->tape->used = 0;
  return tape_zero(c);
}

static int copy_dev(Ctx *c, double **dst, const double *src, size_t n) {
  if (!*dst) EIF_HIP(hipMalloc(dst, n * sizeof(double)));
  EIF_HIP(hipMemcpyAsync(*dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  return EIF_OK;
}

// Stage records, taken by eif_gpu_integrate while recording: the copies
// are queued on the main stream where the data is current (u^k before the
// predictor, u* after it, u^{k+1} after the projection, the particles
// after the stage-start reset and before each sub-iteration).
int rec_stage_begin(Ctx *c, int rk) {
  AdjTape *t = c->tape;
  if (t->used == t->recs.size()) t->recs.emplace_back();
  StageRec &r = t->recs[t->used++];
  r.rk = rk;
  r.dt = c->hscal[S_DT];
  r.S = 0;
  return EIF_OK;
}

int rec_fields(Ctx *c, int which, const double *ux, const double *uy) {
  StageRec &r = c->tape->recs[c->tape->used - 1];
  const size_t nu = c->L.ux_len(), npl = c->L.p_len();
  double **dx = which == 0 ? &r.uk_x : which == 1 ? &r.us_x : &r.u1_x;
  double **dy = which == 0 ? &r.uk_y : which == 1 ? &r.us_y : &r.u1_y;
  TRY(copy_dev(c, dx, ux, nu));
  return copy_dev(c, dy, uy, npl);
}

int rec_p0(Ctx *c) {
  if (c->np == 0) return EIF_OK;
  StageRec &r = c->tape->recs[c->tape->used - 1];
  return copy_dev(c, &r.p0, c->prt, (size_t)c->np * R_N);
}

int rec_sub(Ctx *c, int s) {
  if (c->np == 0) return EIF_OK;
  StageRec &r = c->tape->recs[c->tape->used - 1];
  while ((int)r.ps.size() <= s) r.ps.push_back(nullptr);
  return copy_dev(c, &r.ps[s], c->prt, (size_t)c->np * R_N);
}

// any colliding pair or wall contact of the last collision call
__global__ void k_adj_contacts(const int *row_cnt, const int *wall_hit, int np, int *flag) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < np && (row_cnt[n] || wall_hit[2 * n] || wall_hit[2 * n + 1])) atomicOr(flag, 1);
}

int rec_stage_end(Ctx *c, int S) {
  c->tape->recs[c->tape->used - 1].S = S;
  if (c->np == 0) return EIF_OK;
  hipLaunchKernelGGL(k_adj_contacts, dim3((c->np + 255) / 256), dim3(256), 0, c->stream,
                     c->row_cnt, c->wall_hit, c->np, c->tape->contact);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// collisions are forward-only: a recorded run must be contact-free
int rec_check_contacts(Ctx *c) {
  int h = 0;
  EIF_HIP(hipMemcpyAsync(&h, c->tape->contact, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  if (h) {
    set_error("adjoint: a contact occurred in a recorded step (collisions are forward-only)");
    return EIF_ERR_STATE;
  }
  return EIF_OK;
}

static IbmAdj ibm_adj(Ctx *c, const double *prt, int with_dx, const double *ux,
                      const double *uy, int exchange, double dt) {
  IbmAdj a;
  a.L = c->L;
  a.nx = c->nx;
  a.ny = c->ny;
  a.np = c->np;
  a.ly = c->ly;
  a.dx = c->dx;
  a.dy = c->dy;
  a.grid_size = fmin(c->dx, c->dy);
  a.dt = dt;
  a.xc = c->xc;
  a.yc = c->yc;
  a.prt = prt;
  a.with_dx = with_dx;
  a.ux = ux;
  a.uy = uy;
  a.seed = c->tape->seed;
  a.gx = a.gy = nullptr;
  a.sx = a.sy = nullptr;
  a.out = c->tape->pout;
  a.exchange = exchange;
  return a;
}

// one IBM adjoint pass: seeds and outputs in the tape's per-particle arrays
static int ibm_adj_run(Ctx *c, const IbmAdj &a) {
  hipLaunchKernelGGL(k_adj_ibm, dim3(c->tape->chunks, c->np), dim3(256), 0, c->stream, a);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

static int chain(Ctx *c, int mode, double gdt, double dt) {
  AdjTape *t = c->tape;
  Chain ch{c->np, gdt, dt, t->prt_b, t->force_b, t->X0b, t->U0b, t->dUb, t->Fb, t->I0b,
           t->seed, t->pout};
  hipLaunchKernelGGL(k_adj_chain, dim3((3 * c->np + 255) / 256), dim3(256), 0, c->stream, ch,
                     mode);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

int stage_adjoint(Ctx *c) {
  AdjTape *t = c->tape;
  if (!t || t->used == 0) {
    set_error("adjoint: no recorded stage left");
    return EIF_ERR_STATE;
  }
  StageRec &r = t->recs[t->used - 1];
  const StageRec *prev = (r.rk > 0 && t->used >= 2) ? &t->recs[t->used - 2] : nullptr;
  const int rk = r.rk, np = c->np;
  const double dt = r.dt, ga = c->gamma[rk];
  // the projection kernels read dt on the device: the recorded step's
  c->hscal[S_DT] = dt;
  EIF_HIP(hipMemcpyAsync(c->dscal + S_DT, c->hscal + S_DT, sizeof(double),
                         hipMemcpyHostToDevice, c->stream));
  const Layout &L = c->L;
  const long nu = L.ux_len(), npl = L.p_len();
  // ---- the sub-iteration loop and the particle update
  if (np > 0) {
    TRY(chain(c, CH_START, ga * dt, dt));
    for (int s = r.S - 1; s >= 0; s--) {
      TRY(chain(c, CH_PRE, ga * dt, dt));
      IbmAdj a = ibm_adj(c, r.ps[s], 1, r.u1_x, r.u1_y, 0, dt);
      a.sx = t->a_ux;
      a.sy = t->a_uy;
      TRY(ibm_adj_run(c, a));
      TRY(chain(c, s > 0 ? CH_POST : CH_POST_LAST, ga * dt, dt));
    }
  }
  // ---- projection and pressure update: u** bar = P(u bar) - G L^-1 p bar / (g dt)
  hipLaunchKernelGGL(k_adj_halo, dim3((c->nx + 2 + 255) / 256), dim3(256), 0, c->stream, L,
                     t->a_ux, t->a_uy);
  {
    double *sx = c->ux, *sy = c->uy;
    c->ux = t->a_ux;
    c->uy = t->a_uy;
    int rc = launch_poisson(c, rk, t->a_p);
    if (!rc) rc = launch_correct_pressure(c, rk, true, false);
    c->ux = sx;
    c->uy = sy;
    TRY(rc);
  }
  // ---- exchange and deposit: u* bar = u** bar + (field terms), into w
  EIF_HIP(hipMemcpyAsync(t->w_ux, t->a_ux, nu * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  EIF_HIP(hipMemcpyAsync(t->w_uy, t->a_uy, npl * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (np > 0) {
    TRY(chain(c, CH_SEED_F, ga * dt, dt));
    IbmAdj a = ibm_adj(c, r.p0, 0, r.us_x, r.us_y, 1, dt);
    a.gx = t->a_ux;
    a.gy = t->a_uy;
    a.sx = t->w_ux;
    a.sy = t->w_uy;
    TRY(ibm_adj_run(c, a));
    TRY(chain(c, CH_ADD_F, ga * dt, dt));
  }
  // ---- predictor: u^k bar = u* bar (+ the next stage's beta-term
  // cotangent, pending) + transposed stencils; the beta term of this stage
  // becomes the pending cotangent of the previous stage's u^k
  hipLaunchKernelGGL(k_adj_sum_clear, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0,
                     c->stream, t->a_ux, t->w_ux, t->pend_x, nu);
  hipLaunchKernelGGL(k_adj_sum_clear, dim3((unsigned)((npl + 255) / 256)), dim3(256), 0,
                     c->stream, t->a_uy, t->w_uy, t->pend_y, npl);
  {
    PredAdj a;
    a.L = L;
    a.sx = t->w_ux;
    a.sy = t->w_uy;
    a.ux = r.uk_x;
    a.uy = r.uk_y;
    const bool with_b = prev && c->beta[rk] != 0.;
    a.uxo = with_b ? prev->uk_x : nullptr;
    a.uyo = with_b ? prev->uk_y : nullptr;
    a.ox = t->a_ux;
    a.oy = t->a_uy;
    a.oxo = t->pend_x;
    a.oyo = t->pend_y;
    a.op = t->a_p;
    a.ore = t->a_re;
    a.dxf = c->dxf;
    a.dxc = c->dxc;
    a.dy = c->dy;
    a.rre = 1. / c->cfg.Re;
    a.aldt = c->alpha[rk] * dt;
    a.bedt = c->beta[rk] * dt;
    a.gadt = ga * dt;
    hipLaunchKernelGGL(k_adj_predictor, dim3((c->nx + 255) / 256, c->ny), dim3(256), 0,
                       c->stream, a);
    EIF_HIP(hipGetLastError());
  }
  // ---- inertia(0) at the stage start
  if (np > 0) {
    TRY(chain(c, CH_SEED_I0, ga * dt, dt));
    IbmAdj a = ibm_adj(c, r.p0, 0, r.uk_x, r.uk_y, 0, dt);
    a.sx = t->a_ux;
    a.sy = t->a_uy;
    TRY(ibm_adj_run(c, a));
    TRY(chain(c, CH_ADD_I0, ga * dt, dt));
    TRY(chain(c, CH_END, ga * dt, dt));
  }
  // the stage input's cotangent on the unknowns only
  hipLaunchKernelGGL(k_adj_clean, dim3((c->nx + 2 + 255) / 256), dim3(256), 0, c->stream, L,
                     t->a_ux, t->a_uy);
  EIF_HIP(hipGetLastError());
  t->used--;
  return EIF_OK;
}

// d/d uy of weight * fluid_gpu_wall_shear, added to the uy cotangent
__global__ void k_adj_shear_seed(Layout L, double *ay, const double *dxc, double dy, double rre,
                                 double w) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (j > L.ny) return;
  ay[L.p(1, j)] += w * rre * dy / dxc[0];
  ay[L.p(L.nx, j)] += w * rre * dy / dxc[L.nx];
}

int adjoint_add_wall_shear(Ctx *c, double weight) {
  if (!c->tape) {
    set_error("no tape: call eif_gpu_tape_begin first");
    return EIF_ERR_STATE;
  }
  hipLaunchKernelGGL(k_adj_shear_seed, dim3((c->ny + 255) / 256), dim3(256), 0, c->stream, c->L,
                     c->tape->a_uy, c->dxc, c->dy, 1. / c->cfg.Re, weight);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

int adjoint_access(Ctx *c, int which, double *host, bool add) {
  AdjTape *t = c->tape;
  if (!t) {
    set_error("no tape: call eif_gpu_tape_begin first");
    return EIF_ERR_STATE;
  }
  const size_t nu = c->L.ux_len(), npl = c->L.p_len(), np = (size_t)c->np;
  auto dev = [&](double *d, size_t n) -> int {
    std::vector<double> tmp(n);
    EIF_HIP(hipMemcpyAsync(tmp.data(), d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EIF_HIP(hipStreamSynchronize(c->stream));
    if (add) {
      for (size_t k = 0; k < n; k++) tmp[k] += host[k];
      EIF_HIP(hipMemcpyAsync(d, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      EIF_HIP(hipStreamSynchronize(c->stream));
    } else {
      memcpy(host, tmp.data(), n * sizeof(double));
    }
    return EIF_OK;
  };
  switch (which) {
    case EIF_ADJ_UX: return dev(t->a_ux, nu);
    case EIF_ADJ_UY: return dev(t->a_uy, npl);
    case EIF_ADJ_P: return dev(t->a_p, npl);
    case EIF_ADJ_PARTICLES: return np ? dev(t->prt_b, 6 * np) : EIF_OK;
    case EIF_ADJ_FORCE: return np ? dev(t->force_b, 3 * np) : EIF_OK;
    case EIF_ADJ_RE: {
      double v = 0.;
      EIF_HIP(hipMemcpyAsync(&v, t->a_re, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      EIF_HIP(hipStreamSynchronize(c->stream));
      const double re = c->cfg.Re;
      if (add) {
        // d/dRe = -1/Re^2 d/d(1/Re)
        v += host[0] * (-re * re);
        EIF_HIP(hipMemcpyAsync(t->a_re, &v, sizeof(double), hipMemcpyHostToDevice, c->stream));
        EIF_HIP(hipStreamSynchronize(c->stream));
      } else {
        host[0] = -v / (re * re);
      }
      return EIF_OK;
    }
  }
  set_error("unknown adjoint selector");
  return EIF_ERR_ARG;
}

}  // namespace eif
