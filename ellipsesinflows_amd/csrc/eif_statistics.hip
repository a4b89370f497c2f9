// eif_statistics.hip -- time-averaged statistics on the device
// (src/statistics/collect.c:10-95, src/statistics/init.c:8-44).
//
// The reference adds ux, ux^2, uy, uy^2 and a particle indicator phi to
// five host accumulators at every sample.  Here the accumulators stay in
// HBM for the whole run; a sample is one streaming pass over ux and uy
// (HBM-bound, 2 reads + 2 read-modify-writes per point) plus a box-local
// pass per particle image for phi.  The reference evaluates the indicator
// over the whole grid for every particle and image (collect.c:62-81);
// every point inside an ellipse lies within max(a,b) of its centre, so the
// box around the centre used here visits every point that can count.
//
// Sums per point are in sample order, as in the reference, so ux1, ux2,
// uy1, uy2 are bit-identical.  phi counts are integers held in doubles and
// are exact in any order; overlapping images add with atomics.
#include "eif_internal.hpp"

namespace eif {

// collect.c:10-46: UX1 += UX, UX2 += UX^2 for i = 1..nx+1, and
// UY1 += UY, UY2 += UY^2 for i = 0..nx+1, rows j = 1..ny.  pow(v, 1.) is v
// and pow(v, 2.) is the correctly rounded product v*v.
__global__ __launch_bounds__(256) void k_stat_fluid(Layout L, const double *ux,
                                                    const double *uy, double *ux1,
                                                    double *ux2, double *uy1,
                                                    double *uy2) {
  const int nx = L.nx, ny = L.ny;
  const long nux = (long)(nx + 1) * ny;
  const long nuy = (long)(nx + 2) * ny;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nux + nuy; t += stride) {
    if (t < nux) {
      const long k = t + (nx + 1);  // skip the halo row j = 0
      const double v = ux[k];
      ux1[k] += v;
      ux2[k] += v * v;
    } else {
      const long k = (t - nux) + (nx + 2);
      const double v = uy[k];
      uy1[k] += v;
      uy2[k] += v * v;
    }
  }
}

// collect.c:48-83: one workgroup per (particle, image).  The test of a
// point (XC(i), YC(j)) is the reference's, operand for operand:
//   x_ = c dx - s dy, y_ = s dx + c dy with c = cos(-az), s = sin(-az),
//   inside  <=>  1 - (x_/a)^2 - (y_/b)^2 > 0.
__global__ __launch_bounds__(256) void k_stat_phi(Layout L, const double *prt,
                                                  const double *xc, const double *yc,
                                                  double dx, double dy, double ly,
                                                  double *phi) {
  const int n = blockIdx.x / 3;
  const int periodic = (int)(blockIdx.x % 3) - 1;
  const double *q = prt + (long)n * R_N;
  const double pa = q[R_A], pb = q[R_B], px = q[R_X];
  const double py = q[R_Y] + ly * periodic;
  const double paz = q[R_AZ];
  const double r = fmax(pa, pb);
  const int nx = L.nx, ny = L.ny;
  // cell centres XC(i) = (i - 1/2) dx; one extra cell of margin each side
  int imin = (int)floor((px - r) / dx) - 1, imax = (int)ceil((px + r) / dx) + 2;
  int jmin = (int)floor((py - r) / dy) - 1, jmax = (int)ceil((py + r) / dy) + 2;
  imin = imin < 1 ? 1 : imin;
  imax = imax > nx ? nx : imax;
  jmin = jmin < 1 ? 1 : jmin;
  jmax = jmax > ny ? ny : jmax;
  if (imin > imax || jmin > jmax) return;
  const int w = imax - imin + 1;
  const long cells = (long)w * (jmax - jmin + 1);
  const double c = cos(-paz);
  const double s = sin(-paz);
  for (long t = threadIdx.x; t < cells; t += blockDim.x) {
    const int i = imin + (int)(t % w), j = jmin + (int)(t / w);
    const double ddx = xc[i] - px;
    const double ddy = yc[j - 1] - py;
    const double x_ = c * ddx - s * ddy;
    const double y_ = s * ddx + c * ddy;
    const double ex = x_ / pa, ey = y_ / pb;
    if (1. - ex * ex - ey * ey > 0.) atomicAdd(&phi[L.p(i, j)], 1.);
  }
}

int stat_alloc(Ctx *c) {
  if (c->st_ux1) return EIF_OK;
  const Layout &L = c->L;
  double **ux_like[] = {&c->st_ux1, &c->st_ux2};
  double **p_like[] = {&c->st_uy1, &c->st_uy2, &c->st_phi};
  for (auto b : ux_like) {
    EIF_HIP(hipMalloc(b, L.ux_len() * sizeof(double)));
    EIF_HIP(hipMemsetAsync(*b, 0, L.ux_len() * sizeof(double), c->stream));
  }
  for (auto b : p_like) {
    EIF_HIP(hipMalloc(b, L.p_len() * sizeof(double)));
    EIF_HIP(hipMemsetAsync(*b, 0, L.p_len() * sizeof(double), c->stream));
  }
  c->st_num = 0;
  return EIF_OK;
}

int stat_reset(Ctx *c) {
  if (!c->st_ux1) return EIF_OK;
  const Layout &L = c->L;
  for (double *b : {c->st_ux1, c->st_ux2})
    EIF_HIP(hipMemsetAsync(b, 0, L.ux_len() * sizeof(double), c->stream));
  for (double *b : {c->st_uy1, c->st_uy2, c->st_phi})
    EIF_HIP(hipMemsetAsync(b, 0, L.p_len() * sizeof(double), c->stream));
  c->st_num = 0;
  return EIF_OK;
}

int launch_statistics(Ctx *c) {
  int rc = stat_alloc(c);
  if (rc) return rc;
  const Layout &L = c->L;
  const long work = (long)(2 * L.nx + 3) * L.ny;
  long nb = (work + 255) / 256;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(k_stat_fluid, dim3((unsigned)nb), dim3(256), 0, c->stream, L, c->ux,
                     c->uy, c->st_ux1, c->st_ux2, c->st_uy1, c->st_uy2);
  EIF_HIP(hipGetLastError());
  if (c->np > 0) {
    hipLaunchKernelGGL(k_stat_phi, dim3(3 * c->np), dim3(256), 0, c->stream, L, c->prt,
                       c->xc, c->yc, c->dx, c->dy, c->ly, c->st_phi);
    EIF_HIP(hipGetLastError());
  }
  c->st_num += 1;
  return EIF_OK;
}

}  // namespace eif
