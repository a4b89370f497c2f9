// eif_poisson.hip -- pressure Poisson solve of EllipsesInFlows on gfx950
// (src/fluid/compute_potential.c:48-138, src/tdm.c:37-131).
//
// Data layout: the spectral field qhat[j][k] keeps x-wavenumbers k
// contiguous (the reference's x-aligned QX layout, compute_potential.c:45),
// so both DCT passes read and write whole rows, and the y-direction solve
// runs one lane per wavenumber with loads coalesced across lanes.  The
// reference's two MPI transposes (:94, :118) have no counterpart.
//
// DCT-II / DCT-III (FFTW REDFT10 / REDFT01): Makhoul's n-point complex FFT
// (radix-2, decimation in time) in LDS for powers of two, the direct sum
// otherwise.  Every trigonometric factor comes from the table
// C[m] = cos(pi m/(2n)) built on the host.
//
// y-direction periodic tridiagonal solve, two modes:
//   EIF_POISSON_THOMAS     the bordering method of src/tdm.c:64-108 with its
//                          Thomas sweeps (:37-59); the factors that depend
//                          only on the matrix (1/pivot, modified upper
//                          diagonal, border solution) are tabulated once per
//                          context with the reference's arithmetic, so each
//                          solve performs exactly the reference's operations
//                          on the data.
//   EIF_POISSON_PARTITION  the same linear system split into P chunks per
//                          wavenumber (spike method): chunk-local Thomas
//                          solves in parallel, a reduced periodic system for
//                          the chunk-boundary values, then a fix-up fused
//                          into the DCT-III pass.
#include "eif_internal.hpp"
#include <cmath>
#include <cfloat>

namespace eif {

constexpr int DCT_THREADS = 256;

__device__ inline int bitrev(int v, int logn) {
  return (int)(__brev((unsigned)v) >> (32 - logn));
}

// radix-2 DIT on LDS arrays (input already in bit-reversed order),
// twiddles from C[], same butterfly arithmetic as the host restatement.
__device__ void lds_fft(double *re, double *im, int n, int logn,
                        const double *__restrict__ C, double sign) {
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1;
    const int stride = n / len;
    for (int b = threadIdx.x; b < n / 2; b += blockDim.x) {
      const int k = b & (half - 1);
      const int blk = b / half;
      const int i0 = blk * len + k, i1 = i0 + half;
      const int m = k * stride;
      const double wr = C[4 * m];
      const double wi = sign * C[n >= 4 * m ? n - 4 * m : 4 * m - n];
      const double r1 = re[i1], m1 = im[i1];
      const double tr = wr * r1 - wi * m1;
      const double ti = wr * m1 + wi * r1;
      const double r0 = re[i0], m0 = im[i0];
      re[i1] = r0 - tr;
      im[i1] = m0 - ti;
      re[i0] = r0 + tr;
      im[i0] = m0 + ti;
    }
    __syncthreads();
  }
}

__device__ inline double cos_lookup(const double *C, int n, long m) {
  m %= 4 * (long)n;
  return m <= 2 * n ? C[m] : C[4 * (long)n - m];
}

// ---------------------------------------------------------------------
// RHS (compute_potential.c:70-82) + DCT-II along x (:84-92), one row per
// workgroup.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(DCT_THREADS) void k_rhs_dct2(
    Layout L, const double *ux, const double *uy, const double *dscal, double ga,
    double dx, double dy, const double *__restrict__ C, int logn, double *qhat) {
  extern __shared__ double sh[];
  const int nx = L.nx;
  const int j = blockIdx.x + 1;
  double *re = sh, *im = sh + nx;
  const double dt = dscal[S_DT];
  const double fac = 1. / (ga * dt);
  if (logn >= 0) {
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const double q = fac * (+(ux[L.ux(i + 2, j)] - ux[L.ux(i + 1, j)]) / dx
                              + (uy[L.p(i + 1, j + 1)] - uy[L.p(i + 1, j)]) / dy);
      const int m = (i & 1) ? nx - 1 - (i >> 1) : (i >> 1);  // Makhoul order
      const int pos = bitrev(m, logn);
      re[pos] = q;
      im[pos] = 0.;
    }
    __syncthreads();
    lds_fft(re, im, nx, logn, C, -1.);
    double *out = qhat + (long)(j - 1) * nx;
    for (int k = threadIdx.x; k < nx; k += blockDim.x)
      out[k] = 2. * (C[k] * re[k] + C[nx - k] * im[k]);
  } else {
    for (int i = threadIdx.x; i < nx; i += blockDim.x)
      re[i] = fac * (+(ux[L.ux(i + 2, j)] - ux[L.ux(i + 1, j)]) / dx
                     + (uy[L.p(i + 1, j + 1)] - uy[L.p(i + 1, j)]) / dy);
    __syncthreads();
    double *out = qhat + (long)(j - 1) * nx;
    for (int k = threadIdx.x; k < nx; k += blockDim.x) {
      double acc = 0.;
      for (int jj = 0; jj < nx; jj++)
        acc = acc + 2. * re[jj] * cos_lookup(C, nx, (long)(2 * jj + 1) * k);
      out[k] = acc;
    }
  }
}

// ---------------------------------------------------------------------
// DCT-III along x (:120-128) + normalisation into psi (:130-134); the
// y-solve's final combination (tdm.c:104-106, or the spike fix-up) is
// applied as the row is loaded.
// ---------------------------------------------------------------------
struct CombineArgs {
  int mode;  // 0: Thomas border combine, 1: partition fix-up
  const double *q1;     // Thomas: [ny-1][nx]
  const double *last;   // Thomas: [nx]
  // partition
  int L, P;
  const double *v, *w;  // [L][nx] spikes
  const double *bnd;    // [P][2][nx] chunk-boundary neighbour values
  double coef;          // l (= u) of the matrix
};

__device__ inline double combined(const CombineArgs &a, const double *qhat,
                                  int nx, int ny, int j0, int k) {
  const double q = qhat[(long)j0 * nx + k];
  if (a.mode == 0) {
    if (j0 == ny - 1) return a.last[k];
    return q + a.last[k] * a.q1[(long)j0 * nx + k];
  }
  const int c = j0 / a.L, r = j0 - c * a.L;
  const double xl = a.bnd[((long)c * 2 + 0) * nx + k];
  const double xr = a.bnd[((long)c * 2 + 1) * nx + k];
  return q - a.coef * xl * a.v[(long)r * nx + k] - a.coef * xr * a.w[(long)r * nx + k];
}

__global__ __launch_bounds__(DCT_THREADS) void k_dct3_psi(
    Layout L, const double *qhat, CombineArgs ca, const double *__restrict__ C,
    int logn, double *psi) {
  extern __shared__ double sh[];
  const int nx = L.nx, ny = L.ny;
  const int j0 = blockIdx.x;
  double *re = sh, *im = sh + nx;
  const double norm = 2. * nx;
  if (logn >= 0) {
    for (int k = threadIdx.x; k < nx; k += blockDim.x) {
      const double xr = combined(ca, qhat, nx, ny, j0, k);
      const double xi = k == 0 ? 0. : -combined(ca, qhat, nx, ny, j0, nx - k);
      const int pos = bitrev(k, logn);
      re[pos] = C[k] * xr - C[nx - k] * xi;
      im[pos] = C[k] * xi + C[nx - k] * xr;
    }
    __syncthreads();
    lds_fft(re, im, nx, logn, C, +1.);
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const double v = (i & 1) ? re[nx - 1 - (i >> 1)] : re[i >> 1];
      psi[L.p(i + 1, j0 + 1)] = v / norm;
    }
  } else {
    for (int k = threadIdx.x; k < nx; k += blockDim.x)
      re[k] = combined(ca, qhat, nx, ny, j0, k);
    __syncthreads();
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      double acc = re[0];
      for (int jj = 1; jj < nx; jj++)
        acc = acc + 2. * re[jj] * cos_lookup(C, nx, (long)jj * (2 * i + 1));
      psi[L.p(i + 1, j0 + 1)] = acc / norm;
    }
  }
}

// ---------------------------------------------------------------------
// Thomas mode: one lane per wavenumber k, sweeps along j
// ---------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_thomas(int nx, int ny, double *q,
                                               const double *__restrict__ rv,
                                               const double *__restrict__ um,
                                               const unsigned char *__restrict__ flag,
                                               const double *__restrict__ cdiag,
                                               const double *__restrict__ den,
                                               double coef, double *last) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nx) return;
  const int n1 = ny - 1;  // size of the bordered (n-1) system
  const double l = coef;  // l = u = 1/dy/dy everywhere
  // forward sweep (tdm.c:40-53) on rows 0..n-2
  double prev = q[k] / cdiag[k];
  q[k] = prev;
  for (int i = 1; i <= n1 - 1; i++) {
    const long o = (long)i * nx + k;
    const double qi = q[o];
    double v;
    if (flag[o]) {
      v = 0.;
    } else {
      v = rv[o] * (qi - l * prev);
    }
    q[o] = v;
    prev = v;
  }
  // backward sweep (:55-57)
  const double q_last_fwd = prev;  // q0[n-2]
  double next = prev;
  for (int i = n1 - 2; i >= 0; i--) {
    const long o = (long)i * nx + k;
    const double v = q[o] - um[o] * next;
    q[o] = v;
    next = v;
  }
  // bottom solution (tdm.c:99-102)
  const double q0_0 = next;
  const double num = q[(long)(ny - 1) * nx + k] - l * q0_0 - l * q_last_fwd;
  const double d = den[k];
  const double xl = fabs(d) < DBL_EPSILON ? 0. : num / d;
  last[k] = xl;
}

int poisson_init(Ctx *c) {
  const int nx = c->nx, ny = c->ny;
  // cosine table, same expression as the host restatement
  std::vector<double> C(2 * (size_t)nx + 1);
  for (int m = 0; m <= 2 * nx; m++) C[m] = cos(EIF_PI * m / (2. * nx));
  EIF_HIP(hipMalloc(&c->ctab, C.size() * sizeof(double)));
  EIF_HIP(hipMemcpy(c->ctab, C.data(), C.size() * sizeof(double), hipMemcpyHostToDevice));
  EIF_HIP(hipMalloc(&c->qhat, (size_t)nx * ny * sizeof(double)));
  const double dx = c->dx, dy = c->dy;
  const double l = 1. / dy / dy;  // compute_potential.c:108-111
  const int n1 = ny - 1;
  std::vector<double> cdiag(nx), den(nx);
  std::vector<double> rv((size_t)n1 * nx), um((size_t)n1 * nx), q1((size_t)n1 * nx);
  std::vector<unsigned char> flg((size_t)n1 * nx);
  std::vector<double> uu(n1), qq(n1);
  for (int k = 0; k < nx; k++) {
    // eigenvalue, compute_potential.c:103-106 (pow(., 2.) is a product)
    const double sn = sin(EIF_PI * k / (2. * nx));
    const double eig = -4. / (dx * dx) * (sn * sn);
    const double cc = -l - l + eig;
    cdiag[k] = cc;
    // Thomas on the (n-1) system of the border vector (tdm.c:86-98)
    for (int i = 0; i < n1; i++) {
      uu[i] = l;
      qq[i] = i == 0 ? -l : (i == n1 - 1 ? -l : 0.);
    }
    uu[0] = uu[0] / cc;
    qq[0] = qq[0] / cc;
    rv[k] = 0.;
    um[k] = uu[0];
    flg[k] = 0;
    for (int i = 1; i <= n1 - 1; i++) {
      double val = cc - l * uu[i - 1];
      const size_t o = (size_t)i * nx + k;
      if (fabs(val) < DBL_EPSILON) {
        uu[i] = 0.;
        qq[i] = 0.;
        flg[o] = 1;
        rv[o] = 0.;
      } else {
        val = 1. / val;
        uu[i] = val * uu[i];
        qq[i] = val * (qq[i] - l * qq[i - 1]);
        flg[o] = 0;
        rv[o] = val;
      }
      um[o] = uu[i];
    }
    for (int i = n1 - 2; i >= 0; i--) qq[i] -= uu[i] * qq[i + 1];
    for (int i = 0; i < n1; i++) q1[(size_t)i * nx + k] = qq[i];
    den[k] = cc + l * qq[0] + l * qq[n1 - 1];
  }
  EIF_HIP(hipMalloc(&c->th_val, rv.size() * sizeof(double)));
  EIF_HIP(hipMalloc(&c->th_u, um.size() * sizeof(double)));
  EIF_HIP(hipMalloc(&c->th_q1, q1.size() * sizeof(double)));
  EIF_HIP(hipMalloc(&c->th_flag, flg.size()));
  EIF_HIP(hipMalloc(&c->th_den, 2 * (size_t)nx * sizeof(double)));
  EIF_HIP(hipMalloc(&c->th_last, (size_t)nx * sizeof(double)));
  EIF_HIP(hipMemcpy(c->th_val, rv.data(), rv.size() * sizeof(double), hipMemcpyHostToDevice));
  EIF_HIP(hipMemcpy(c->th_u, um.data(), um.size() * sizeof(double), hipMemcpyHostToDevice));
  EIF_HIP(hipMemcpy(c->th_q1, q1.data(), q1.size() * sizeof(double), hipMemcpyHostToDevice));
  EIF_HIP(hipMemcpy(c->th_flag, flg.data(), flg.size(), hipMemcpyHostToDevice));
  EIF_HIP(hipMemcpy(c->th_den, den.data(), nx * sizeof(double), hipMemcpyHostToDevice));
  EIF_HIP(hipMemcpy(c->th_den + nx, cdiag.data(), nx * sizeof(double), hipMemcpyHostToDevice));
  return EIF_OK;
}

static int ilog2_exact(int n) {
  if (n < 2 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}

int launch_poisson(Ctx *c, int rk) {
  const int nx = c->nx, ny = c->ny;
  const int logn = ilog2_exact(nx);
  const size_t shmem = 2 * (size_t)nx * sizeof(double);
  {
    ProfScope ps(c, "rhs_dct2");
    hipLaunchKernelGGL(k_rhs_dct2, dim3(ny), dim3(DCT_THREADS), shmem, c->stream, c->L,
                       c->ux, c->uy, c->dscal, c->gamma[rk], c->dx, c->dy, c->ctab,
                       logn, c->qhat);
    EIF_HIP(hipGetLastError());
  }
  const double l = 1. / c->dy / c->dy;
  CombineArgs ca{};
  ca.coef = l;
  {
    ProfScope ps(c, "thomas");
    hipLaunchKernelGGL(k_thomas, dim3((nx + 63) / 64), dim3(64), 0, c->stream, nx, ny,
                       c->qhat, c->th_val, c->th_u, c->th_flag, c->th_den + nx,
                       c->th_den, l, c->th_last);
    EIF_HIP(hipGetLastError());
    ca.mode = 0;
    ca.q1 = c->th_q1;
    ca.last = c->th_last;
  }
  {
    ProfScope ps(c, "dct3_psi");
    hipLaunchKernelGGL(k_dct3_psi, dim3(ny), dim3(DCT_THREADS), shmem, c->stream, c->L,
                       c->qhat, ca, c->ctab, logn, c->psi);
    EIF_HIP(hipGetLastError());
  }
  return launch_bc_p(c, c->psi);
}

}  // namespace eif
