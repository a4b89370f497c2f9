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
// (radix-2, decimation in time) for powers of two, the direct sum
// otherwise.  For 64 <= n <= 8192 the FFT runs in registers and two rows
// share one complex transform (real and imaginary part); other powers of
// two use one row per transform in LDS.  Every trigonometric factor comes
// from the table C[m] = cos(pi m/(2n)) built on the host.
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

#define TRY_P(x)          \
  do {                    \
    const int _r = (x);   \
    if (_r) return _r;    \
  } while (0)

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

// Register form of the same radix-2 DIT: n/2^LOGR threads each hold 2^LOGR
// complex values; stages are taken LOGR at a time (groups of positions
// closed under those stages), with LDS used only to regroup between phases.
// Every butterfly sees the same inputs and performs the same operations as
// lds_fft, so the results are identical.
struct RegMap {
  int lg, lh0, lgpt;  // log2 of group size, first butterfly distance, groups/thread
  __device__ int pos(int t, int slot) const {
    const int G = (t << lgpt) + (slot >> lg);
    return ((G >> lh0) << (lg + lh0)) + (G & ((1 << lh0) - 1)) + ((slot & ((1 << lg) - 1)) << lh0);
  }
};

// the position map of phase p (stages LOGR p .. LOGR p + lg - 1)
template <int LOGR>
__device__ inline RegMap reg_phase_r(int p, int logn) {
  RegMap m;
  m.lg = min(LOGR, logn - LOGR * p);
  m.lh0 = LOGR * p;
  m.lgpt = LOGR - m.lg;
  return m;
}

// LDS address of element p: XOR-swizzles the 16-double bank group so that
// the bit-reversed, strided and natural access patterns of the transform
// hit at most two-way bank conflicts
__device__ inline int sw(int p) { return p ^ ((p >> 4) & 15) ^ ((p >> 8) & 15); }

// The transform with n = 2^LOGN known at compile time: phases and stages
// unroll, so the slot -> position maps and twiddle offsets fold to constants
// plus a per-thread base.  W holds the per-stage twiddles, stage len in
// entries [len/2-1, len-1): W[2e] = C[4m], W[2e+1] = C[|n-4m|] with
// m = k n/len (the table values lds_fft reads from C).
// the twiddles of phase p's butterflies (entries of butterflies that do not
// exist are left unset and never read)
template <int LOGN, int LOGR>
__device__ inline void phase_twiddles(int p, int t, const double *__restrict__ W,
                                      double (&twr)[LOGR][1 << LOGR],
                                      double (&twi)[LOGR][1 << LOGR]) {
  constexpr int R = 1 << LOGR;
  const RegMap mp = reg_phase_r<LOGR>(p, LOGN);
  const int r = mp.lg;
  const int gmask = (1 << r) - 1;
#pragma unroll
  for (int ss = 0; ss < LOGR; ss++) {
    if (ss >= r) break;
    const int half = 1 << (mp.lh0 + ss), len = half << 1;
    const double *Ws = W + 2 * (half - 1);
#pragma unroll
    for (int s = 0; s < R; s++) {
      const int e = s & gmask;
      if (e & (1 << ss)) continue;
      if ((e | (1 << ss)) > gmask) continue;
      const int kk = mp.pos(t, s) & (len - 1);
      twr[ss][s] = Ws[2 * kk];
      twi[ss][s] = Ws[2 * kk + 1];
    }
  }
}

// t: the calling thread's index in the workgroup
template <int LOGN, int LOGR, int SIGN>
__device__ inline void reg_fft_t(double (&xr)[1 << LOGR], double (&xi)[1 << LOGR], double *lre,
                                 double *lim, const double *__restrict__ W, int t) {
  constexpr int R = 1 << LOGR;
  constexpr int NP = (LOGN + LOGR - 1) / LOGR;
  // each phase's twiddles are loaded before the LDS exchange that precedes
  // the phase, so their latency overlaps the exchange
  double twr[LOGR][R], twi[LOGR][R];
  phase_twiddles<LOGN, LOGR>(0, t, W, twr, twi);
#pragma unroll
  for (int p = 0; p < NP; p++) {
    const RegMap mp = reg_phase_r<LOGR>(p, LOGN);
    if (p > 0) {
      const RegMap mq = reg_phase_r<LOGR>(p - 1, LOGN);
      phase_twiddles<LOGN, LOGR>(p, t, W, twr, twi);
      __syncthreads();
#pragma unroll
      for (int s = 0; s < R; s++) {
        const int q = sw(mq.pos(t, s));
        lre[q] = xr[s];
        lim[q] = xi[s];
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < R; s++) {
        const int q = sw(mp.pos(t, s));
        xr[s] = lre[q];
        xi[s] = lim[q];
      }
    }
    const int r = mp.lg;
    const int gmask = (1 << r) - 1;
#pragma unroll
    for (int ss = 0; ss < LOGR; ss++) {
      if (ss >= r) break;
#pragma unroll
      for (int s = 0; s < R; s++) {
        const int e = s & gmask;
        if (e & (1 << ss)) continue;
        if ((e | (1 << ss)) > gmask) continue;
        const int s1 = s + (1 << ss);
        const double wr = twr[ss][s];
        const double wi = SIGN * twi[ss][s];
        const double r1 = xr[s1], m1 = xi[s1];
        const double tr = wr * r1 - wi * m1;
        const double ti = wr * m1 + wi * r1;
        const double r0 = xr[s], m0 = xi[s];
        xr[s1] = r0 - tr;
        xi[s1] = m0 - ti;
        xr[s] = r0 + tr;
        xi[s] = m0 + ti;
      }
    }
  }
}

// log2 of the values per thread of the register FFT: 8 by default (the
// tuner may take 4 or 16 per context, eif_gpu_autotune; every choice
// performs the same butterflies)
constexpr int FFT_LOGR_DEFAULT = 3;
template <int LOGN>
constexpr int dct_logr() {
  return FFT_LOGR_DEFAULT;
}

__device__ inline double cos_lookup(const double *C, int n, long m) {
  m %= 4 * (long)n;
  return m <= 2 * n ? C[m] : C[4 * (long)n - m];
}

// ---------------------------------------------------------------------
// RHS (compute_potential.c:70-82) + DCT-II along x (:84-92), one row per
// workgroup: mode 1 LDS FFT (powers of two outside the register-FFT
// range), mode 0 direct sum (nx not a power of two).
// ---------------------------------------------------------------------
template <bool P2 = false>
__device__ inline double rhs_value(const Layout &L, const double *ux, const double *uy,
                                   double fac, double dx, double dy, int i, int j) {
  return fac * (+qdiv<P2>(ux[L.ux(i + 2, j)] - ux[L.ux(i + 1, j)], dx, 1. / dx)
                + qdiv<P2>(uy[L.p(i + 1, j + 1)] - uy[L.p(i + 1, j)], dy, 1. / dy));
}

// the right-hand side with an additive cell field (the adjoint projection's
// pressure cotangent): fac * (div u + fac * add(i, j)); add == null is the
// forward right-hand side, operation for operation
template <bool P2 = false>
__device__ inline double rhs_value_add(const Layout &L, const double *ux, const double *uy,
                                       const double *add, double fac, double dx, double dy,
                                       int i, int j) {
  if (!add) return rhs_value<P2>(L, ux, uy, fac, dx, dy, i, j);
  const double div = qdiv<P2>(ux[L.ux(i + 2, j)] - ux[L.ux(i + 1, j)], dx, 1. / dx)
                     + qdiv<P2>(uy[L.p(i + 1, j + 1)] - uy[L.p(i + 1, j)], dy, 1. / dy);
  return fac * (div + fac * add[L.p(i + 1, j)]);
}

// How a one-row kernel transforms a row of n values: logn >= 0, the
// radix-2 LDS FFT of n = 2^logn; otherwise the Stockham FFT with the stage
// radices r[0..nr) (n's prime factors in increasing order)
struct DctPlan {
  int logn;
  int nr;
  int r[24];
};

// exp(sign 2 pi i e / n) from the table C[m] = cos(pi m / (2n)):
// cos(2 pi e / n) = C(4e), sin(2 pi e / n) = C(|n - 4e|)
__device__ inline void unit_root(const double *C, int n, long e, double sign, double &wr,
                                 double &wi) {
  const long d = (long)n - 4 * e;
  wr = cos_lookup(C, n, 4 * e);
  wi = sign * cos_lookup(C, n, d < 0 ? -d : d);
}

// Stockham autosort FFT of any length n in LDS (sign -1 forward, +1
// backward, unnormalised), from (xr, xi) with (yr, yi) as the second
// buffer; *outr, *outi: the pair that holds the result.  In the stage of
// radix R after radices of product Ns, butterfly j (k = j mod Ns) takes
// v_r = x[j + r n/R] w^(k r), w = e^(sign 2 pi i / (Ns R)), and stores
// sum_r v_r e^(sign 2 pi i q r / R) at (j / Ns) Ns R + k + q Ns.  Factors
// with exponent 0 are not multiplied; the oracle's fft_mixed performs the
// same operations in the same order.
__device__ void lds_fft_mixed(double *xr, double *xi, double *yr, double *yi, int n,
                              const DctPlan &pl, const double *__restrict__ C, double sign,
                              double **outr, double **outi) {
  int Ns = 1;
  for (int st = 0; st < pl.nr; st++) {
    const int R = pl.r[st], nR = n / R, tw = n / (Ns * R);
    for (int j = threadIdx.x; j < nR; j += blockDim.x) {
      const int k = j % Ns;
      const int o = (j / Ns) * Ns * R + k;
      for (int q = 0; q < R; q++) {
        double sr = 0., si = 0.;
        for (int r = 0; r < R; r++) {
          double vr = xr[j + r * nR], vi = xi[j + r * nR];
          const long e = (long)k * r * tw;
          if (e != 0) {
            double wr, wi;
            unit_root(C, n, e, sign, wr, wi);
            const double tr = vr * wr - vi * wi, ti = vr * wi + vi * wr;
            vr = tr;
            vi = ti;
          }
          if (r == 0) {
            sr = vr;
            si = vi;
            continue;
          }
          const long e2 = (long)((q * r) % R) * nR;
          if (e2 == 0) {
            sr = sr + vr;
            si = si + vi;
          } else {
            double cr, ci;
            unit_root(C, n, e2, sign, cr, ci);
            sr = sr + (vr * cr - vi * ci);
            si = si + (vr * ci + vi * cr);
          }
        }
        yr[o + q * Ns] = sr;
        yi[o + q * Ns] = si;
      }
    }
    __syncthreads();
    double *t = xr;
    xr = yr;
    yr = t;
    t = xi;
    xi = yi;
    yi = t;
    Ns *= R;
  }
  *outr = xr;
  *outi = xi;
}

struct RhsDct2B {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, 
    Layout L, const double *ux, const double *uy, const double *dscal, double ga,
    double dx, double dy, const double *__restrict__ C, const double *__restrict__ W,
    DctPlan pl, int mode, double *qhat, const double *add) {
  extern __shared__ double sh[];
  const int nx = L.nx;
  const int j = vb.bx + 1;
  double *re = sh, *im = sh + nx;
  const double dt = dscal[S_DT];
  const double fac = 1. / (ga * dt);
  double *out = qhat + (long)(j - 1) * nx;
  if (mode == 3) {  // Makhoul order, then the Stockham FFT (4 nx values of LDS)
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const int m = (i & 1) ? nx - 1 - (i >> 1) : (i >> 1);
      re[m] = rhs_value_add(L, ux, uy, add, fac, dx, dy, i, j);
      im[m] = 0.;
    }
    __syncthreads();
    double *Xr, *Xi;
    lds_fft_mixed(re, im, sh + 2 * nx, sh + 3 * nx, nx, pl, C, -1., &Xr, &Xi);
    for (int k = threadIdx.x; k < nx; k += blockDim.x)
      out[k] = 2. * (C[k] * Xr[k] + C[nx - k] * Xi[k]);
  } else if (mode == 1) {
    const int logn = pl.logn;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const int m = (i & 1) ? nx - 1 - (i >> 1) : (i >> 1);
      const int pos = bitrev(m, logn);
      re[pos] = rhs_value_add(L, ux, uy, add, fac, dx, dy, i, j);
      im[pos] = 0.;
    }
    __syncthreads();
    lds_fft(re, im, nx, logn, C, -1.);
    for (int k = threadIdx.x; k < nx; k += blockDim.x)
      out[k] = 2. * (C[k] * re[k] + C[nx - k] * im[k]);
  } else {
    for (int i = threadIdx.x; i < nx; i += blockDim.x)
      re[i] = rhs_value_add(L, ux, uy, add, fac, dx, dy, i, j);
    __syncthreads();
    for (int k = threadIdx.x; k < nx; k += blockDim.x) {
      double acc = 0.;
      for (int jj = 0; jj < nx; jj++)
        acc = acc + 2. * re[jj] * cos_lookup(C, nx, (long)(2 * jj + 1) * k);
      out[k] = acc;
    }
  }
}
};
__global__ __launch_bounds__(1024) void k_rhs_dct2(
    Layout L, const double *ux, const double *uy, const double *dscal, double ga,
    double dx, double dy, const double *__restrict__ C, const double *__restrict__ W,
    DctPlan pl, int mode, double *qhat, const double *add) {
  RhsDct2B::run(VB::here(), L, ux, uy, dscal, ga, dx, dy, C, W, pl, mode, qhat, add);
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
  const double *v;      // [L][nx] spike A_L^-1 e_0 (the other is its mirror)
  const double *bnd;    // [P][2][nx] chunk-boundary neighbour values
  double coef;          // l (= u) of the matrix
  const int *halt;      // graph-step stop flag: psi is left alone
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
  return q - a.coef * xl * a.v[(long)r * nx + k] - a.coef * xr * a.v[(long)(a.L - 1 - r) * nx + k];
}

struct Dct3PsiB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, 
    Layout L, const double *qhat, CombineArgs ca, const double *__restrict__ C,
    const double *__restrict__ W, DctPlan pl, int mode, double *psi) {
  if (halted(ca.halt)) return;
  extern __shared__ double sh[];
  const int nx = L.nx, ny = L.ny;
  const int j0 = vb.bx;
  double *re = sh, *im = sh + nx;
  const double norm = 2. * nx;
  if (mode == 3) {  // pre-twiddled spectrum, Stockham FFT, Makhoul order back
    for (int k = threadIdx.x; k < nx; k += blockDim.x) {
      const double xr = combined(ca, qhat, nx, ny, j0, k);
      const double xi = k == 0 ? 0. : -combined(ca, qhat, nx, ny, j0, nx - k);
      re[k] = C[k] * xr - C[nx - k] * xi;
      im[k] = C[k] * xi + C[nx - k] * xr;
    }
    __syncthreads();
    double *Xr, *Xi;
    lds_fft_mixed(re, im, sh + 2 * nx, sh + 3 * nx, nx, pl, C, +1., &Xr, &Xi);
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      const double v = (i & 1) ? Xr[nx - 1 - (i >> 1)] : Xr[i >> 1];
      psi[L.p(i + 1, j0 + 1)] = v / norm;
    }
  } else if (mode == 1) {
    const int logn = pl.logn;
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
};
__global__ __launch_bounds__(1024) void k_dct3_psi(
    Layout L, const double *qhat, CombineArgs ca, const double *__restrict__ C,
    const double *__restrict__ W, DctPlan pl, int mode, double *psi) {
  Dct3PsiB::run(VB::here(), L, qhat, ca, C, W, pl, mode, psi);
}

// Register-FFT kernels specialised for nx = 2^LOGN (nx / 2^LOGR threads),
// two rows per workgroup through one complex FFT: row a is the real and
// row b the imaginary part.  With Z the transform and kk = (n - k) mod n,
// row a's spectrum is (Z[k] + conj Z[kk]) / 2 and row b's
// (Z[k] - conj Z[kk]) / 2i (the halves cancel against REDFT10's factor 2);
// the inverse takes the two pre-twiddled spectra as Va + i Vb.  An odd last
// row pairs with a zero row.
template <int LOGN, bool P2, bool ADD = false, int LOGR_ = dct_logr<LOGN>()>
struct RhsDct2TB {
  static constexpr int T = (1 << LOGN) >> LOGR_, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, 
    Layout L, const double *ux, const double *uy, const double *dscal, double ga,
    double dx, double dy, const double *__restrict__ C, const double *__restrict__ W,
    double *qhat, const double *add = nullptr) {
  extern __shared__ double sh[];
  constexpr int LOGR = LOGR_, R = 1 << LOGR;
  constexpr int nx = 1 << LOGN;
  constexpr int T = nx / R;
  const int ja = 2 * vb.bx + 1, jb = ja + 1;
  const bool hasb = jb <= L.ny;
  double *re = sh, *im = sh + nx;
  const double dt = dscal[S_DT];
  const double fac = 1. / (ga * dt);
  // cells in memory order (coalesced loads), each stored at its Makhoul
  // position m: i = 2m for m < nx/2, i = 2(nx-1-m)+1 above.  Without a
  // second row, row ja is evaluated again and discarded, so that no load
  // sits behind a branch and all of them issue together.
  const int jbl = hasb ? jb : ja;
#pragma unroll
  for (int s = 0; s < R; s++) {
    const int i = threadIdx.x + s * T;
    const int m = (i & 1) ? nx - 1 - (i >> 1) : i >> 1;
    const int q = sw(bitrev(m, LOGN));
    const double vb = ADD ? rhs_value_add<P2>(L, ux, uy, add, fac, dx, dy, i, jbl)
                          : rhs_value<P2>(L, ux, uy, fac, dx, dy, i, jbl);
    re[q] = ADD ? rhs_value_add<P2>(L, ux, uy, add, fac, dx, dy, i, ja)
                : rhs_value<P2>(L, ux, uy, fac, dx, dy, i, ja);
    im[q] = hasb ? vb : 0.;
  }
  __syncthreads();
  double xr[R], xi[R];
#pragma unroll
  for (int s = 0; s < R; s++) {
    xr[s] = re[sw(R * threadIdx.x + s)];
    xi[s] = im[sw(R * threadIdx.x + s)];
  }
  reg_fft_t<LOGN, LOGR, -1>(xr, xi, re, im, W, threadIdx.x);
  constexpr int NP = (LOGN + LOGR - 1) / LOGR;
  const RegMap ml = reg_phase_r<LOGR>(NP - 1, LOGN);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < R; s++) {
    const int q = sw(ml.pos(threadIdx.x, s));
    re[q] = xr[s];
    im[q] = xi[s];
  }
  __syncthreads();
  double *outa = qhat + (long)(ja - 1) * nx, *outb = outa + nx;
#pragma unroll
  for (int s = 0; s < R; s++) {
    const int k = threadIdx.x + s * T, kk = (nx - k) & (nx - 1);
    const double zr = re[sw(k)], zi = im[sw(k)], wr = re[sw(kk)], wi = im[sw(kk)];
    outa[k] = C[k] * (zr + wr) + C[nx - k] * (zi - wi);
    if (hasb) outb[k] = C[k] * (zi + wi) + C[nx - k] * (wr - zr);
  }
}
};
template <int LOGN, bool P2, bool ADD = false, int LOGR_ = dct_logr<LOGN>()>
__global__ __launch_bounds__((1 << LOGN) >> LOGR_) void k_rhs_dct2_t(
    Layout L, const double *ux, const double *uy, const double *dscal, double ga,
    double dx, double dy, const double *__restrict__ C, const double *__restrict__ W,
    double *qhat, const double *add = nullptr) {
  RhsDct2TB<LOGN, P2, ADD, LOGR_>::run(VB::here(), L, ux, uy, dscal, ga, dx, dy, C, W, qhat,
                                       add);
}

// The DCT-III (compute_potential.c:118-137) of two rows packed as one
// complex transform: re, im hold the rows' combined coefficients at sw(k)
// (written and synchronised by the caller); on return they hold 2 nx psi
// of the two rows in the Makhoul order, cell i at sw(i / 2) for even i and
// at sw(nx - 1 - i / 2) for odd i (see psi_at).
// ORD: the factors C[k], C[nx - k] of the pre-processing come from D, the
// same values stored in the order the threads read them (D[s T + t] and
// D[nx + s T + t] for k = bitrev(R t + s)), so that a wave's loads are
// contiguous instead of bit-reversed across the table
template <int LOGN, int LOGR_ = dct_logr<LOGN>(), bool ORD = false>
__device__ __forceinline__ void dct3_pair_lds(double *re, double *im,
                                              const double *__restrict__ C,
                                              const double *__restrict__ W, int t,
                                              const double *__restrict__ D = nullptr) {
  constexpr int LOGR = LOGR_, R = 1 << LOGR;
  constexpr int nx = 1 << LOGN;
  constexpr int T = nx / R;
  double xr[R], xi[R];
#pragma unroll
  for (int s = 0; s < R; s++) {
    const int k = bitrev(R * t + s, LOGN);
    const double ck = ORD ? D[s * T + t] : C[k], cnk = ORD ? D[nx + s * T + t] : C[nx - k];
    const double ar = re[sw(k)], ai = k == 0 ? 0. : -re[sw(nx - k)];
    const double br = im[sw(k)], bi = k == 0 ? 0. : -im[sw(nx - k)];
    const double var = ck * ar - cnk * ai, vai = ck * ai + cnk * ar;
    const double vbr = ck * br - cnk * bi, vbi = ck * bi + cnk * br;
    xr[s] = var - vbi;
    xi[s] = vai + vbr;
  }
  reg_fft_t<LOGN, LOGR, +1>(xr, xi, re, im, W, t);
  constexpr int NP = (LOGN + LOGR - 1) / LOGR;
  const RegMap ml = reg_phase_r<LOGR>(NP - 1, LOGN);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < R; s++) {
    const int q = sw(ml.pos(t, s));
    re[q] = xr[s];
    im[q] = xi[s];
  }
  __syncthreads();
}

// psi of cell i (0-based) of a row left by dct3_pair_lds
template <int LOGN>
__device__ __forceinline__ double psi_at(const double *src, int i) {
  constexpr int nx = 1 << LOGN;
  constexpr double inv_norm = 1. / (2. * nx);  // a power of two: exact product
  const double v = (i & 1) ? src[sw(nx - 1 - (i >> 1))] : src[sw(i >> 1)];
  return v * inv_norm;
}

// up to 8 values per thread bounded to 4 waves per SIMD (128 VGPRs, no
// spills; 3 waves otherwise); 16 values per thread need more registers
template <int LOGN, int LOGR_ = dct_logr<LOGN>()>
struct Dct3PsiTB {
  static constexpr int T = (1 << LOGN) >> LOGR_, M = LOGR_ <= 3 ? 4 : 0;
  static __device__ __forceinline__ void run(const VB &vb, 
    Layout L, const double *qhat, CombineArgs ca, const double *__restrict__ C,
    const double *__restrict__ W, const double *__restrict__ D, double *psi) {
  if (halted(ca.halt)) return;
  extern __shared__ double sh[];
  constexpr int LOGR = LOGR_, R = 1 << LOGR;
  constexpr int nx = 1 << LOGN;
  constexpr int T = nx / R;
  const int ny = L.ny;
  // the two rows of this workgroup: consecutive rows, or in partition mode
  // mirror rows r, L-1-r of one chunk, whose fix-ups read the same two spike
  // rows v[r], v[L-1-r] and boundary values (half the table loads; the
  // pairing changes only the rounding inside the shared transform)
  int ja0 = 2 * vb.bx, jb0 = ja0 + 1;
  if (ca.mode == 1) {
    const int h = ca.L >> 1, cc = vb.bx / h, t = vb.bx - cc * h;
    ja0 = cc * ca.L + t;
    jb0 = cc * ca.L + ca.L - 1 - t;
  }
  const bool hasb = jb0 < ny;
  double *re = sh, *im = sh + nx;
  // division by 2 nx, a power of two, is the exact product with its inverse
  constexpr double inv_norm = 1. / (2. * nx);
  if (ca.mode == 1) {  // both rows from one load of the spike rows and boundary values
    const int cc = ja0 / ca.L, ra = ja0 - cc * ca.L, rb = ca.L - 1 - ra;
    const double *xlp = ca.bnd + ((long)cc * 2 + 0) * nx, *xrp = ca.bnd + ((long)cc * 2 + 1) * nx;
    const double *va = ca.v + (long)ra * nx, *vb = ca.v + (long)rb * nx;
#pragma unroll
    for (int s = 0; s < R; s++) {
      const int k = threadIdx.x + s * T;
      const double qa = qhat[(long)ja0 * nx + k], qb = qhat[(long)jb0 * nx + k];
      const double xl = xlp[k], xr = xrp[k], v0 = va[k], v1 = vb[k];
      re[sw(k)] = qa - ca.coef * xl * v0 - ca.coef * xr * v1;
      im[sw(k)] = qb - ca.coef * xl * v1 - ca.coef * xr * v0;
    }
  } else {
#pragma unroll
    for (int s = 0; s < R; s++) {
      const int k = threadIdx.x + s * T;
      re[sw(k)] = combined(ca, qhat, nx, ny, ja0, k);
      im[sw(k)] = hasb ? combined(ca, qhat, nx, ny, jb0, k) : 0.;
    }
  }
  __syncthreads();
  dct3_pair_lds<LOGN, LOGR, true>(re, im, C, W, threadIdx.x, D);
  // psi rows and their boundary values (the Neumann walls, and the periodic
  // halo rows 0 and ny+1 copied from rows ny and 1: k_bc_p's values)
#pragma unroll
  for (int r = 0; r < 2; r++) {
    if (r == 1 && !hasb) break;
    const double *src = r == 0 ? re : im;
    const int j = (r == 0 ? ja0 : jb0) + 1;
#pragma unroll
    for (int s = 0; s < R; s++) {
      const int i = threadIdx.x + s * T;
      const double v = (i & 1) ? src[sw(nx - 1 - (i >> 1))] : src[sw(i >> 1)];
      const double w = v * inv_norm;
      psi[L.p(i + 1, j)] = w;
      if (j == ny) psi[L.p(i + 1, 0)] = w;
      if (j == 1) psi[L.p(i + 1, ny + 1)] = w;
      if (i == 0 || i == nx - 1) {
        const int iw = i == 0 ? 0 : nx + 1;
        psi[L.p(iw, j)] = w;
        if (j == ny) psi[L.p(iw, 0)] = w;
        if (j == 1) psi[L.p(iw, ny + 1)] = w;
      }
    }
  }
}
};
template <int LOGN, int LOGR_ = dct_logr<LOGN>()>
__global__ __launch_bounds__((1 << LOGN) >> LOGR_, LOGR_ <= 3 ? 4 : 1) void k_dct3_psi_t(
    Layout L, const double *qhat, CombineArgs ca, const double *__restrict__ C,
    const double *__restrict__ W, const double *__restrict__ D, double *psi) {
  Dct3PsiTB<LOGN, LOGR_>::run(VB::here(), L, qhat, ca, C, W, D, psi);
}

// ---------------------------------------------------------------------
// DCT-III fused with the projection and the pressure update
// (compute_potential.c:118-137, correct_velocity.c:16-41,
// update_pressure.c:14-23): one workgroup per strip of S rows (S even).
// Rows are transformed in the pairs (2q, 2q+1) of k_dct3_psi_t's Thomas
// mode, so psi has the same bits; each row's psi stays in LDS and
// registers and is never stored.  A thread owns cells i = tid + s T of
// every row: psi(i - 1, j) comes from LDS, psi(i, j - 1) from the row
// before, kept in registers (for the strip's first row, the pair holding
// row j - 1, periodic, is transformed first).  The arithmetic and the
// boundary values written are those of k_correct_pressure; with DT the
// advective time-step limit is left as one partial minimum per strip.
// ---------------------------------------------------------------------
template <int LOGN, bool P2, bool DT>
struct ProjectTB {
  // at least 4 waves per SIMD (at most 128 VGPRs), as the unfused DCT kernels
  static constexpr int T = (1 << LOGN) >> dct_logr<LOGN>(), M = 4;

  // the pair of rows (pb, pb + 1) into LDS
  static __device__ __forceinline__ void pair(const CombineArgs &ca, const double *qhat,
                                              int ny, int pb, double *re, double *im,
                                              const double *__restrict__ C,
                                              const double *__restrict__ W, int t) {
    constexpr int LOGR = dct_logr<LOGN>(), R = 1 << LOGR;
    constexpr int nx = 1 << LOGN;
    constexpr int TT = nx / R;
    const bool hasb = pb + 1 < ny;
    __syncthreads();  // the previous pair's values are read
#pragma unroll
    for (int s = 0; s < R; s++) {
      const int k = t + s * TT;
      re[sw(k)] = combined(ca, qhat, nx, ny, pb, k);
      im[sw(k)] = hasb ? combined(ca, qhat, nx, ny, pb + 1, k) : 0.;
    }
    __syncthreads();
    dct3_pair_lds<LOGN>(re, im, C, W, t);
  }

  // the projection and pressure update of row j (k_correct_pressure's
  // arithmetic and boundary values; ny >= 2, so that row j has at most one
  // periodic halo copy: row ny to row 0, row 1 to row ny + 1)
  static __device__ __forceinline__ void row(const Layout &L, const double *src, int t, int j,
                                             double *ux, double *uy, double *p,
                                             const double *dxc, double ga, double dt,
                                             double dy, double (&prev)[1 << dct_logr<LOGN>()],
                                             double &m) {
    constexpr int LOGR = dct_logr<LOGN>(), R = 1 << LOGR;
    constexpr int nx = 1 << LOGN;
    constexpr int TT = nx / R;
    const int ny = L.ny;
    // the row's fields from uniform row bases (32-bit cell offsets):
    // ux(i, j) = uxr[i], uy(i, j) = uyr[i], p(i, j) = pr[i]; *h the halo copy
    const int jh = j == ny ? 0 : (j == 1 ? ny + 1 : -1);
    double *uxr = ux + L.ux(0, j), *uyr = uy + L.p(0, j), *pr = p + L.p(0, j);
    double *uxh = jh >= 0 ? ux + L.ux(0, jh) : nullptr;
    double *uyh = jh >= 0 ? uy + L.p(0, jh) : nullptr;
    double *ph = jh >= 0 ? p + L.p(0, jh) : nullptr;
#pragma unroll
    for (int s = 0; s < R; s++) {
      const int i = t + s * TT + 1;  // 1..nx
      const double ps_c = psi_at<LOGN>(src, i - 1);
      if (i >= 2) {
        const double ps_w = psi_at<LOGN>(src, i - 2);
        const double dxi = 1. / dxc[i - 1];
        const double v = uxr[i] - ga * dt * (ps_c - ps_w) * dxi;
        uxr[i] = v;
        if (uxh) uxh[i] = v;
        if (DT) m = fmin(m, dxc[i - 1] / fmax(fabs(v), 1.e-8));
      }
      {
        const double v = uyr[i] - qdiv<P2>(ga * dt * (ps_c - prev[s]), dy, 1. / dy);
        uyr[i] = v;
        if (uyh) uyh[i] = v;
        if (DT) m = fmin(m, dy / fmax(fabs(v), 1.e-8));
      }
      {
        const double v = pr[i] + ps_c;
        pr[i] = v;
        if (ph) ph[i] = v;
        // the Neumann walls (with their halo corners)
        if (i == 1) {
          pr[0] = v;
          if (ph) ph[0] = v;
        }
        if (i == nx) {
          pr[nx + 1] = v;
          if (ph) ph[nx + 1] = v;
        }
      }
      prev[s] = ps_c;
    }
  }

  static __device__ __forceinline__ void run(const VB &vb, Layout L, const double *qhat,
                                             CombineArgs ca, const double *__restrict__ C,
                                             const double *__restrict__ W, int S, double *ux,
                                             double *uy, double *p, const double *dxc,
                                             const double *dscal, double ga, double dy,
                                             double *dt_part) {
    if (halted(ca.halt)) return;
    extern __shared__ double sh[];
    constexpr int LOGR = dct_logr<LOGN>(), R = 1 << LOGR;
    constexpr int nx = 1 << LOGN;
    constexpr int TT = nx / R;
    const int ny = L.ny;
    double *re = sh, *im = sh + nx;
    const double dt = dscal[S_DT];
    const int r0 = vb.bx * S, r1 = min(r0 + S, ny);  // rows r0 .. r1-1 (0-based)
    double m = 5.e-2;
    double prev[R];
    // the pair holding row r0 - 1 (periodic) first, for prev only; then the
    // strip's pairs (one call site of the transform)
    const int rp = r0 == 0 ? ny - 1 : r0 - 1;
    int pb = rp & ~1;
    bool first = true;
    for (;;) {
      // the tables' addresses and the thread index made opaque per pair:
      // the transform's table loads and index arithmetic are redone per
      // pair instead of hoisted out of the loop (every phase's twiddles and
      // positions would be held in registers at once)
      const double *Cl = C, *Wl = W;
      int t = threadIdx.x;
      asm volatile("" : "+s"(Cl), "+s"(Wl), "+v"(t));
      pair(ca, qhat, ny, pb, re, im, Cl, Wl, t);
      if (first) {
        const double *src = (rp & 1) ? im : re;
#pragma unroll
        for (int s = 0; s < R; s++) prev[s] = psi_at<LOGN>(src, t + s * TT);
        first = false;
        pb = r0;
        continue;
      }
#pragma unroll 1
      for (int h = 0; h < 2; h++) {
        const int r = pb + h;
        if (r >= r1) break;
        const double *src = h ? im : re;
        const int j = r + 1;
        row(L, src, t, j, ux, uy, p, dxc, ga, dt, dy, prev, m);
      }
      pb += 2;
      if (pb >= r1) break;
    }
    if (DT) {
      __syncthreads();  // sh is free
      double *red = sh;
      red[threadIdx.x] = m;
      __syncthreads();
      for (int st = TT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
      }
      if (threadIdx.x == 0) dt_part[vb.bx] = red[0];
    }
  }
};
template <int LOGN, bool P2, bool DT>
__global__ __launch_bounds__((1 << LOGN) >> dct_logr<LOGN>(), 4) void k_project_t(
    Layout L, const double *qhat, CombineArgs ca, const double *__restrict__ C,
    const double *__restrict__ W, int S, double *ux, double *uy, double *p, const double *dxc,
    const double *dscal, double ga, double dy, double *dt_part) {
  ProjectTB<LOGN, P2, DT>::run(VB::here(), L, qhat, ca, C, W, S, ux, uy, p, dxc, dscal, ga, dy,
                               dt_part);
}

// ---------------------------------------------------------------------
// Thomas mode: one lane per wavenumber k, sweeps along j
// ---------------------------------------------------------------------
struct ThomasB {
  static constexpr int T = 64, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, int nx, int ny, double *q,
                                               const double *__restrict__ rv,
                                               const double *__restrict__ um,
                                               const unsigned char *__restrict__ flag,
                                               const double *__restrict__ cdiag,
                                               const double *__restrict__ den,
                                               double coef, double *last) {
  const int k = vb.bx * blockDim.x + threadIdx.x;
  if (k >= nx) return;
  const int n1 = ny - 1;  // size of the bordered (n-1) system
  const double l = coef;  // l = u = 1/dy/dy everywhere
  // forward sweep (tdm.c:40-53) on rows 0..n-2.  The sweeps are one
  // dependent chain per wavenumber: the rows' loads are issued a batch of
  // TH_U rows ahead of the arithmetic (same operations, same order).
  constexpr int TH_U = 16;
  double prev = q[k] / cdiag[k];
  q[k] = prev;
  {
    const int iend = n1 - 1;  // last row of the forward sweep
    double qa[TH_U], ra[TH_U];
    bool fa[TH_U];
    auto load = [&](int i0, double *qv, double *rw, bool *fl) {
#pragma unroll
      for (int u = 0; u < TH_U; u++) {
        const int ii = i0 + u;
        if (ii <= iend) {
          const long o = (long)ii * nx + k;
          qv[u] = q[o];
          rw[u] = rv[o];
          fl[u] = flag[o] != 0;
        }
      }
    };
    load(1, qa, ra, fa);
    for (int i = 1; i <= iend; i += TH_U) {
      double qb[TH_U] = {}, rb[TH_U] = {};
      bool fb[TH_U] = {};
      if (i + TH_U <= iend) load(i + TH_U, qb, rb, fb);  // later rows: not yet written
#pragma unroll
      for (int u = 0; u < TH_U; u++) {
        const int ii = i + u;
        if (ii > iend) break;
        const double v = fa[u] ? 0. : ra[u] * (qa[u] - l * prev);
        q[(long)ii * nx + k] = v;
        prev = v;
      }
#pragma unroll
      for (int u = 0; u < TH_U; u++) {
        qa[u] = qb[u];
        ra[u] = rb[u];
        fa[u] = fb[u];
      }
    }
  }
  // backward sweep (:55-57), loads a batch ahead in the same way
  const double q_last_fwd = prev;  // q0[n-2]
  double next = prev;
  {
    double qa[TH_U], ua[TH_U];
    auto load = [&](int i0, double *qv, double *uw) {
#pragma unroll
      for (int u = 0; u < TH_U; u++) {
        const int ii = i0 - u;
        if (ii >= 0) {
          const long o = (long)ii * nx + k;
          qv[u] = q[o];
          uw[u] = um[o];
        }
      }
    };
    load(n1 - 2, qa, ua);
    for (int i = n1 - 2; i >= 0; i -= TH_U) {
      double qb[TH_U] = {}, ub[TH_U] = {};
      if (i - TH_U >= 0) load(i - TH_U, qb, ub);  // earlier rows: not yet written
#pragma unroll
      for (int u = 0; u < TH_U; u++) {
        const int ii = i - u;
        if (ii < 0) break;
        const double v = qa[u] - ua[u] * next;
        q[(long)ii * nx + k] = v;
        next = v;
      }
#pragma unroll
      for (int u = 0; u < TH_U; u++) {
        qa[u] = qb[u];
        ua[u] = ub[u];
      }
    }
  }
  // bottom solution (tdm.c:99-102)
  const double q0_0 = next;
  const double num = q[(long)(ny - 1) * nx + k] - l * q0_0 - l * q_last_fwd;
  const double d = den[k];
  const double xl = fabs(d) < DBL_EPSILON ? 0. : num / d;
  last[k] = xl;
}
};
__global__ __launch_bounds__(64) void k_thomas(int nx, int ny, double *q,
                                               const double *__restrict__ rv,
                                               const double *__restrict__ um,
                                               const unsigned char *__restrict__ flag,
                                               const double *__restrict__ cdiag,
                                               const double *__restrict__ den,
                                               double coef, double *last) {
  ThomasB::run(VB::here(), nx, ny, q, rv, um, flag, cdiag, den, coef, last);
}

// ---------------------------------------------------------------------
// Partition mode (spike method).  Chunk c holds rows [cL, cL+L).  With the
// chunk-local solution y = A_L^-1 d, v = A_L^-1 e_0 and w = A_L^-1 e_{L-1}
// (w[r] = v[L-1-r] by symmetry), the periodic solution is
//   x[cL+r] = y_c[r] - l*beta_{c-1}*v[r] - l*alpha_{c+1}*w[r]
// where alpha_c, beta_c are the first/last values of chunk c.  The pairs
// u_c = (beta_c, alpha_{c+1}) satisfy a periodic block-tridiagonal system
//   Lb u_{c-1} + D u_c + Ub u_{c+1} = (y_c[L-1], y_{c+1}[0])
//   D = [[1, l v0], [l v0, 1]], Lb = [[l vL, 0], [0, 0]], Ub = [[0, 0], [0, l vL]]
// solved per wavenumber by block elimination with u_{P-1} as the border
// unknown; all matrix-only factors are tabulated per context.
// ---------------------------------------------------------------------
constexpr int PART_LANES = 64;

template <int L>
struct PartLocalB {
  static constexpr int T = PART_LANES, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, 
    int nx, double *q, const double *__restrict__ fwd,
    const double *__restrict__ um, double coef, double *ends) {
  const int k = vb.bx * PART_LANES + threadIdx.x;
  const int c = vb.by;
  if (k >= nx) return;
  const long j0 = (long)c * L;
  const double l = coef;
  // the forward sweep keeps y in LDS (one column per lane) for the
  // backward sweep, so q is read and written once
  __shared__ double ys[L][PART_LANES];
  const int lane = threadIdx.x;
  double prev = 0.;
#pragma unroll 16
  for (int r = 0; r < L; r++) {
    const double d = q[(j0 + r) * nx + k];
    const double v = fwd[(long)r * nx + k] * (d - l * prev);
    ys[r][lane] = v;
    prev = v;
  }
  const double ylast = prev;
  q[(j0 + L - 1) * nx + k] = prev;
  double next = prev;
#pragma unroll 16
  for (int r = L - 2; r >= 0; r--) {
    next = ys[r][lane] - um[(long)r * nx + k] * next;
    q[(j0 + r) * nx + k] = next;
  }
  ends[((long)c * 2 + 0) * nx + k] = next;   // y_c[0]
  ends[((long)c * 2 + 1) * nx + k] = ylast;  // y_c[L-1]
}
};
template <int L>
__global__ __launch_bounds__(PART_LANES) void k_part_local(
    int nx, double *q, const double *__restrict__ fwd,
    const double *__restrict__ um, double coef, double *ends) {
  PartLocalB<L>::run(VB::here(), nx, q, fwd, um, coef, ends);
}

// Reduced system: the blocks are the same for every chunk, so the periodic
// block system is block-circulant and the length-P DFT over chunks
// diagonalises it.  For mode m (theta = 2 pi m / P):
//   M_m = [[1 + l vL e^{-i theta}, l v0], [l v0, 1 + l vL e^{i theta}]]
//   det_m = 1 + 2 l vL cos(theta) + (l vL)^2 - (l v0)^2   (real)
// Mode 0 of wavenumber 0 is singular (constant null vector); its
// minimum-norm solution fixes the gauge.  One workgroup handles
// PART_KG wavenumbers with 64 threads each.
constexpr int PART_KG = 4;

// the 2x2 system of one chunk-DFT mode (cos/sin of its angle ct, st) for
// one wavenumber; gauge: the singular mode 0 of wavenumber 0
__device__ inline void mode_solve(double lv0, double lvL, bool gauge, double ct, double st,
                                  double a0r, double a0i, double a1r, double a1i, double &x0r,
                                  double &x0i, double &x1r, double &x1i) {
  if (gauge) {
    const double f = 1. / (2. * ((1. + lvL) - lv0));
    x0r = (a0r - a1r) * f;
    x0i = (a0i - a1i) * f;
    x1r = -x0r;
    x1i = -x0i;
  } else {
    const double det = 1. + 2. * lvL * ct + lvL * lvL - lv0 * lv0;
    // adj(M) = [[1 + lvL e^{i th}, -lv0], [-lv0, 1 + lvL e^{-i th}]]
    const double p0r = 1. + lvL * ct, p0i = lvL * st;   // 1 + lvL e^{i th}
    const double p1r = 1. + lvL * ct, p1i = -lvL * st;  // 1 + lvL e^{-i th}
    x0r = (p0r * a0r - p0i * a0i - lv0 * a1r) / det;
    x0i = (p0r * a0i + p0i * a0r - lv0 * a1i) / det;
    x1r = (-lv0 * a0r + p1r * a1r - p1i * a1i) / det;
    x1i = (-lv0 * a0i + p1r * a1i + p1i * a1r) / det;
  }
}

struct PartReducedB {
  static constexpr int T = 64 * PART_KG, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, 
    int nx, int P, const double *__restrict__ ends, const double *__restrict__ lv,
    const double *__restrict__ twc, const double *__restrict__ tws,
    double *__restrict__ bnd) {
  extern __shared__ double sh[];
  const int kk = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int k = vb.bx * PART_KG + kk;
  // chunk-DFT twiddles staged in LDS: the sums below read them P times per
  // lane
  double *tc = sh + (long)PART_KG * 6 * P, *ts = tc + P;
  for (int e = threadIdx.x; e < P; e += blockDim.x) {
    tc[e] = twc[e];
    ts[e] = tws[e];
  }
  double *r0 = sh + (long)kk * 6 * P, *r1 = r0 + P;
  double *u0r = r1 + P, *u0i = u0r + P, *u1r = u0i + P, *u1i = u1r + P;
  const bool on = k < nx;
  if (on)
    for (int c = lane; c < P; c += 64) {
      r0[c] = ends[((long)c * 2 + 1) * nx + k];             // y_c[L-1]
      r1[c] = ends[((long)((c + 1) % P) * 2 + 0) * nx + k];  // y_{c+1}[0]
    }
  __syncthreads();
  if (on) {
    const double lv0 = lv[k], lvL = lv[nx + k];
    const bool gauge = lv[2 * (long)nx + k] != 0.;
    for (int m = lane; m < P; m += 64) {
      double a0r = 0., a0i = 0., a1r = 0., a1i = 0.;
      for (int c = 0, t = 0; c < P; c++) {  // t = m c mod P
        const double cs = tc[t], sn = ts[t];
        a0r += r0[c] * cs;
        a0i -= r0[c] * sn;
        a1r += r1[c] * cs;
        a1i -= r1[c] * sn;
        t += m;
        if (t >= P) t -= P;
      }
      double x0r, x0i, x1r, x1i;
      mode_solve(lv0, lvL, gauge && m == 0, tc[m], ts[m], a0r, a0i, a1r, a1i, x0r, x0i, x1r,
                 x1i);
      u0r[m] = x0r; u0i[m] = x0i; u1r[m] = x1r; u1i[m] = x1i;
    }
  }
  __syncthreads();
  if (on) {
    const double invP = 1. / P;
    for (int c = lane; c < P; c += 64) {
      double b0 = 0., b1 = 0.;
      for (int m = 0, t = 0; m < P; m++) {  // t = m c mod P
        const double cs = tc[t], sn = ts[t];
        b0 += u0r[m] * cs - u0i[m] * sn;
        b1 += u1r[m] * cs - u1i[m] * sn;
        t += c;
        if (t >= P) t -= P;
      }
      // u_c = (beta_c, alpha_{c+1}): chunk c+1's left value, chunk c's right
      bnd[((long)((c + 1) % P) * 2 + 0) * nx + k] = b0 * invP;
      bnd[((long)c * 2 + 1) * nx + k] = b1 * invP;
    }
  }
}
};
__global__ __launch_bounds__(64 * PART_KG) void k_part_reduced(
    int nx, int P, const double *__restrict__ ends, const double *__restrict__ lv,
    const double *__restrict__ twc, const double *__restrict__ tws,
    double *__restrict__ bnd) {
  PartReducedB::run(VB::here(), nx, P, ends, lv, twc, tws, bnd);
}

// The same reduced system with both chunk DFTs as FP64 matrix products on
// the matrix cores (v_mfma_f64_16x16x4f64), for P a multiple of 16 up to
// 64: a workgroup takes 16 wavenumbers and P/16 waves.  Forward: wave w
// forms the modes m in [16w, 16w+16) of a0 = C r0 - i S r0 and a1 (C, S:
// cos and sin of 2 pi m c / P); the 2x2 solves run on the accumulators
// (every lane holds the four values of its (m, k)); inverse: wave w forms
// the chunks c in [16w, 16w+16) of b = C x_r - S x_i.  Lane layout of the
// 16x16x4 product: lane l holds A[l%16][l/16], B[l/16][l%16], and
// D[4i + l/16][l%16] in element i.  The products are summed in another
// order than the lane loop of k_part_reduced (rounding only).
typedef double pr_d4 __attribute__((ext_vector_type(4)));
constexpr int PRM_K = 16;

struct PartReducedMfmaB {
  static constexpr int T = 256, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, 
    int nx, int P, const double *__restrict__ ends, const double *__restrict__ lv,
    const double *__restrict__ twc, const double *__restrict__ tws,
    double *__restrict__ bnd) {
  extern __shared__ double sh[];
  double *tc = sh, *ts = tc + P;
  double *r0 = ts + P, *r1 = r0 + P * PRM_K;  // [c][kk]
  double *xs = r1 + P * PRM_K;                 // [4][m][kk]: x0r, x0i, x1r, x1i
  const int k0 = vb.bx * PRM_K;
  for (int e = threadIdx.x; e < P; e += blockDim.x) {
    tc[e] = twc[e];
    ts[e] = tws[e];
  }
  for (int e = threadIdx.x; e < P * PRM_K; e += blockDim.x) {
    const int c = e / PRM_K, kk = e - c * PRM_K, k = k0 + kk;
    const bool on = k < nx;
    r0[e] = on ? ends[((long)c * 2 + 1) * nx + k] : 0.;             // y_c[L-1]
    r1[e] = on ? ends[((long)((c + 1) % P) * 2 + 0) * nx + k] : 0.;  // y_{c+1}[0]
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int lo = l & 15, hi = l >> 4;
  {
    const int m0 = 16 * w;
    pr_d4 a0r = {0., 0., 0., 0.}, a0i = a0r, a1r = a0r, a1i = a0r;
    for (int c0 = 0; c0 < P; c0 += 4) {
      const int t = ((m0 + lo) * (c0 + hi)) % P;
      const double cw = tc[t], sw = -ts[t];
      const double v0 = r0[(c0 + hi) * PRM_K + lo], v1 = r1[(c0 + hi) * PRM_K + lo];
      a0r = __builtin_amdgcn_mfma_f64_16x16x4f64(cw, v0, a0r, 0, 0, 0);
      a0i = __builtin_amdgcn_mfma_f64_16x16x4f64(sw, v0, a0i, 0, 0, 0);
      a1r = __builtin_amdgcn_mfma_f64_16x16x4f64(cw, v1, a1r, 0, 0, 0);
      a1i = __builtin_amdgcn_mfma_f64_16x16x4f64(sw, v1, a1i, 0, 0, 0);
    }
    const int k = k0 + lo;
    const bool on = k < nx;
    const double lv0 = on ? lv[k] : 0., lvL = on ? lv[nx + k] : 0.;
    const bool gauge = on && lv[2 * (long)nx + k] != 0.;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int m = m0 + 4 * i + hi;
      double x0r, x0i, x1r, x1i;
      mode_solve(lv0, lvL, gauge && m == 0, tc[m], ts[m], a0r[i], a0i[i], a1r[i], a1i[i], x0r,
                 x0i, x1r, x1i);
      const int e = m * PRM_K + lo;
      xs[e] = x0r;
      xs[P * PRM_K + e] = x0i;
      xs[2 * P * PRM_K + e] = x1r;
      xs[3 * P * PRM_K + e] = x1i;
    }
  }
  __syncthreads();
  {
    const int cb = 16 * w;
    pr_d4 b0 = {0., 0., 0., 0.}, b1 = b0;
    for (int m0 = 0; m0 < P; m0 += 4) {
      const int t = ((cb + lo) * (m0 + hi)) % P;
      const double cw = tc[t], sw = -ts[t];
      const int e = (m0 + hi) * PRM_K + lo;
      b0 = __builtin_amdgcn_mfma_f64_16x16x4f64(cw, xs[e], b0, 0, 0, 0);
      b0 = __builtin_amdgcn_mfma_f64_16x16x4f64(sw, xs[P * PRM_K + e], b0, 0, 0, 0);
      b1 = __builtin_amdgcn_mfma_f64_16x16x4f64(cw, xs[2 * P * PRM_K + e], b1, 0, 0, 0);
      b1 = __builtin_amdgcn_mfma_f64_16x16x4f64(sw, xs[3 * P * PRM_K + e], b1, 0, 0, 0);
    }
    const double invP = 1. / P;
    const int k = k0 + lo;
    if (k < nx) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        // u_c = (beta_c, alpha_{c+1}): chunk c+1's left value, chunk c's right
        const int c = cb + 4 * i + hi;
        bnd[((long)((c + 1) % P) * 2 + 0) * nx + k] = b0[i] * invP;
        bnd[((long)c * 2 + 1) * nx + k] = b1[i] * invP;
      }
    }
  }
}
};
__global__ __launch_bounds__(256) void k_part_reduced_mfma(
    int nx, int P, const double *__restrict__ ends, const double *__restrict__ lv,
    const double *__restrict__ twc, const double *__restrict__ tws,
    double *__restrict__ bnd) {
  PartReducedMfmaB::run(VB::here(), nx, P, ends, lv, twc, tws, bnd);
}

// The lane-loop reduced system (k_part_reduced) keeps (6 PART_KG + 2) P
// doubles in LDS; P is capped so that fits the 64 KiB a workgroup gets
// without an opt-in.  Longer lines use the Thomas solve.
constexpr int PART_P_MAX = 65536 / ((6 * PART_KG + 2) * (int)sizeof(double));

static int part_chunk(int ny) {
  for (int L = 64; L >= 8; L >>= 1)
    if (ny % L == 0 && ny / L >= 2 && ny / L <= PART_P_MAX) return L;
  return 0;
}

int partition_setup(Ctx *c, int L);

int partition_layout(int ny, int *chunk, int *nchunks) {
  const int L = ny >= 3 ? part_chunk(ny) : 0;
  if (chunk) *chunk = L;
  if (nchunks) *nchunks = L ? ny / L : 0;
  return L ? 1 : 0;
}

static void part_free(Ctx *c) {
  double **bufs[] = {&c->pt_fwd, &c->pt_u, &c->pt_v, &c->pt_red, &c->pt_w, &c->pt_lvl,
                     &c->pt_bnd, &c->pt_ends};
  for (auto b : bufs)
    if (*b) {
      (void)hipFree(*b);
      *b = nullptr;
    }
  c->part_L = c->part_P = 0;
}

static int part_init(Ctx *c, int Lreq = 0) {
  const int nx = c->nx, ny = c->ny;
  const int L = Lreq ? Lreq : part_chunk(ny);
  if (!L) return EIF_OK;  // partition mode unavailable for this jtot
  const int P = ny / L;
  c->part_L = L;
  c->part_P = P;
  const double dx = c->dx, dy = c->dy;
  const double l = 1. / dy / dy;
  std::vector<double> fwd((size_t)L * nx), um((size_t)L * nx), v((size_t)L * nx);
  std::vector<double> lv(3 * (size_t)nx), twc(P), tws(P);
  std::vector<long double> vv(L), uu(L), ff(L);
  for (int k = 0; k < nx; k++) {
    const double sn = sin(EIF_PI * k / (2. * nx));
    const double eig = -4. / (dx * dx) * (sn * sn);
    const double cc = -l - l + eig;
    // chunk-local Thomas factors (Dirichlet truncation A_L)
    double up = 0.;
    for (int r = 0; r < L; r++) {
      const double f = 1. / (cc - l * up);
      fwd[(size_t)r * nx + k] = f;
      up = f * l;
      um[(size_t)r * nx + k] = up;
    }
    // v = A_L^-1 e_0 in long double
    {
      long double u = 0., q = 0.;
      for (int r = 0; r < L; r++) {
        const long double f = 1.0L / ((long double)cc - (long double)l * u);
        ff[r] = f;
        q = f * ((r == 0 ? 1.0L : 0.0L) - (long double)l * q);
        vv[r] = q;
        u = f * (long double)l;
        uu[r] = u;
      }
      for (int r = L - 2; r >= 0; r--) vv[r] = vv[r] - uu[r] * vv[r + 1];
      for (int r = 0; r < L; r++) v[(size_t)r * nx + k] = (double)vv[r];
    }
    lv[k] = (double)((long double)l * vv[0]);
    lv[nx + k] = (double)((long double)l * vv[L - 1]);
    lv[2 * (size_t)nx + k] = (k == 0 && eig == 0.) ? 1. : 0.;
  }
  for (int t = 0; t < P; t++) {
    twc[t] = cos(2. * EIF_PI * t / P);
    tws[t] = sin(2. * EIF_PI * t / P);
  }
  auto up = [&](double **d, const std::vector<double> &h) -> int {
    EIF_HIP(hipMalloc(d, h.size() * sizeof(double)));
    EIF_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return EIF_OK;
  };
  int rc;
  if ((rc = up(&c->pt_fwd, fwd)) || (rc = up(&c->pt_u, um)) || (rc = up(&c->pt_v, v)) ||
      (rc = up(&c->pt_red, twc)) || (rc = up(&c->pt_w, tws)) || (rc = up(&c->pt_lvl, lv)))
    return rc;
  EIF_HIP(hipMalloc(&c->pt_bnd, (size_t)P * 2 * nx * sizeof(double)));
  EIF_HIP(hipMalloc(&c->pt_ends, (size_t)P * 2 * nx * sizeof(double)));
  return EIF_OK;
}

int partition_setup(Ctx *c, int L) {
  if (L && !(L >= 8 && L <= 64 && (L & (L - 1)) == 0 && c->ny % L == 0 && c->ny / L >= 2 &&
             c->ny / L <= PART_P_MAX)) {
    set_error("partition chunk does not split jtot within the limits");
    return EIF_ERR_ARG;
  }
  EIF_HIP(hipStreamSynchronize(c->stream));
  part_free(c);
  return part_init(c, L);
}

int poisson_init(Ctx *c) {
  const int nx = c->nx, ny = c->ny;
  // cosine table, same expression as the host restatement
  std::vector<double> C(2 * (size_t)nx + 1);
  for (int m = 0; m <= 2 * nx; m++) C[m] = cos(EIF_PI * m / (2. * nx));
  EIF_HIP(hipMalloc(&c->ctab, C.size() * sizeof(double)));
  EIF_HIP(hipMemcpy(c->ctab, C.data(), C.size() * sizeof(double), hipMemcpyHostToDevice));
  // the DCT-III pre-processing factors in thread order, one block of 2 nx
  // per register-FFT width R = 4, 8, 16: D[s T + t] = C[k] and
  // D[nx + s T + t] = C[nx - k] for k = bitrev(R t + s), T = nx / R
  if ((nx & (nx - 1)) == 0 && nx >= 16) {
    int logn = 0;
    while ((1 << logn) < nx) logn++;
    std::vector<double> D(3 * 2 * (size_t)nx);
    for (int lr = 2; lr <= 4; lr++) {
      const int R = 1 << lr, T = nx / R;
      double *blk = D.data() + (size_t)(lr - 2) * 2 * nx;
      for (int t = 0; t < T; t++)
        for (int s = 0; s < R; s++) {
          const unsigned v = (unsigned)(R * t + s);
          unsigned k = 0;
          for (int b = 0; b < logn; b++) k |= ((v >> b) & 1u) << (logn - 1 - b);
          blk[(size_t)s * T + t] = C[k];
          blk[(size_t)nx + (size_t)s * T + t] = C[nx - k];
        }
    }
    EIF_HIP(hipMalloc(&c->ctab3, D.size() * sizeof(double)));
    EIF_HIP(hipMemcpy(c->ctab3, D.data(), D.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  // per-stage twiddles of the radix-2 transform (copies of C entries)
  {
    std::vector<double> W(2 * (size_t)(nx > 1 ? nx - 1 : 1), 0.);
    if ((nx & (nx - 1)) == 0)
      for (int len = 2; len <= nx; len <<= 1)
        for (int k = 0; k < len / 2; k++) {
          const int m = k * (nx / len);
          const size_t e = (size_t)(len / 2 - 1 + k);
          W[2 * e] = C[4 * m];
          W[2 * e + 1] = C[nx >= 4 * m ? nx - 4 * m : 4 * m - nx];
        }
    EIF_HIP(hipMalloc(&c->twid, W.size() * sizeof(double)));
    EIF_HIP(hipMemcpy(c->twid, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice));
  }
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
  return part_init(c);
}

static int ilog2_exact(int n) {
  if (n < 2 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}

// lengths that are not powers of two: the Stockham FFT up to this many
// values per row (4 n doubles of LDS), the direct sum beyond
constexpr int DCT_MIXED_MAX = 5120;

// 2: register FFT; 1: radix-2 LDS FFT; 3: Stockham FFT; 0: direct sum
static int dct_mode(int nx) {
  const int logn = ilog2_exact(nx);
  if (logn >= 0) return (nx >= 64 && nx <= 8192) ? 2 : 1;
  return nx <= DCT_MIXED_MAX ? 3 : 0;
}

static DctPlan dct_plan(int n) {
  DctPlan pl{};
  pl.logn = ilog2_exact(n);
  pl.nr = 0;
  for (int p = 2; n > 1 && pl.nr < 24;) {
    if (n % p == 0) {
      pl.r[pl.nr++] = p;
      n /= p;
    } else {
      p++;
    }
  }
  return pl;
}

// values per thread (2^LOGR) of the register FFT of 2^LG-point rows: the
// default 8, or the tuner's 4 or 16 where the workgroup keeps 32..1024
// threads (every choice performs the same butterflies: same bits)
constexpr bool fft_logr_ok(int lg, int lr) {
  return lr == FFT_LOGR_DEFAULT || (lg - lr >= 5 && lg - lr <= 10);
}
// whether the context's rows use the register FFT, and whether 2^lr values
// per thread are available there (host, for the tuner)
bool fft_tunable(const Ctx *c) { return dct_mode(c->nx) == 2; }
bool fft_logr_valid(const Ctx *c, int lr) {
  return fft_tunable(c) && fft_logr_ok(ilog2_exact(c->nx), lr);
}
static int fft_logr(const Ctx *c, int lg) {
  const int lr = c->fft_logr;
  return (lr == 2 || lr == 4) && fft_logr_ok(lg, lr) ? lr : FFT_LOGR_DEFAULT;
}

// the register-FFT right-hand side of a power-of-two row length 2^LG
template <int LG, bool P2, bool ADD, int LR>
static void rhs_tr(Ctx *c, int rk, size_t shmem, const double *rhs_add) {
  if constexpr (fft_logr_ok(LG, LR))
    go<RhsDct2TB<LG, P2, ADD, LR>>(k_rhs_dct2_t<LG, P2, ADD, LR>, dim3((c->ny + 1) / 2),
                                   dim3(c->nx >> LR), shmem, c->stream, c->L, c->ux, c->uy,
                                   c->dscal, c->gamma[rk], c->dx, c->dy, c->ctab, c->twid,
                                   c->qhat, rhs_add);
}
template <int LG, bool P2, bool ADD>
static void rhs_t(Ctx *c, int rk, size_t shmem, const double *rhs_add) {
  switch (fft_logr(c, LG)) {
    case 2: rhs_tr<LG, P2, ADD, 2>(c, rk, shmem, rhs_add); break;
    case 4: rhs_tr<LG, P2, ADD, 4>(c, rk, shmem, rhs_add); break;
    default: rhs_tr<LG, P2, ADD, dct_logr<LG>()>(c, rk, shmem, rhs_add); break;
  }
}

// the block of ctab3 for register-FFT width 2^lr
static inline size_t dct3_table_offset(const Ctx *c, int lr) {
  return (size_t)(lr - 2) * 2 * c->nx;
}

// the register-FFT DCT-III of 2^LG-point rows into psi
template <int LG, int LR>
static void dct3_tr(Ctx *c, size_t shmem, const CombineArgs &ca) {
  if constexpr (fft_logr_ok(LG, LR))
    go<Dct3PsiTB<LG, LR>>(k_dct3_psi_t<LG, LR>, dim3((c->ny + 1) / 2), dim3(c->nx >> LR),
                          shmem, c->stream, c->L, c->qhat, ca, c->ctab, c->twid,
                          c->ctab3 + dct3_table_offset(c, LR), c->psi);
}
template <int LG>
static void dct3_t(Ctx *c, size_t shmem, const CombineArgs &ca) {
  switch (fft_logr(c, LG)) {
    case 2: dct3_tr<LG, 2>(c, shmem, ca); break;
    case 4: dct3_tr<LG, 4>(c, shmem, ca); break;
    default: dct3_tr<LG, dct_logr<LG>()>(c, shmem, ca); break;
  }
}

// the right-hand side and its DCT-II along x, then the solves along y;
// *ca: how the DCT-III stage combines the solves' results
static int poisson_lines(Ctx *c, int rk, const double *rhs_add, CombineArgs *out) {
  const int nx = c->nx, ny = c->ny;
  const int logn = ilog2_exact(nx);
  // register FFT for 64 <= nx <= 8192, LDS FFT for other powers of two,
  // the Stockham FFT for other lengths up to DCT_MIXED_MAX (its two buffer
  // pairs in LDS), the direct sum beyond
  const int mode = dct_mode(nx);
  const DctPlan pl = dct_plan(nx);
  const size_t shmem = (mode == 3 ? 4 : 2) * (size_t)nx * sizeof(double);
  const int threads = DCT_THREADS;  // of the one-row kernels (modes 0, 1)
  {
    ProfScope ps(c, "rhs_dct2");
#define EIF_RHS_T(LG)                                                                    \
  case LG:                                                                               \
    if (rhs_add)                                                                         \
      c->pow2_grid ? rhs_t<LG, true, true>(c, rk, shmem, rhs_add)                        \
                   : rhs_t<LG, false, true>(c, rk, shmem, rhs_add);                      \
    else                                                                                 \
      c->pow2_grid ? rhs_t<LG, true, false>(c, rk, shmem, rhs_add)                       \
                   : rhs_t<LG, false, false>(c, rk, shmem, rhs_add);                     \
    break;
    switch (mode == 2 ? logn : -1) {
      EIF_RHS_T(6) EIF_RHS_T(7) EIF_RHS_T(8) EIF_RHS_T(9) EIF_RHS_T(10) EIF_RHS_T(11)
      EIF_RHS_T(12) EIF_RHS_T(13)
      default:
        go<RhsDct2B>(k_rhs_dct2, dim3(ny), dim3(threads), shmem, c->stream, c->L, c->ux, c->uy,
                     c->dscal, c->gamma[rk], c->dx, c->dy, c->ctab, c->twid, pl, mode,
                     c->qhat, rhs_add);
    }
#undef EIF_RHS_T
    EIF_HIP(hipGetLastError());
  }
  const double l = 1. / c->dy / c->dy;
  CombineArgs ca{};
  ca.coef = l;
  ca.halt = c->halt;
  if (c->cfg.poisson_mode == EIF_POISSON_PARTITION && c->part_L) {
    ProfScope ps(c, "ysolve_partition");
    const int Lc = c->part_L, P = c->part_P;
    dim3 g1((nx + PART_LANES - 1) / PART_LANES, P);
    switch (Lc) {
      case 64: go<PartLocalB<64>>(k_part_local<64>, g1, dim3(PART_LANES), 0, c->stream, nx, c->qhat, c->pt_fwd, c->pt_u, l, c->pt_ends); break;
      case 32: go<PartLocalB<32>>(k_part_local<32>, g1, dim3(PART_LANES), 0, c->stream, nx, c->qhat, c->pt_fwd, c->pt_u, l, c->pt_ends); break;
      case 16: go<PartLocalB<16>>(k_part_local<16>, g1, dim3(PART_LANES), 0, c->stream, nx, c->qhat, c->pt_fwd, c->pt_u, l, c->pt_ends); break;
      default: go<PartLocalB<8>>(k_part_local<8>, g1, dim3(PART_LANES), 0, c->stream, nx, c->qhat, c->pt_fwd, c->pt_u, l, c->pt_ends); break;
    }
    if (P % 16 == 0 && P <= 64)  // the chunk DFTs on the matrix cores
      go<PartReducedMfmaB>(k_part_reduced_mfma, dim3((nx + PRM_K - 1) / PRM_K), dim3(4 * P),
                         (size_t)(2 + 6 * PRM_K) * P * sizeof(double), c->stream, nx, P,
                         c->pt_ends, c->pt_lvl, c->pt_red, c->pt_w, c->pt_bnd);
    else
      go<PartReducedB>(k_part_reduced, dim3((nx + PART_KG - 1) / PART_KG), dim3(64 * PART_KG),
                         (size_t)(PART_KG * 6 + 2) * P * sizeof(double), c->stream, nx, P,
                         c->pt_ends, c->pt_lvl, c->pt_red, c->pt_w, c->pt_bnd);
    EIF_HIP(hipGetLastError());
    ca.mode = 1;
    ca.L = Lc;
    ca.P = P;
    ca.v = c->pt_v;
    ca.bnd = c->pt_bnd;
  } else {
    ProfScope ps(c, "thomas");
    go<ThomasB>(k_thomas, dim3((nx + 63) / 64), dim3(64), 0, c->stream, nx, ny,
                       c->qhat, c->th_val, c->th_u, c->th_flag, c->th_den + nx,
                       c->th_den, l, c->th_last);
    EIF_HIP(hipGetLastError());
    ca.mode = 0;
    ca.q1 = c->th_q1;
    ca.last = c->th_last;
  }
  *out = ca;
  return EIF_OK;
}

int launch_poisson(Ctx *c, int rk, const double *rhs_add) {
  const int nx = c->nx, ny = c->ny;
  const int logn = ilog2_exact(nx);
  const int mode = dct_mode(nx);
  const DctPlan pl = dct_plan(nx);
  const size_t shmem = (mode == 3 ? 4 : 2) * (size_t)nx * sizeof(double);
  const int threads = DCT_THREADS;
  CombineArgs ca;
  TRY_P(poisson_lines(c, rk, rhs_add, &ca));
  bool halo_done = false;  // the register-FFT kernel writes psi's boundary values
  {
    ProfScope ps(c, "dct3_psi");
#define EIF_DCT3_T(LG)                                                                  \
  case LG:                                                                              \
    halo_done = true;                                                                   \
    dct3_t<LG>(c, shmem, ca);                                                           \
    break;
    switch (mode == 2 ? logn : -1) {
      EIF_DCT3_T(6) EIF_DCT3_T(7) EIF_DCT3_T(8) EIF_DCT3_T(9) EIF_DCT3_T(10) EIF_DCT3_T(11)
      EIF_DCT3_T(12) EIF_DCT3_T(13)
      default:
        go<Dct3PsiB>(k_dct3_psi, dim3(ny), dim3(threads), shmem, c->stream, c->L,
                           c->qhat, ca, c->ctab, c->twid, pl, mode, c->psi);
    }
#undef EIF_DCT3_T
    EIF_HIP(hipGetLastError());
  }
  return halo_done ? EIF_OK : launch_bc_p(c, c->psi);
}

// the fused DCT-III + projection + pressure update of 2^LG-point rows
template <int LG, bool P2, bool DT>
static void project_t(Ctx *c, int rk, const CombineArgs &ca, int S, double *ux, double *uy,
                      double *p) {
  go<ProjectTB<LG, P2, DT>>(k_project_t<LG, P2, DT>, dim3((c->ny + S - 1) / S),
                            dim3(c->nx >> dct_logr<LG>()), 2 * (size_t)c->nx * sizeof(double),
                            c->stream, c->L, c->qhat, ca, c->ctab, c->twid, S, ux, uy, p,
                            c->dxc, c->dscal, c->gamma[rk], c->dy, c->dt_part);
}

template <int LG>
static void project_lg(Ctx *c, int rk, const CombineArgs &ca, int S, bool dt, double *ux,
                       double *uy, double *p) {
  if (c->pow2_grid)
    dt ? project_t<LG, true, true>(c, rk, ca, S, ux, uy, p)
       : project_t<LG, true, false>(c, rk, ca, S, ux, uy, p);
  else
    dt ? project_t<LG, false, true>(c, rk, ca, S, ux, uy, p)
       : project_t<LG, false, false>(c, rk, ca, S, ux, uy, p);
}

// rows per strip of the fused kernel: 2, 4 or 8 (0: the unfused kernels,
// the default: on MI355X at 2048 x 4096 the fused kernel's extra transforms
// and its serial transform/stream phases cost more than the psi round trip
// it saves; the tuner times both)
int project_strip(const Ctx *c) { return c->proj_strip; }

bool project_fusable(const Ctx *c) {
  const int logn = ilog2_exact(c->nx);
  return logn >= 6 && logn <= 13 && c->ny >= 2;
}

int launch_project_on(Ctx *c, int rk, const CombineArgs &ca, bool dt_fused, double *ux,
                      double *uy, double *p) {
  ProfScope ps(c, "dct3_project");
  const int S = project_strip(c);
  const int parts = (c->ny + S - 1) / S;
  const bool dt = dt_fused && parts <= DT_PARTS_CAP;
  switch (ilog2_exact(c->nx)) {
    case 6: project_lg<6>(c, rk, ca, S, dt, ux, uy, p); break;
    case 7: project_lg<7>(c, rk, ca, S, dt, ux, uy, p); break;
    case 8: project_lg<8>(c, rk, ca, S, dt, ux, uy, p); break;
    case 9: project_lg<9>(c, rk, ca, S, dt, ux, uy, p); break;
    case 10: project_lg<10>(c, rk, ca, S, dt, ux, uy, p); break;
    case 11: project_lg<11>(c, rk, ca, S, dt, ux, uy, p); break;
    case 12: project_lg<12>(c, rk, ca, S, dt, ux, uy, p); break;
    case 13: project_lg<13>(c, rk, ca, S, dt, ux, uy, p); break;
    default:
      set_error("fused projection: nx must be a power of two in 64..8192");
      return EIF_ERR_ARG;
  }
  EIF_HIP(hipGetLastError());
  c->dt_adv_parts = dt ? parts : 0;
  return EIF_OK;
}

// compute_potential + correct_velocity + update_pressure of stage rk (the
// stage's whole projection), fused where the row length allows; psi is
// then not stored
int launch_poisson_project(Ctx *c, int rk, bool dt_fused) {
  if (c->proj_strip <= 0 || !project_fusable(c)) {
    TRY_P(launch_poisson(c, rk));
    return launch_correct_pressure(c, rk, true, true, dt_fused);
  }
  CombineArgs ca;
  TRY_P(poisson_lines(c, rk, nullptr, &ca));
  return launch_project_on(c, rk, ca, dt_fused, c->ux, c->uy, c->p);
}

// The tuner's choice between the unfused DCT-III + projection kernels and
// the fused kernel with strips of 2, 4 or 8 rows: each candidate's whole
// potential + projection + pressure update, 10 timed repetitions on the
// given copies of the fields (events without the system-scope fence).
// Every candidate gives the same psi bits in Thomas mode; in partition mode
// the fused kernel pairs consecutive rows (rounding only).  Appends
// "projection: unfused=..ms 2=..ms 4=..ms 8=..ms -> S; " (S = 0: unfused).
int project_tune(Ctx *c, double *ux, double *uy, double *p, std::string *rep) {
  if (!project_fusable(c)) {
    c->proj_strip = 0;
    *rep += "projection: unfused only; ";
    return EIF_OK;
  }
  hipEvent_t a, b;
  EIF_HIP(hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
  EIF_HIP(hipEventCreateWithFlags(&b, hipEventDisableSystemFence));
  double best = 1e30;
  int pick = 0, rc = EIF_OK;
  char buf[64];
  *rep += "projection:";
  auto once = [&](int S) -> int {
    if (S == 0) {
      TRY_P(launch_poisson(c, 1));
      return launch_correct_pressure_on(c, 1, true, true, true, ux, uy, p);
    }
    CombineArgs ca;
    TRY_P(poisson_lines(c, 1, nullptr, &ca));
    return launch_project_on(c, 1, ca, true, ux, uy, p);
  };
  for (int S : {0, 2, 4, 8}) {
    c->proj_strip = S;
    rc = once(S);  // untimed first run
    if (!rc) rc = hipEventRecord(a, c->stream) == hipSuccess ? EIF_OK : EIF_ERR_HIP;
    for (int k = 0; k < 10 && !rc; k++) rc = once(S);
    if (!rc) rc = hipEventRecord(b, c->stream) == hipSuccess ? EIF_OK : EIF_ERR_HIP;
    float el = 0.f;
    if (!rc && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&el, a, b) != hipSuccess))
      rc = hip_fail(hipErrorUnknown, "timing events", __FILE__, __LINE__);
    if (rc) break;
    const double ms = el / 10.;
    snprintf(buf, sizeof(buf), " %s=%.4fms", S ? std::to_string(S).c_str() : "unfused", ms);
    *rep += buf;
    if (ms < best) {
      best = ms;
      pick = S;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  c->proj_strip = pick;
  snprintf(buf, sizeof(buf), " -> %s; ", pick ? std::to_string(pick).c_str() : "unfused");
  *rep += buf;
  return rc;
}

// eif_gpu_set_projection: 0 (unfused) or a strip height of 2, 4 or 8
int set_projection(Ctx *c, int strip) {
  if (strip != 0 && strip != 2 && strip != 4 && strip != 8) {
    set_error("projection strip must be 0 (unfused), 2, 4 or 8");
    return EIF_ERR_ARG;
  }
  if (strip && !project_fusable(c)) {
    set_error("the fused projection needs nx a power of two in 64..8192 and ny >= 2");
    return EIF_ERR_ARG;
  }
  c->proj_strip = strip;
  return EIF_OK;
}

}  // namespace eif
