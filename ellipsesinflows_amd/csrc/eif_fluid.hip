// eif_fluid.hip -- boundary conditions, RK predictor, projection/pressure
// update, time-step size and diagnostics of the EllipsesInFlows fluid
// solver on gfx950.
//
// Arithmetic follows NaokiHori/EllipsesInFlows expression by expression
// (operand order, no contraction), so these kernels reproduce the
// reference's rounding:
//   boundary conditions  src/fluid/boundary_conditions.c:7-62
//   predictor            src/fluid/update_velocity.c:9-237
//   projection           src/fluid/correct_velocity.c:6-42
//   pressure             src/fluid/update_pressure.c:7-24
//   time-step size       src/param/decide_dt.c:11-81
//   diagnostics          src/logging.c:40-172
#include "eif_internal.hpp"
#include <mutex>

namespace eif {

// ---------------------------------------------------------------------
// boundary conditions (one process: the y-halo exchange is a periodic
// copy of rows ny -> 0 and 1 -> ny+1, src/parallel/halo.c:7-31)
// ---------------------------------------------------------------------
struct BcUxB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, Layout L, double *ux) {
  const int nx = L.nx, ny = L.ny;
  const int t = vb.bx * blockDim.x + threadIdx.x;
  if (t >= 2 && t <= nx) {  // halo rows, i = 2..nx
    ux[L.ux(t, 0)] = ux[L.ux(t, ny)];
    ux[L.ux(t, ny + 1)] = ux[L.ux(t, 1)];
  }
  if (t <= ny + 1) {  // impermeable walls
    ux[L.ux(1, t)] = 0.;
    ux[L.ux(nx + 1, t)] = 0.;
  }
}
};
__global__ void k_bc_ux(Layout L, double *ux) {
  BcUxB::run(VB::here(), L, ux);
}

struct BcUyB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, Layout L, double *uy, double wxm, double wxp) {
  const int nx = L.nx, ny = L.ny;
  const int t = vb.bx * blockDim.x + threadIdx.x;
  if (t >= 1 && t <= nx) {
    uy[L.p(t, 0)] = uy[L.p(t, ny)];
    uy[L.p(t, ny + 1)] = uy[L.p(t, 1)];
  }
  if (t <= ny + 1) {  // no-slip (possibly moving) walls
    uy[L.p(0, t)] = wxm;
    uy[L.p(nx + 1, t)] = wxp;
  }
}
};
__global__ void k_bc_uy(Layout L, double *uy, double wxm, double wxp) {
  BcUyB::run(VB::here(), L, uy, wxm, wxp);
}

// Neumann walls after the halo copy; the wall values of the halo rows
// read the interior rows they were copied from, so one pass suffices.
struct BcPB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, Layout L, double *p) {
  const int nx = L.nx, ny = L.ny;
  const int t = vb.bx * blockDim.x + threadIdx.x;
  if (t >= 1 && t <= nx) {
    p[L.p(t, 0)] = p[L.p(t, ny)];
    p[L.p(t, ny + 1)] = p[L.p(t, 1)];
  }
  if (t <= ny + 1) {
    const int src = t == 0 ? ny : (t == ny + 1 ? 1 : t);
    p[L.p(0, t)] = p[L.p(1, src)];
    p[L.p(nx + 1, t)] = p[L.p(nx, src)];
  }
}
};
__global__ void k_bc_p(Layout L, double *p) {
  BcPB::run(VB::here(), L, p);
}

static int bc_grid(const Ctx *c) {
  const int n = (c->nx + 2 > c->ny + 2 ? c->nx + 2 : c->ny + 2);
  return (n + 255) / 256;
}

int launch_bc_ux(Ctx *c, double *ux) {
  ProfScope ps(c, "bc");
  go<BcUxB>(k_bc_ux, dim3(bc_grid(c)), dim3(256), 0, c->stream, c->L, ux);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}
int launch_bc_uy(Ctx *c, double *uy) {
  ProfScope ps(c, "bc");
  go<BcUyB>(k_bc_uy, dim3(bc_grid(c)), dim3(256), 0, c->stream, c->L, uy,
                     c->cfg.uy_wall_xm, c->cfg.uy_wall_xp);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}
int launch_bc_p(Ctx *c, double *p) {
  ProfScope ps(c, "bc");
  go<BcPB>(k_bc_p, dim3(bc_grid(c)), dim3(256), 0, c->stream, c->L, p);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// ---------------------------------------------------------------------
// predictor: compute_src_ux/uy + update_ux/uy + boundary conditions,
// fused.  Each thread owns column i and marches over PRED_ROWS rows,
// keeping the three-row window of ux/uy in registers.  The explicit terms
// A of the previous stage (the beta term) come from one of two places:
//   PRED_STORE: srcuxa/srcuya, which this call overwrites with its own A
//               (the reference's data flow; reads ux, uy, p, srcA and
//               writes ux_new, uy_new, srcA);
//   PRED_PREV:  recomputed from the previous stage's velocity, which the
//               caller keeps (A depends on the velocity alone, so the
//               values are the same bits); srcA is neither read nor
//               written, which saves 2N values per stage.
// PRED_SRC only evaluates A of (ux, uy) into srcA: srcuxa/srcuya on demand
// after PRED_PREV stages.  PRED_NOB is PRED_PREV of a stage without a beta
// term (the first RK stage): no previous-velocity window, so fewer
// registers and more waves per SIMD.
// ---------------------------------------------------------------------
constexpr int PRED_BX = 256;
constexpr int PRED_ROWS = 16;  // default rows per workgroup (the tuner may pick 8 or 32)
enum { PRED_STORE = 0, PRED_PREV = 1, PRED_SRC = 2, PRED_NOB = 3 };
// bit MODE set: that mode loads a row ahead.  The row loop is unrolled
// over ROWS, so that the compiler's waits on the loads are per row, not
// for every load in flight at the loop's back edge.  Measured on the
// 2048x4096 grid (graph steps 498 -> 508 /s): the first stage's kernel
// (PRED_NOB, look-ahead, 112 VGPRs) 108 -> 89 us; PRED_PREV unrolled
// without look-ahead 135 -> 129 us (140 VGPRs, 3 waves per SIMD); with the
// look-ahead of both windows 135 us (178 VGPRs, 2 waves), of the current
// window alone 156 us
constexpr int PRED_LOOKAHEAD = 1 << PRED_NOB;

struct PredArgs {
  Layout L;
  const double *ux, *uy, *p;
  const double *uxo, *uyo;  // previous stage's velocity (PRED_PREV)
  double *uxn, *uyn, *sa_x, *sa_y;
  const double *dxf, *dxc;
  const double *dscal;
  double dy, Re, ext, al, be, ga, wxm, wxp;
  int use_b;
  const int *halt;
};

// the velocity stencil around (i, j): ux columns i-1, i, i+1 and uy
// columns i-1, i, i+1 on rows j-1 (m), j (c), j+1 (p), as far as the
// explicit terms read them
struct PredWin {
  double uxm_c, uxm_e, uxc_w, uxc_c, uxc_e, uxp_w, uxp_c, uxp_e;
  double uym_c, uyc_w, uyc_c, uyc_e, uyp_w, uyp_c, uyp_e;
};

__device__ inline void win_start(const Layout &L, const double *ux, const double *uy, int i,
                                 int j0, bool hx, PredWin &w) {
  w.uxm_c = ux[L.ux(i, j0 - 1)];
  w.uxm_e = ux[L.ux(i + 1, j0 - 1)];
  w.uxc_w = hx ? ux[L.ux(i - 1, j0)] : 0.;
  w.uxc_c = ux[L.ux(i, j0)];
  w.uxc_e = ux[L.ux(i + 1, j0)];
  w.uym_c = uy[L.p(i, j0 - 1)];
  w.uyc_w = uy[L.p(i - 1, j0)];
  w.uyc_c = uy[L.p(i, j0)];
  w.uyc_e = uy[L.p(i + 1, j0)];
}
__device__ inline void win_row(const Layout &L, const double *ux, const double *uy, int i,
                               int j, bool hx, PredWin &w) {
  w.uxp_w = hx ? ux[L.ux(i - 1, j + 1)] : 0.;
  w.uxp_c = ux[L.ux(i, j + 1)];
  w.uxp_e = ux[L.ux(i + 1, j + 1)];
  w.uyp_w = uy[L.p(i - 1, j + 1)];
  w.uyp_c = uy[L.p(i, j + 1)];
  w.uyp_e = uy[L.p(i + 1, j + 1)];
}
// one row's stencil values (ux and uy at columns i-1, i, i+1) loaded ahead
// of the row that reads them
struct PredRow {
  double uxw, uxc, uxe, uyw, uyc, uye;
};
__device__ inline void row_load(const Layout &L, const double *ux, const double *uy, int i,
                                int jr, bool hx, PredRow &r) {
  r.uxw = hx ? ux[L.ux(i - 1, jr)] : 0.;
  r.uxc = ux[L.ux(i, jr)];
  r.uxe = ux[L.ux(i + 1, jr)];
  r.uyw = uy[L.p(i - 1, jr)];
  r.uyc = uy[L.p(i, jr)];
  r.uye = uy[L.p(i + 1, jr)];
}
__device__ inline void row_into(const PredRow &r, PredWin &w) {
  w.uxp_w = r.uxw;
  w.uxp_c = r.uxc;
  w.uxp_e = r.uxe;
  w.uyp_w = r.uyw;
  w.uyp_c = r.uyc;
  w.uyp_e = r.uye;
}

__device__ inline void win_slide(PredWin &w) {
  w.uxm_c = w.uxc_c; w.uxm_e = w.uxc_e;
  w.uxc_w = w.uxp_w; w.uxc_c = w.uxp_c; w.uxc_e = w.uxp_e;
  w.uym_c = w.uyc_c;
  w.uyc_w = w.uyp_w; w.uyc_c = w.uyp_c; w.uyc_e = w.uyp_e;
}

// grid factors of column i
struct PredCol {
  double dxf_im, dxf_i, dxc_i, dxc_ip, idxf_im, idxf_i, idxc_i, idxc_ip, dy, idy, rre;
};

// explicit terms of the x-momentum at face (i, j), src/fluid/update_velocity.c:28-88
template <bool P2>
__device__ inline double src_x(const PredWin &w, const PredCol &g) {
  const double g_w = qdiv<P2>(-w.uxc_w + w.uxc_c, g.dxf_im, g.idxf_im);
  const double g_e = qdiv<P2>(-w.uxc_c + w.uxc_e, g.dxf_i, g.idxf_i);
  const double g_s = qdiv<P2>(-w.uxm_c + w.uxc_c, g.dy, g.idy);
  const double g_n = qdiv<P2>(-w.uxc_c + w.uxp_c, g.dy, g.idy);
  const double cw = g.dxf_im / (2. * g.dxc_i);
  const double ce = g.dxf_i / (2. * g.dxc_i);
  const double aw = 0.5 * w.uxc_w + 0.5 * w.uxc_c;
  const double ae = 0.5 * w.uxc_c + 0.5 * w.uxc_e;
  const double adv1 = -cw * aw * g_w - ce * ae * g_e;
  const double vs = cw * w.uyc_w + ce * w.uyc_c;
  const double vn = cw * w.uyp_w + ce * w.uyp_c;
  const double adv2 = -0.5 * vs * g_s - 0.5 * vn * g_n;
  const double dif1 = qdiv<P2>(g.rre * (g_e - g_w), g.dxc_i, g.idxc_i);
  const double dif2 = qdiv<P2>(g.rre * (g_n - g_s), g.dy, g.idy);
  return (adv1 + adv2) + (dif1 + dif2);
}

// explicit terms of the y-momentum at face (i, j), src/fluid/update_velocity.c:112-171
template <bool P2>
__device__ inline double src_y(const PredWin &w, const PredCol &g, double ext) {
  const double g_w = qdiv<P2>(-w.uyc_w + w.uyc_c, g.dxc_i, g.idxc_i);
  const double g_e = qdiv<P2>(-w.uyc_c + w.uyc_e, g.dxc_ip, g.idxc_ip);
  const double g_s = qdiv<P2>(-w.uym_c + w.uyc_c, g.dy, g.idy);
  const double g_n = qdiv<P2>(-w.uyc_c + w.uyp_c, g.dy, g.idy);
  const double cw = g.dxc_i / (2. * g.dxf_i);
  const double ce = g.dxc_ip / (2. * g.dxf_i);
  const double uw = 0.5 * w.uxm_c + 0.5 * w.uxc_c;
  const double ue = 0.5 * w.uxm_e + 0.5 * w.uxc_e;
  const double adv1 = -cw * uw * g_w - ce * ue * g_e;
  const double as = 0.5 * w.uym_c + 0.5 * w.uyc_c;
  const double an = 0.5 * w.uyc_c + 0.5 * w.uyp_c;
  const double adv2 = -0.5 * as * g_s - 0.5 * an * g_n;
  const double dif1 = qdiv<P2>(g.rre * (g_e - g_w), g.dxf_i, g.idxf_i);
  const double dif2 = qdiv<P2>(g.rre * (g_n - g_s), g.dy, g.idy);
  return (adv1 + adv2) + (dif1 + dif2) + ext;
}

template <bool P2, int MODE, int ROWS = PRED_ROWS>
struct PredictorB {
  static constexpr int T = PRED_BX, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, PredArgs a) {
  if (halted(a.halt)) return;
  const Layout L = a.L;
  const int nx = L.nx, ny = L.ny;
  const int i = vb.bx * PRED_BX + threadIdx.x + 1;  // 1..nx
  const int j0 = vb.by * ROWS + 1;
  if (i > nx || j0 > ny) return;
  const int j1 = min(j0 + ROWS - 1, ny);
  const double dt = a.dscal[S_DT];
  const double aldt = a.al * dt, bedt = a.be * dt, gadt = a.ga * dt;
  const bool hx = i >= 2;  // ux(i) is an unknown for i = 2..nx
  PredCol g;
  g.dy = a.dy;
  g.rre = 1. / a.Re;
  g.dxf_im = hx ? a.dxf[i - 2] : 0.;  // DXF(i-1)
  g.dxf_i = a.dxf[i - 1];             // DXF(i)
  g.dxc_i = a.dxc[i - 1];             // DXC(i)
  g.dxc_ip = a.dxc[i];                // DXC(i+1)
  g.idxf_im = hx ? 1. / g.dxf_im : 0.;
  g.idxf_i = 1. / g.dxf_i;
  g.idxc_i = 1. / g.dxc_i;
  g.idxc_ip = 1. / g.dxc_ip;
  g.idy = 1. / g.dy;
  const bool prev = MODE == PRED_PREV && a.use_b;
  const double *p = a.p;
  // LA: the next row's loads are issued a row ahead, so that each wave
  // keeps two rows of loads in flight (the march waits on its loads more
  // than on its arithmetic)
  constexpr bool LA = (PRED_LOOKAHEAD >> MODE) & 1;
  PredWin w, wo;
  win_start(L, a.ux, a.uy, i, j0, hx, w);
  if (prev) win_start(L, a.uxo, a.uyo, i, j0, hx, wo);
  double pm_c = MODE == PRED_SRC ? 0. : p[L.p(i, j0 - 1)];
  // one row's update from the window (rows j-1, j, j+1) and p(i-1, j), p(i, j)
  auto row = [&](int j, double pc_w, double pc_c) {
    // ---- x-momentum at face (i, j)
    if (hx) {
      const double A = src_x<P2>(w, g);
      const long si = L.sx(i, j);
      if (MODE == PRED_SRC) {
        a.sa_x[si] = A;
      } else {
        const double pre = qdiv<P2>(-(pc_c - pc_w), g.dxc_i, g.idxc_i);
        double bterm = 0.;
        if (MODE == PRED_STORE) {
          bterm = a.use_b ? bedt * a.sa_x[si] : 0.;
          a.sa_x[si] = A;
        } else if (prev) {
          bterm = bedt * src_x<P2>(wo, g);
        }
        const double v = w.uxc_c + (aldt * A + bterm + gadt * pre);
        a.uxn[L.ux(i, j)] = v;
        if (j == ny) a.uxn[L.ux(i, 0)] = v;
        if (j == 1) a.uxn[L.ux(i, ny + 1)] = v;
      }
    }
    // ---- y-momentum at face (i, j)
    {
      const double A = src_y<P2>(w, g, a.ext);
      const long si = L.sy(i, j);
      if (MODE == PRED_SRC) {
        a.sa_y[si] = A;
      } else {
        const double pre = qdiv<P2>(-(pc_c - pm_c), g.dy, g.idy);
        double bterm = 0.;
        if (MODE == PRED_STORE) {
          bterm = a.use_b ? bedt * a.sa_y[si] : 0.;
          a.sa_y[si] = A;
        } else if (prev) {
          bterm = bedt * src_y<P2>(wo, g, a.ext);
        }
        const double v = w.uyc_c + (aldt * A + bterm + gadt * pre);
        a.uyn[L.p(i, j)] = v;
        if (j == ny) a.uyn[L.p(i, 0)] = v;
        if (j == 1) a.uyn[L.p(i, ny + 1)] = v;
      }
    }
    // walls of the new fields (boundary_conditions.c:19-22, :38-41)
    if (MODE != PRED_SRC && i == 1) {
      a.uxn[L.ux(1, j)] = 0.;
      a.uyn[L.p(0, j)] = a.wxm;
      if (j == 1) {
        a.uxn[L.ux(1, 0)] = 0.;
        a.uxn[L.ux(1, ny + 1)] = 0.;
        a.uyn[L.p(0, 0)] = a.wxm;
        a.uyn[L.p(0, ny + 1)] = a.wxm;
      }
    }
    if (MODE != PRED_SRC && i == nx) {
      a.uxn[L.ux(nx + 1, j)] = 0.;
      a.uyn[L.p(nx + 1, j)] = a.wxp;
      if (j == 1) {
        a.uxn[L.ux(nx + 1, 0)] = 0.;
        a.uxn[L.ux(nx + 1, ny + 1)] = 0.;
        a.uyn[L.p(nx + 1, 0)] = a.wxp;
        a.uyn[L.p(nx + 1, ny + 1)] = a.wxp;
      }
    }
    win_slide(w);
    if (prev) win_slide(wo);
    pm_c = pc_c;
  };
  if (!LA) {
#pragma unroll
    for (int jj = 0; jj < ROWS; jj++) {
      const int j = j0 + jj;
      if (j > j1) break;
      win_row(L, a.ux, a.uy, i, j, hx, w);
      if (prev) win_row(L, a.uxo, a.uyo, i, j, hx, wo);
      double pc_w = 0., pc_c = 0.;
      if (MODE != PRED_SRC) {
        pc_w = p[L.p(i - 1, j)];
        pc_c = p[L.p(i, j)];
      }
      row(j, pc_w, pc_c);
    }
    return;
  }
  // LA: rows in pairs with two register sets (A, B) taking turns.  Each
  // half issues the loads of the row after next into the idle set (rows
  // j + 2 of the velocity and j + 1 of p, clamped to the last halo row:
  // values loaded past the march are not used), then moves its own set,
  // loaded a row earlier, into the window.  No set is copied into the
  // other, so a row's loads are first waited for a row after they issue.
  PredRow ra, rao, rb, rbo;
  double pa_w = 0., pa_c = 0., pb_w = 0., pb_c = 0.;
  auto fetch = [&](int jv, int jp, PredRow &r, PredRow &ro, double &pw, double &pc) {
    row_load(L, a.ux, a.uy, i, min(jv, ny + 1), hx, r);
    if (prev) row_load(L, a.uxo, a.uyo, i, min(jv, ny + 1), hx, ro);
    if (MODE != PRED_SRC) {
      pw = p[L.p(i - 1, min(jp, ny + 1))];
      pc = p[L.p(i, min(jp, ny + 1))];
    }
  };
  fetch(j0 + 1, j0, ra, rao, pa_w, pa_c);
#pragma unroll
  for (int jj = 0; jj < ROWS; jj += 2) {
    const int j = j0 + jj;
    if (j > j1) break;
    fetch(j + 2, j + 1, rb, rbo, pb_w, pb_c);
    row_into(ra, w);
    if (prev) row_into(rao, wo);
    row(j, pa_w, pa_c);
    if (j + 1 > j1) break;
    fetch(j + 3, j + 2, ra, rao, pa_w, pa_c);
    row_into(rb, w);
    if (prev) row_into(rbo, wo);
    row(j + 1, pb_w, pb_c);
  }
}
};
template <bool P2, int MODE, int ROWS = PRED_ROWS>
__global__ __launch_bounds__(PRED_BX) void k_predictor(PredArgs a) {
  PredictorB<P2, MODE, ROWS>::run(VB::here(), a);
}

template <int MODE, int ROWS>
static void pred_launch_rows(Ctx *c, const PredArgs &a, hipStream_t st) {
  if constexpr (MODE == PRED_PREV) {
    if (!a.use_b) {  // the same operations without the unused window
      pred_launch_rows<PRED_NOB, ROWS>(c, a, st);
      return;
    }
  }
  dim3 grid((c->nx + PRED_BX - 1) / PRED_BX, (c->ny + ROWS - 1) / ROWS);
  if (c->pow2_grid)
    go<PredictorB<true, MODE, ROWS>>(k_predictor<true, MODE, ROWS>, grid, dim3(PRED_BX), 0, st, a);
  else
    go<PredictorB<false, MODE, ROWS>>(k_predictor<false, MODE, ROWS>, grid, dim3(PRED_BX), 0, st, a);
}

// rows per workgroup from the context (tuned, or the default); every
// choice performs the same operations per cell
template <int MODE>
static void pred_launch(Ctx *c, const PredArgs &a, hipStream_t st) {
  switch (c->pred_rows) {
    case 8: pred_launch_rows<MODE, 8>(c, a, st); break;
    case 32: pred_launch_rows<MODE, 32>(c, a, st); break;
    default: pred_launch_rows<MODE, PRED_ROWS>(c, a, st); break;
  }
}


static PredArgs pred_args(Ctx *c, int rk) {
  PredArgs a;
  a.L = c->L;
  a.ux = c->ux; a.uy = c->uy; a.p = c->p;
  a.uxo = nullptr; a.uyo = nullptr;
  a.uxn = nullptr; a.uyn = nullptr;
  a.sa_x = c->srcuxa; a.sa_y = c->srcuya;
  a.dxf = c->dxf; a.dxc = c->dxc;
  a.dscal = c->dscal;
  a.dy = c->dy; a.Re = c->cfg.Re; a.ext = c->cfg.extfrcy;
  a.al = c->alpha[rk]; a.be = c->beta[rk]; a.ga = c->gamma[rk];
  a.wxm = c->cfg.uy_wall_xm; a.wxp = c->cfg.uy_wall_xp;
  a.use_b = c->beta[rk] != 0.;
  a.halt = c->halt;
  return a;
}

// one predictor of a stage > 0 (the beta term from the kept pair) into the
// spare pair, pointers untouched: the tuner's timing launch
int launch_predictor_probe(Ctx *c) {
  PredArgs a = pred_args(c, 1);
  a.uxo = c->ux_alt;
  a.uyo = c->uy_alt;
  a.uxn = c->ux_spare;
  a.uyn = c->uy_spare;
  pred_launch<PRED_PREV>(c, a, c->stream);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// srcuxa/srcuya of the velocity they were last evaluated at, if the
// PRED_PREV stages left them unevaluated
int launch_materialise_src(Ctx *c) {
  if (!c->src_lazy) return EIF_OK;
  PredArgs a = pred_args(c, 0);
  a.ux = c->ux_alt;  // the last stage's start velocity
  a.uy = c->uy_alt;
  pred_launch<PRED_SRC>(c, a, c->stream);
  EIF_HIP(hipGetLastError());
  c->src_lazy = false;
  return EIF_OK;
}

int launch_predictor(Ctx *c, int rk, hipStream_t st) {
  if (!st) st = c->stream;
  ProfScope ps(c, "predictor", st);
  if (c->beta[rk] != 0.) {
    const int rc = launch_materialise_src(c);
    if (rc) return rc;
  }
  PredArgs a = pred_args(c, rk);
  a.uxn = c->ux_alt; a.uyn = c->uy_alt;
  pred_launch<PRED_STORE>(c, a, st);
  EIF_HIP(hipGetLastError());
  c->src_lazy = false;
  std::swap(c->ux, c->ux_alt);
  std::swap(c->uy, c->uy_alt);
  return EIF_OK;
}

// The predictor of a stage of eif_gpu_integrate: the beta term from the
// previous stage's velocity (ux_alt, kept since that stage's predictor),
// the new velocity into the spare pair.  Afterwards ux holds the new
// velocity, ux_alt this stage's start velocity and the spare the previous
// one (free).  srcuxa/srcuya are evaluated only if read.
int launch_predictor_prev(Ctx *c, int rk, hipStream_t st) {
  if (!st) st = c->stream;
  ProfScope ps(c, "predictor", st);
  PredArgs a = pred_args(c, rk);
  a.uxo = c->ux_alt; a.uyo = c->uy_alt;
  a.uxn = c->ux_spare; a.uyn = c->uy_spare;
  pred_launch<PRED_PREV>(c, a, st);
  EIF_HIP(hipGetLastError());
  c->src_lazy = true;
  double *ox = c->ux_alt, *oy = c->uy_alt;
  c->ux_alt = c->ux; c->uy_alt = c->uy;
  c->ux = c->ux_spare; c->uy = c->uy_spare;
  c->ux_spare = ox; c->uy_spare = oy;
  return EIF_OK;
}

// ---------------------------------------------------------------------
// projection + pressure update, fused, in place
// (correct_velocity.c:16-41, update_pressure.c:14-23) with the boundary
// values of ux, uy and p refreshed by the threads that own the source
// interior values.
// ---------------------------------------------------------------------
constexpr int CP_ROWS = 8;   // default rows per workgroup of k_correct_pressure

template <bool P2, bool DT, int ROWS = CP_ROWS>
struct CorrectPressureB {
  static constexpr int T = 256, M = 0;
  // ux, uy, p and psi are distinct arrays (restrict): the loads of a row
  // need not wait for the stores of the field before them
  static __device__ __forceinline__ void run(const VB &vb, 
    Layout L, double *__restrict__ ux, double *__restrict__ uy, double *__restrict__ p,
    const double *__restrict__ psi, const double *__restrict__ dxc,
    const double *dscal, double ga, double dy, int do_correct, int do_pressure,
    double *dt_part, const int *halt) {
  if (halted(halt)) return;
  const int nx = L.nx, ny = L.ny;
  const int i = vb.bx * blockDim.x + threadIdx.x + 1;
  const double dt = dscal[S_DT];
  // with DT: the advective time-step limit of decide_dt.c:26-48 over the
  // corrected velocities, as one partial minimum per workgroup
  double m = 5.e-2;
  // rows in chunks of up to 4: the chunk's loads first (threads past nx
  // and rows past ny load clamped addresses inside the arrays, and do not
  // use them), then its updates and stores
  constexpr int CH = ROWS < 4 ? ROWS : 4;
  const int ic = min(i, nx);
  const double dxc_i = dxc[ic - 1], dxi = 1. / dxc_i;
#pragma unroll
  for (int r0 = 0; r0 < ROWS; r0 += CH) {
    double psc[CH], psw[CH], pss[CH], uxv[CH], uyv[CH], pv[CH];
#pragma unroll
    for (int q = 0; q < CH; q++) {
      const int jc = min((int)vb.by * ROWS + 1 + r0 + q, ny);
      psc[q] = psi[L.p(ic, jc)];
      psw[q] = pss[q] = uxv[q] = uyv[q] = pv[q] = 0.;
      if (do_correct) {
        psw[q] = psi[L.p(ic - 1, jc)];
        pss[q] = psi[L.p(ic, jc - 1)];
        uxv[q] = ux[L.ux(ic, jc)];
        uyv[q] = uy[L.p(ic, jc)];
      }
      if (do_pressure) pv[q] = p[L.p(ic, jc)];
    }
#pragma unroll
    for (int q = 0; q < CH; q++) {
      const int j = vb.by * ROWS + 1 + r0 + q;
      if (j > ny || i > nx) break;
      const double ps_c = psc[q];
      if (do_correct) {
        if (i >= 2) {
          const long k = L.ux(i, j);
          const double v = uxv[q] - ga * dt * (ps_c - psw[q]) * dxi;
          ux[k] = v;
          if (j == ny) ux[L.ux(i, 0)] = v;
          if (j == 1) ux[L.ux(i, ny + 1)] = v;
          if (DT) m = fmin(m, dxc_i / fmax(fabs(v), 1.e-8));
        }
        {
          const long k = L.p(i, j);
          const double v = uyv[q] - qdiv<P2>(ga * dt * (ps_c - pss[q]), dy, 1. / dy);
          uy[k] = v;
          if (j == ny) uy[L.p(i, 0)] = v;
          if (j == 1) uy[L.p(i, ny + 1)] = v;
          if (DT) m = fmin(m, dy / fmax(fabs(v), 1.e-8));
        }
      }
      if (do_pressure) {
        const long k = L.p(i, j);
        const double v = pv[q] + ps_c;
        p[k] = v;
        if (j == ny) p[L.p(i, 0)] = v;
        if (j == 1) p[L.p(i, ny + 1)] = v;
        if (i == 1) {
          p[L.p(0, j)] = v;
          if (j == ny) p[L.p(0, 0)] = v;
          if (j == 1) p[L.p(0, ny + 1)] = v;
        }
        if (i == nx) {
          p[L.p(nx + 1, j)] = v;
          if (j == ny) p[L.p(nx + 1, 0)] = v;
          if (j == 1) p[L.p(nx + 1, ny + 1)] = v;
        }
      }
    }
  }
  if (DT) {
    __shared__ double red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) dt_part[vb.by * vb.gx + vb.bx] = red[0];
  }
}
};
template <bool P2, bool DT, int ROWS = CP_ROWS>
__global__ __launch_bounds__(256) void k_correct_pressure(
    Layout L, double *ux, double *uy, double *p, const double *psi, const double *dxc,
    const double *dscal, double ga, double dy, int do_correct, int do_pressure,
    double *dt_part, const int *halt) {
  CorrectPressureB<P2, DT, ROWS>::run(VB::here(), L, ux, uy, p, psi, dxc, dscal, ga, dy, do_correct, do_pressure, dt_part, halt);
}

// a projection kernel and its body, as one type for the launcher below
template <bool P2, bool DT, int ROWS>
struct CpTag {
  using type = CorrectPressureB<P2, DT, ROWS>;
  static constexpr auto kern = &k_correct_pressure<P2, DT, ROWS>;
};

template <int ROWS>
static int cp_launch(Ctx *c, int rk, bool correct, bool pressure, bool dt_fused, double *ux,
                     double *uy, double *p) {
  dim3 grid((c->nx + 255) / 256, (c->ny + ROWS - 1) / ROWS);
  const bool dt = dt_fused && correct && (long)grid.x * grid.y <= DT_PARTS_CAP;
  auto run = [&](auto kern_tag) {
    using B = typename decltype(kern_tag)::type;
    go<B>(decltype(kern_tag)::kern, grid, dim3(256), 0, c->stream, c->L, ux, uy, p, c->psi,
          c->dxc, c->dscal, c->gamma[rk], c->dy, (int)correct, (int)pressure, c->dt_part,
          c->halt);
  };
  if (c->pow2_grid) {
    if (dt) run(CpTag<true, true, ROWS>{}); else run(CpTag<true, false, ROWS>{});
  } else {
    if (dt) run(CpTag<false, true, ROWS>{}); else run(CpTag<false, false, ROWS>{});
  }
  EIF_HIP(hipGetLastError());
  c->dt_adv_parts = dt ? (int)(grid.x * grid.y) : 0;
  return EIF_OK;
}

// on explicit buffers (the tuner's timing launch uses copies)
int launch_correct_pressure_on(Ctx *c, int rk, bool correct, bool pressure, bool dt_fused,
                               double *ux, double *uy, double *p) {
  ProfScope ps(c, "correct_pressure");
  switch (c->cp_rows) {
    case 4: return cp_launch<4>(c, rk, correct, pressure, dt_fused, ux, uy, p);
    case 16: return cp_launch<16>(c, rk, correct, pressure, dt_fused, ux, uy, p);
    default: return cp_launch<CP_ROWS>(c, rk, correct, pressure, dt_fused, ux, uy, p);
  }
}

int launch_correct_pressure(Ctx *c, int rk, bool correct, bool pressure, bool dt_fused) {
  return launch_correct_pressure_on(c, rk, correct, pressure, dt_fused, c->ux, c->uy, c->p);
}

// ---------------------------------------------------------------------
// time-step size, src/param/decide_dt.c:11-81.  Every candidate is a
// positive finite double, so the IEEE bit pattern orders like the value
// and an integer atomicMin gives the exact minimum in any order.
// ---------------------------------------------------------------------
__device__ inline void atomic_min_pos(double *addr, double v) {
  atomicMin(reinterpret_cast<unsigned long long *>(addr),
            (unsigned long long)__double_as_longlong(v));
}
__device__ inline void atomic_max_pos(double *addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long *>(addr),
            (unsigned long long)__double_as_longlong(v));
}

struct DtInitB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, double *dred) { dred[0] = 5.e-2; }
};
__global__ void k_dt_init(double *dred) {
  DtInitB::run(VB::here(), dred);
}

struct DtAdvB {
  static constexpr int T = 256, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, Layout L, const double *ux,
                                                const double *uy,
                                                const double *dxc, double dy,
                                                double *dred) {
  const int nx = L.nx, ny = L.ny;
  const double small = 1.e-8;
  double m = 5.e-2;
  const long n = (long)nx * ny;
  for (long t = vb.bx * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)vb.gx * blockDim.x) {
    const int i = (int)(t % nx) + 1, j = (int)(t / nx) + 1;
    if (i >= 2) {
      const double vel = fmax(fabs(ux[L.ux(i, j)]), small);
      m = fmin(m, dxc[i - 1] / vel);
    }
    const double vel = fmax(fabs(uy[L.p(i, j)]), small);
    m = fmin(m, dy / vel);
  }
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomic_min_pos(dred, red[0]);
}
};
__global__ __launch_bounds__(256) void k_dt_adv(Layout L, const double *ux,
                                                const double *uy,
                                                const double *dxc, double dy,
                                                double *dred) {
  DtAdvB::run(VB::here(), L, ux, uy, dxc, dy, dred);
}

// nparts > 0: the advective limit comes as partial minima part[0..nparts)
// from the projection kernel of the previous step's last stage
struct DtFinalB {
  static constexpr int T = 1024, M = 0;
  static __device__ __forceinline__ void run(const VB &vb, const double *dred, const double *part, const double *prt, int np,
                           double sf0, double sf2, double dt_dif_x,
                           double dt_dif_y, double delta, double *dscal, int nparts,
                           const int *halt) {
  if (halted(halt)) return;
  __shared__ double red[256];
  __shared__ double adv;
  if (nparts > 0) {
    double m = 5.e-2;
    // eight loads in flight per round; an index past the end reads the
    // last partial again, which leaves the minimum unchanged
    for (int b0 = threadIdx.x; b0 < nparts; b0 += 8 * blockDim.x) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = part[min(b0 + u * (int)blockDim.x, nparts - 1)];
#pragma unroll
      for (int u = 0; u < 8; u++) m = fmin(m, v[u]);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) adv = red[0];
    __syncthreads();
  } else if (threadIdx.x == 0) {
    adv = dred[0];
  }
  __syncthreads();
  const double small = 1.e-8;
  double dt_par = 1.7976931348623157e308;
  for (int n = threadIdx.x; n < np; n += blockDim.x) {
    const double *q = prt + (long)n * R_N;
    const double pux = fmax(fabs(q[R_UX]) + fabs(fmax(q[R_A], q[R_B]) * q[R_VZ]), small);
    const double puy = fmax(fabs(q[R_UY]) + fabs(fmax(q[R_A], q[R_B]) * q[R_VZ]), small);
    const double val = fmax(pux, puy);
    dt_par = fmin(dt_par, delta / val);
  }
  red[threadIdx.x] = dt_par;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double dt_adv = adv;
    dt_adv *= sf0;
    double dtp = red[0];
    dtp *= sf2;
    dscal[S_DT] = fmin(fmin(dt_adv, fmin(dt_dif_x, dt_dif_y)), dtp);
  }
}
};
__global__ void k_dt_final(const double *dred, const double *part, const double *prt, int np,
                           double sf0, double sf2, double dt_dif_x,
                           double dt_dif_y, double delta, double *dscal, int nparts,
                           const int *halt) {
  DtFinalB::run(VB::here(), dred, part, prt, np, sf0, sf2, dt_dif_x, dt_dif_y, delta, dscal, nparts, halt);
}

int launch_decide_dt(Ctx *c) {
  ProfScope ps(c, "decide_dt");
  const int nparts = c->dt_adv_valid ? c->dt_adv_parts : 0;
  c->dt_adv_valid = false;
  if (nparts == 0) {
    go<DtInitB>(k_dt_init, dim3(1), dim3(1), 0, c->stream, c->dred);
    const long n = (long)c->nx * c->ny;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    go<DtAdvB>(k_dt_adv, dim3(blocks), dim3(256), 0, c->stream, c->L, c->ux,
                       c->uy, c->dxc, c->dy, c->dred);
  }
  // diffusive limits are grid constants (decide_dt.c:50-62)
  double dxmin = c->lx;
  for (int i = 2; i <= c->nx; i++) dxmin = fmin(dxmin, c->h_dxc[i - 1]);
  double dt_dif_x = 0.25 * c->cfg.Re * (dxmin * dxmin);
  double dt_dif_y = 0.25 * c->cfg.Re * (c->dy * c->dy);
  dt_dif_x *= c->cfg.safefactors[1];
  dt_dif_y *= c->cfg.safefactors[1];
  const double delta = fmin(c->dx, c->dy);
  go<DtFinalB>(k_dt_final, dim3(1), dim3(256), 0, c->stream, c->dred, c->dt_part, c->prt,
                     c->np, c->cfg.safefactors[0], c->cfg.safefactors[2], dt_dif_x,
                     dt_dif_y, delta, c->dscal, nparts, c->halt);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// ---------------------------------------------------------------------
// diagnostics of src/logging.c (block sums in a fixed tree, then one
// ordered pass over the block partials)
// ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_diag(Layout L, const double *ux,
                                              const double *uy,
                                              const double *dxf,
                                              const double *dxc, double dy,
                                              int which, double *part) {
  __shared__ double s0[256], s1[256];
  const int nx = L.nx, ny = L.ny;
  double a0 = 0., a1 = 0.;
  const long n = (long)(nx + 1) * ny;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % (nx + 1)) + 1, j = (int)(t / (nx + 1)) + 1;
    if (which == 0) {  // divergence max / sum over i = 1..nx
      if (i <= nx) {
        const double div = (ux[L.ux(i + 1, j)] - ux[L.ux(i, j)]) / dxf[i - 1]
                           + (uy[L.p(i, j + 1)] - uy[L.p(i, j)]) / dy;
        a0 = fmax(a0, fabs(div));
        a1 += div;
      }
    } else if (which == 1) {  // momentum
      a0 += ux[L.ux(i, j)] * (dxc[i - 1] * dy);
      if (i <= nx) a1 += uy[L.p(i, j)] * (dxf[i - 1] * dy);
    } else {  // energy
      const double u = ux[L.ux(i, j)];
      a0 += 0.5 * (u * u) * (dxc[i - 1] * dy);
      if (i <= nx) {
        const double v = uy[L.p(i, j)];
        a1 += 0.5 * (v * v) * (dxf[i - 1] * dy);
      }
    }
  }
  s0[threadIdx.x] = a0;
  s1[threadIdx.x] = a1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s0[threadIdx.x] = which == 0 ? fmax(s0[threadIdx.x], s0[threadIdx.x + s])
                                   : s0[threadIdx.x] + s0[threadIdx.x + s];
      s1[threadIdx.x] += s1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

__global__ void k_diag_final(const double *part, int nb, int which, double *out) {
  if (threadIdx.x != 0) return;
  double a0 = 0., a1 = 0.;
  for (int b = 0; b < nb; b++) {
    a0 = which == 0 ? fmax(a0, part[2 * b]) : a0 + part[2 * b];
    a1 += part[2 * b + 1];
  }
  out[0] = a0;
  out[1] = a1;
}

// Partial sums of the no-allocation diagnostics (fluid_gpu_diagnostics_noalloc):
// a file-scope device array instead of context memory, so that call
// allocates nothing.  Each device holds its own copy; host code uses only
// the per-device address from hipGetSymbolAddress, never this name.
static __device__ __align__(256) double scratch[2 * DIAG_SCRATCH_BLOCKS];

int diag_scratch_address(int device, double **out) {
  static std::mutex mtx;
  static std::map<int, double *> addr;
  std::lock_guard<std::mutex> lock(mtx);
  auto it = addr.find(device);
  if (it != addr.end()) {
    *out = it->second;
    return EIF_OK;
  }
  int cur = -1;
  EIF_HIP(hipGetDevice(&cur));
  if (cur != device) {
    set_error("diagnostics scratch: the context's device is not current");
    return EIF_ERR_STATE;
  }
  void *ptr = nullptr;
  EIF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  addr[device] = static_cast<double *>(ptr);
  *out = static_cast<double *>(ptr);
  return EIF_OK;
}

int launch_diagnostics_noalloc(Ctx *c, int which, double *out_dev) {
  const int nb = c->nred;
  if (nb > DIAG_SCRATCH_BLOCKS) {
    set_error("diagnostics: more partial sums than the static scratch holds");
    return EIF_ERR_CAPACITY;
  }
  if (!c->diag_scratch) {
    set_error("diagnostics scratch address not resolved");
    return EIF_ERR_STATE;
  }
  // every partial is written before it is read: nothing to clear
  hipLaunchKernelGGL(k_diag, dim3(nb), dim3(256), 0, c->stream, c->L, c->ux, c->uy,
                     c->dxf, c->dxc, c->dy, which, c->diag_scratch);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, c->stream, c->diag_scratch, nb,
                     which, out_dev);
  EIF_HIP(hipGetLastError());
  return EIF_OK;
}

// wall shear stress summed over both walls:
// (1/Re) sum_j dy [(UY(1,j) - uy_wall_xm) / DXC(1) + (UY(nx,j) - uy_wall_xp) / DXC(nx+1)]
__global__ __launch_bounds__(256) void k_wall_shear(Layout L, const double *uy,
                                                    const double *dxc, double dy, double rre,
                                                    double wxm, double wxp, double *out) {
  double a = 0.;
  for (int j = threadIdx.x + 1; j <= L.ny; j += blockDim.x)
    a += dy * ((uy[L.p(1, j)] - wxm) / dxc[0] + (uy[L.p(L.nx, j)] - wxp) / dxc[L.nx]);
  __shared__ double red[256];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = rre * red[0];
}

int launch_wall_shear(Ctx *c, double *out) {
  hipLaunchKernelGGL(k_wall_shear, dim3(1), dim3(256), 0, c->stream, c->L, c->uy, c->dxc, c->dy,
                     1. / c->cfg.Re, c->cfg.uy_wall_xm, c->cfg.uy_wall_xp, c->dscal + S_DIVMAX);
  EIF_HIP(hipGetLastError());
  EIF_HIP(hipMemcpyAsync(c->hscal + S_DIVMAX, c->dscal + S_DIVMAX, sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  out[0] = c->hscal[S_DIVMAX];
  return EIF_OK;
}

int launch_diagnostics(Ctx *c, int which, double *out) {
  const int nb = c->nred;
  hipLaunchKernelGGL(k_diag, dim3(nb), dim3(256), 0, c->stream, c->L, c->ux, c->uy,
                     c->dxf, c->dxc, c->dy, which, c->dred + 16);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, c->stream, c->dred + 16, nb,
                     which, c->dscal + S_DIVMAX);
  EIF_HIP(hipGetLastError());
  EIF_HIP(hipMemcpyAsync(c->hscal, c->dscal + S_DIVMAX, 2 * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  out[0] = c->hscal[0];
  out[1] = c->hscal[1];
  return EIF_OK;
}

}  // namespace eif
