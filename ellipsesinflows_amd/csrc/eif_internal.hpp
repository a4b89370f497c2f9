// eif_internal.hpp -- shared definitions of the gfx950 EllipsesInFlows path.
//
// Everything is FP64, compiled with -ffp-contract=off so that expressions
// round exactly like the reference's C99 build (ISO mode, no contraction).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include <map>
#include "eif_gpu.h"
#include "eif_batch.hpp"

// the reference's value of pi (include/common.h:7)
#define EIF_PI 3.1415926535897932

namespace eif {

// ---------------------------------------------------------------------
// Layouts of the reference (include/arrays/fluid.h, param.h,
// suspensions.h) for one process owning all jtot rows.
// ---------------------------------------------------------------------
struct Layout {
  int nx, ny;
  __host__ __device__ long ux(int i, int j) const { return (long)j * (nx + 1) + (i - 1); }
  __host__ __device__ long p(int i, int j) const { return (long)j * (nx + 2) + i; }
  __host__ __device__ long sx(int i, int j) const { return (long)(j - 1) * (nx - 1) + (i - 2); }
  __host__ __device__ long sy(int i, int j) const { return (long)(j - 1) * nx + (i - 1); }
  long ux_len() const { return (long)(nx + 1) * (ny + 2); }
  long p_len() const { return (long)(nx + 2) * (ny + 2); }
  long sx_len() const { return (long)(nx - 1) * ny; }
  long sy_len() const { return (long)nx * ny; }
};

// particle record member offsets (include/suspensions.h:8-31)
enum {
  R_DEN = 0, R_A, R_B, R_X, R_Y, R_AZ, R_DX, R_DY, R_DAZ,
  R_UX, R_UY, R_VZ, R_DUX, R_DUY, R_DVZ, R_FUX, R_FUY, R_TVZ,
  R_IUX = 18, R_IUY = 20, R_IVZ = 22, R_CFX = 24, R_CFY = 26, R_CTZ = 28,
  R_N = 30
};

// device scalars (one small array, so graph-captured work reads them)
enum { S_DT = 0, S_RESIDUAL = 1, S_DIVMAX = 2, S_DIVSUM = 3, S_NSCAL = 8 };

// IBM box of one (particle, periodic image)
struct Box {
  int imin, imax, jmin, jmax;
  long off;  // first cell of this box in the per-cell scratch arrays
};

// tile of the box-local deposit/apply step
constexpr int TILE_X = 64;
constexpr int TILE_Y = 4;
constexpr int TILE_CAP = 48;     // boxes registered per tile
constexpr int CONTACT_CAP = 64;  // colliding partners per particle
constexpr int COLL_CELLS_MIN = 4096;  // particles from which the cell list is chosen
constexpr int COLL_CELLS_CAP = 14336; // cells at most: their counts and a scan in 60 KiB of LDS
constexpr int EQC_ITER_CAP = 100000;  // guard of the uncapped fixed point
constexpr int DT_PARTS_CAP = 16384;   // workgroups of the projection with fused dt
constexpr int RES_SLOTS = 128;        // sub-iterations of one stage queued at most
constexpr int DIAG_SCRATCH_BLOCKS = 4096;  // partial-sum pairs of the static scratch

struct Profiler;

// A graph-captured step stops early when a stage's Crank-Nicolson loop
// needs more sub-iterations than were captured: every kernel that changes
// state across the step tests this device flag first (null outside graph
// steps' contexts never happens: the flag always exists, and is zero
// except between such a stop and the host's resumption).
__device__ inline bool halted(const int *h) { return h && *h != 0; }

// x / d; with POW2 the divisor is known to be a power of two and the
// division is the exact product with inv = 1/d (identical result)
template <bool POW2>
__host__ __device__ inline double qdiv(double x, double d, double inv) {
  return POW2 ? x * inv : x / d;
}

// what the host reads after each batch of sub-iterations, written by one
// kernel straight into pinned host memory (one readback per batch)
struct HostStatus {
  double residual;  // of the last sub-iteration run
  double dt;        // of the current step
  int substeps;     // sub-iterations run in this stage
  int flags;        // device capacity flags
  unsigned seq;     // written last (system-scope release): record number
};

// a graph step's record for the host (eif_gpu_run_steps)
struct GraphStatus {
  double dt;
  double resid[3];  // last residual of each stage's loop
  int counts[3];    // sub-iterations each stage ran (within the captured ones)
  int halt;         // 0, or the stage + 1 whose loop needed more
  int flags;        // device capacity flags
  unsigned seq;     // written last (system-scope release)
};

struct GraphCache;
struct AdjTape;

struct Ctx {
  eif_config cfg;
  int nx, ny;
  double lx, ly, dx, dy;
  Layout L;
  hipStream_t stream = nullptr;
  // second stream (default priority) for particle work that can overlap the
  // fluid kernels, and the events that join it back
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_i0 = nullptr, ev_c0 = nullptr;
  int device = 0;
  // host copies of the grid (exact reference arithmetic)
  std::vector<double> h_xf, h_xc, h_dxf, h_dxc, h_yf, h_yc;
  double alpha[3], beta[3], gamma[3];
  bool pow2_grid = false;  // dx, dy and every DXF, DXC are powers of two
  // ux, uy and p carry current boundary values (every kernel of a stage
  // refreshes the ones it changes); host-side changes clear it
  bool bc_valid = false;
  // the advective dt limit of the current velocities is already reduced
  // into dt_adv_parts partial minima (dred[16...])
  bool dt_adv_valid = false;
  int dt_adv_parts = 0;
  // device grid
  double *xf = nullptr, *xc = nullptr, *dxf = nullptr, *dxc = nullptr,
         *yf = nullptr, *yc = nullptr;
  // fluid fields (reference layouts); *_alt are ping-pong partners
  double *ux = nullptr, *uy = nullptr, *p = nullptr, *psi = nullptr;
  double *ux_alt = nullptr, *uy_alt = nullptr;
  // the third velocity pair of eif_gpu_integrate's stages (see
  // launch_predictor_prev); src_lazy: srcuxa/srcuya are due the values of
  // A at (ux_alt, uy_alt)
  double *ux_spare = nullptr, *uy_spare = nullptr;
  bool src_lazy = false;
  double *srcuxa = nullptr, *srcuya = nullptr;
  double *dux = nullptr, *duy = nullptr;  // materialised only on request
  // Poisson
  double *qhat = nullptr;      // spectral field, x-contiguous rows [ny][nx]
  double *ctab = nullptr;      // C[m] = cos(pi m/(2 nx)), m=0..2nx
  double *ctab3 = nullptr;     // DCT-III pre-processing factors in thread order (3 x 2nx)
  double *twid = nullptr;      // per-stage radix-2 twiddles (entries of C)
  double *th_val = nullptr;    // Thomas tables [ny-1][nx] (forward factors)
  double *th_u = nullptr;      // [ny-1][nx] modified upper diagonal
  double *th_q1 = nullptr;     // [ny-1][nx] border solution
  double *th_den = nullptr;    // [nx]
  double *th_last = nullptr;   // [nx] q[n-1] of the current solve
  unsigned char *th_flag = nullptr;  // [ny-1][nx] zero-pivot flags
  // partition solver tables
  int part_L = 0, part_P = 0;
  // tuned launch shapes (eif_gpu_autotune); 0 is the built-in default
  int pred_rows = 0, cp_rows = 0;
  // rows per strip of the fused DCT-III + projection (2, 4, 8), or 0: the
  // unfused kernels (the default)
  int proj_strip = 0;
  // log2 of the register FFT's values per thread (2, 4; 0: the default 3)
  int fft_logr = 0;
  double *pt_fwd = nullptr, *pt_u = nullptr, *pt_v = nullptr, *pt_w = nullptr;
  double *pt_bnd = nullptr, *pt_red = nullptr, *pt_lvl = nullptr;
  double *pt_ends = nullptr, *pt_g = nullptr;
  // scalars
  double *dscal = nullptr;
  double *hscal = nullptr;  // pinned
  double *dred = nullptr;   // reduction partials
  int *halt = nullptr;      // graph-step stop flag (see halted())
  // graph steps: per-stage counts and residuals, the step counter, the
  // pinned two-slot record, and the captured graphs
  int *g_counts = nullptr;
  double *g_resid = nullptr;
  unsigned *g_seq = nullptr;
  GraphStatus *gstat = nullptr;
  GraphCache *graphs = nullptr;
  bool graph_on = false;
  AdjTape *tape = nullptr;  // adjoint records and cotangents
  bool recording = false;   // eif_gpu_integrate copies the stage records
  double *diag_scratch = nullptr;  // this device's static diagnostics scratch
  double *dt_part = nullptr;  // DT_PARTS_CAP partial minima of the fused dt
  int nred = 0;
  // particles
  int np = 0;
  double *prt = nullptr;       // np x R_N records
  double *pconst = nullptr;    // per particle: mass, moi
  Box *boxes = nullptr;        // np x 3
  long *box_off = nullptr;     // np+1 offsets into per-cell scratch
  long scratch_cells = 0;
  int box_cap_x = 0, box_cap_y = 0;
  double *sc_fx = nullptr, *sc_fy = nullptr;  // per box cell forcing*dt
  double *sc_dist = nullptr;                  // cached tanh(BETA d), d the signed distance
  double *sc_t = nullptr;                     // foot parameter per box cell
  int *sc_key = nullptr;                      // np x 8: boxes of sc_t
  double *sc_pos = nullptr;                   // np x 3: positions of cache
  int *slice_first = nullptr;  // IBM work slices
  long *slice_info = nullptr;
  int *slice_owner = nullptr;
  double *slice_part = nullptr;
  double *slice_part0 = nullptr;  // inertia(0) sums of the fused exchange
  void *heads = nullptr;          // np per-particle pass headers (SliceHead)
  int nslices = 0;
  int *tile_cnt = nullptr, *tile_list = nullptr;
  int ntx = 0, nty = 0;
  int *flags = nullptr;    // device error flags
  unsigned long long *ibm_stats = nullptr;  // [8] measurement counters
  unsigned long long *ibm_pp = nullptr;     // np x 4 IBM work counters, then calls
  double *res_slots = nullptr;  // RES_SLOTS: residual of each sub-iteration of the stage
  HostStatus *hstat = nullptr;  // pinned, written by k_status
  unsigned hstat_seq = 0;       // number of the last status record requested
  int substep_guess = 3; // sub-iterations queued before the first check
  // collisions
  // broad phase: 0 chooses (cell list from COLL_CELLS_MIN particles, where
  // the grid of cells allows it), 1 all pairs, 2 cell list where allowed
  int coll_mode = 0;
  // the cell list: ncx x ncy cells of cl_hx x cl_hy, each at least the
  // widest pair reach 2 max(a, b) (0 cells: not allowed for these sizes)
  int cl_ncx = 0, cl_ncy = 0;
  double cl_hx = 0., cl_hy = 0.;
  int *cl_start = nullptr;  // ncells + 1 offsets into cl_items
  int *cl_items = nullptr;  // np particle indices, by cell
  int *cl_cell = nullptr;   // np: each particle's cell
  int *cl_slot = nullptr;   // np: its place within the cell
  int *cl_count = nullptr;  // COLL_CELLS_CAP per-cell counts, zero between builds
  // np x 4: x + dx, y + dy, max(a, b) of each particle (and a pad), as the
  // broad phase compares them, packed so that a wave's reads are contiguous
  double *cpos = nullptr;
  int *rows = nullptr;      // np x CONTACT_CAP colliding partners, ascending
  int *row_cnt = nullptr;   // np
  int *wall_hit = nullptr;  // np x 2
  // statistics accumulators (allocated on first use), reference layouts
  double *st_ux1 = nullptr, *st_ux2 = nullptr;                   // like ux
  double *st_uy1 = nullptr, *st_uy2 = nullptr, *st_phi = nullptr;  // like uy
  int st_num = 0;
  Profiler *prof = nullptr;
};

// error state
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define EIF_HIP(call)                                                     \
  do {                                                                    \
    hipError_t _e = (call);                                               \
    if (_e != hipSuccess) return ::eif::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

// profiling helpers (host)
// prof_begin: whether this launch is timed (then prof_end closes it)
bool prof_begin(Ctx *c, const char *name, hipStream_t st);
void prof_end(Ctx *c, const char *name, hipStream_t st);
bool prof_kernel_events(Ctx *c, const char *name, hipEvent_t *start, hipEvent_t *stop);
struct ProfScope {  // events on the stream the scoped kernels run on
  Ctx *c;
  const char *n;
  hipStream_t st;
  bool timed;
  ProfScope(Ctx *c_, const char *n_) : c(c_), n(n_), st(c_->stream) {
    timed = prof_begin(c, n, st);
  }
  ProfScope(Ctx *c_, const char *n_, hipStream_t s_) : c(c_), n(n_), st(s_) {
    timed = prof_begin(c, n, st);
  }
  ~ProfScope() {
    if (timed) prof_end(c, n, st);
  }
};

// launchers implemented in the .hip files
int launch_bc_ux(Ctx *c, double *ux);
int launch_bc_uy(Ctx *c, double *uy);
int launch_bc_p(Ctx *c, double *p);
// swaps ux, uy with their ping-pong partners at launch
int launch_predictor(Ctx *c, int rk, hipStream_t st = nullptr);
int launch_predictor_prev(Ctx *c, int rk, hipStream_t st = nullptr);
int launch_materialise_src(Ctx *c);
// dt_fused: also leave the next decide_dt's advective limit as partial minima
int launch_correct_pressure(Ctx *c, int rk, bool correct, bool pressure,
                            bool dt_fused = false);
int launch_correct_pressure_on(Ctx *c, int rk, bool correct, bool pressure, bool dt_fused,
                               double *ux, double *uy, double *p);
int launch_predictor_probe(Ctx *c);
// the partitioned y-solve's tables for chunks of L rows (0: the default
// choice); returns EIF_ERR_ARG when L does not split jtot within the limits
int partition_setup(Ctx *c, int L);
int launch_decide_dt(Ctx *c);
int launch_diagnostics(Ctx *c, int which, double *out);
int launch_wall_shear(Ctx *c, double *out);
// the same sums into caller device memory, partials in the static scratch
int launch_diagnostics_noalloc(Ctx *c, int which, double *out_dev);
int diag_scratch_address(int device, double **out);
int poisson_init(Ctx *c);
// chunk length and count of the partitioned y-solve for jtot = ny (0 when
// the line is solved by the Thomas sweep instead); host only
int partition_layout(int ny, int *chunk, int *nchunks);
// rhs_add (cell field, P layout): solve for div u / (gamma dt) + rhs_add / (gamma dt)^2
int launch_poisson(Ctx *c, int rk, const double *rhs_add = nullptr);
// the stage's potential, projection and pressure update (launch_poisson +
// launch_correct_pressure(correct, pressure)), in one fused kernel after the
// solves where nx is a power of two in 64..8192; psi is then not stored
int launch_poisson_project(Ctx *c, int rk, bool dt_fused);
bool project_fusable(const Ctx *c);
// the tuner's choice of unfused kernels or fused strip height, timed on the
// given copies of the fields
int project_tune(Ctx *c, double *ux, double *uy, double *p, std::string *rep);
int set_projection(Ctx *c, int strip);
// the collision broad phase's cells for the particles' sizes (records on
// the host, np x R_N); called whenever the records are set
int coll_cells_setup(Ctx *c, const double *records);
int set_collision_search(Ctx *c, int mode);
// the register FFT's values per thread: tunable for this context's rows,
// and 2^lr available there
bool fft_tunable(const Ctx *c);
bool fft_logr_valid(const Ctx *c, int lr);
int susp_init(Ctx *c);
// defer_finalize: the slice sums are added by the next launch_increment
// with fin_inertia (one launch fewer per sub-iteration)
int launch_inertia(Ctx *c, int cnstep, hipStream_t st = nullptr, bool defer_finalize = false);
// cached_known: the distance cache holds the exchange position (it follows
// inertia(0) of the same stage), so the distance pass is skipped
// tiles_clear: the tile registrations were cleared by launch_reset(clear_tiles);
// defer_finalize: the per-particle sums are added by launch_exchange_finalize
int launch_exchange(Ctx *c, bool cached_known = false, bool tiles_clear = false,
                    bool defer_finalize = false);
int launch_exchange_finalize(Ctx *c, hipStream_t st);
int launch_ibm_dist(Ctx *c, hipStream_t st = nullptr);
// the exchange with inertia(0) of the stage-start velocity in the same pass
int launch_exchange_inertia0(Ctx *c, const double *us_x, const double *us_y);
int launch_update_momentum(Ctx *c);
int launch_update_momentum_x0(Ctx *c);
int launch_materialise_du(Ctx *c);
int launch_collision(Ctx *c, int cnstep, hipStream_t st = nullptr);
// sub >= 0: sub-iteration `sub` of the stage, gated on the residual slots;
// it also finalises the preceding launch_substep_forces
int launch_increment(Ctx *c, int rk, int sub = -1);
// inertia(1) and collision(1) of sub-iteration `sub`, gated, in one launch
int launch_substep_forces(Ctx *c, int sub);
int launch_update_particles(Ctx *c, hipStream_t st = nullptr);
// clear_tiles: also clear the IBM tile registrations for the next exchange
int launch_reset(Ctx *c, hipStream_t st = nullptr, bool clear_tiles = false);
int launch_status(Ctx *c, int queued);
int launch_stage_gate(Ctx *c, int stage, int K);
int launch_graph_status(Ctx *c);
void graphs_free(Ctx *c);
// adjoint (eif_adjoint.hip)
int tape_begin(Ctx *c);
void tape_free(Ctx *c);
// stage records taken by integrate_from while c->recording
int rec_stage_begin(Ctx *c, int rk);
int rec_fields(Ctx *c, int which, const double *ux, const double *uy);  // 0 u^k, 1 u*, 2 u^{k+1}
int rec_p0(Ctx *c);
int rec_sub(Ctx *c, int s);
int rec_stage_end(Ctx *c, int S);
int rec_check_contacts(Ctx *c);
int stage_adjoint(Ctx *c);
int adjoint_access(Ctx *c, int which, double *host, bool add);
int adjoint_add_wall_shear(Ctx *c, double weight);
int check_flags(Ctx *c);
int flags_error(int h);
int launch_statistics(Ctx *c);
int launch_ref_iterations(Ctx *c, unsigned long long *dout);
int stat_alloc(Ctx *c);
int stat_reset(Ctx *c);

}  // namespace eif
