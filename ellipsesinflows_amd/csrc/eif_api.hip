// eif_api.cpp -- the C-ABI of include/eif_gpu.h: context set-up, transfers,
// the time-step orchestration of src/main.c:13-76, errors and profiling.
#include "eif_internal.hpp"
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <memory>
#include <mutex>

namespace eif {

thread_local Recorder *g_rec = nullptr;

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file,
           line);
  g_err = buf;
  return EIF_ERR_HIP;
}

// ---------------------------------------------------------------------
// profiling: HIP events around each named launch on the context stream
// ---------------------------------------------------------------------
static const char *k_prof_names =
    "bc,predictor,rhs_dct2,thomas,ysolve_partition,dct3_psi,correct_pressure,dct3_project,"
    "decide_dt,"
    "ibm_inertia,ibm_substep,ibm_exchange,ibm_apply,collision";

// Events that only time work: without the system-scope fence, recording one
// does not write back and invalidate the L2 around the timed launch, which
// would otherwise add that writeback (the launch's dirty output) to the
// measured duration.
static hipError_t timing_event(hipEvent_t *e) {
  return hipEventCreateWithFlags(e, hipEventDisableSystemFence);
}

struct Profiler {
  bool on = false;
  std::string only;
  int every = 1;  // time one launch in `every` (each event pair idles the stream briefly)
  struct Rec {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t used = 0;
    long seen = 0;
    long count = 0;
    double total_ms = 0.;
  };
  std::map<std::string, Rec> recs;
  std::vector<std::string> open;
};

static int prof_flush(Ctx *c) {
  if (!c->prof) return EIF_OK;
  for (auto &kv : c->prof->recs) {
    auto &r = kv.second;
    for (size_t k = 0; k < r.used; k++) {
      float ms = 0.f;
      EIF_HIP(hipEventSynchronize(r.ev[k].second));
      EIF_HIP(hipEventElapsedTime(&ms, r.ev[k].first, r.ev[k].second));
      r.total_ms += ms;
      r.count++;
    }
    r.used = 0;
  }
  return EIF_OK;
}

bool prof_begin(Ctx *c, const char *name, hipStream_t st) {
  if (!c->prof || !c->prof->on) return false;
  if (!c->prof->only.empty() && c->prof->only != name) return false;
  auto &r = c->prof->recs[name];
  if (r.seen++ % c->prof->every != 0) return false;
  if (r.used == r.ev.size()) {
    hipEvent_t a, b;
    if (timing_event(&a) != hipSuccess || timing_event(&b) != hipSuccess) return false;
    r.ev.push_back({a, b});
  }
  return hipEventRecord(r.ev[r.used].first, st) == hipSuccess;
}

// For a scope that is one kernel launch: the event pair that the launch
// itself stamps (hipExtLaunchKernel), so the elapsed time is the kernel's
// and no marker packet sits between it and its neighbours.  False when
// this launch is not sampled.
bool prof_kernel_events(Ctx *c, const char *name, hipEvent_t *start, hipEvent_t *stop) {
  if (!c->prof || !c->prof->on) return false;
  if (!c->prof->only.empty() && c->prof->only != name) return false;
  auto &r = c->prof->recs[name];
  if (r.seen++ % c->prof->every != 0) return false;
  // pairs already stamped are read back before the store is reused (the
  // pair handed out below is stamped only by the caller's launch)
  if (r.used >= 4096 && prof_flush(c) != EIF_OK) return false;
  if (r.used == r.ev.size()) {
    hipEvent_t a, b;
    if (timing_event(&a) != hipSuccess || timing_event(&b) != hipSuccess) return false;
    r.ev.push_back({a, b});
  }
  *start = r.ev[r.used].first;
  *stop = r.ev[r.used].second;
  r.used++;
  return true;
}

void prof_end(Ctx *c, const char *name, hipStream_t st) {
  auto &r = c->prof->recs[name];
  if (r.used >= r.ev.size()) return;
  (void)hipEventRecord(r.ev[r.used].second, st);
  r.used++;
  if (r.used >= 4096) prof_flush(c);
}

// message for nonzero device capacity flags
int flags_error(int h) {
  std::string m = "device capacity exceeded:";
  if (h & 1) m += " IBM box scratch;";
  if (h & 2) m += " tile box list;";
  if (h & 4) m += " colliding pair list;";
  if (h & 8) m += " equivalent-circle iteration cap;";
  if (h & 16) m += " contacts per particle;";
  set_error(m);
  return EIF_ERR_CAPACITY;
}

int check_flags(Ctx *c) {
  EIF_HIP(hipMemcpyAsync(c->hscal + S_NSCAL - 1, c->flags, sizeof(int), hipMemcpyDeviceToHost,
                         c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  int h;
  memcpy(&h, c->hscal + S_NSCAL - 1, sizeof(int));
  if (h) {
    EIF_HIP(hipMemsetAsync(c->flags, 0, sizeof(int), c->stream));
    return flags_error(h);
  }
  return EIF_OK;
}

}  // namespace eif

using namespace eif;

#define CHECK_CTX(c)                    \
  do {                                  \
    if (!(c)) {                         \
      set_error("null context");        \
      return EIF_ERR_ARG;               \
    }                                   \
  } while (0)

#define TRY(x)            \
  do {                    \
    int _r = (x);         \
    if (_r) return _r;    \
  } while (0)

struct eif_ctx : public Ctx {};

// ux, uy or p changed outside eif_gpu_integrate: boundary values and the
// fused advective dt limit are no longer known to be current
static void fields_changed(Ctx *c) {
  c->bc_valid = false;
  c->dt_adv_valid = false;
}

extern "C" {

const char *eif_gpu_last_error(void) { return g_err.c_str(); }

int poisson_gpu_partition_layout(int jtot, int *chunk, int *nchunks) {
  return partition_layout(jtot, chunk, nchunks);
}
const char *eif_gpu_profile_names(void) { return k_prof_names; }

static void free_ctx(Ctx *c) {
  double **bufs[] = {&c->xf, &c->xc, &c->dxf, &c->dxc, &c->yf, &c->yc, &c->ux, &c->uy,
                     &c->p, &c->psi, &c->ux_alt, &c->uy_alt, &c->ux_spare, &c->uy_spare, &c->srcuxa, &c->srcuya,
                     &c->dux, &c->duy, &c->qhat, &c->ctab, &c->ctab3, &c->twid, &c->cpos, &c->th_val, &c->th_u,
                     &c->th_q1, &c->th_den, &c->th_last, &c->pt_fwd, &c->pt_u,
                     &c->pt_v, &c->pt_w, &c->pt_bnd, &c->pt_red, &c->pt_lvl, &c->pt_ends, &c->pt_g, &c->dscal, &c->dred,
                     &c->prt, &c->pconst, &c->dt_part, &c->sc_fx, &c->sc_fy, &c->sc_dist, &c->sc_pos,
                     &c->sc_t, &c->res_slots, &c->st_ux1, &c->st_ux2, &c->st_uy1,
                     &c->st_uy2, &c->st_phi, &c->slice_part, &c->slice_part0};
  for (auto b : bufs)
    if (*b) (void)hipFree(*b);
  void *others[] = {c->th_flag, c->boxes, c->box_off, c->tile_cnt, c->tile_list,
                    c->flags, c->rows, c->row_cnt, c->wall_hit,
                    c->ibm_stats, c->slice_first, c->slice_info, c->ibm_pp, c->slice_owner,
                    c->sc_key, c->heads, c->cl_start, c->cl_items, c->cl_cell, c->cl_slot, c->cl_count};
  for (auto o : others)
    if (o) (void)hipFree(o);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->hstat) (void)hipHostFree(c->hstat);
  if (c->gstat) (void)hipHostFree(c->gstat);
  graphs_free(c);
  tape_free(c);
  void *gbufs[] = {c->halt, c->g_counts, c->g_resid, c->g_seq};
  for (auto b : gbufs)
    if (b) (void)hipFree(b);
  if (c->prof) {
    for (auto &kv : c->prof->recs)
      for (auto &e : kv.second.ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
      }
    delete c->prof;
  }
  if (c->side) (void)hipStreamSynchronize(c->side);
  hipEvent_t evs[] = {c->ev_fork, c->ev_i0, c->ev_c0};
  for (auto ev : evs)
    if (ev) (void)hipEventDestroy(ev);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

int fluid_gpu_create(const eif_config *cfg, eif_ctx **out) {
  if (!cfg || !out) {
    set_error("null argument");
    return EIF_ERR_ARG;
  }
  *out = nullptr;
  if (cfg->itot < 2 || cfg->jtot < 3 || !(cfg->ly > 0.) || !(cfg->Re > 0.)) {
    set_error("invalid grid: need itot >= 2, jtot >= 3, ly > 0, Re > 0");
    return EIF_ERR_ARG;
  }
  if ((long)cfg->itot * cfg->jtot > (1L << 31) / 4) {
    set_error("grid too large for one context");
    return EIF_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("no HIP device available");
    return EIF_ERR_NODEVICE;
  }
  eif_ctx *c = new eif_ctx();
  c->cfg = *cfg;
  const int nx = cfg->itot, ny = cfg->jtot;
  c->nx = nx;
  c->ny = ny;
  c->L = Layout{nx, ny};
  c->lx = 1.;
  c->ly = cfg->ly;
  int rc = EIF_OK;
  auto fail = [&](int r) {
    free_ctx(c);
    delete c;
    return r;
  };
  if (hipGetDevice(&c->device) != hipSuccess) return fail(hip_fail(hipErrorUnknown, "hipGetDevice", __FILE__, __LINE__));
  {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    hipEvent_t *evs[] = {&c->ev_fork, &c->ev_i0, &c->ev_c0};
    for (auto ev : evs)
      if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e != hipSuccess) return fail(hip_fail(e, "hipStreamCreate", __FILE__, __LINE__));
  }
  // grid, src/param/init.c:105-167 (one process)
  {
    const double lx = c->lx, ly = c->ly;
    const double dx = lx / nx;
    c->dx = dx;
    c->h_xf.assign(nx + 1, 0.);
    c->h_xc.assign(nx + 2, 0.);
    c->h_dxf.assign(nx, 0.);
    c->h_dxc.assign(nx + 1, 0.);
    c->h_yf.assign(ny + 1, 0.);
    c->h_yc.assign(ny, 0.);
    for (int i = 1; i <= nx + 1; i++) c->h_xf[i - 1] = 1. * (i - 1) * dx;
    c->h_xf[0] = 0.;
    c->h_xf[nx] = lx;
    for (int i = 1; i <= nx; i++) c->h_dxf[i - 1] = c->h_xf[i] - c->h_xf[i - 1];
    c->h_xc[0] = c->h_xf[0];
    for (int i = 1; i <= nx; i++) c->h_xc[i] = 0.5 * (c->h_xf[i - 1] + c->h_xf[i]);
    c->h_xc[nx + 1] = c->h_xf[nx];
    for (int i = 1; i <= nx + 1; i++) c->h_dxc[i - 1] = c->h_xc[i] - c->h_xc[i - 1];
    const double dy = ly / ny;
    c->dy = dy;
    const double yoffset = dy * 0;
    for (int j = 1; j <= ny + 1; j++) c->h_yf[j - 1] = yoffset + 1. * (j - 1) * dy;
    for (int j = 1; j <= ny; j++) c->h_yc[j - 1] = yoffset + 0.5 * (2 * j - 1) * dy;
  }
  // uniform power-of-two spacings make every grid division exact as a
  // product with the reciprocal
  {
    auto p2 = [](double v) {
      int e;
      return v > 0. && frexp(v, &e) == 0.5;
    };
    bool ok = p2(c->dx) && p2(c->dy);
    for (double v : c->h_dxf) ok = ok && p2(v);
    for (double v : c->h_dxc) ok = ok && p2(v);
    c->pow2_grid = ok;
  }
  // RK coefficients, src/param/init.c:169-191
  c->alpha[0] = 32. / 60.;
  c->beta[0] = 0. / 60.;
  c->alpha[1] = 25. / 60.;
  c->beta[1] = -17. / 60.;
  c->alpha[2] = 45. / 60.;
  c->beta[2] = -25. / 60.;
  for (int k = 0; k < 3; k++) c->gamma[k] = c->alpha[k] + c->beta[k];
  auto dalloc = [&](double **p, size_t n, const double *src) -> int {
    hipError_t e = hipMalloc(p, n * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    if (src)
      e = hipMemcpy(*p, src, n * sizeof(double), hipMemcpyHostToDevice);
    else
      e = hipMemset(*p, 0, n * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
    return EIF_OK;
  };
  const Layout &L = c->L;
  if ((rc = dalloc(&c->xf, nx + 1, c->h_xf.data())) ||
      (rc = dalloc(&c->xc, nx + 2, c->h_xc.data())) ||
      (rc = dalloc(&c->dxf, nx, c->h_dxf.data())) ||
      (rc = dalloc(&c->dxc, nx + 1, c->h_dxc.data())) ||
      (rc = dalloc(&c->yf, ny + 1, c->h_yf.data())) ||
      (rc = dalloc(&c->yc, ny, c->h_yc.data())) ||
      (rc = dalloc(&c->ux, L.ux_len(), nullptr)) ||
      (rc = dalloc(&c->ux_alt, L.ux_len(), nullptr)) ||
      (rc = dalloc(&c->uy, L.p_len(), nullptr)) ||
      (rc = dalloc(&c->uy_alt, L.p_len(), nullptr)) ||
      (rc = dalloc(&c->ux_spare, L.ux_len(), nullptr)) ||
      (rc = dalloc(&c->uy_spare, L.p_len(), nullptr)) ||
      (rc = dalloc(&c->p, L.p_len(), nullptr)) ||
      (rc = dalloc(&c->psi, L.p_len(), nullptr)) ||
      (rc = dalloc(&c->dux, L.p_len(), nullptr)) ||
      (rc = dalloc(&c->duy, L.p_len(), nullptr)) ||
      (rc = dalloc(&c->srcuxa, L.sx_len(), nullptr)) ||
      (rc = dalloc(&c->srcuya, L.sy_len(), nullptr)) ||
      (rc = dalloc(&c->dscal, S_NSCAL, nullptr)))
    return fail(rc);
  c->nred = 1024;
  if ((rc = dalloc(&c->dred, 16 + 2 * (size_t)c->nred, nullptr))) return fail(rc);
  if ((rc = dalloc(&c->dt_part, DT_PARTS_CAP, nullptr))) return fail(rc);
  {
    hipError_t e = hipHostMalloc((void **)&c->hscal, S_NSCAL * sizeof(double), 0);
    if (e != hipSuccess) return fail(hip_fail(e, "hipHostMalloc", __FILE__, __LINE__));
    e = hipMalloc(&c->flags, 4 * sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->flags, 0, 4 * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&c->res_slots, RES_SLOTS * sizeof(double));
    if (e == hipSuccess) e = hipMemset(c->res_slots, 0, RES_SLOTS * sizeof(double));
    if (e == hipSuccess)
      e = hipHostMalloc((void **)&c->hstat, sizeof(HostStatus), hipHostMallocCoherent);
    if (e == hipSuccess) memset(c->hstat, 0, sizeof(HostStatus));
    if (e == hipSuccess) e = hipMalloc(&c->halt, sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->halt, 0, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&c->g_counts, 4 * sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->g_counts, 0, 4 * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&c->g_resid, 4 * sizeof(double));
    if (e == hipSuccess) e = hipMemset(c->g_resid, 0, 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->g_seq, sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(c->g_seq, 0, sizeof(unsigned));
    if (e == hipSuccess)
      e = hipHostMalloc((void **)&c->gstat, 2 * sizeof(GraphStatus), hipHostMallocCoherent);
    if (e == hipSuccess) memset(c->gstat, 0, 2 * sizeof(GraphStatus));
    if (e == hipSuccess) e = hipMalloc(&c->ibm_stats, 8 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->ibm_stats, 0, 8 * sizeof(unsigned long long));
    if (e != hipSuccess) return fail(hip_fail(e, "hipMalloc", __FILE__, __LINE__));
  }
  if ((rc = poisson_init(c))) return fail(rc);
  if ((rc = diag_scratch_address(c->device, &c->diag_scratch))) return fail(rc);
  c->prof = new Profiler();
  // initial boundary values (src/fluid/init.c:33-85)
  if ((rc = launch_bc_ux(c, c->ux)) || (rc = launch_bc_uy(c, c->uy)) ||
      (rc = launch_bc_ux(c, c->ux_alt)) || (rc = launch_bc_uy(c, c->uy_alt)) ||
      (rc = launch_bc_ux(c, c->ux_spare)) || (rc = launch_bc_uy(c, c->uy_spare)) ||
      (rc = launch_bc_p(c, c->p)))
    return fail(rc);
  if (hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(hip_fail(hipErrorUnknown, "hipStreamSynchronize", __FILE__, __LINE__));
  *out = c;
  return EIF_OK;
}

int fluid_gpu_destroy(eif_ctx *c) {
  if (!c) return EIF_OK;
  (void)hipStreamSynchronize(c->stream);
  free_ctx(c);
  delete c;
  return EIF_OK;
}

int fluid_gpu_get_grid(const eif_ctx *c, double *xf, double *xc, double *dxf,
                       double *dxc, double *yf, double *yc) {
  CHECK_CTX(c);
  auto cp = [](double *d, const std::vector<double> &s) {
    if (d) memcpy(d, s.data(), s.size() * sizeof(double));
  };
  cp(xf, c->h_xf);
  cp(xc, c->h_xc);
  cp(dxf, c->h_dxf);
  cp(dxc, c->h_dxc);
  cp(yf, c->h_yf);
  cp(yc, c->h_yc);
  return EIF_OK;
}

int fluid_gpu_upload(eif_ctx *c, const double *ux, const double *uy, const double *p) {
  CHECK_CTX(c);
  fields_changed(c);
  const Layout &L = c->L;
  if (ux) EIF_HIP(hipMemcpyAsync(c->ux, ux, L.ux_len() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (uy) EIF_HIP(hipMemcpyAsync(c->uy, uy, L.p_len() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (p) EIF_HIP(hipMemcpyAsync(c->p, p, L.p_len() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  return EIF_OK;
}

int fluid_gpu_download(eif_ctx *c, double *ux, double *uy, double *p) {
  CHECK_CTX(c);
  const Layout &L = c->L;
  if (ux) EIF_HIP(hipMemcpyAsync(ux, c->ux, L.ux_len() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (uy) EIF_HIP(hipMemcpyAsync(uy, c->uy, L.p_len() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (p) EIF_HIP(hipMemcpyAsync(p, c->p, L.p_len() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  return EIF_OK;
}

static int field_ptr(eif_ctx *c, int which, double **ptr, size_t *n) {
  const Layout &L = c->L;
  switch (which) {
    case EIF_FIELD_PSI: *ptr = c->psi; *n = L.p_len(); return EIF_OK;
    case EIF_FIELD_SRCUXA: *ptr = c->srcuxa; *n = L.sx_len(); return EIF_OK;
    case EIF_FIELD_SRCUYA: *ptr = c->srcuya; *n = L.sy_len(); return EIF_OK;
    case EIF_FIELD_DUX: *ptr = c->dux; *n = L.p_len(); return EIF_OK;
    case EIF_FIELD_DUY: *ptr = c->duy; *n = L.p_len(); return EIF_OK;
  }
  set_error("unknown field selector");
  return EIF_ERR_ARG;
}

int fluid_gpu_download_field(eif_ctx *c, int which, double *host) {
  CHECK_CTX(c);
  if (!host) {
    set_error("null host buffer");
    return EIF_ERR_ARG;
  }
  double *ptr;
  size_t n;
  TRY(field_ptr(c, which, &ptr, &n));
  if (which == EIF_FIELD_DUX || which == EIF_FIELD_DUY) TRY(launch_materialise_du(c));
  if (which == EIF_FIELD_SRCUXA || which == EIF_FIELD_SRCUYA) TRY(launch_materialise_src(c));
  EIF_HIP(hipMemcpyAsync(host, ptr, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  return EIF_OK;
}

int fluid_gpu_upload_field(eif_ctx *c, int which, const double *host) {
  CHECK_CTX(c);
  fields_changed(c);
  if (!host || which == EIF_FIELD_DUX || which == EIF_FIELD_DUY) {
    set_error("field cannot be uploaded");
    return EIF_ERR_ARG;
  }
  double *ptr;
  size_t n;
  TRY(field_ptr(c, which, &ptr, &n));
  // the other source array keeps its values
  if (which == EIF_FIELD_SRCUXA || which == EIF_FIELD_SRCUYA) TRY(launch_materialise_src(c));
  EIF_HIP(hipMemcpyAsync(ptr, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  return EIF_OK;
}

int fluid_gpu_update_boundaries_ux(eif_ctx *c) {
  CHECK_CTX(c);
  fields_changed(c);
  return launch_bc_ux(c, c->ux);
}
int fluid_gpu_update_boundaries_uy(eif_ctx *c) {
  CHECK_CTX(c);
  fields_changed(c);
  return launch_bc_uy(c, c->uy);
}
int fluid_gpu_update_boundaries_p(eif_ctx *c) {
  CHECK_CTX(c);
  fields_changed(c);
  return launch_bc_p(c, c->p);
}

int fluid_gpu_get_dt(eif_ctx *c, double *dt_out) {
  CHECK_CTX(c);
  if (!dt_out) return EIF_OK;
  EIF_HIP(hipMemcpyAsync(c->hscal, c->dscal + S_DT, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  *dt_out = c->hscal[0];
  return EIF_OK;
}

int fluid_gpu_decide_dt(eif_ctx *c, double *dt_out) {
  CHECK_CTX(c);
  TRY(launch_decide_dt(c));
  if (dt_out) return fluid_gpu_get_dt(c, dt_out);
  return EIF_OK;
}

int fluid_gpu_set_dt(eif_ctx *c, double dt) {
  CHECK_CTX(c);
  if (!(dt > 0.) || !std::isfinite(dt)) {
    set_error("dt must be positive and finite");
    return EIF_ERR_ARG;
  }
  c->hscal[S_DT] = dt;
  EIF_HIP(hipMemcpyAsync(c->dscal + S_DT, c->hscal + S_DT, sizeof(double), hipMemcpyHostToDevice, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  return EIF_OK;
}

static int check_rk(int rk) {
  if (rk < 0 || rk > 2) {
    set_error("rkstep must be 0, 1 or 2");
    return EIF_ERR_ARG;
  }
  return EIF_OK;
}

int fluid_gpu_update_velocity(eif_ctx *c, int rk) {
  CHECK_CTX(c);
  fields_changed(c);
  TRY(check_rk(rk));
  return launch_predictor(c, rk);
}

int poisson_gpu_compute_potential(eif_ctx *c, int rk) {
  CHECK_CTX(c);
  TRY(check_rk(rk));
  return launch_poisson(c, rk);
}

int fluid_gpu_correct_velocity(eif_ctx *c, int rk) {
  CHECK_CTX(c);
  fields_changed(c);
  TRY(check_rk(rk));
  return launch_correct_pressure(c, rk, true, false);
}

int fluid_gpu_update_pressure(eif_ctx *c) {
  CHECK_CTX(c);
  fields_changed(c);
  return launch_correct_pressure(c, 0, false, true);
}

int fluid_gpu_divergence(eif_ctx *c, double *divmax, double *divsum) {
  CHECK_CTX(c);
  double o[2];
  TRY(launch_diagnostics(c, 0, o));
  if (divmax) *divmax = o[0];
  if (divsum) *divsum = o[1];
  return EIF_OK;
}

int fluid_gpu_momentum(eif_ctx *c, double moms[2]) {
  CHECK_CTX(c);
  return launch_diagnostics(c, 1, moms);
}

int fluid_gpu_energy(eif_ctx *c, double q[2]) {
  CHECK_CTX(c);
  return launch_diagnostics(c, 2, q);
}

int fluid_gpu_diagnostics_noalloc(eif_ctx *c, int which, double *out_dev) {
  CHECK_CTX(c);
  if (which < 0 || which > 2 || !out_dev) {
    set_error("diagnostics: which must be 0, 1 or 2 and out_dev non-null");
    return EIF_ERR_ARG;
  }
  return launch_diagnostics_noalloc(c, which, out_dev);
}

int statistics_gpu_collect(eif_ctx *c) {
  CHECK_CTX(c);
  return launch_statistics(c);
}

int statistics_gpu_reset(eif_ctx *c) {
  CHECK_CTX(c);
  return stat_reset(c);
}

int statistics_gpu_download(eif_ctx *c, double *ux1, double *ux2, double *uy1,
                            double *uy2, double *phi, int *num) {
  CHECK_CTX(c);
  TRY(stat_alloc(c));
  const size_t nu = c->L.ux_len() * sizeof(double), np = c->L.p_len() * sizeof(double);
  if (ux1) EIF_HIP(hipMemcpyAsync(ux1, c->st_ux1, nu, hipMemcpyDeviceToHost, c->stream));
  if (ux2) EIF_HIP(hipMemcpyAsync(ux2, c->st_ux2, nu, hipMemcpyDeviceToHost, c->stream));
  if (uy1) EIF_HIP(hipMemcpyAsync(uy1, c->st_uy1, np, hipMemcpyDeviceToHost, c->stream));
  if (uy2) EIF_HIP(hipMemcpyAsync(uy2, c->st_uy2, np, hipMemcpyDeviceToHost, c->stream));
  if (phi) EIF_HIP(hipMemcpyAsync(phi, c->st_phi, np, hipMemcpyDeviceToHost, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  if (num) *num = c->st_num;
  return EIF_OK;
}

// drops captured graphs (defined with the graph steps)
static int graphs_recapture(eif_ctx *c);

int suspensions_gpu_create(eif_ctx *c, int n, const double *den, const double *a,
                           const double *b, const double *x, const double *y,
                           const double *az, const double *ux, const double *uy,
                           const double *vz) {
  CHECK_CTX(c);
  if (n < 0 || (n > 0 && !(den && a && b && x && y && az && ux && uy && vz))) {
    set_error("invalid particle arrays");
    return EIF_ERR_ARG;
  }
  if (c->np > 0 || c->prt) {
    set_error("particles already created for this context");
    return EIF_ERR_STATE;
  }
  for (int k = 0; k < n; k++)
    if (!(a[k] > 0.) || !(b[k] > 0.) || !(den[k] > 0.)) {
      set_error("particle semi-axes and densities must be positive");
      return EIF_ERR_ARG;
    }
  std::vector<double> rec((size_t)(n > 0 ? n : 1) * R_N, 0.);
  for (int k = 0; k < n; k++) {
    double *r = rec.data() + (size_t)k * R_N;
    r[R_DEN] = den[k]; r[R_A] = a[k]; r[R_B] = b[k];
    r[R_X] = x[k]; r[R_Y] = y[k]; r[R_AZ] = az[k];
    r[R_UX] = ux[k]; r[R_UY] = uy[k]; r[R_VZ] = vz[k];
  }
  EIF_HIP(hipMalloc(&c->prt, rec.size() * sizeof(double)));
  EIF_HIP(hipMemcpy(c->prt, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice));
  c->np = n;
  return susp_init(c);
}

int suspensions_gpu_n_particles(const eif_ctx *c) { return c ? c->np : 0; }

int suspensions_gpu_download(eif_ctx *c, double *records) {
  CHECK_CTX(c);
  if (c->np == 0) return EIF_OK;
  EIF_HIP(hipMemcpyAsync(records, c->prt, (size_t)c->np * R_N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  return EIF_OK;
}

int suspensions_gpu_upload(eif_ctx *c, const double *records) {
  CHECK_CTX(c);
  if (c->np == 0) return EIF_OK;
  EIF_HIP(hipMemcpyAsync(c->prt, records, (size_t)c->np * R_N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // the records may change shapes and positions: drop the per-particle
  // distance caches
  EIF_HIP(hipMemsetAsync(c->sc_pos, 0, 4 * (size_t)c->np * sizeof(double), c->stream));
  EIF_HIP(hipMemsetAsync(c->sc_key, 0, 8 * (size_t)c->np * sizeof(int), c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  // new sizes may change the cells of the collision broad phase, which
  // captured graphs hold
  const int ncx = c->cl_ncx, ncy = c->cl_ncy;
  const double hx = c->cl_hx, hy = c->cl_hy;
  TRY(coll_cells_setup(c, records));
  if (ncx != c->cl_ncx || ncy != c->cl_ncy || hx != c->cl_hx || hy != c->cl_hy)
    TRY(graphs_recapture(c));
  return EIF_OK;
}

int suspensions_gpu_reset_particle_increments(eif_ctx *c) {
  CHECK_CTX(c);
  return launch_reset(c);
}

int suspensions_gpu_compute_inertia(eif_ctx *c, int cnstep) {
  CHECK_CTX(c);
  if (cnstep != 0 && cnstep != 1) {
    set_error("cnstep must be 0 or 1");
    return EIF_ERR_ARG;
  }
  return launch_inertia(c, cnstep);
}

int suspensions_gpu_exchange_momentum(eif_ctx *c) {
  CHECK_CTX(c);
  return launch_exchange(c);
}

int suspensions_gpu_update_momentum_fleid(eif_ctx *c) {
  CHECK_CTX(c);
  fields_changed(c);
  return launch_update_momentum(c);
}

int suspensions_gpu_compute_collision_force(eif_ctx *c, int cnstep) {
  CHECK_CTX(c);
  if (cnstep != 0 && cnstep != 1) {
    set_error("cnstep must be 0 or 1");
    return EIF_ERR_ARG;
  }
  return launch_collision(c, cnstep);
}

int suspensions_gpu_increment_particles(eif_ctx *c, int rk, double *residual) {
  CHECK_CTX(c);
  TRY(check_rk(rk));
  TRY(launch_increment(c, rk));
  if (residual) {
    EIF_HIP(hipMemcpyAsync(c->hscal + S_RESIDUAL, c->dscal + S_RESIDUAL, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EIF_HIP(hipStreamSynchronize(c->stream));
    *residual = c->hscal[S_RESIDUAL];
    TRY(check_flags(c));
  }
  return EIF_OK;
}

int suspensions_gpu_update_particles(eif_ctx *c) {
  CHECK_CTX(c);
  return launch_update_particles(c);
}

int suspensions_gpu_colliding_pairs(eif_ctx *c, int capacity, int *pairs, int *n_pairs) {
  CHECK_CTX(c);
  std::vector<std::pair<int, int>> out;
  if (c->np > 0) {
    // pairs n < m from the per-particle partner rows of the last collision call
    std::vector<int> cnt(c->np), rows((size_t)c->np * CONTACT_CAP), hits(2 * (size_t)c->np);
    EIF_HIP(hipMemcpyAsync(cnt.data(), c->row_cnt, (size_t)c->np * sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
    EIF_HIP(hipMemcpyAsync(rows.data(), c->rows, rows.size() * sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
    EIF_HIP(hipMemcpyAsync(hits.data(), c->wall_hit, hits.size() * sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
    EIF_HIP(hipStreamSynchronize(c->stream));
    for (int n = 0; n < c->np; n++)
      for (int k = 0; k < std::min(cnt[n], CONTACT_CAP); k++)
        if (rows[(size_t)n * CONTACT_CAP + k] > n)
          out.push_back({n, rows[(size_t)n * CONTACT_CAP + k]});
    std::sort(out.begin(), out.end());
    for (int w = 0; w < 2; w++)
      for (int n = 0; n < c->np; n++)
        if (hits[2 * n + w]) out.push_back({n, -1 - w});
  }
  if (n_pairs) *n_pairs = (int)out.size();
  if (pairs)
    for (int s = 0; s < (int)out.size() && s < capacity; s++) {
      pairs[2 * s] = out[s].first;
      pairs[2 * s + 1] = out[s].second;
    }
  if ((int)out.size() > capacity && pairs) {
    set_error("pair buffer too small");
    return EIF_ERR_CAPACITY;
  }
  return EIF_OK;
}

}  // extern "C"

namespace eif {

// integrate(), src/main.c:13-55
// Sub-iterations first .. first+n-1 of the stage's Crank-Nicolson loop
// (src/main.c:37-50), queued speculatively: each kernel checks the
// previous sub-iteration's residual (sub_open), so the ones queued past
// convergence do nothing.  Then the status record for the host.
static int queue_substeps(Ctx *c, int rk, int first, int n) {
  if (first + n > RES_SLOTS) {
    set_error("sub-iteration slots exhausted");
    return EIF_ERR_CAPACITY;
  }
  for (int s = first; s < first + n; s++) {
    if (c->recording) TRY(rec_sub(c, s));
    TRY(launch_substep_forces(c, s));
    TRY(launch_increment(c, rk, s));
  }
  TRY(launch_status(c, first + n));
  return EIF_OK;
}

// Waits for the status record queued last.  The host polls the pinned
// record (an event record between the batch and the kernels queued behind
// it would idle the stream).  Only a long wait queries the stream (a query
// can itself cost the stream a pause), so a failed kernel still ends the
// wait with its error.
static int wait_seq(Ctx *c, const unsigned *seq, unsigned want, const char *what) {
  for (long spin = 1;; spin++) {
    if (__atomic_load_n(seq, __ATOMIC_ACQUIRE) == want) return EIF_OK;
    if (spin >= (1L << 20) && spin % (1L << 14) == 0) {
      const hipError_t e = hipStreamQuery(c->stream);
      if (e == hipSuccess) {
        if (__atomic_load_n(seq, __ATOMIC_ACQUIRE) == want) return EIF_OK;
        set_error(std::string(what) + " record missing after the stream drained");
        return EIF_ERR_HIP;
      }
      if (e != hipErrorNotReady) return hip_fail(e, "hipStreamQuery", __FILE__, __LINE__);
    }
  }
}

static int wait_status(Ctx *c, HostStatus *out) {
  TRY(wait_seq(c, &c->hstat->seq, c->hstat_seq, "sub-iteration status"));
  out->residual = c->hstat->residual;
  out->dt = c->hstat->dt;
  out->substeps = c->hstat->substeps;
  out->flags = c->hstat->flags;
  return EIF_OK;
}

// the velocity buffers swapped with their ping-pong partners (the stage's
// own fields while the next stage's predictor is already queued)
static void swap_fields(Ctx *c) {
  std::swap(c->ux, c->ux_alt);
  std::swap(c->uy, c->uy_alt);
}

// The stage start up to the Crank-Nicolson loop (src/main.c:15-36).
// pred_queued: the stage's predictor was queued behind the previous
// stage's sub-iterations; side_start: ... and the previous loop ended at
// ev_fork, so the particle update, reset and distances may run beside it.
static int stage_start(Ctx *c, int rk, bool pred_queued, bool side_start) {
  // Two streams.  Main: fluid kernels, exchange (with inertia(0) in the
  // same pass), deposit (with the exchange's particle sums in the same
  // launch) and the sub-iterations (inertia(1) with collision(1) in one
  // launch, then the increment).  Side: the particle update, reset and
  // stage-start distances beside the stage's predictor (stage 0, and later
  // stages whose previous loop ended with its first batch), then
  // collision(0), which overlaps the fluid kernels.  Events join the side
  // work back before anything that reads its results or reuses its
  // buffers.
  if (c->recording) TRY(rec_stage_begin(c, rk));
  if (!pred_queued) {
    TRY(launch_reset(c, c->stream, true));
    // src/main.c:16-19: the kernels of every stage keep the boundary and
    // halo values they change up to date, so this only runs after a
    // host-side change of the fields
    if (!c->bc_valid) {
      TRY(launch_bc_ux(c, c->ux));
      TRY(launch_bc_uy(c, c->uy));
      TRY(launch_bc_p(c, c->p));
      c->bc_valid = true;
    }
    if (c->recording) TRY(rec_fields(c, 0, c->ux, c->uy));  // u^k
    // the distances at the stage-start positions and collision(0) beside
    // the predictor (queued first: at a step's start the device waits
    // for these submissions)
    EIF_HIP(hipEventRecord(c->ev_fork, c->stream));
    TRY(launch_predictor_prev(c, rk));  // writes the spare ux, uy buffers
    if (c->recording) TRY(rec_fields(c, 1, c->ux, c->uy));  // u*
    EIF_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    TRY(launch_ibm_dist(c, c->side));
    EIF_HIP(hipEventRecord(c->ev_i0, c->side));
    TRY(launch_collision(c, 0, c->side));
    EIF_HIP(hipEventRecord(c->ev_c0, c->side));
    EIF_HIP(hipStreamWaitEvent(c->stream, c->ev_i0, 0));
  } else if (side_start) {
    if (c->recording) {  // the stage-start velocity and the queued predictor's
      TRY(rec_fields(c, 0, c->ux_alt, c->uy_alt));
      TRY(rec_fields(c, 1, c->ux, c->uy));
    }
    // the previous stage's loop ended with its first batch, which ev_fork
    // closes: the particle update, this stage's reset and the distances
    // at the new positions run on the side stream beside the predictor
    EIF_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    TRY(launch_update_particles(c, c->side));
    TRY(launch_reset(c, c->side, true));
    TRY(launch_ibm_dist(c, c->side));
    EIF_HIP(hipEventRecord(c->ev_i0, c->side));
    TRY(launch_collision(c, 0, c->side));
    EIF_HIP(hipEventRecord(c->ev_c0, c->side));
    EIF_HIP(hipStreamWaitEvent(c->stream, c->ev_i0, 0));
  } else {
    if (c->recording) {
      TRY(rec_fields(c, 0, c->ux_alt, c->uy_alt));
      TRY(rec_fields(c, 1, c->ux, c->uy));
    }
    // behind the predictor (and the batches queued after it): the previous
    // stage's particle update, this stage's reset and the distances at the
    // new positions; collision(0) on the side stream
    TRY(launch_update_particles(c));
    TRY(launch_reset(c, c->stream, true));
    TRY(launch_ibm_dist(c));
    EIF_HIP(hipEventRecord(c->ev_fork, c->stream));
    EIF_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    TRY(launch_collision(c, 0, c->side));
    EIF_HIP(hipEventRecord(c->ev_c0, c->side));
  }
  // the exchange with the predicted velocity and, in the same pass,
  // inertia(0) with the stage-start velocity: the pair the predictor read.
  // The reset cleared the tile registrations.
  TRY(launch_exchange_inertia0(c, c->ux_alt, c->uy_alt));
  // the deposit, with the exchange's and inertia(0)'s particle sums in the
  // same launch
  TRY(launch_update_momentum_x0(c));
  // the potential, the projection and the pressure update, fused where the
  // row length allows; the last stage also reduces the advective dt limit
  // of the final velocities for the next param_decide_dt
  TRY(launch_poisson_project(c, rk, rk == 2));
  if (c->recording) TRY(rec_fields(c, 2, c->ux, c->uy));  // u^{k+1}
  // the sub-iterations read collision(0)
  EIF_HIP(hipStreamWaitEvent(c->stream, c->ev_c0, 0));
  if (c->recording) TRY(rec_p0(c));  // the stage-start particles (increments reset)
  return EIF_OK;
}

// integrate() from stage rk0; sub0 > 0 resumes that stage's loop at
// sub-iteration sub0 (its start and first sub0 sub-iterations done, the
// next stage's predictor not queued).
//
// The predictor of stage rk+1 reads only the fluid fields, which the
// sub-iterations of stage rk leave alone.  It is queued right behind the
// first batch of sub-iterations, so the GPU runs it while the host waits
// for that batch's count.  (Beside the sub-iterations, on the side stream,
// it slows them more than it gains.)
static int integrate_from(Ctx *c, int rk0, int sub0, int substeps[3]) {
  bool pred_queued = false;  // this stage's predictor is already queued
  bool side_start = false;   // ... and the previous loop ended at ev_fork
  for (int rk = rk0; rk < 3; rk++) {
    const bool resume = rk == rk0 && sub0 > 0;
    if (!resume) TRY(stage_start(c, rk, pred_queued, side_start));
    int queued = resume ? sub0 : 0, done = queued;
    int batch = std::min(std::max(c->substep_guess - queued, 1), 102);
    bool early = false;  // predictor of stage rk+1 queued
    int batches = 0;
    for (;;) {
      if (early) swap_fields(c);  // the stage's own velocity for inertia(1)
      const int r = queue_substeps(c, rk, queued, batch);
      if (early) swap_fields(c);
      TRY(r);
      queued += batch;
      batches++;
      if (!early && rk < 2) {
        EIF_HIP(hipEventRecord(c->ev_fork, c->stream));  // the end of the first batch
        TRY(launch_predictor_prev(c, rk + 1));
        early = true;
      }
      HostStatus st;
      TRY(wait_status(c, &st));
      if (st.flags) {
        EIF_HIP(hipStreamSynchronize(c->side));
        EIF_HIP(hipStreamSynchronize(c->stream));
        EIF_HIP(hipMemsetAsync(c->flags, 0, sizeof(int), c->stream));
        EIF_HIP(hipStreamSynchronize(c->stream));
        return flags_error(st.flags);
      }
      c->hscal[S_DT] = st.dt;
      done = st.substeps;
      if (done < queued || st.residual < 1.e-8 || done - 1 > 100) break;
      batch = 1;
    }
    c->substep_guess = done;
    if (substeps) substeps[rk] = done;
    if (c->recording) TRY(rec_stage_end(c, done));
    pred_queued = early;
    side_start = early && batches == 1 && !resume;
  }
  TRY(launch_update_particles(c));
  c->dt_adv_valid = c->dt_adv_parts > 0;
  return EIF_OK;
}

// ---------------------------------------------------------------------
// Whole-step graphs.  One time step (decide_dt + the three stages) is
// captured with K[rk] sub-iterations queued per stage; a gate after them
// (k_stage_gate) stops the rest of the step, through the device flag every
// state-changing kernel tests, if a loop needs more.  The host then resumes
// that stage eagerly (integrate_from) and captures a graph with more
// sub-iterations.  A step maps its three velocity pairs onto themselves
// (three predictor rotations), so one graph per starting assignment serves
// every later step.
// ---------------------------------------------------------------------
namespace {
struct PtrState {
  double *p[6];
  bool src_lazy;
};

PtrState save_ptrs(const Ctx *c) {
  return PtrState{{c->ux, c->uy, c->ux_alt, c->uy_alt, c->ux_spare, c->uy_spare}, c->src_lazy};
}

void load_ptrs(Ctx *c, const PtrState &s) {
  c->ux = s.p[0]; c->uy = s.p[1]; c->ux_alt = s.p[2]; c->uy_alt = s.p[3];
  c->ux_spare = s.p[4]; c->uy_spare = s.p[5];
  c->src_lazy = s.src_lazy;
}

bool same_ptrs(const PtrState &a, const PtrState &b) {
  for (int k = 0; k < 6; k++)
    if (a.p[k] != b.p[k]) return false;
  return true;
}
}  // namespace

struct GraphEntry {
  int K[3];
  PtrState start;
  PtrState at_gate[3];  // host pointer state right after each stage's gate
  int dt_parts;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;  // kept: an ensemble step nests it as a child graph
};

struct GraphCache {
  std::vector<std::unique_ptr<GraphEntry>> entries;  // stable addresses
  int K[3] = {0, 0, 0};  // sub-iterations captured per stage
  int low_run[3] = {0, 0, 0};  // consecutive steps that needed 2+ fewer
  long launches = 0, resumes = 0, captures = 0;
  unsigned seq = 0;  // number of the last graph step launched (device counter)
};

void batch_forget(Ctx *c);

void graphs_free(Ctx *c) {
  batch_forget(c);
  if (!c->graphs) return;
  for (auto &e : c->graphs->entries) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
  }
  delete c->graphs;
  c->graphs = nullptr;
}

// The kernels of one graph step queued on the context's streams: decide_dt,
// the three stages with K[rk] sub-iterations each and a gate after them, the
// particle update and the step's record.  at_gate: the host pointer state
// right after each stage's gate.
static int queue_graph_step(Ctx *c, const int K[3], PtrState at_gate[3]) {
  TRY(launch_decide_dt(c));
  for (int rk = 0; rk < 3; rk++) {
    // one batch per stage: the next predictor follows it, so stages 1 and
    // 2 start on the side stream (stage_start's side_start form)
    TRY(stage_start(c, rk, rk > 0, rk > 0));
    for (int s = 0; s < K[rk]; s++) {
      TRY(launch_substep_forces(c, s));
      TRY(launch_increment(c, rk, s));
    }
    TRY(launch_stage_gate(c, rk, K[rk]));
    at_gate[rk] = save_ptrs(c);
    if (rk < 2) {
      EIF_HIP(hipEventRecord(c->ev_fork, c->stream));
      TRY(launch_predictor_prev(c, rk + 1));
    }
  }
  TRY(launch_update_particles(c));
  return launch_graph_status(c);
}

static int capture_step(Ctx *c, const int K[3], GraphEntry *out) {
  // a graph step follows a step whose last projection left the fused dt
  // partials
  c->dt_adv_valid = c->dt_adv_parts > 0;
  const PtrState start = save_ptrs(c);
  out->start = start;
  for (int k = 0; k < 3; k++) out->K[k] = K[k];
  const bool prof_was = c->prof->on;
  c->prof->on = false;
  hipGraph_t graph = nullptr;
  EIF_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  out->dt_parts = 0;
  const int rc = queue_graph_step(c, K, out->at_gate);
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  c->prof->on = prof_was;
  const PtrState end = save_ptrs(c);
  load_ptrs(c, start);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  if (!same_ptrs(start, end)) {
    (void)hipGraphDestroy(graph);
    set_error("graph step: the velocity buffers do not return to their assignment");
    return EIF_ERR_STATE;
  }
  out->dt_parts = c->dt_adv_parts;
  const hipError_t ei = hipGraphInstantiate(&out->exec, graph, nullptr, nullptr, 0);
  out->graph = graph;
  if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate", __FILE__, __LINE__);
  c->graphs->captures++;
  return EIF_OK;
}

static int graph_for(Ctx *c, GraphEntry **out) {
  GraphCache *g = c->graphs;
  const PtrState now = save_ptrs(c);
  for (auto &e : g->entries)
    if (same_ptrs(e->start, now) && e->K[0] == g->K[0] && e->K[1] == g->K[1] &&
        e->K[2] == g->K[2] && e->dt_parts == c->dt_adv_parts) {
      *out = e.get();
      return EIF_OK;
    }
  std::unique_ptr<GraphEntry> e(new GraphEntry());
  TRY(capture_step(c, g->K, e.get()));
  g->entries.push_back(std::move(e));
  *out = g->entries.back().get();
  return EIF_OK;
}

// the record of the graph step numbered seq
static int graph_wait(Ctx *c, unsigned seq, GraphStatus *out) {
  GraphStatus *slot = c->gstat + (seq & 1);
  TRY(wait_seq(c, &slot->seq, seq, "graph step status"));
  *out = *slot;
  return EIF_OK;
}

// decide_dt + integrate, eagerly
static int eager_step(Ctx *c, int sub[3]) {
  TRY(launch_decide_dt(c));
  return integrate_from(c, 0, 0, sub);
}

// Before graph steps: the cache trimmed (between runs every graph is done),
// and, when the fields were changed from the host or no counts are known
// yet, one eager step (*ran) that leaves current boundary values, the fused
// dt partials and the first sub-iteration counts.
static int graph_prepare(Ctx *c, bool *ran, int sub[3]) {
  GraphCache *g = c->graphs;
  *ran = false;
  if (g->entries.size() >= 32) {
    EIF_HIP(hipStreamSynchronize(c->stream));
    for (auto &e : g->entries) {
      if (e->exec) (void)hipGraphExecDestroy(e->exec);
      if (e->graph) (void)hipGraphDestroy(e->graph);
    }
    g->entries.clear();
  }
  if (!c->bc_valid || !c->dt_adv_valid || g->K[0] == 0) {
    TRY(eager_step(c, sub));
    for (int r = 0; r < 3; r++)
      if (g->K[r] == 0) g->K[r] = sub[r];
    *ran = true;
  }
  return EIF_OK;
}

// one graph step queued on the context's stream
static int graph_submit(Ctx *c, GraphEntry **e, unsigned *seq) {
  TRY(graph_for(c, e));
  EIF_HIP(hipGraphLaunch((*e)->exec, c->stream));
  *seq = ++c->graphs->seq;
  c->graphs->launches++;
  return EIF_OK;
}

// A graph step's record: capacity flags become the error; a stop is
// resumed eagerly where it stopped (nothing may be queued behind it that
// changes state), and K follows the counts.  dt, sub: the step's.
static int graph_finish(Ctx *c, GraphEntry *e, const GraphStatus &st, double *dt, int sub[3]) {
  GraphCache *g = c->graphs;
  if (st.flags) {
    EIF_HIP(hipStreamSynchronize(c->stream));
    EIF_HIP(hipMemsetAsync(c->flags, 0, sizeof(int), c->stream));
    EIF_HIP(hipStreamSynchronize(c->stream));
    return flags_error(st.flags);
  }
  for (int r = 0; r < 3; r++) sub[r] = st.counts[r];
  if (st.halt) {
    const int rk = st.halt - 1;
    EIF_HIP(hipStreamSynchronize(c->stream));
    load_ptrs(c, e->at_gate[rk]);
    EIF_HIP(hipMemsetAsync(c->halt, 0, sizeof(int), c->stream));
    c->substep_guess = e->K[rk] + 1;
    TRY(integrate_from(c, rk, e->K[rk], sub));
    g->resumes++;
    *dt = c->hscal[S_DT];
    g->K[rk] = sub[rk];
    g->low_run[rk] = 0;
    return EIF_OK;
  }
  *dt = st.dt;
  c->hscal[S_DT] = st.dt;
  // K follows the counts: a stage captured with one sub-iteration too many
  // for 16 steps in a row loses it (a stop costs the host one resumption;
  // an idle sub-iteration costs two launches every step)
  for (int r = 0; r < 3; r++) {
    if (st.counts[r] < g->K[r])
      g->low_run[r]++;
    else
      g->low_run[r] = 0;
    if (g->low_run[r] >= 16) {
      g->K[r]--;
      g->low_run[r] = 0;
    }
  }
  return EIF_OK;
}

// ---------------------------------------------------------------------
// Batched ensemble steps.  Every case's graph step is recorded launch by
// launch (go<B> with a Recorder set: the batched kernel, the case's grid,
// block and LDS bytes, and its arguments), then replayed as batched
// launches: at each position of the cases' launch sequences, the cases whose
// launch there runs the same kernel body with the same block and LDS size
// share one launch (k_batch), each reading its own arguments from a
// descriptor array.  Each case's launches keep their order, cases are
// independent, and a kernel body does per case exactly what its own launch
// does, so every case's results are those of its own graph step.  The
// positions are captured on one stream as one graph per ensemble step.
// ---------------------------------------------------------------------
struct BatchGraph {
  std::vector<Ctx *> cases;
  std::vector<GraphEntry> entries;  // per case: K, start and gate pointer states, dt parts
  int K[3] = {0, 0, 0};
  char *arena = nullptr;            // the descriptor arrays of every launch
  hipStream_t stream = nullptr;     // where the batched step runs
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nodes = 0;                    // batched launches per ensemble step
};

static void batch_free(BatchGraph *b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->arena) (void)hipFree(b->arena);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->exec = nullptr;
  b->graph = nullptr;
  b->arena = nullptr;
  b->stream = nullptr;
}

// the batched steps built so far, kept between calls (a context's graph
// state reset drops every batch it belongs to)
static std::mutex g_batch_mu;
static std::vector<std::unique_ptr<BatchGraph>> g_batches;
constexpr size_t BATCH_CACHE = 16;

void batch_forget(Ctx *c) {
  std::lock_guard<std::mutex> lk(g_batch_mu);
  for (size_t k = 0; k < g_batches.size();) {
    BatchGraph *b = g_batches[k].get();
    if (std::find(b->cases.begin(), b->cases.end(), c) != b->cases.end()) {
      batch_free(b);
      g_batches.erase(g_batches.begin() + k);
    } else {
      k++;
    }
  }
}

// one case's graph step as recorded launches.  The step is queued inside a
// stream capture that receives nothing but the streams' event joins: a
// launch that went past the recorder would show there as a node.
static int record_case(Ctx *c, const int K[3], GraphEntry *ge, std::vector<LaunchRec> *out) {
  c->dt_adv_valid = c->dt_adv_parts > 0;
  const PtrState start = save_ptrs(c);
  ge->start = start;
  for (int k = 0; k < 3; k++) ge->K[k] = K[k];
  const bool prof_was = c->prof->on;
  c->prof->on = false;
  Recorder rec;
  EIF_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  g_rec = &rec;
  const int rc = queue_graph_step(c, K, ge->at_gate);
  g_rec = nullptr;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->prof->on = prof_was;
  const PtrState end = save_ptrs(c);
  load_ptrs(c, start);
  size_t nn = 0;
  hipError_t eg = g ? hipGraphGetNodes(g, nullptr, &nn) : hipSuccess;
  std::vector<hipGraphNode_t> nodes(nn);
  if (eg == hipSuccess && nn) eg = hipGraphGetNodes(g, nodes.data(), &nn);
  int stray = 0;
  for (size_t k = 0; k < nn && eg == hipSuccess; k++) {
    hipGraphNodeType t;
    if ((eg = hipGraphNodeGetType(nodes[k], &t)) != hipSuccess) break;
    if (t == hipGraphNodeTypeKernel || t == hipGraphNodeTypeMemset ||
        t == hipGraphNodeTypeMemcpy)
      stray++;
  }
  if (g) (void)hipGraphDestroy(g);
  TRY(rc);
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  if (eg != hipSuccess) return hip_fail(eg, "hipGraphGetNodes", __FILE__, __LINE__);
  if (stray) {
    set_error("batched step: " + std::to_string(stray) + " operation(s) bypassed the recorder");
    return EIF_ERR_STATE;
  }
  if (!same_ptrs(start, end)) {
    set_error("batched step: the velocity buffers do not return to their assignment");
    return EIF_ERR_STATE;
  }
  ge->dt_parts = c->dt_adv_parts;
  *out = std::move(rec.recs);
  return EIF_OK;
}

static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// records every case's step with K sub-iterations per stage and captures
// the batched replay on stream bs
static int batch_build(const std::vector<Ctx *> &cs, const int K[3], BatchGraph *b) {
  const int n = (int)cs.size();
  if (n < 1 || n > BATCH_MAX) {
    set_error("batched step: 1 to 64 cases per batch");
    return EIF_ERR_ARG;
  }
  b->cases = cs;
  b->entries.assign(n, GraphEntry());
  for (int r = 0; r < 3; r++) b->K[r] = K[r];
  std::vector<std::vector<LaunchRec>> recs(n);
  size_t maxlen = 0;
  for (int k = 0; k < n; k++) {
    TRY(record_case(cs[k], K, &b->entries[k], &recs[k]));
    maxlen = std::max(maxlen, recs[k].size());
  }
  struct Node {
    const void *fn;
    dim3 block;
    size_t shm, off;
    BatchMap map;
    unsigned total;
  };
  std::vector<Node> nodes;
  std::vector<char> host;
  std::vector<char> used(n);
  for (size_t pos = 0; pos < maxlen; pos++) {
    std::fill(used.begin(), used.end(), 0);
    for (int k = 0; k < n; k++) {
      if (used[k] || pos >= recs[k].size()) continue;
      const LaunchRec &r0 = recs[k][pos];
      std::vector<int> grp;
      for (int j = k; j < n; j++) {
        if (used[j] || pos >= recs[j].size()) continue;
        const LaunchRec &r = recs[j][pos];
        if (r.fn == r0.fn && r.block.x == r0.block.x && r.block.y == r0.block.y &&
            r.block.z == r0.block.z && r.shm == r0.shm && r.pack.size() == r0.pack.size()) {
          grp.push_back(j);
          used[j] = 1;
        }
      }
      Node nd;
      std::memset(&nd.map, 0, sizeof nd.map);
      nd.fn = r0.fn;
      nd.block = r0.block;
      nd.shm = r0.shm;
      const size_t stride = round_up(r0.pack.size(), 16);
      nd.off = round_up(host.size(), 256);
      host.resize(nd.off + stride * grp.size());
      nd.map.n = (int)grp.size();
      nd.map.stride = (unsigned)stride;
      unsigned long long total = 0;
      for (size_t g = 0; g < grp.size(); g++) {
        const LaunchRec &r = recs[grp[g]][pos];
        std::memcpy(host.data() + nd.off + g * stride, r.pack.data(), r.pack.size());
        nd.map.start[g] = (unsigned)total;
        nd.map.gx[g] = r.grid.x;
        nd.map.gy[g] = r.grid.y;
        nd.map.gz[g] = r.grid.z;
        total += (unsigned long long)r.grid.x * r.grid.y * r.grid.z;
      }
      if (total >= (1ull << 31)) {
        set_error("batched step: too many workgroups in one launch");
        return EIF_ERR_CAPACITY;
      }
      nd.map.start[grp.size()] = (unsigned)total;
      nd.total = (unsigned)total;
      if (total > 0) nodes.push_back(nd);
    }
  }
  EIF_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  const hipStream_t bs = b->stream;
  EIF_HIP(hipMalloc(&b->arena, std::max<size_t>(host.size(), 256)));
  EIF_HIP(hipMemcpy(b->arena, host.data(), host.size(), hipMemcpyHostToDevice));
  EIF_HIP(hipStreamBeginCapture(bs, hipStreamCaptureModeThreadLocal));
  hipError_t el = hipSuccess;
  for (size_t k = 0; k < nodes.size() && el == hipSuccess; k++) {
    Node &nd = nodes[k];
    const char *packs = b->arena + nd.off;
    void *args[] = {(void *)&packs, (void *)&nd.map};
    el = hipLaunchKernel(nd.fn, dim3(nd.total), nd.block, args, nd.shm, bs);
  }
  const hipError_t e = hipStreamEndCapture(bs, &b->graph);
  if (el != hipSuccess) return hip_fail(el, "hipLaunchKernel", __FILE__, __LINE__);
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  EIF_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->nodes = (int)nodes.size();
  return EIF_OK;
}

// whether b replays the cases' next step: the same cases, K and states
static bool batch_matches(const BatchGraph &b, const std::vector<Ctx *> &cs, const int K[3]) {
  if (b.cases != cs) return false;
  for (int r = 0; r < 3; r++)
    if (b.K[r] != K[r]) return false;
  for (size_t k = 0; k < cs.size(); k++)
    if (!same_ptrs(b.entries[k].start, save_ptrs(cs[k])) ||
        b.entries[k].dt_parts != cs[k]->dt_adv_parts)
      return false;
  return true;
}

// One batched ensemble step launched: the cases of `run` in batches of at
// most BATCH_MAX, each batch's graph (cached, or recorded and captured now)
// on its own stream; every launched case's record number is taken.  `busy`:
// graphs still running, which the cache keeps.
struct StepLaunch {
  std::vector<BatchGraph *> graphs;
  std::vector<unsigned> seq;  // per case of each graph, in order
};

// Batches of one step launched side by side.  Measured on the 32-case
// ensemble (14 cases of 1024x2048, 11 of 512x1024, 7 of 256x512), in
// case-steps/s: cases dealt round robin into 1 batch 2399, 2 batches
// 2594-2667, 4 batches 2489; one batch per grid shape 2740-2749 (the
// 14-case batch split in two 2722-2750, every shape of more than 8 cases
// split 2652-2655).  An ensemble of two grid shapes or more is batched
// per shape; an ensemble of one shape is dealt round robin into
// BATCH_CONCURRENT batches.
constexpr size_t BATCH_CONCURRENT = 2;
constexpr bool BATCH_BY_SHAPE = true;

static int batch_launch_step(const std::vector<Ctx *> &run, const StepLaunch &busy,
                             StepLaunch *out) {
  out->graphs.clear();
  out->seq.clear();
  std::vector<Ctx *> order(run);
  std::stable_sort(order.begin(), order.end(), [](const Ctx *a, const Ctx *b) {
    return (long)a->nx * a->ny > (long)b->nx * b->ny;
  });
  // the batches, each on its own stream so that their kernels interleave on
  // the GPU (the VALU-bound IBM launches of one beside the memory-bound
  // fluid launches of another)
  std::vector<std::vector<Ctx *>> batches;
  if (BATCH_BY_SHAPE) {
    // one batch per grid shape (chunks of at most BATCH_MAX): kernels
    // specialised on the grid (its DCT length, power-of-two spacing) are
    // then one launch per position of the step, not one per shape in turn
    for (size_t k = 0; k < order.size();) {
      size_t e = k;
      while (e < order.size() && order[e]->nx == order[k]->nx && order[e]->ny == order[k]->ny) e++;
      for (size_t c0 = k; c0 < e; c0 += BATCH_MAX)
        batches.emplace_back(order.begin() + c0, order.begin() + std::min(e, c0 + BATCH_MAX));
      k = e;
    }
  }
  if (batches.size() < 2) {
    // at least BATCH_CONCURRENT batches: cases dealt round robin in order
    // of decreasing grid size, so the batches carry similar work
    batches.clear();
    const size_t nbat = std::min(order.size(), std::max<size_t>(
        (order.size() + BATCH_MAX - 1) / BATCH_MAX, BATCH_CONCURRENT));
    batches.resize(nbat);
    for (size_t k = 0; k < order.size(); k++) batches[k % nbat].push_back(order[k]);
  }
  for (const std::vector<Ctx *> &cs : batches) {
    // each batch captures as many sub-iterations per stage as its cases last needed
    int K[3] = {0, 0, 0};
    for (Ctx *c : cs)
      for (int r = 0; r < 3; r++) K[r] = std::max(K[r], c->graphs->K[r]);
    BatchGraph *b = nullptr;
    for (auto &x : g_batches)
      if (batch_matches(*x, cs, K)) b = x.get();
    if (!b) {
      // a cached step that is not running is idle
      auto in_use = [&](BatchGraph *x) {
        return std::find(busy.graphs.begin(), busy.graphs.end(), x) != busy.graphs.end() ||
               std::find(out->graphs.begin(), out->graphs.end(), x) != out->graphs.end();
      };
      for (size_t j = 0; j < g_batches.size() && g_batches.size() >= BATCH_CACHE;) {
        if (!in_use(g_batches[j].get())) {
          batch_free(g_batches[j].get());
          g_batches.erase(g_batches.begin() + j);
        } else {
          j++;
        }
      }
      g_batches.emplace_back(new BatchGraph());
      b = g_batches.back().get();
      const int rc = batch_build(cs, K, b);
      if (rc) {
        batch_free(b);
        g_batches.pop_back();
        return rc;
      }
      for (Ctx *c : cs) c->graphs->captures++;
    }
    // the chunk's cases are idle on their own streams (host-synchronised)
    EIF_HIP(hipGraphLaunch(b->exec, b->stream));
    out->graphs.push_back(b);
    for (Ctx *c : b->cases) {
      c->graphs->launches++;
      out->seq.push_back(++c->graphs->seq);
    }
  }
  return EIF_OK;
}

// nsteps graph steps of every case, as batched ensemble steps of up to
// BATCH_MAX cases per batch (batches on their own streams).  Steps are pipelined
// like eif_gpu_run_steps: the next batched step is launched before the
// records of the previous one are read.  A case that stopped at a gate
// (its loop needed more sub-iterations than captured) did nothing in the
// step behind it; after that step has drained, its record is consumed, the
// case is resumed alone and its step is run again later.
static int ensemble_batched(eif_ctx **ctxs, int ncases, int nsteps, int *nodes_out) {
  std::vector<int> queued(ncases, 0);  // steps launched (or run) per case
  for (int k = 0; k < ncases; k++) {
    bool ran = false;
    int sub[3];
    if (nsteps > 0) TRY(graph_prepare(ctxs[k], &ran, sub));
    queued[k] = ran ? 1 : 0;
  }
  // the batched launches run on their own streams, after each case's queued work
  for (int k = 0; k < ncases; k++) EIF_HIP(hipStreamSynchronize(ctxs[k]->stream));
  auto index = [&](Ctx *c) { return (int)(std::find(ctxs, ctxs + ncases, c) - ctxs); };
  std::lock_guard<std::mutex> lk(g_batch_mu);
  StepLaunch prev, cur;
  std::vector<char> skip(ncases, 0);  // case's record in `prev` already consumed
  int rc = EIF_OK;
  for (;;) {
    std::vector<Ctx *> run;
    for (int k = 0; k < ncases; k++)
      if (queued[k] < nsteps) run.push_back(ctxs[k]);
    cur.graphs.clear();
    cur.seq.clear();
    if (!run.empty()) {
      if ((rc = batch_launch_step(run, prev, &cur))) break;
      for (Ctx *c : run) queued[index(c)]++;
      if (nodes_out && !cur.graphs.empty()) *nodes_out = cur.graphs.back()->nodes;
    }
    if (prev.graphs.empty() && cur.graphs.empty()) break;
    // the records of the previous step
    bool drained = false;  // cur has finished (before any resumption)
    size_t q = 0;
    for (BatchGraph *b : prev.graphs) {
      if (!rc && hipStreamSynchronize(b->stream) != hipSuccess)
        rc = hip_fail(hipErrorUnknown, "hipStreamSynchronize", __FILE__, __LINE__);
      for (size_t j = 0; j < b->cases.size() && !rc; j++, q++) {
        Ctx *c = b->cases[j];
        const int k = index(c);
        if (skip[k]) continue;
        GraphStatus st;
        if ((rc = graph_wait(c, prev.seq[q], &st))) break;
        if (st.halt) {
          // the case's step in cur did nothing; once cur has drained, its
          // record (the same stop) is consumed and the step queued again
          if (!drained) {
            for (BatchGraph *g : cur.graphs)
              if (!rc && hipStreamSynchronize(g->stream) != hipSuccess)
                rc = hip_fail(hipErrorUnknown, "hipStreamSynchronize", __FILE__, __LINE__);
            drained = true;
          }
          size_t qc = 0;
          for (BatchGraph *g : cur.graphs)
            for (Ctx *x : g->cases) {
              if (x == c && !rc) {
                GraphStatus skipped;
                rc = graph_wait(c, cur.seq[qc], &skipped);
                c->graphs->launches--;
                queued[k]--;
                skip[k] = 2;  // consumed in cur
              }
              qc++;
            }
          if (rc) break;
        }
        double dt;
        int sub[3];
        if ((rc = graph_finish(c, &b->entries[j], st, &dt, sub))) break;
        if (st.halt && hipStreamSynchronize(c->stream) != hipSuccess)
          rc = hip_fail(hipErrorUnknown, "hipStreamSynchronize", __FILE__, __LINE__);
      }
    }
    if (rc) break;
    // cases whose cur record was consumed are skipped when cur is read
    for (int k = 0; k < ncases; k++) skip[k] = skip[k] == 2 ? 1 : 0;
    prev = cur;
  }
  for (auto &x : g_batches)
    if (x->stream) (void)hipStreamSynchronize(x->stream);
  for (int k = 0; k < ncases; k++) ctxs[k]->dt_adv_valid = ctxs[k]->dt_adv_parts > 0;
  return rc;
}

// mean time of `reps` launches queued by f, by events around them
template <typename Fn>
static int time_launches(Ctx *c, Fn f, int reps, double *ms) {
  hipEvent_t a, b;
  EIF_HIP(timing_event(&a));
  EIF_HIP(timing_event(&b));
  int rc = f();  // first launch untimed
  if (!rc) rc = hipEventRecord(a, c->stream) == hipSuccess ? EIF_OK : EIF_ERR_HIP;
  for (int k = 0; k < reps && !rc; k++) rc = f();
  if (!rc) rc = hipEventRecord(b, c->stream) == hipSuccess ? EIF_OK : EIF_ERR_HIP;
  float el = 0.f;
  if (!rc && hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&el, a, b) == hipSuccess)
    *ms = el / reps;
  else if (!rc)
    rc = hip_fail(hipErrorUnknown, "timing events", __FILE__, __LINE__);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

}  // namespace eif

extern "C" {

int eif_gpu_integrate(eif_ctx *c, int substeps[3]) {
  CHECK_CTX(c);
  return integrate_from(c, 0, 0, substeps);
}

int eif_gpu_step(eif_ctx *c, double *dt_out) {
  CHECK_CTX(c);
  TRY(launch_decide_dt(c));
  TRY(eif_gpu_integrate(c, nullptr));
  // dt is read back with the first sub-iteration count of the step
  if (dt_out) *dt_out = c->hscal[S_DT];
  return EIF_OK;
}

int eif_gpu_tape_begin(eif_ctx *c) {
  CHECK_CTX(c);
  return tape_begin(c);
}

int eif_gpu_step_fixed_dt(eif_ctx *c, double dt, int record) {
  CHECK_CTX(c);
  if (!(dt > 0.) || !std::isfinite(dt)) {
    set_error("dt must be positive and finite");
    return EIF_ERR_ARG;
  }
  if (record && !c->tape) {
    set_error("no tape: call eif_gpu_tape_begin first");
    return EIF_ERR_STATE;
  }
  // integrate() itself at the given dt; recorded: with the stage records
  // copied as it goes (rec_*)
  c->hscal[S_DT] = dt;
  EIF_HIP(hipMemcpyAsync(c->dscal + S_DT, c->hscal + S_DT, sizeof(double),
                         hipMemcpyHostToDevice, c->stream));
  c->recording = record != 0;
  const int rc = integrate_from(c, 0, 0, nullptr);
  c->recording = false;
  TRY(rc);
  if (record) TRY(rec_check_contacts(c));
  return EIF_OK;
}

int fluid_gpu_wall_shear(eif_ctx *c, double *out) {
  CHECK_CTX(c);
  if (!out) {
    set_error("null output");
    return EIF_ERR_ARG;
  }
  return launch_wall_shear(c, out);
}

int eif_gpu_stage_adjoint(eif_ctx *c) {
  CHECK_CTX(c);
  return stage_adjoint(c);
}

int eif_gpu_adjoint_get(eif_ctx *c, int which, double *host) {
  CHECK_CTX(c);
  if (!host) {
    set_error("null host buffer");
    return EIF_ERR_ARG;
  }
  return adjoint_access(c, which, host, false);
}

int eif_gpu_adjoint_add_wall_shear(eif_ctx *c, double weight) {
  CHECK_CTX(c);
  return adjoint_add_wall_shear(c, weight);
}

int eif_gpu_adjoint_add(eif_ctx *c, int which, const double *host) {
  CHECK_CTX(c);
  if (!host) {
    set_error("null host buffer");
    return EIF_ERR_ARG;
  }
  return adjoint_access(c, which, const_cast<double *>(host), true);
}

int eif_gpu_autotune(eif_ctx *c, int flags, char *report, int report_len) {
  CHECK_CTX(c);
  EIF_HIP(hipStreamSynchronize(c->stream));
  const bool prof_was = c->prof->on;
  c->prof->on = false;
  const int parts_was = c->dt_adv_parts;
  const bool dtv_was = c->dt_adv_valid;
  std::string rep;
  char buf[160];
  int rc = EIF_OK;
  // predictor rows per workgroup (same operations per cell for each)
  {
    double best = 1e30;
    int pick = 0;
    rep += "predictor_rows:";
    for (int rows : {8, 16, 32}) {
      c->pred_rows = rows;
      double ms = 0.;
      if ((rc = time_launches(c, [&] { return launch_predictor_probe(c); }, 10, &ms))) break;
      snprintf(buf, sizeof(buf), " %d=%.4fms", rows, ms);
      rep += buf;
      if (ms < best) {
        best = ms;
        pick = rows;
      }
    }
    c->pred_rows = pick;
    snprintf(buf, sizeof(buf), " -> %d; ", pick);
    rep += buf;
  }
  // projection + pressure rows per workgroup, on copies
  double *ptmp = nullptr;
  if (!rc && hipMalloc(&ptmp, c->L.p_len() * sizeof(double)) != hipSuccess)
    rc = hip_fail(hipErrorOutOfMemory, "hipMalloc", __FILE__, __LINE__);
  if (!rc) {
    double best = 1e30;
    int pick = 0;
    rep += "projection_rows:";
    for (int rows : {4, 8, 16}) {
      c->cp_rows = rows;
      double ms = 0.;
      if ((rc = time_launches(c, [&] {
             return launch_correct_pressure_on(c, 1, true, true, true, c->ux_spare, c->uy_spare,
                                               ptmp);
           }, 10, &ms)))
        break;
      snprintf(buf, sizeof(buf), " %d=%.4fms", rows, ms);
      rep += buf;
      if (ms < best) {
        best = ms;
        pick = rows;
      }
    }
    c->cp_rows = pick;
    snprintf(buf, sizeof(buf), " -> %d; ", pick);
    rep += buf;
  }
  if (ptmp) (void)hipFree(ptmp);
  // the y-solve: chunk length of the partitioned solve (and, with flags &
  // 1, the Thomas sweep too); Thomas-mode contexts stay bit-exact unless
  // flags allow the switch
  if (!rc && (c->cfg.poisson_mode == EIF_POISSON_PARTITION || (flags & 1))) {
    const int mode_was = c->cfg.poisson_mode, L_was = c->part_L;
    double best = 1e30;
    int pick_mode = mode_was, pick_L = L_was;
    rep += "ysolve:";
    std::vector<std::pair<int, int>> cands;
    for (int L : {16, 32, 64})
      if (c->ny % L == 0 && c->ny / L >= 2 && c->ny / L <= 315) cands.push_back({1, L});
    if (flags & 1) cands.push_back({0, 0});
    for (auto &cd : cands) {
      c->cfg.poisson_mode = cd.first;
      if (cd.first == 1 && (rc = partition_setup(c, cd.second))) break;
      double ms = 0.;
      if ((rc = time_launches(c, [&] { return launch_poisson(c, 1); }, 10, &ms))) break;
      snprintf(buf, sizeof(buf), " %s=%.4fms", cd.first ? ("partition L=" + std::to_string(cd.second)).c_str() : "thomas", ms);
      rep += buf;
      if (ms < best) {
        best = ms;
        pick_mode = cd.first;
        pick_L = cd.second;
      }
    }
    if (!rc) {
      c->cfg.poisson_mode = pick_mode;
      if (pick_mode == 1 && pick_L != c->part_L) rc = partition_setup(c, pick_L);
      snprintf(buf, sizeof(buf), " -> %s %d", pick_mode ? "partition" : "thomas", pick_L);
      rep += buf;
    } else {
      c->cfg.poisson_mode = mode_was;
    }
  }
  // values per thread of the register FFT (same butterflies, same bits)
  if (!rc && fft_tunable(c)) {
    double best = 1e30;
    int pick = 0;
    rep += "; fft_values_per_thread:";
    for (int lr : {3, 2, 4}) {
      if (!fft_logr_valid(c, lr)) continue;
      c->fft_logr = lr == 3 ? 0 : lr;
      double ms = 0.;
      if ((rc = time_launches(c, [&] { return launch_poisson(c, 1); }, 10, &ms))) break;
      snprintf(buf, sizeof(buf), " %d=%.4fms", 1 << lr, ms);
      rep += buf;
      if (ms < best) {
        best = ms;
        pick = c->fft_logr;
      }
    }
    c->fft_logr = pick;
    snprintf(buf, sizeof(buf), " -> %d", 1 << (pick ? pick : 3));
    rep += buf;
  }
  // strip height of the fused DCT-III + projection, for the chosen y-solve
  if (!rc) {
    double *pt = nullptr;
    if (hipMalloc(&pt, c->L.p_len() * sizeof(double)) != hipSuccess) {
      rc = hip_fail(hipErrorOutOfMemory, "hipMalloc", __FILE__, __LINE__);
    } else {
      rep += "; ";
      rc = project_tune(c, c->ux_spare, c->uy_spare, pt, &rep);
      (void)hipFree(pt);
    }
  }
  c->dt_adv_parts = parts_was;
  c->dt_adv_valid = dtv_was;
  c->prof->on = prof_was;
  if (report && report_len > 0) {
    strncpy(report, rep.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = 0;
  }
  return rc;
}

int eif_gpu_tune_copy(eif_ctx *dst, const eif_ctx *src) {
  CHECK_CTX(dst);
  CHECK_CTX(const_cast<eif_ctx *>(src));
  if (dst->nx != src->nx || dst->ny != src->ny) {
    set_error("tuned shapes apply to a context of the same grid");
    return EIF_ERR_ARG;
  }
  EIF_HIP(hipStreamSynchronize(dst->stream));
  dst->pred_rows = src->pred_rows;
  dst->cp_rows = src->cp_rows;
  dst->fft_logr = src->fft_logr;
  dst->proj_strip = src->proj_strip;
  if (src->cfg.poisson_mode != dst->cfg.poisson_mode || src->part_L != dst->part_L) {
    dst->cfg.poisson_mode = src->cfg.poisson_mode;
    if (dst->cfg.poisson_mode == EIF_POISSON_PARTITION && src->part_L)
      TRY(partition_setup(dst, src->part_L));
  }
  // graphs captured with the old shapes are dropped
  return graphs_recapture(dst);
}

int eif_gpu_set_projection(eif_ctx *c, int strip) {
  CHECK_CTX(c);
  TRY(set_projection(c, strip));
  return graphs_recapture(c);
}

int eif_gpu_set_collision_search(eif_ctx *c, int mode) {
  CHECK_CTX(c);
  TRY(set_collision_search(c, mode));
  return graphs_recapture(c);
}

int eif_gpu_graph_enable(eif_ctx *c, int enable) {
  CHECK_CTX(c);
  EIF_HIP(hipStreamSynchronize(c->stream));
  graphs_free(c);
  c->graph_on = enable != 0;
  if (c->graph_on) {
    c->graphs = new GraphCache();
    EIF_HIP(hipMemcpy(&c->graphs->seq, c->g_seq, sizeof(unsigned), hipMemcpyDeviceToHost));
  }
  return EIF_OK;
}

// graphs captured with the kernels or arguments a setting changes are
// dropped; the sub-iterations per stage are kept
static int graphs_recapture(eif_ctx *c) {
  if (!c->graph_on || !c->graphs) return EIF_OK;
  int K[3];
  for (int r = 0; r < 3; r++) K[r] = c->graphs->K[r];
  TRY(eif_gpu_graph_enable(c, 1));
  for (int r = 0; r < 3; r++) c->graphs->K[r] = K[r];
  return EIF_OK;
}

int eif_gpu_graph_set_subiterations(eif_ctx *c, const int K[3]) {
  CHECK_CTX(c);
  if (!c->graphs) {
    set_error("graph mode is off");
    return EIF_ERR_STATE;
  }
  for (int r = 0; r < 3; r++)
    if (K[r] < 1 || K[r] > 102) {
      set_error("sub-iterations per stage must be within 1..102");
      return EIF_ERR_ARG;
    }
  for (int r = 0; r < 3; r++) {
    c->graphs->K[r] = K[r];
    c->graphs->low_run[r] = 0;
  }
  return EIF_OK;
}

int eif_gpu_graph_stats(eif_ctx *c, long out[4]) {
  CHECK_CTX(c);
  GraphCache *g = c->graphs;
  out[0] = g ? g->launches : 0;
  out[1] = g ? g->resumes : 0;
  out[2] = g ? g->captures : 0;
  out[3] = g ? (long)g->K[0] * 10000 + g->K[1] * 100 + g->K[2] : 0;
  return EIF_OK;
}

int eif_gpu_run_steps(eif_ctx *c, int nsteps, double *dts, int *substeps) {
  CHECK_CTX(c);
  if (nsteps < 0) {
    set_error("nsteps must be non-negative");
    return EIF_ERR_ARG;
  }
  int n = 0;
  auto record = [&](int k, double dt, const int sub[3]) {
    if (dts) dts[k] = dt;
    if (substeps)
      for (int r = 0; r < 3; r++) substeps[3 * k + r] = sub[r];
  };
  if (!c->graph_on) {
    for (; n < nsteps; n++) {
      int sub[3];
      TRY(eager_step(c, sub));
      record(n, c->hscal[S_DT], sub);
    }
    return EIF_OK;
  }
  if (n < nsteps) {
    bool ran = false;
    int sub[3];
    TRY(graph_prepare(c, &ran, sub));
    if (ran) record(n++, c->hscal[S_DT], sub);
  }
  // graph launches run one step ahead of the host's reading of their
  // records; a stopped step stops the one behind it too (the flag stays
  // set), which is then launched again
  GraphEntry *prev = nullptr;
  unsigned prev_seq = 0;
  int k_prev = -1;
  while (n < nsteps || prev) {
    GraphEntry *cur = nullptr;
    unsigned cur_seq = 0;
    if (n < nsteps) TRY(graph_submit(c, &cur, &cur_seq));
    if (prev) {
      GraphStatus st;
      TRY(graph_wait(c, prev_seq, &st));
      if (st.halt && cur) {
        // the step behind it did nothing: its record (a stop) is consumed,
        // and it is launched again after the resumption
        EIF_HIP(hipStreamSynchronize(c->stream));
        GraphStatus skip;
        TRY(graph_wait(c, cur_seq, &skip));
        c->graphs->launches--;
        cur = nullptr;
      }
      double dt;
      int sub[3];
      TRY(graph_finish(c, prev, st, &dt, sub));
      record(k_prev, dt, sub);
      if (st.halt) {
        prev = nullptr;
        continue;
      }
    }
    prev = cur;
    prev_seq = cur_seq;
    if (cur) k_prev = n++;
  }
  c->dt_adv_valid = c->dt_adv_parts > 0;
  return EIF_OK;
}

// Every case's step graph nested as a child of one graph: one launch per
// ensemble step.  The children are independent; a stop in one case is
// resumed on that case's own stream, and the merged graph is rebuilt only
// when some case's captured graph changes.
static int ensemble_merged(eif_ctx **ctxs, int ncases, int nsteps) {
  std::vector<int> done(ncases, 0);
  for (int k = 0; k < ncases; k++) {
    bool ran = false;
    int sub[3];
    if (nsteps > 0) TRY(graph_prepare(ctxs[k], &ran, sub));
    done[k] = ran ? 1 : 0;
  }
  // the merged launch runs on one stream: every case's own queued work
  // (the eager steps above) is finished first
  for (int k = 0; k < ncases; k++) EIF_HIP(hipStreamSynchronize(ctxs[k]->stream));
  hipGraph_t parent = nullptr;
  hipGraphExec_t pexec = nullptr;
  std::vector<GraphEntry *> ent(ncases, nullptr), built(ncases, nullptr);
  auto release = [&]() {
    if (pexec) (void)hipGraphExecDestroy(pexec);
    if (parent) (void)hipGraphDestroy(parent);
    pexec = nullptr;
    parent = nullptr;
  };
  int rc = EIF_OK;
  for (;;) {
    // cases still running; all launch together, so they share the count
    bool any = false, changed = false;
    for (int k = 0; k < ncases && !rc; k++) {
      ent[k] = nullptr;
      if (done[k] >= nsteps) {
        changed = changed || built[k] != nullptr;
        continue;
      }
      rc = graph_for(ctxs[k], &ent[k]);
      any = true;
      changed = changed || ent[k] != built[k];
    }
    if (rc || !any) break;
    if (changed || !pexec) {
      release();
      if (hipGraphCreate(&parent, 0) != hipSuccess) {
        rc = hip_fail(hipErrorUnknown, "hipGraphCreate", __FILE__, __LINE__);
        break;
      }
      for (int k = 0; k < ncases && !rc; k++) {
        built[k] = ent[k];
        if (!ent[k]) continue;
        hipGraphNode_t node;
        const hipError_t e = hipGraphAddChildGraphNode(&node, parent, nullptr, 0, ent[k]->graph);
        if (e != hipSuccess) rc = hip_fail(e, "hipGraphAddChildGraphNode", __FILE__, __LINE__);
      }
      if (!rc) {
        const hipError_t e = hipGraphInstantiate(&pexec, parent, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
      }
      if (rc) break;
    }
    // one launch for the whole ensemble step, on the first running case's stream
    int first = 0;
    while (!ent[first]) first++;
    if (hipGraphLaunch(pexec, ctxs[first]->stream) != hipSuccess) {
      rc = hip_fail(hipErrorUnknown, "hipGraphLaunch", __FILE__, __LINE__);
      break;
    }
    // every case's record is written by the end of the merged launch
    if (hipStreamSynchronize(ctxs[first]->stream) != hipSuccess) {
      rc = hip_fail(hipErrorUnknown, "hipStreamSynchronize", __FILE__, __LINE__);
      break;
    }
    for (int k = 0; k < ncases && !rc; k++) {
      if (!ent[k]) continue;
      Ctx *c = ctxs[k];
      c->graphs->launches++;
      const unsigned seq = ++c->graphs->seq;
      GraphStatus st;
      if ((rc = graph_wait(c, seq, &st))) break;
      double dt;
      int sub[3];
      if ((rc = graph_finish(c, ent[k], st, &dt, sub))) break;
      // a resumption ran on this case's stream, ahead of the next merged launch
      if (st.halt && hipStreamSynchronize(c->stream) != hipSuccess)
        rc = hip_fail(hipErrorUnknown, "hipStreamSynchronize", __FILE__, __LINE__);
      done[k]++;
    }
    if (rc) break;
  }
  for (int k = 0; k < ncases; k++) (void)hipStreamSynchronize(ctxs[k]->stream);
  release();
  for (int k = 0; k < ncases; k++) ctxs[k]->dt_adv_valid = ctxs[k]->dt_adv_parts > 0;
  return rc;
}

int eif_gpu_ensemble_run_steps(eif_ctx **ctxs, int ncases, int nsteps, int concurrent) {
  if (!ctxs || ncases < 0 || nsteps < 0) {
    set_error("invalid ensemble arguments");
    return EIF_ERR_ARG;
  }
  for (int k = 0; k < ncases; k++) CHECK_CTX(ctxs[k]);
  for (int k = 0; k < ncases; k++)
    for (int j = 0; j < k; j++)
      if (ctxs[j] == ctxs[k]) {
        set_error("a context appears twice in the ensemble");
        return EIF_ERR_ARG;
      }
  if (concurrent < 0 || concurrent > 3) {
    set_error("concurrent must be 0 (sequential), 1 (streams), 2 (one graph) or 3 (batched)");
    return EIF_ERR_ARG;
  }
  if (!concurrent) {
    for (int k = 0; k < ncases; k++) TRY(eif_gpu_run_steps(ctxs[k], nsteps, nullptr, nullptr));
    return EIF_OK;
  }
  for (int k = 0; k < ncases; k++)
    if (!ctxs[k]->graph_on) TRY(eif_gpu_graph_enable(ctxs[k], 1));
  if (concurrent == 2) return ensemble_merged(ctxs, ncases, nsteps);
  if (concurrent == 3) return ensemble_batched(ctxs, ncases, nsteps, nullptr);
  std::vector<int> done(ncases, 0);
  for (int k = 0; k < ncases; k++) {
    bool ran = false;
    int sub[3];
    if (nsteps > 0) TRY(graph_prepare(ctxs[k], &ran, sub));
    done[k] = ran ? 1 : 0;
  }
  // step by step: every context's next graph step is launched on its own
  // stream before the host waits for any of them
  std::vector<GraphEntry *> ent(ncases, nullptr);
  std::vector<unsigned> seq(ncases, 0);
  for (;;) {
    bool any = false;
    for (int k = 0; k < ncases; k++) {
      ent[k] = nullptr;
      if (done[k] >= nsteps) continue;
      TRY(graph_submit(ctxs[k], &ent[k], &seq[k]));
      any = true;
    }
    if (!any) break;
    for (int k = 0; k < ncases; k++) {
      if (!ent[k]) continue;
      GraphStatus st;
      TRY(graph_wait(ctxs[k], seq[k], &st));
      double dt;
      int sub[3];
      TRY(graph_finish(ctxs[k], ent[k], st, &dt, sub));
      done[k]++;
    }
  }
  for (int k = 0; k < ncases; k++)
    ctxs[k]->dt_adv_valid = ctxs[k]->dt_adv_parts > 0;
  return EIF_OK;
}

int eif_gpu_synchronize(eif_ctx *c) {
  CHECK_CTX(c);
  EIF_HIP(hipStreamSynchronize(c->stream));
  return check_flags(c);
}

int eif_gpu_ibm_stats(eif_ctx *c, unsigned long long out[5], int reset) {
  CHECK_CTX(c);
  unsigned long long h[5] = {0, 0, 0, 0, 0};
  if (c->np > 0) {
    std::vector<unsigned long long> pp(4 * (size_t)c->np + 1);
    EIF_HIP(hipMemcpyAsync(pp.data(), c->ibm_pp, pp.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
    EIF_HIP(hipStreamSynchronize(c->stream));
    for (int n = 0; n < c->np; n++) {
      h[0] += pp[4 * (size_t)n];
      h[1] += pp[4 * (size_t)n + 1];
      h[2] += pp[4 * (size_t)n + 2];
      h[4] += pp[4 * (size_t)n + 3];
    }
    h[3] = pp[4 * (size_t)c->np];
  }
  if (out)
    for (int k = 0; k < 5; k++) out[k] = h[k];
  if (reset) {
    if (c->np > 0) {
      EIF_HIP(hipMemsetAsync(c->ibm_pp, 0, (4 * (size_t)c->np + 1) * sizeof(unsigned long long),
                             c->stream));
      EIF_HIP(hipStreamSynchronize(c->stream));
    }
  }
  return EIF_OK;
}

int eif_gpu_ibm_reference_iterations(eif_ctx *c, unsigned long long out[2]) {
  CHECK_CTX(c);
  if (!out) {
    set_error("null output");
    return EIF_ERR_ARG;
  }
  // two counters in the unused tail of the statistics array
  unsigned long long *d = c->ibm_stats + 6;
  EIF_HIP(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), c->stream));
  TRY(launch_ref_iterations(c, d));
  unsigned long long h[2];
  EIF_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  EIF_HIP(hipStreamSynchronize(c->stream));
  out[0] = h[0];
  out[1] = h[1];
  return check_flags(c);
}

int eif_gpu_profile_enable(eif_ctx *c, int enable) {
  CHECK_CTX(c);
  c->prof->on = enable != 0;
  return EIF_OK;
}

int eif_gpu_profile_select(eif_ctx *c, const char *name) {
  CHECK_CTX(c);
  c->prof->only = name ? name : "";
  return EIF_OK;
}

int eif_gpu_profile_reset(eif_ctx *c) {
  CHECK_CTX(c);
  TRY(prof_flush(c));
  for (auto &kv : c->prof->recs) {
    kv.second.count = 0;
    kv.second.total_ms = 0.;
    kv.second.seen = 0;
  }
  return EIF_OK;
}

int eif_gpu_profile_sample(eif_ctx *c, int every) {
  CHECK_CTX(c);
  if (every < 1) {
    set_error("every must be at least 1");
    return EIF_ERR_ARG;
  }
  c->prof->every = every;
  return EIF_OK;
}

int eif_gpu_profile_query(eif_ctx *c, const char *name, long *count, double *total_ms) {
  CHECK_CTX(c);
  TRY(prof_flush(c));
  auto it = c->prof->recs.find(name ? name : "");
  if (count) *count = it == c->prof->recs.end() ? 0 : it->second.count;
  if (total_ms) *total_ms = it == c->prof->recs.end() ? 0. : it->second.total_ms;
  return EIF_OK;
}

}  // extern "C"
