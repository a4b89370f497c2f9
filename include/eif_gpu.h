/*
 * eif_gpu.h -- C-ABI of the MI355X (gfx950) EllipsesInFlows time-step path.
 *
 * Drop-in boundary for the reference's internal C API
 * (NaokiHori/EllipsesInFlows include/fluid.h:27-38,
 * include/suspensions.h:44-63, include/param.h:55-58), called the way
 * src/main.c:13-118 calls those functions.  One context holds one whole
 * domain on one GPU; the reference's MPI halo exchanges, transposes and
 * Allreduces have no counterpart here.
 *
 * Conventions
 *   - every function returns EIF_OK (0) or a negative EIF_ERR_* code and
 *     never aborts; eif_gpu_last_error() describes the last failure
 *     (thread-local);
 *   - host arrays use the reference's halo-padded layouts
 *     (include/arrays/fluid.h:5-43, include/arrays/param.h:5-45):
 *       ux          [(jtot+2) x (itot+1)], UX(i,j) = ux[j*(itot+1)+i-1]
 *       uy, p, psi  [(jtot+2) x (itot+2)], P(i,j)  = p[j*(itot+2)+i]
 *       srcuxa      [jtot x (itot-1)],     S(i,j)  = s[(j-1)*(itot-1)+i-2]
 *       srcuya      [jtot x itot],         S(i,j)  = s[(j-1)*itot+i-1]
 *   - particle records are 30 doubles each, in the member order of
 *     particle_t (include/suspensions.h:8-31);
 *   - calls are asynchronous on the context's stream; only calls that
 *     return host values synchronise;
 *   - dt lives in device memory: fluid_gpu_decide_dt() writes it there and
 *     every kernel reads it there.
 */
#ifndef EIF_GPU_H
#define EIF_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define EIF_OK 0
#define EIF_ERR_ARG (-1)      /* invalid argument or size */
#define EIF_ERR_HIP (-2)      /* HIP runtime failure (message has detail) */
#define EIF_ERR_CAPACITY (-3) /* a fixed device capacity was exceeded */
#define EIF_ERR_STATE (-4)    /* call out of order (e.g. no particles) */
#define EIF_ERR_NODEVICE (-5) /* no HIP device available */

#define EIF_PARTICLE_NDOUBLES 30

/* y-direction solve of the pressure Poisson equation */
#define EIF_POISSON_THOMAS 0    /* src/tdm.c bordering Thomas, same arithmetic */
#define EIF_POISSON_PARTITION 1 /* partitioned Thomas (spike), parallel in y */

typedef struct eif_ctx eif_ctx;

/* Physical and numerical configuration; field names and defaults follow
 * the reference's environment variables (src/param/init.c:67-103). */
typedef struct {
  int itot, jtot;          /* grid points in x (wall-normal) and y          */
  double ly;               /* domain length in y (lx is 1, param.h:14)      */
  double Re, Fr, extfrcy;  /* Reynolds, Froude, body force in y             */
  double uy_wall_xm;       /* uy on the wall x=0  (reference default 0)     */
  double uy_wall_xp;       /* uy on the wall x=lx (reference default 0)     */
  double safefactors[3];   /* dt safety factors: advective, diffusive, par. */
  int poisson_mode;        /* EIF_POISSON_*                                 */
} eif_config;

/* Replaces param_init's grid/RK set-up (src/param/init.c:105-191) and
 * fluid_init's allocation (src/fluid/init.c:87-93).  Fields start at zero
 * with boundary values applied, as src/fluid/init.c:33-85 without restart. */
int fluid_gpu_create(const eif_config *cfg, eif_ctx **out);
/* Replaces fluid_finalise + suspensions_finalise
 * (src/fluid/finalise.c:19-33, src/suspensions/finalise.c:5-20). */
int fluid_gpu_destroy(eif_ctx *ctx);

/* Grid vectors in the reference layouts (include/arrays/param.h:5-45):
 * xf[itot+1], xc[itot+2], dxf[itot], dxc[itot+1], yf[jtot+1], yc[jtot]. */
int fluid_gpu_get_grid(const eif_ctx *ctx, double *xf, double *xc, double *dxf,
                       double *dxc, double *yf, double *yc);

/* Host <-> device copies of ux, uy, p in the reference layouts (any pointer
 * may be NULL to skip that field).  Used where the reference reads/writes
 * restart files (src/fluid/init.c:33-85, src/save.c:41-46). */
int fluid_gpu_upload(eif_ctx *ctx, const double *ux, const double *uy,
                     const double *p);
int fluid_gpu_download(eif_ctx *ctx, double *ux, double *uy, double *p);

/* Auxiliary fields for inspection, reference layouts.  After eif_gpu_step
 * the source arrays hold the reference's values (A of the last stage's start
 * velocity); the step itself does not store them, so the first read after a
 * step evaluates them (one kernel). */
#define EIF_FIELD_PSI 0    /* fluid_t.psi   (include/fluid.h:22)        */
#define EIF_FIELD_SRCUXA 1 /* fluid_t.srcuxa (include/fluid.h:23)       */
#define EIF_FIELD_SRCUYA 2 /* fluid_t.srcuya (include/fluid.h:24)       */
#define EIF_FIELD_DUX 3    /* suspensions_t.dux (include/suspensions.h:38) */
#define EIF_FIELD_DUY 4    /* suspensions_t.duy                          */
int fluid_gpu_download_field(eif_ctx *ctx, int which, double *host);
int fluid_gpu_upload_field(eif_ctx *ctx, int which, const double *host);

/* src/fluid/boundary_conditions.c:7-24, :26-43, :45-62 */
int fluid_gpu_update_boundaries_ux(eif_ctx *ctx);
int fluid_gpu_update_boundaries_uy(eif_ctx *ctx);
int fluid_gpu_update_boundaries_p(eif_ctx *ctx);

/* param_decide_dt, src/param/decide_dt.c:11-81.  dt stays on the device;
 * dt_out (may be NULL) receives it after a synchronisation. */
int fluid_gpu_decide_dt(eif_ctx *ctx, double *dt_out);
int fluid_gpu_set_dt(eif_ctx *ctx, double dt);
int fluid_gpu_get_dt(eif_ctx *ctx, double *dt_out);

/* fluid_update_velocity, src/fluid/update_velocity.c:229-237 */
int fluid_gpu_update_velocity(eif_ctx *ctx, int rkstep);
/* fluid_compute_potential, src/fluid/compute_potential.c:48-138 */
int poisson_gpu_compute_potential(eif_ctx *ctx, int rkstep);
/* Layout of the partitioned y-solve (EIF_POISSON_PARTITION) for jtot
 * (no counterpart in the reference, whose y-solve is the sequential
 * src/tdm.c:64-111 after a transpose): chunk rows per chunk and nchunks
 * chunks; returns 1, or 0 (both set to 0) when the partitioned solve does
 * not apply and the context falls back to the Thomas sweep.  Host only:
 * needs no device. */
int poisson_gpu_partition_layout(int jtot, int *chunk, int *nchunks);
/* fluid_correct_velocity, src/fluid/correct_velocity.c:6-42 */
int fluid_gpu_correct_velocity(eif_ctx *ctx, int rkstep);
/* fluid_update_pressure, src/fluid/update_pressure.c:7-24 */
int fluid_gpu_update_pressure(eif_ctx *ctx);

/* Diagnostics of src/logging.c:40-172 (divergence, momentum, energy),
 * computed on the device; synchronising. */
int fluid_gpu_divergence(eif_ctx *ctx, double *divmax, double *divsum);
int fluid_gpu_momentum(eif_ctx *ctx, double moms[2]);
int fluid_gpu_energy(eif_ctx *ctx, double quantities[2]);
/* No counterpart in the reference: the three diagnostics above for callers
 * that cannot allocate or synchronise (code under HIP graph capture, tight
 * serving loops).  which = 0 divergence (max, sum), 1 momentum, 2 energy;
 * the two results go to out_dev (device memory, 2 doubles), asynchronously
 * on the context's stream, equal bit for bit to the synchronising calls.
 * The partial sums live in one static device array per GPU (4096 pairs,
 * 64 KiB), not in memory the call allocates: the call allocates nothing,
 * and calls that overlap on one GPU (other contexts, other threads) race
 * on that array. */
int fluid_gpu_diagnostics_noalloc(eif_ctx *ctx, int which, double *out_dev);

/* statistics_collect, src/statistics/collect.c:85-95: adds ux, ux^2, uy,
 * uy^2 and the particle indicator phi (collect.c:48-83) to accumulators
 * kept on the device (allocated and zeroed on first use, as
 * src/statistics/init.c:8-44). */
int statistics_gpu_collect(eif_ctx *ctx);
/* Zero the accumulators and the sample count. */
int statistics_gpu_reset(eif_ctx *ctx);
/* The accumulators written by statistics_output
 * (src/statistics/output.c:45-57) in the reference layouts
 * (include/arrays/statistics.h): ux1, ux2 like ux; uy1, uy2, phi like uy.
 * num receives the number of samples.  Any pointer may be NULL. */
int statistics_gpu_download(eif_ctx *ctx, double *ux1, double *ux2,
                            double *uy1, double *uy2, double *phi, int *num);

/* suspensions_init's particle load (src/suspensions/init.c:24-86): the nine
 * SoA arrays of the initparticles .npy files. */
int suspensions_gpu_create(eif_ctx *ctx, int n_particles, const double *den,
                           const double *a, const double *b, const double *x,
                           const double *y, const double *az, const double *ux,
                           const double *uy, const double *vz);
int suspensions_gpu_n_particles(const eif_ctx *ctx);
/* whole particle_t records, n x 30 doubles */
int suspensions_gpu_download(eif_ctx *ctx, double *records);
int suspensions_gpu_upload(eif_ctx *ctx, const double *records);

/* src/suspensions/reset.c:7-18 */
int suspensions_gpu_reset_particle_increments(eif_ctx *ctx);
/* src/suspensions/inertia.c:96-102 */
int suspensions_gpu_compute_inertia(eif_ctx *ctx, int cnstep);
/* src/suspensions/exchange.c:114-120 */
int suspensions_gpu_exchange_momentum(eif_ctx *ctx);
/* src/suspensions/update.c:36-59 (name keeps the reference's spelling) */
int suspensions_gpu_update_momentum_fleid(eif_ctx *ctx);
/* src/suspensions/collision.c:306-367 */
int suspensions_gpu_compute_collision_force(eif_ctx *ctx, int cnstep);
/* src/suspensions/increment.c:66-72; residual (may be NULL) is the
 * max-norm change of the velocity increments, after a synchronisation */
int suspensions_gpu_increment_particles(eif_ctx *ctx, int rkstep,
                                        double *residual);
/* src/suspensions/update.c:61-65 */
int suspensions_gpu_update_particles(eif_ctx *ctx);
/* Colliding pairs of the last collision call (overlap > 0 at
 * src/suspensions/collision.c:232 and :292), sorted: (n0, n1) with n0<n1,
 * then (n, -1) for the wall x=0 and (n, -2) for x=lx. */
int suspensions_gpu_colliding_pairs(eif_ctx *ctx, int capacity, int *pairs,
                                    int *n_pairs);

/* integrate(), src/main.c:13-55: three RK stages with the CN
 * sub-iteration.  substeps (may be NULL) receives the sub-iteration count
 * of each stage. */
int eif_gpu_integrate(eif_ctx *ctx, int substeps[3]);
/* One time step of src/main.c:71-76: decide_dt + integrate.  dt_out (may
 * be NULL) receives the dt used. */
int eif_gpu_step(eif_ctx *ctx, double *dt_out);

/* nsteps time steps of src/main.c:71-76 (decide_dt + integrate each);
 * dts (may be NULL) receives each step's dt, substeps (may be NULL) the
 * sub-iteration counts, 3 per step.  Synchronising at the end.  With the
 * graph mode on, each step is one launch of a captured hipGraph: the host
 * reads one record per step, one step behind the device, instead of one
 * per batch of sub-iterations. */
int eif_gpu_run_steps(eif_ctx *ctx, int nsteps, double *dts, int *substeps);
/* Whole-step hipGraph mode for eif_gpu_run_steps (no counterpart in the
 * reference).  The graph of a step holds K sub-iterations per stage, K
 * following the counts the loops need; a stage whose loop needs more stops
 * the rest of the captured step on the device, and the host resumes it
 * where it stopped, so results equal the eager path's bit for bit.  A step
 * is captured again only when K or the buffer assignment changes. */
int eif_gpu_graph_enable(eif_ctx *ctx, int enable);
/* Sets K of the three stages (1..102) for the next graph steps; the
 * counts the loops need adjust it again afterwards. */
int eif_gpu_graph_set_subiterations(eif_ctx *ctx, const int K[3]);
/* out: graph launches, resumed stops, captures, and the current K of the
 * three stages as K0*10000 + K1*100 + K2 */
int eif_gpu_graph_stats(eif_ctx *ctx, long out[4]);

/* Autotuner (no counterpart in the reference): times the candidate launch
 * shapes on this context's grid and keeps the fastest.  Rows per workgroup
 * of the predictor (8, 16, 32) and of the projection (4, 8, 16) and the
 * register FFT's values per thread (4, 8, 16: the same butterflies) change
 * no result.  The chunk length of the partitioned y-solve (16, 32, 64 rows
 * where they divide jtot), and in partition mode the fused projection (see
 * eif_gpu_set_projection), change rounding only; flags & 1 also lets it
 * switch between the Thomas sweep and the partitioned solve.  report (may
 * be NULL) receives the timings and the choices.  Synchronising; leaves psi
 * holding its last timing solve. */
int eif_gpu_autotune(eif_ctx *ctx, int flags, char *report, int report_len);

/* The launch shapes eif_gpu_autotune chose for src (rows per workgroup,
 * y-solve and its chunk length, FFT values per thread, projection form)
 * applied to dst, a context of the same grid: an ensemble tunes one case
 * per grid shape.  No counterpart in the reference.  Drops dst's captured
 * graphs. */
int eif_gpu_tune_copy(eif_ctx *dst, const eif_ctx *src);

/* How eif_gpu_integrate (and graph steps) run each stage's potential,
 * projection and pressure update (compute_potential.c:48-138,
 * correct_velocity.c:6-42, update_pressure.c:7-24; no counterpart in the
 * reference's interface): strip = 0, the DCT-III kernel storing psi and then
 * the projection kernel (the default); 2, 4 or 8, one kernel that fuses the
 * DCT-III with the projection and the pressure update over strips of that
 * many rows and never stores psi (psi is then undefined after a step).
 * Needs nx a power of two in 64..8192.  In Thomas mode both give the same
 * bits; eif_gpu_autotune times both. */
int eif_gpu_set_projection(eif_ctx *ctx, int strip);

/* The broad phase of the collisions, which finds the pairs whose bounding
 * circles overlap before the contacts are solved (it replaces the pair-matrix
 * loop of collision.c:320-332 with its early exit, collision.c:172-189; no
 * counterpart in the reference's interface): mode 1 tests all pairs; mode 2
 * a cell list rebuilt before each collision launch, cells at least the
 * widest pair reach (2 max(a, b) over the particles) and at most 14336 of
 * them, each particle tested against the particles of its cell and the eight
 * neighbours (periodic in y); mode 0 (the default) takes the cell list from
 * 4096 particles on.  Both find the same partners in the same order: results
 * are bit-identical.  The cell list needs at least three cells across y
 * (ly >= 6 max(a, b)); otherwise all pairs are tested.  Drops captured
 * graphs. */
int eif_gpu_set_collision_search(eif_ctx *ctx, int mode);

/* Ensembles of independent cases (no counterpart in the reference, which
 * runs one case per process): advances ncases contexts, of any grid sizes
 * and particle counts, by nsteps steps each.  concurrent = 0: one context
 * after the other through eif_gpu_run_steps (the plain path run
 * sequentially); 1: step by step across the ensemble, each context's step
 * one captured graph on its own stream, all of a step's graphs launched
 * before the host waits for any (graph mode is switched on for them); 2:
 * as 1, with every context's step graph nested in one graph, so that an
 * ensemble step is a single launch; 3: batched launches, each kernel of the
 * step launched once for up to 64 contexts (per-context argument records in
 * device memory), the whole ensemble step one captured graph.  A context
 * appears at most once.  Results equal each context's plain run bit for
 * bit.  Synchronising. */
int eif_gpu_ensemble_run_steps(eif_ctx **ctxs, int ncases, int nsteps, int concurrent);

/* Adjoint of one RK stage (no counterpart in the reference).  The stage
 * is integrate()'s body (src/main.c:15-53) with dt fixed; its adjoint
 * covers the transposed advection/diffusion stencils
 * (src/fluid/update_velocity.c:9-237), the projection and pressure update
 * (compute_potential.c:48-138, correct_velocity.c:6-42,
 * update_pressure.c:7-24), the IBM weights, forcing and inertia integrals
 * (src/suspensions/auxiliary.c:19-59, exchange.c:13-88, inertia.c:9-70,
 * update.c:36-59) and the rigid-body update with its Crank-Nicolson
 * sub-iterations as run (increment.c:10-72, update.c:8-34).  Collisions
 * are forward-only: a recorded step that meets a contact returns
 * EIF_ERR_STATE.
 *
 * eif_gpu_tape_begin clears the records and zeroes every cotangent.
 * eif_gpu_step_fixed_dt runs one time step with the given dt (no
 * decide_dt): record == 0 through eif_gpu_integrate; record != 0 through
 * the per-function calls in the reference's order, pushing its three
 * stages.  eif_gpu_stage_adjoint replaces the cotangents of the last
 * recorded stage's outputs by those of its inputs and pops it.  Cotangent
 * access (host arrays; get copies, add accumulates):
 *   EIF_ADJ_UX, EIF_ADJ_UY, EIF_ADJ_P  fields in the reference layouts
 *                                      (meaningful on the unknowns; halo
 *                                      rows and wall values are zero)
 *   EIF_ADJ_PARTICLES  n x 6: x, y, az, ux, uy, vz of each particle
 *   EIF_ADJ_FORCE      n x 3: cotangents of fux, fuy, tvz of the exchange
 *                      of the next stage reversed (consumed by it)
 *   EIF_ADJ_RE         1: cotangent of Re (through the diffusion terms) */
#define EIF_ADJ_UX 0
#define EIF_ADJ_UY 1
#define EIF_ADJ_P 2
#define EIF_ADJ_PARTICLES 3
#define EIF_ADJ_FORCE 4
#define EIF_ADJ_RE 5
/* The wall shear stress of the adjoint runs' objective (no counterpart in
 * the reference), summed over both walls:
 * (1/Re) sum_j dy [(UY(1,j) - uy_wall_xm)/DXC(1) + (UY(itot,j) - uy_wall_xp)/DXC(itot+1)].
 * Synchronising. */
int fluid_gpu_wall_shear(eif_ctx *ctx, double *out);
int eif_gpu_tape_begin(eif_ctx *ctx);
int eif_gpu_step_fixed_dt(eif_ctx *ctx, double dt, int record);
int eif_gpu_stage_adjoint(eif_ctx *ctx);
int eif_gpu_adjoint_get(eif_ctx *ctx, int which, double *host);
int eif_gpu_adjoint_add(eif_ctx *ctx, int which, const double *host);
/* adds weight * d(fluid_gpu_wall_shear)/d(uy) to the uy cotangent on the
 * device */
int eif_gpu_adjoint_add_wall_shear(eif_ctx *ctx, double weight);

int eif_gpu_synchronize(eif_ctx *ctx);

/* IBM work counters since the last reset: out[0] box cells visited
 * (the reference evaluates every one, inertia.c:45-63 / exchange.c:55-79),
 * out[1] cells whose signed distance was computed (the rest were served by
 * the per-particle distance cache), out[2] FP64 iterations of the foot
 * point search, out[3] IBM kernel calls, out[4] FP32 iterations of the
 * foot point search.  reset != 0 clears them. */
int eif_gpu_ibm_stats(eif_ctx *ctx, unsigned long long out[5], int reset);
/* Measurement of the reference's work (no counterpart in the reference):
 * runs the reference's own FP64 iteration of find_normal_t
 * (src/ellipse.c:30-65: start pi/4, stop at |dt| < 1e-8) at every cell of
 * the particles' current IBM boxes and returns out[0] cells and out[1]
 * iterations.  Synchronising; does not change the simulation state. */
int eif_gpu_ibm_reference_iterations(eif_ctx *ctx, unsigned long long out[2]);
const char *eif_gpu_last_error(void);

/* Per-kernel timing with HIP events on the stream each kernel runs on.
 * Names are those listed by eif_gpu_profile_names() ("ibm_substep": the
 * launch of a sub-iteration's inertia(1) and collision(1)). */
int eif_gpu_profile_enable(eif_ctx *ctx, int enable);
/* restrict timing to one kernel name (NULL or "" times all of them) */
int eif_gpu_profile_select(eif_ctx *ctx, const char *name);
int eif_gpu_profile_reset(eif_ctx *ctx);
/* time one launch in `every` of each kernel name (default 1: all); an
 * event pair briefly idles the stream it is recorded on */
int eif_gpu_profile_sample(eif_ctx *ctx, int every);
int eif_gpu_profile_query(eif_ctx *ctx, const char *name, long *count,
                          double *total_ms);
const char *eif_gpu_profile_names(void);

#ifdef __cplusplus
}
#endif
#endif
