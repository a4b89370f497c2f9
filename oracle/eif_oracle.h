/*
 * eif_oracle.h -- CPU restatement of the EllipsesInFlows time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path
 * (libeif_gpu.so) never links, loads or calls it.
 *
 * Scope: one process owning the whole domain (the reference run with
 * `mpirun -n 1`), so every MPI halo exchange becomes a periodic wrap and
 * every MPI_Allreduce becomes the identity.  All arrays use the reference's
 * halo-padded layouts (include/arrays/fluid.h:5-103,
 * include/arrays/param.h:5-45, include/arrays/suspensions.h:5-23).
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * comment of eif_oracle.c.
 */
#ifndef EIF_ORACLE_H
#define EIF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* field selectors for orc_field() */
enum {
  ORC_UX = 0, ORC_UY, ORC_P, ORC_PSI,
  ORC_SRCUXA, ORC_SRCUXB, ORC_SRCUXG,
  ORC_SRCUYA, ORC_SRCUYB, ORC_SRCUYG,
  ORC_DUX, ORC_DUY,
  ORC_NFIELDS
};
/* grid selectors for orc_grid() */
enum { ORC_XF = 0, ORC_XC, ORC_DXF, ORC_DXC, ORC_YF, ORC_YC, ORC_NGRIDS };

/* number of doubles per particle record (30), member order of particle_t
 * (include/suspensions.h:8-31):
 * den a b x y az dx dy daz ux uy vz dux duy dvz fux fuy tvz
 * iux[2] iuy[2] ivz[2] cfx[2] cfy[2] ctz[2] */
#define ORC_PARTICLE_NDOUBLES 30

orc_t *orc_create(int itot, int jtot, double ly, double Re, double Fr,
                  double extfrcy, double uy_wall_xm, double uy_wall_xp,
                  const double safefactors[3]);
void orc_destroy(orc_t *s);
int orc_set_threads(int nthreads);

double *orc_field(orc_t *s, int which);
long orc_field_len(orc_t *s, int which);
double *orc_grid(orc_t *s, int which);
double orc_get_dt(const orc_t *s);
void orc_set_dt(orc_t *s, double dt);
double orc_get_time(const orc_t *s);
int orc_get_step(const orc_t *s);

int orc_set_particles(orc_t *s, int n, const double *den, const double *a,
                      const double *b, const double *x, const double *y,
                      const double *az, const double *ux, const double *uy,
                      const double *vz);
int orc_n_particles(const orc_t *s);
/* full records, n x ORC_PARTICLE_NDOUBLES, row-major */
int orc_get_particle_records(const orc_t *s, double *out);
int orc_set_particle_records(orc_t *s, const double *in);

/* the reference's public functions, one process */
int orc_update_boundaries_ux(orc_t *s, double *ux);
int orc_update_boundaries_uy(orc_t *s, double *uy);
int orc_update_boundaries_p(orc_t *s, double *p);
int orc_decide_dt(orc_t *s);
int orc_reset_particle_increments(orc_t *s);
int orc_compute_inertia(orc_t *s, int cnstep);
int orc_update_velocity(orc_t *s, int rkstep);
int orc_exchange_momentum(orc_t *s);
int orc_update_momentum_field(orc_t *s);
int orc_compute_potential(orc_t *s, int rkstep);
int orc_correct_velocity(orc_t *s, int rkstep);
int orc_update_pressure(orc_t *s);
int orc_compute_collision_force(orc_t *s, int cnstep);
int orc_increment_particles(orc_t *s, int rkstep, double *residual);
int orc_update_particles(orc_t *s);
int orc_integrate(orc_t *s);
/* decide_dt + integrate + step/time advance (src/main.c:71-76) */
int orc_step(orc_t *s);
/* sub-iteration counts of the last integrate(), one per RK stage */
int orc_last_substeps(const orc_t *s, int out[3]);

/* colliding pairs recorded by the last orc_compute_collision_force():
 * (n0, n1) with overlap > 0, n1 == -1 / -2 for the walls x=0 / x=lx */
int orc_n_colliding(const orc_t *s);
int orc_get_colliding(const orc_t *s, int *pairs);

/* logging diagnostics (src/logging.c:40-172) */
int orc_divergence(const orc_t *s, double *divmax, double *divsum);
int orc_momentum(const orc_t *s, double moms[2]);
int orc_energy(const orc_t *s, double quantities[2]);

/* statistics_collect (src/statistics/collect.c:10-95): adds the current
 * sample to caller-owned accumulators in the reference layouts
 * (include/arrays/statistics.h): ux1, ux2 like ux; uy1, uy2, phi like uy */
int orc_statistics_collect(const orc_t *s, double *ux1, double *ux2, double *uy1,
                           double *uy2, double *phi);

/* pure kernels (no state) */
double orc_find_normal_t(double a, double b, double x, double y);
double orc_s_weight(double grid_size, double pa, double pb, double px,
                    double py, double paz, double x, double y);
double orc_v_weight(double grid_size, double pa, double pb, double px,
                    double py, double paz, double x, double y);
int orc_decide_loop_size(int lbound, int ubound, double grid_size,
                         double radius, double grav_center, int *min, int *max);
/* FFTW_REDFT10 / FFTW_REDFT01 (unnormalised DCT-II / DCT-III), in place */
int orc_dct2(int n, double *x);
int orc_dct3(int n, double *x);
/* the same for two rows at once through one complex FFT (n a power of two;
 * the Poisson solve's form for 64 <= n <= 8192) */
int orc_dct2_pair(int n, double *xa, double *xb);
int orc_dct3_pair(int n, double *xa, double *xb);
/* periodic tridiagonal solve, bordering method of src/tdm.c:64-108 */
int orc_tdm_periodic(int n, const double *l, const double *c, const double *u,
                     double *q);
/* collision between two particles given as 30-double records; updates the
 * records' cf*[cnstep] exactly as src/suspensions/collision.c:156-247 */
int orc_collide_pair(double ly, int cnstep, double *p0, double *p1);
int orc_collide_wall(double wallx, int cnstep, double *p);

#ifdef __cplusplus
}
#endif
#endif
