"""Closed-form pins of the oracle's particle coupling and of the order of
integrate(), the functions whose reference sources include <fftw3.h> or
live in translation units that do not build here:

  increment   src/suspensions/increment.c:10-72
  update      src/suspensions/update.c:8-34 (particles), :36-59 (field)
  exchange    src/suspensions/exchange.c:13-88
  pressure    src/fluid/update_pressure.c:7-24
  integrate   src/main.c:13-55

Each test states a property that the reference's algorithm satisfies by
construction (a constant acceleration integrated exactly by the stage
trapezoid, the rigid-body forcing of a body moving through fluid at rest,
conservation of the interpolated momentum increment, gauge invariance of
the pressure) and is chosen so that a plausible misreading -- a wrong
factor, a dropped grid spacing, a flipped sign, a reordered call -- fails.
"""
import numpy as np
import pytest

import cases
from oracle import Oracle, oracle_lib, ref_lib

# particle record offsets (particle_t member order, include/suspensions.h:8-31)
DEN, A, B, X, Y, AZ, DX, DY, DAZ, UX, UY, VZ, DUX, DUY, DVZ, FUX, FUY, TVZ = range(18)
IUX, IUY, IVZ, CFX, CFY, CTZ = 18, 20, 22, 24, 26, 28
RK_GAMMA = [32. / 60., 25. / 60. - 17. / 60., 45. / 60. - 25. / 60.]

_REF = ref_lib()


def _weights():
    """s_weight / v_weight: the reference's own compiled auxiliary.c where it
    builds, else the oracle's (bit-exact to it, test_oracle_vs_reference)."""
    if _REF is not None:
        return _REF.ref_s_weight, _REF.ref_v_weight
    lib = oracle_lib()
    return lib.orc_s_weight, lib.orc_v_weight


def _box(lb, ub, h, r, c):
    """decide_loop_size, src/suspensions/auxiliary.c:11-17 (NEXTRA = 3)."""
    lo = max(lb, int(np.floor((c - r) / h)) - 3)
    hi = min(ub, int(np.ceil((c + r) / h)) + 3)
    return lo, hi


# --------------------------------------------------------------- increment
@pytest.mark.parametrize("den,Fr", [(1.0, np.finfo(float).max), (3.0, 0.5), (0.4, 2.0)])
def test_free_fall_with_constant_contact_force_is_integrated_exactly(den, Fr):
    """A particle under a constant acceleration (buoyancy (1-den)/Fr^2 in x,
    increment.c:27, plus a constant contact force per unit mass held in both
    collision slots) and no fluid forcing: over the three RK stages (the
    gammas sum to 1) the velocity gains a*dt and the position U0*dt +
    a*dt^2/2 -- the trapezoid of each stage (increment.c:46-63) is exact for
    a velocity linear in time.  The second sub-iteration of a stage sees the
    same forces, so its residual is zero (increment.c:40-42)."""
    o = Oracle(16, 16, ly=1.0, Fr=Fr)
    U0 = np.array([0.3, -0.7, 1.9])
    o.set_particles(den=[den], a=[0.05], b=[0.03], x=[0.5], y=[0.5], az=[0.2],
                    ux=[U0[0]], uy=[U0[1]], vz=[U0[2]])
    cf = np.array([4.0, -2.5, 7.0])  # cfx, cfy, ctz per unit mass / moment
    dt = 0.02
    o.set_dt(dt)
    rec = o.particles()
    for k, off in enumerate((CFX, CFY, CTZ)):
        rec[0, off] = rec[0, off + 1] = cf[k]
    o.set_particle_records(rec)
    x0 = rec[0, [X, Y, AZ]].copy()
    for rk in range(3):
        o.reset_particle_increments()
        o.increment_particles(rk)
        assert o.increment_particles(rk) == 0.0
        o.update_particles()
    r = o.particles()[0]
    gb = 0.0 if den == 1.0 else (1.0 - den) / Fr ** 2  # pow(DBL_MAX, 2.) is inf
    acc = cf + np.array([gb, 0.0, 0.0])
    np.testing.assert_allclose(r[[UX, UY, VZ]], U0 + acc * dt, rtol=2e-15, atol=1e-15)
    np.testing.assert_allclose(r[[X, Y]], x0[:2] + U0[:2] * dt + 0.5 * acc[:2] * dt ** 2,
                               rtol=0, atol=4e-16)
    # the angle wraps into [0, 2 pi] (update.c:29-30): still inside here
    assert r[AZ] == pytest.approx(x0[2] + U0[2] * dt + 0.5 * acc[2] * dt ** 2, abs=4e-16)


def test_spin_up_by_the_enclosed_fluid():
    """The internal-inertia terms (increment.c:23-38, iu*[1]-iu*[0]) pass the
    change of the momentum of the fluid inside the particle to the particle.
    A neutrally buoyant particle whose enclosed fluid goes from rest
    (cnstep 0) to a rigid motion U + Omega x r (cnstep 1) takes that motion,
    up to the smoothing of the indicator (inertia.c:9-70)."""
    nx = ny = 128
    o = Oracle(nx, ny, ly=1.0)
    px, py, az = 0.47, 0.52, 0.6
    o.set_particles(den=[1.0], a=[0.12], b=[0.07], x=[px], y=[py], az=[az])
    o.set_dt(1e-3)
    g = o.grid()
    o.compute_inertia(0)  # fluid at rest
    U, V, W = 0.4, -0.25, 1.5
    xf = np.broadcast_to(g["xf"][None, :], (ny + 2, nx + 1))
    xc = np.broadcast_to(g["xc"][None, :], (ny + 2, nx + 2))
    yc = np.concatenate(([g["yc"][0] - 1.0 / ny], g["yc"], [g["yc"][-1] + 1.0 / ny]))
    o.view("ux")[...] = U - W * (yc[:, None] - py) + 0 * xf
    o.view("uy")[...] = V + W * (xc - px)
    o.compute_inertia(1)
    inc = o.increment_particles(0)
    r = o.particles()[0]
    assert inc > 0.1
    assert r[DUX] == pytest.approx(U, rel=0.02)
    assert r[DUY] == pytest.approx(V, rel=0.02)
    assert r[DVZ] == pytest.approx(W, rel=0.05)


# --------------------------------------------------------------- exchange
def _exchange_sums(o, rec, nx, ny, ly):
    """Sum_cells s_w * dx * dy, and its first and second moments about the
    particle centre, over the reference's boxes and periodic images
    (exchange.c:40-79)."""
    sw, _ = _weights()
    g = o.grid()
    dx, dy = 1.0 / nx, ly / ny
    h = min(dx, dy)
    pa, pb, px, py, paz = rec[A], rec[B], rec[X], rec[Y], rec[AZ]
    S = Sx = Sy = Srr = 0.0
    cells = {}
    for per in (-1, 0, 1):
        py_ = py + ly * per
        i0, i1 = _box(1, nx, dx, max(pa, pb), px)
        j0, j1 = _box(1, ny, dy, max(pa, pb), py_ - g["yf"][0])
        for j in range(j0, j1 + 1):
            y = g["yc"][j - 1]
            for i in range(i0, i1 + 1):
                x = g["xc"][i]
                w = sw(h, pa, pb, px, py_, paz, x, y)
                S += w * dx * dy
                Sx += w * (x - px) * dx * dy
                Sy += w * (y - py_) * dx * dy
                Srr += w * ((x - px) ** 2 + (y - py_) ** 2) * dx * dy
                cells[(i, j)] = cells.get((i, j), 0.0) + w
    return S, Sx, Sy, Srr, cells


@pytest.mark.parametrize("ly,ny", [(1.0, 64), (2.0, 96)])
def test_exchange_of_a_translating_and_a_rotating_body_in_fluid_at_rest(ly, ny):
    """Fluid at rest, the particle moving rigidly with (U, V, Omega): the
    forcing density is s_w * u_p / dt (exchange.c:64-72), so with
    m = den*pi*a*b and I = m (a^2+b^2)/4 (auxiliary.c:63-75)
        fux = -(U S - Omega Sy) / (m dt),   fuy = -(V S + Omega Sx) / (m dt),
        tvz = -(-U Sy + V Sx + Omega Srr) / (I dt)
    with S, Sx, Sy, Srr the sums of s_w dx dy times 1, x-px, y-py and r^2;
    a rotating body in fluid at rest feels a torque against its spin, a
    translating one a force against its motion.  The field increment at
    each cell is s_w * u_p (dux = fx * dt)."""
    nx = 64
    o = Oracle(nx, ny, ly=ly)
    U, V, W = 0.35, -0.2, 2.5
    den = 1.3
    o.set_particles(den=[den], a=[0.1], b=[0.06], x=[0.45], y=[0.03 * ly], az=[0.7],
                    ux=[U], uy=[V], vz=[W])
    dt = 2e-3
    o.set_dt(dt)
    o.exchange_momentum()
    r = o.particles()[0]
    S, Sx, Sy, Srr, cells = _exchange_sums(o, r, nx, ny, ly)
    m = den * 3.1415926535897932 * 0.1 * 0.06
    I = 0.25 * m * (0.1 ** 2 + 0.06 ** 2)
    assert r[FUX] == pytest.approx(-(U * S - W * Sy) / (m * dt), rel=1e-12)
    assert r[FUY] == pytest.approx(-(V * S + W * Sx) / (m * dt), rel=1e-12)
    assert r[TVZ] == pytest.approx(-(-U * Sy + V * Sx + W * Srr) / (I * dt), rel=1e-12)
    # the pure parts: a force against the motion, a torque against the spin
    assert r[TVZ] * W < 0 and r[FUX] * U < 0
    # the field increment where no image overlaps another (ux_p = U - W (y - py))
    dux = o.field("dux")
    g = o.grid()
    for (i, j), w in list(cells.items())[::37]:
        y = g["yc"][j - 1]
        # the nearest image owns the cell's rigid velocity
        yy = min((y - (r[Y] + ly * k) for k in (-1, 0, 1)), key=abs)
        assert dux[j, i] == pytest.approx(w * (U - W * yy), rel=1e-12, abs=1e-300)


# ------------------------------------------------------- momentum field
def test_momentum_field_update_conserves_and_splits_the_increment():
    """update.c:36-59 interpolates the cell-centred increments DUX, DUY to
    the faces: a unit increment in one cell adds 1/2 to each of its two
    faces, so summed over faces the momentum added equals the sum of the
    increments (away from the walls, and exactly with the periodic y-wrap)."""
    nx, ny = 24, 20
    o = Oracle(nx, ny, ly=1.0)
    rng = np.random.default_rng(3)
    ux0, uy0, p0 = cases.random_fields(nx, ny, 1)
    o.upload(ux0, uy0, p0)
    o.update_boundaries()
    before = o.download()
    dux = np.zeros((ny + 2, nx + 2))
    duy = np.zeros((ny + 2, nx + 2))
    dux[1:ny + 1, 2:nx] = rng.standard_normal((ny, nx - 2))
    duy[1:ny + 1, 1:nx + 1] = rng.standard_normal((ny, nx))
    # halo rows as the exchange leaves them (periodic in y)
    for f in (dux, duy):
        f[0] = f[ny]
        f[ny + 1] = f[1]
    o.set_field("dux", dux)
    o.set_field("duy", duy)
    o.update_momentum_field()
    after = o.download()
    dUX = after["ux"][1:ny + 1] - before["ux"][1:ny + 1]
    dUY = after["uy"][1:ny + 1] - before["uy"][1:ny + 1]
    np.testing.assert_allclose(dUX.sum(axis=1), dux[1:ny + 1].sum(axis=1), rtol=0, atol=1e-12)
    np.testing.assert_allclose(dUY.sum(axis=0), duy[1:ny + 1].sum(axis=0), rtol=0, atol=1e-12)
    # a unit impulse
    o.upload(before["ux"], before["uy"], before["p"])
    imp = np.zeros((ny + 2, nx + 2))
    imp[7, 9] = 1.0
    o.set_field("dux", imp)
    o.set_field("duy", imp)
    o.update_momentum_field()
    after = o.download()
    d = after["ux"] - before["ux"]
    # UX(i,j) = ux[j, i-1]: faces i=9 and i=10 of cell (9, 7)
    assert d[7, 8] == pytest.approx(0.5, abs=1e-14) and d[7, 9] == pytest.approx(0.5, abs=1e-14)
    assert np.count_nonzero(np.abs(d) > 1e-14) == 2
    d = after["uy"] - before["uy"]
    assert d[7, 9] == pytest.approx(0.5, abs=1e-14) and d[8, 9] == pytest.approx(0.5, abs=1e-14)
    assert np.count_nonzero(np.abs(d[1:ny + 1]) > 1e-14) == 2


# --------------------------------------------------------------- pressure
def test_pressure_accumulates_the_potential_and_is_gauge_invariant():
    """update_pressure.c:7-24 adds psi to p.  Two consequences pin it:
    (1) p_new - p_old is the potential just solved for; (2) p enters the
    step only through its gradient, so a whole step (with particles)
    started from p + c gives the same velocities and particles, and p + c
    at the end."""
    nx, ny = 32, 48
    cfg = dict(itot=nx, jtot=ny, ly=1.5, Re=150.0, extfrcy=0.02)
    parts = dict(den=[1.2, 1.0], a=[0.1, 0.08], b=[0.05, 0.06], x=[0.3, 0.7], y=[0.4, 1.1],
                 az=[0.3, 2.0], ux=[0.1, 0.0], uy=[0.0, -0.2], vz=[0.5, 0.0])
    ux, uy, p = cases.random_fields(nx, ny, 7, amp=0.5)
    runs = []
    for c in (0.0, 3.25):
        o = Oracle(**cfg)
        o.upload(ux, uy, p + c)
        o.update_boundaries()
        o.set_particles(**parts)
        o.step()
        runs.append((o.download(), o.particles()))
    (f0, p0), (f1, p1) = runs
    np.testing.assert_allclose(f1["ux"], f0["ux"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(f1["uy"], f0["uy"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(p1, p0, rtol=1e-12, atol=1e-13)
    # the additive constant of psi is round-off (tdm.c:100-102 divides two
    # round-off-level numbers for wavenumber 0), so p is compared less its mean
    d = (f1["p"] - f0["p"])[1:ny + 1, 1:nx + 1]
    np.testing.assert_allclose(d - d.mean(), 0.0, rtol=0, atol=1e-12)
    # (1): one projection
    o = Oracle(**cfg)
    o.upload(ux, uy, p)
    o.update_boundaries()
    o.set_dt(1e-3)
    o.update_velocity(0)
    o.compute_potential(0)
    before = o.field("p")
    o.correct_velocity(0)
    o.update_pressure()
    psi = o.field("psi")
    assert np.abs(psi[1:ny + 1, 1:nx + 1]).max() > 1e-6
    np.testing.assert_array_equal(o.field("p")[1:ny + 1, 1:nx + 1],
                                  before[1:ny + 1, 1:nx + 1] + psi[1:ny + 1, 1:nx + 1])


# --------------------------------------------------------------- integrate
def _integrate_in_reference_order(o, record):
    """src/main.c:13-55, call for call, through the oracle's pinned
    functions; returns the sub-iteration counts and residual sequences."""
    counts = []
    for rk in range(3):
        o.reset_particle_increments()
        o.update_boundaries()
        o.compute_inertia(0)
        o.update_velocity(rk)
        o.exchange_momentum()
        o.update_momentum_field()
        o.compute_potential(rk)
        o.correct_velocity(rk)
        o.update_pressure()
        o.compute_collision_force(0)
        substep = 0
        while True:
            o.compute_inertia(1)
            o.compute_collision_force(1)
            res = o.increment_particles(rk)
            record.append(res)
            if res < 1.e-8:
                break
            if substep > 100:
                break
            substep += 1
        counts.append(substep + 1)
        o.update_particles()
    return counts


def _contact_case():
    cfg = dict(itot=64, jtot=96, ly=1.5, Re=200.0, extfrcy=8.0 / 200.0)
    parts = dict(den=[1.0, 1.4, 1.2, 1.0], a=[0.06, 0.05, 0.05, 0.04], b=[0.03, 0.04, 0.02, 0.04],
                 x=[0.30, 0.37, 0.045, 0.7], y=[0.40, 0.44, 1.2, 0.02],
                 az=[0.3, 1.2, 0.0, 0.0], ux=[0.1, -0.1, 0.0, 0.0],
                 uy=[0.2, 0.0, 0.0, 0.3], vz=[0.0, 1.0, 0.0, -2.0])
    return cfg, parts


def test_integrate_follows_main_c_call_for_call():
    """orc_integrate against src/main.c:13-55 restated here call by call
    (reset, boundaries, inertia(0), predictor, exchange, field update,
    potential, correction, pressure, collision(0), then inertia(1) +
    collision(1) + increment until residual < 1e-8 or past 100
    sub-iterations, then the particle update): bit-identical fields and
    particles, the same sub-iteration counts.  The case has two touching
    ellipses and one overlapping the wall, so collision(0) matters, and its
    residual sequence passes through [1e-8, 1e-6), so the tolerance does."""
    cfg, parts = _contact_case()
    o = Oracle(**cfg)
    o.set_particles(**parts)
    o.compute_collision_force(0)
    assert len(o.colliding_pairs()) >= 2
    ux, uy, p = cases.random_fields(cfg["itot"], cfg["jtot"], 11, amp=0.3)
    runs = []
    for mode in ("lib", "restated"):
        o = Oracle(**cfg)
        o.upload(ux, uy, p)
        o.update_boundaries()
        o.set_particles(**parts)
        res = []
        counts = []
        for _ in range(2):
            o.decide_dt()
            if mode == "lib":
                counts.append(o.integrate())
            else:
                counts.append(_integrate_in_reference_order(o, res))
        runs.append((o.download(), o.particles(), counts, res, o.colliding_pairs()))
    (fa, pa, ca, _, cpa), (fb, pb, cb, res, cpb) = runs
    assert ca == cb
    assert any(1e-8 <= r < 1e-6 for r in res), res
    assert cpa == cpb
    for k in ("ux", "uy", "p"):
        np.testing.assert_array_equal(fa[k], fb[k])
    np.testing.assert_array_equal(pa, pb)


def test_contact_force_acts_over_the_whole_stage():
    """A heavy pair pressed together in fluid at rest: collisions dominate,
    and over one step the velocity change is the contact force per unit
    mass times dt -- collision(0) at the stage start and collision(1) at the
    sub-iteration positions enter increment.c:23-38 as a trapezoid over the
    stage, so both are needed (force from collision.c:214-244, pinned
    bit-exact against the reference)."""
    cfg = dict(itot=64, jtot=64, ly=1.0, Re=100.0, extfrcy=0.0)
    den = 1e4
    parts = dict(den=[den, den], a=[0.05, 0.05], b=[0.05, 0.05], x=[0.455, 0.545],
                 y=[0.5, 0.5], az=[0.0, 0.0])
    # the contact is stiff (k/m ~ pi^2/r^3, collision.c:214-222): a step
    # short against its period keeps the force nearly constant
    dt = 1e-5
    o = Oracle(**cfg)
    o.set_particles(**parts)
    o.compute_collision_force(0)
    cf0 = o.particles()[:, CFX].copy()
    assert cf0[0] < 0 < cf0[1]
    o2 = Oracle(**cfg)
    o2.set_particles(**parts)
    o2.set_dt(dt)
    o2.integrate()
    r = o2.particles()
    np.testing.assert_allclose(r[:, UX], cf0 * dt, rtol=1e-3)
    assert abs(r[0, UY]) < 1e-6 * abs(r[0, UX])


def test_exchange_and_increment_conserve_momentum():
    """The exchange drives the fluid in the particle's shell towards the
    rigid-body velocity (dux = fx dt, exchange.c:69-72) and hands the
    opposite momentum to the particle (fux, exchange.c:73-76), which the
    increment applies over dt (increment.c:23-38, not gamma*dt: the fluid
    took the whole change in this stage).  Fluid (density 1) plus particle
    momentum is therefore unchanged: m dU + sum dux dx dy = 0, in every RK
    stage."""
    nx, ny, ly = 64, 64, 1.0
    for rk in range(3):
        o = Oracle(nx, ny, ly=ly)
        den, a, b = 1.7, 0.1, 0.05
        o.set_particles(den=[den], a=[a], b=[b], x=[0.52], y=[0.48], az=[1.1],
                        ux=[0.3], uy=[-0.45], vz=[1.2])
        rng = np.random.default_rng(rk)
        ux, uy, p = cases.random_fields(nx, ny, 2 + rk, amp=0.2)
        o.upload(ux, uy, p)
        o.update_boundaries()
        o.set_dt(3e-3)
        o.exchange_momentum()
        o.increment_particles(rk)  # no inertia or contact terms: iu*, cf* are zero
        r = o.particles()[0]
        m = den * 3.1415926535897932 * a * b
        dxdy = (1.0 / nx) * (ly / ny)
        assert r[DUX] != 0.0
        assert m * r[DUX] == pytest.approx(-o.field("dux")[1:ny + 1, 1:nx + 1].sum() * dxdy,
                                           rel=1e-11)
        assert m * r[DUY] == pytest.approx(-o.field("duy")[1:ny + 1, 1:nx + 1].sum() * dxdy,
                                           rel=1e-11)


def test_sub_iterations_are_capped_at_102():
    """src/main.c:37-50 breaks once substep > 100 after the increment, so a
    loop whose residual never falls below 1e-8 runs 102 sub-iterations.  A
    very light ellipse (den 0.05) spinning up in shear does not converge in
    its third stage."""
    o = Oracle(32, 32, ly=1.0, Re=100.0, extfrcy=0.0, uy_wall_xm=-0.5, uy_wall_xp=0.5)
    ux, uy, p = cases.couette_fields(32, 32, o.grid()["xc"])
    o.upload(ux, uy, p)
    o.update_boundaries()
    o.set_particles(den=[0.05], a=[0.12], b=[0.06], x=[0.5], y=[0.5], az=[0.3])
    o.decide_dt()
    counts = o.integrate()
    assert counts[:2] == [7, 9] and counts[2] == 102
