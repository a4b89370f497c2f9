"""Pins the oracle where the reference's own sources cannot be built here
(they include <fftw3.h>, absent from this image): the DCT definitions of
FFTW REDFT10/REDFT01, the periodic tridiagonal solver (src/tdm.c), the
Poisson projection (src/fluid/compute_potential.c, correct_velocity.c),
the predictor stencils (src/fluid/update_velocity.c), the time-step size
(src/param/decide_dt.c), the IBM sums (src/suspensions/inertia.c,
exchange.c) and the particle update (increment.c, update.c) -- each
against mathematics the reference's algorithm must satisfy, chosen so a
dropped term, a wrong sign or index, or a transposed operand fails."""
import ctypes

import numpy as np
import pytest

import cases
from oracle import Oracle, oracle_lib

_PD = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_PD)


# ---------------------------------------------------------------------- DCT
@pytest.mark.parametrize("n", [8, 16, 48, 64, 96, 128, 256, 2048,
                               12, 45, 77, 97, 100, 360, 1000])
def test_dct_definitions_against_long_double_sums(n):
    """REDFT10: Y_k = 2 sum_j X_j cos(pi (j+1/2) k / n);
    REDFT01: Y_k = X_0 + 2 sum_{j>=1} X_j cos(pi j (k+1/2) / n).
    Lengths that are not powers of two go through the Stockham mixed-radix
    FFT: even and odd lengths, a prime (97, one radix-97 stage) and
    radices 2, 3, 5, 7, 11."""
    lib = oracle_lib()
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    j = np.arange(n, dtype=np.longdouble)
    k = j[:, None]
    pi = np.longdouble("3.14159265358979323846264338327950288")
    y2 = x.copy()
    lib.orc_dct2(n, _ptr(y2))
    ref2 = 2 * np.sum(x[None, :].astype(np.longdouble) * np.cos(pi * (j[None, :] + 0.5) * k / n), axis=1)
    y3 = x.copy()
    lib.orc_dct3(n, _ptr(y3))
    ref3 = x[0] + 2 * np.sum(x[None, 1:].astype(np.longdouble) * np.cos(pi * j[None, 1:] * (k + 0.5) / n), axis=1)
    tol = 8e-16 * np.sqrt(n) * np.abs(x).sum()
    assert np.max(np.abs(y2 - ref2.astype(np.float64))) < tol
    assert np.max(np.abs(y3 - ref3.astype(np.float64))) < tol
    # the round trip scales by 2n (compute_potential.c:132 divides by 2 itot)
    z = y2.copy()
    lib.orc_dct3(n, _ptr(z))
    assert np.max(np.abs(z / (2 * n) - x)) < 1e-13 * np.abs(x).max() * np.log2(n)
    if n & (n - 1):
        return
    # two rows through one complex FFT (the Poisson solve's form): each row
    # matches its own sums, and swapping the rows swaps the results exactly
    xb = rng.standard_normal(n)
    refb2 = 2 * np.sum(xb[None, :].astype(np.longdouble) * np.cos(pi * (j[None, :] + 0.5) * k / n), axis=1)
    refb3 = xb[0] + 2 * np.sum(xb[None, 1:].astype(np.longdouble) * np.cos(pi * j[None, 1:] * (k + 0.5) / n), axis=1)
    tolb = 8e-16 * np.sqrt(n) * (np.abs(x).sum() + np.abs(xb).sum())
    for fn, ra, rb in ((lib.orc_dct2_pair, ref2, refb2), (lib.orc_dct3_pair, ref3, refb3)):
        pa, pb = x.copy(), xb.copy()
        assert fn(n, _ptr(pa), _ptr(pb)) == 0
        assert np.max(np.abs(pa - ra.astype(np.float64))) < tolb
        assert np.max(np.abs(pb - rb.astype(np.float64))) < tolb
        qa, qb = xb.copy(), x.copy()
        fn(n, _ptr(qa), _ptr(qb))
        assert np.max(np.abs(qa - rb.astype(np.float64))) < tolb
        assert np.max(np.abs(qb - ra.astype(np.float64))) < tolb
    # a zero partner: the row's result to rounding, the partner's ~0
    pa, pb = x.copy(), np.zeros(n)
    lib.orc_dct2_pair(n, _ptr(pa), _ptr(pb))
    assert np.max(np.abs(pa - ref2.astype(np.float64))) < tol
    assert np.max(np.abs(pb)) < tol


# ------------------------------------------------------- periodic tridiagonal
@pytest.mark.parametrize("n,c", [(16, -4.0), (37, -2.5), (64, -2.0 - 1e-3), (256, -3.0)])
def test_periodic_tdm_against_dense_solve(n, c):
    lib = oracle_lib()
    rng = np.random.default_rng(n)
    l = np.full(n, 1.0)
    u = np.full(n, 1.0)
    cc = np.full(n, c)
    A = np.diag(cc) + np.diag(u[:-1], 1) + np.diag(l[1:], -1)
    A[0, -1] = l[0]
    A[-1, 0] = u[-1]
    b = rng.standard_normal(n)
    q = b.copy()
    lib.orc_tdm_periodic(n, _ptr(l), _ptr(cc), _ptr(u), _ptr(q))
    x = np.linalg.solve(A, b)
    assert np.max(np.abs(q - x)) < 1e-11 * np.abs(x).max()


def test_periodic_tdm_singular_mode_gives_a_solution():
    """wavenumber 0: c = -2l, singular with constant null space; for a
    compatible right-hand side the result solves the system up to a constant."""
    lib = oracle_lib()
    n = 128
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n)
    b -= b.mean()
    l = np.ones(n)
    cc = np.full(n, -2.0)
    q = b.copy()
    lib.orc_tdm_periodic(n, _ptr(l), _ptr(cc), _ptr(l), _ptr(q))
    r = np.roll(q, 1) - 2 * q + np.roll(q, -1) - b
    assert np.max(np.abs(r)) < 1e-9


# -------------------------------------------------------- Poisson / projection
def test_projection_removes_divergence_and_matches_eigen_solution():
    nx, ny, ly = 32, 64, 2.0
    o = Oracle(nx, ny, ly=ly, Re=100.0)
    ux, uy, p = cases.random_fields(nx, ny, 4)
    o.upload(ux, uy, p)
    o.update_boundaries()
    o.set_dt(0.01)
    g = o.grid()
    U = o.download()
    # discrete divergence of the input (compute_potential.c:70-82 RHS form)
    div = ((U["ux"][1:ny + 1, 1:nx + 1] - U["ux"][1:ny + 1, 0:nx]) / (1.0 / nx)
           + (U["uy"][2:ny + 2, 1:nx + 1] - U["uy"][1:ny + 1, 1:nx + 1]) / (ly / ny))
    o.compute_potential(0)
    psi = o.field("psi")[1:ny + 1, 1:nx + 1]
    # independent solve: DCT-II basis in x (Neumann), Fourier in y
    gamma, dt = 32.0 / 60.0, 0.01
    rhs = div / (gamma * dt)
    kx = np.arange(nx)
    ky = np.fft.fftfreq(ny) * ny
    lam_x = -4.0 * nx ** 2 * np.sin(np.pi * kx / (2 * nx)) ** 2
    lam_y = -4.0 * (ny / ly) ** 2 * np.sin(np.pi * ky / ny) ** 2
    i = np.arange(nx)
    Cx = np.cos(np.pi * (i[:, None] + 0.5) * kx[None, :] / nx)  # (i, k)
    coef = np.linalg.solve(Cx, rhs.T)  # (k, j)
    F = np.fft.fft(coef, axis=1)
    lam = lam_x[:, None] + lam_y[None, :]
    lam[0, 0] = 1.0
    F = F / lam
    F[0, 0] = 0.0
    sol = (Cx @ np.real(np.fft.ifft(F, axis=1))).T
    d = (psi - psi.mean()) - (sol - sol.mean())
    assert np.max(np.abs(d)) < 1e-9 * np.abs(sol).max()
    o.correct_velocity(0)
    dmax, _ = o.divergence()
    assert dmax < 1e-10 * np.abs(div).max()


# ----------------------------------------------------------------- predictor
def test_couette_state_has_zero_advection_and_diffusion():
    """uy = -0.5 + x with walls moving at -/+0.5 and ux = 0 (case 3): the
    explicit terms of y-momentum reduce to the body force, those of
    x-momentum vanish."""
    nx, ny = 48, 32
    o = Oracle(nx, ny, ly=1.0, Re=50.0, extfrcy=0.125, uy_wall_xm=-0.5, uy_wall_xp=0.5)
    g = o.grid()
    ux, uy, p = cases.couette_fields(nx, ny, g["xc"])
    o.upload(ux, uy, p)
    o.update_boundaries()
    o.set_dt(1e-3)
    o.update_velocity(0)
    assert np.max(np.abs(o.field("srcuya") - 0.125)) < 1e-12
    assert np.max(np.abs(o.field("srcuxa"))) < 1e-14


def test_poiseuille_profile_is_discrete_steady_state():
    """uy = (Re f/2) x (1-x): (1/Re) d2/dx2 uy = -f exactly for a quadratic
    on the uniform interior cells (update_velocity.c:146-149)."""
    nx, ny, Re, f = 64, 16, 400.0, 0.02
    o = Oracle(nx, ny, ly=0.5, Re=Re, extfrcy=f)
    g = o.grid()
    ux = np.zeros((ny + 2, nx + 1))
    uy = np.zeros((ny + 2, nx + 2))
    xc = g["xc"]
    uy[1:ny + 1, 1:nx + 1] = 0.5 * Re * f * xc[None, 1:nx + 1] * (1 - xc[None, 1:nx + 1])
    o.upload(ux, uy, np.zeros_like(uy))
    o.update_boundaries()
    o.set_dt(1e-3)
    o.update_velocity(0)
    s = o.field("srcuya")[:, 1:nx - 1]  # i = 2..nx-1: uniform neighbours
    assert np.max(np.abs(s)) < 1e-9 * 0.25 * Re * f


def test_y_diffusion_matches_discrete_laplacian_eigenvalue():
    """ux = A sin(2 pi m y / ly), uy = 0: away from the walls the explicit
    term of x-momentum is (1/Re) * (2 cos(2 pi m dy/ly) - 2)/dy^2 * ux."""
    nx, ny, ly, Re, m = 16, 40, 1.25, 7.0, 3
    o = Oracle(nx, ny, ly=ly, Re=Re, extfrcy=0.0)
    dy = ly / ny
    yc = (np.arange(ny + 2) - 0.5) * dy
    ux = np.zeros((ny + 2, nx + 1))
    ux[:, :] = np.sin(2 * np.pi * m * yc / ly)[:, None]
    o.upload(ux, np.zeros((ny + 2, nx + 2)), np.zeros((ny + 2, nx + 2)))
    o.update_boundaries()
    o.set_dt(1e-3)
    o.update_velocity(0)
    s = o.field("srcuxa")[:, 2:nx - 3]  # columns i = 4..nx-2: constant in x
    lam = (2 * np.cos(2 * np.pi * m * dy / ly) - 2) / dy ** 2
    expect = lam / Re * np.sin(2 * np.pi * m * yc[1:ny + 1] / ly)[:, None]
    assert np.max(np.abs(s - expect)) < 1e-10 * np.abs(expect).max()


@pytest.mark.parametrize("seed", [5, 7])
def test_advection_conserves_discrete_kinetic_energy(seed):
    """For a discretely divergence-free field the advective terms transfer no
    kinetic energy: sum ux*A_x*dxc*dy + uy*A_y*dxf*dy = 0 (the reference's
    stencil is energy conserving; a wrong sign, index or weight breaks it)."""
    nx, ny = 32, 48
    o = Oracle(nx, ny, ly=1.5, Re=1e300, extfrcy=0.0)
    ux, uy, p = cases.random_fields(nx, ny, seed)
    o.upload(ux, uy, np.zeros_like(p))
    o.update_boundaries()
    o.set_dt(1.0)
    o.compute_potential(0)
    o.correct_velocity(0)
    U = o.download()
    o.set_field("p", np.zeros_like(U["p"]))
    o.update_boundaries()
    o.update_velocity(0)
    g = o.grid()
    sa, sb = o.field("srcuxa"), o.field("srcuya")
    ux_i = U["ux"][1:ny + 1, 1:nx]
    uy_i = U["uy"][1:ny + 1, 1:nx + 1]
    tx = ux_i * sa * g["dxc"][1:nx][None, :]
    ty = uy_i * sb * g["dxf"][None, :]
    scale = np.abs(tx).sum() + np.abs(ty).sum()
    assert abs(tx.sum() + ty.sum()) < 1e-13 * scale


def test_rk_update_is_alpha_beta_gamma_combination():
    """update_velocity.c:192-195: u += a dt A + b dt B + g dt G with B the
    previous stage's A and G the pressure gradient."""
    nx, ny = 24, 20
    o = Oracle(nx, ny, ly=1.0, Re=30.0, extfrcy=0.1)
    ux, uy, p = cases.random_fields(nx, ny, 9)
    o.upload(ux, uy, p)
    o.update_boundaries()
    dt = 2e-3
    o.set_dt(dt)
    o.update_velocity(0)
    A0 = o.field("srcuxa")
    u1 = o.download()
    o.update_velocity(1)
    A1, B1, G1 = o.field("srcuxa"), o.field("srcuxb"), o.field("srcuxg")
    assert np.array_equal(B1, A0)
    u2 = o.download()
    inc = u2["ux"][1:ny + 1, 1:nx] - u1["ux"][1:ny + 1, 1:nx]
    expect = (25 / 60) * dt * A1 + (-17 / 60) * dt * B1 + (8 / 60) * dt * G1
    assert np.max(np.abs(inc - expect)) < 1e-14 * (1 + np.abs(u1["ux"]).max())
    # G is the pressure gradient at the x-faces
    P = u1["p"]
    dxc = o.grid()["dxc"]
    grad = -(P[1:ny + 1, 2:nx + 1] - P[1:ny + 1, 1:nx]) / dxc[None, 1:nx]
    assert np.allclose(G1, grad, rtol=0, atol=1e-12 * np.abs(grad).max())


# ----------------------------------------------------------------- time step
def test_decide_dt_closed_form():
    nx, ny, ly, Re = 40, 30, 1.5, 80.0
    sf = (0.7, 0.6, 0.9)
    o = Oracle(nx, ny, ly=ly, Re=Re, safefactors=sf)
    ux = np.zeros((ny + 2, nx + 1))
    uy = np.zeros((ny + 2, nx + 2))
    ux[1:ny + 1, 1:nx] = 0.3
    uy[1:ny + 1, 1:nx + 1] = -0.8
    ux[5, 7] = 2.0
    o.upload(ux, uy, np.zeros_like(uy))
    o.update_boundaries()
    o.set_particles(den=[1.0], a=[0.1], b=[0.05], x=[0.5], y=[0.5], az=[0.0],
                    ux=[0.1], uy=[0.2], vz=[3.0])
    dt = o.decide_dt()
    dx, dy = 1.0 / nx, ly / ny
    adv = min(0.05, dx / 2.0, dy / 0.8) * sf[0]
    dif = 0.25 * Re * min(dx, dy) ** 2 * sf[1]
    par = min(dx, dy) / max(0.1 + 0.1 * 3.0, 0.2 + 0.1 * 3.0) * sf[2]
    assert dt == pytest.approx(min(adv, dif, par), rel=1e-15)


# ------------------------------------------------------- IBM and particles
def test_circle_weights_closed_form():
    """For a = b the signed distance is (r - |x - c|)/h, so
    s_w = (1 - tanh^2(2d)) and v_w = (1 + tanh(2d))/2 (auxiliary.c:49-59)."""
    lib = oracle_lib()
    rng = np.random.default_rng(21)
    r, h = 0.05, 1.0 / 256
    for _ in range(300):
        px, py, az = 0.5, 0.5, rng.uniform(0, 6.28)
        x, y = px + rng.uniform(-0.1, 0.1), py + rng.uniform(-0.1, 0.1)
        d = (r - np.hypot(x - px, y - py)) / h
        assert lib.orc_v_weight(h, r, r, px, py, az, x, y) == pytest.approx(0.5 * (1 + np.tanh(2 * d)), abs=2e-12)
        assert lib.orc_s_weight(h, r, r, px, py, az, x, y) == pytest.approx(1 - np.tanh(2 * d) ** 2, abs=2e-12)


def test_find_normal_t_circle_and_brute_force():
    lib = oracle_lib()
    rng = np.random.default_rng(22)
    for _ in range(200):
        x, y = rng.uniform(-0.2, 0.2, 2)
        t = lib.orc_find_normal_t(0.1, 0.1, x, y)
        assert np.cos(t) == pytest.approx(np.cos(np.arctan2(y, x)), abs=1e-9)
        assert np.sin(t) == pytest.approx(np.sin(np.arctan2(y, x)), abs=1e-9)
    # ellipses, points outside: t minimises the distance (dense search)
    ts = np.linspace(-np.pi, np.pi, 400001)
    for _ in range(40):
        a = rng.uniform(0.05, 0.2)
        b = a * rng.uniform(0.3, 1.0)
        ang = rng.uniform(0, 2 * np.pi)
        rad = rng.uniform(1.05, 2.0)
        x, y = rad * a * np.cos(ang), rad * b * np.sin(ang)
        t = lib.orc_find_normal_t(a, b, x, y)
        dist = np.hypot(a * np.cos(ts) - x, b * np.sin(ts) - y)
        got = np.hypot(a * np.cos(t) - x, b * np.sin(t) - y)
        assert got <= dist.min() + 1e-9


def test_inertia_of_uniform_and_rigid_fields():
    """inertia.c:45-63: for uniform (U, V) the translational integrals are
    (U, V) * sum(w) dxdy / m and the rotational one vanishes by symmetry for
    an axis-aligned ellipse centred on a cell corner; for rigid rotation
    u = -W (y - yp), v = W (x - xp) the translational ones vanish."""
    nx = ny = 128
    o = Oracle(nx, ny, ly=1.0)
    g = o.grid()
    xc, yc = g["xc"], np.concatenate([[0.0], g["yc"], [0.0]])
    a, b = 0.15, 0.08
    o.set_particles(den=[1.0], a=[a], b=[b], x=[0.5], y=[0.5], az=[0.0])
    ux = np.full((ny + 2, nx + 1), 0.7)
    uy = np.full((ny + 2, nx + 2), -0.4)
    o.upload(ux, uy, np.zeros_like(uy))
    o.update_boundaries()
    v = o.view("ux")
    v[:, 1:nx] = 0.7  # keep the walls but make the box interior uniform
    o.compute_inertia(0)
    rec = o.particles()[0]
    lib = oracle_lib()
    h = 1.0 / nx
    # box of auxiliary.c:11-17: [floor((c-r)/h)-3, ceil((c+r)/h)+3]
    lo, hi = int(np.floor((0.5 - a) / h)) - 3, int(np.ceil((0.5 + a) / h)) + 3
    W = sum(lib.orc_v_weight(h, a, b, 0.5, 0.5, 0.0, xc[i], yc[j])
            for j in range(lo, hi + 1) for i in range(lo, hi + 1))
    m = np.pi * a * b
    assert rec[18] == pytest.approx(0.7 * W * h * h / m, rel=1e-12)
    assert rec[20] == pytest.approx(-0.4 * W * h * h / m, rel=1e-12)
    # the box is not symmetric about the centre (floor/ceil, auxiliary.c:12-13),
    # so only the far-tail cells contribute to the angular integral
    assert abs(rec[22]) < 1e-7
    # rigid rotation about the centre
    Wz = 2.0
    X = np.broadcast_to(g["xf"][None, :], (ny + 2, nx + 1))
    Y = np.broadcast_to(yc[:, None], (ny + 2, nx + 1))
    o.view("ux")[...] = -Wz * (Y - 0.5)
    o.view("uy")[...] = Wz * (np.broadcast_to(xc[None, :], (ny + 2, nx + 2)) - 0.5)
    o.compute_inertia(1)
    rec = o.particles()[0]
    assert abs(rec[19]) < 1e-8 and abs(rec[21]) < 1e-8
    # I = rho * int r^2 dA, so the angular integral recovers W up to the
    # smoothing of the indicator
    assert rec[23] == pytest.approx(Wz, rel=0.05)


def test_exchange_relaxes_toward_rigid_body_velocity():
    """exchange.c:64-76: with the fluid already moving like the particle
    the forcing vanishes; with the fluid at rest and U != 0 the field forcing
    points along U and the particle force opposes it."""
    nx = ny = 96
    o = Oracle(nx, ny, ly=1.0)
    o.set_particles(den=[1.0], a=[0.1], b=[0.06], x=[0.5], y=[0.5], az=[0.4],
                    ux=[0.3], uy=[0.0], vz=[0.0])
    o.set_dt(1e-3)
    ux = np.full((ny + 2, nx + 1), 0.3)
    o.upload(ux, np.zeros((ny + 2, nx + 2)), np.zeros((ny + 2, nx + 2)))
    o.view("ux")[...] = 0.3
    o.exchange_momentum()
    assert np.max(np.abs(o.field("dux"))) < 1e-15
    assert abs(o.particles()[0, 15]) < 1e-12
    o.view("ux")[...] = 0.0
    o.exchange_momentum()
    assert o.field("dux").min() >= 0.0 and o.field("dux").max() > 0.0
    assert o.particles()[0, 15] < 0.0


def test_particle_at_rest_in_quiescent_fluid_stays_at_rest():
    nx = ny = 64
    o = Oracle(nx, ny, ly=1.0, Re=10.0, extfrcy=0.0)
    o.set_particles(den=[1.0], a=[0.1], b=[0.05], x=[0.5], y=[0.5], az=[0.3])
    for _ in range(3):
        o.step()
    rec = o.particles()[0]
    assert np.max(np.abs(rec[[3, 4, 5]] - [0.5, 0.5, 0.3])) == 0.0
    assert np.max(np.abs(rec[9:12])) == 0.0


def test_update_wraps_position_and_angle():
    o = Oracle(16, 16, ly=1.0)
    o.set_particles(den=[1.0], a=[0.05], b=[0.05], x=[0.5], y=[0.999], az=[6.28],
                    ux=[0.0], uy=[1.0], vz=[1.0])
    rec = o.particles()
    rec[0, 7] = 0.002   # dy
    rec[0, 8] = 0.01    # daz
    o.set_particle_records(rec)
    o.update_particles()
    r = o.particles()[0]
    assert r[4] == pytest.approx(0.001, abs=1e-15)
    assert r[5] == pytest.approx(6.29 - 2 * 3.1415926535897932, abs=1e-15)


def test_statistics_accumulate_and_indicator():
    """statistics_collect (src/statistics/collect.c:10-95): collecting the
    same fields k times gives k*u and k*u^2 (k <= 3 rounds identically),
    and for an axis-aligned ellipse (az = 0, so the rotation is the
    identity) phi is the textbook indicator of cell centres inside
    ((x-px)/a)^2 + ((y-py)/b)^2 < 1, including the periodic image of an
    ellipse that crosses y = 0."""
    nx, ny, ly = 40, 50, 1.0
    o = Oracle(itot=nx, jtot=ny, ly=ly, Re=100.0)
    ux, uy, p = cases.random_fields(nx, ny, 5)
    o.upload(ux, uy, p)
    o.update_boundaries()
    parts = dict(den=[1.0, 1.0], a=[0.2, 0.1], b=[0.1, 0.15], x=[0.4, 0.7], y=[0.5, 0.05],
                 az=[0.0, 0.0])
    o.set_particles(**parts)
    for _ in range(3):
        o.statistics_collect()
    st = o.statistics()
    f = o.download()
    assert st["num"] == 3
    rows = slice(1, ny + 1)
    np.testing.assert_array_equal(st["ux1"][rows], 3 * f["ux"][rows])
    np.testing.assert_array_equal(st["ux2"][rows], 3 * (f["ux"][rows] * f["ux"][rows]))
    np.testing.assert_array_equal(st["uy1"][rows], 3 * f["uy"][rows])
    assert not st["ux1"][0].any() and not st["uy2"][ny + 1].any()
    xc = o.grid()["xc"]
    yc = o.grid()["yc"]
    X, Y = np.meshgrid(xc[1:nx + 1], yc)
    inside = np.zeros((ny, nx))
    for a, b, px, py in zip(parts["a"], parts["b"], parts["x"], parts["y"]):
        for per in (-1, 0, 1):
            inside += (1.0 - ((X - px) / a) ** 2 - ((Y - (py + ly * per)) / b) ** 2 > 0)
    np.testing.assert_array_equal(st["phi"][1:ny + 1, 1:nx + 1], 3 * inside)
    assert inside[:3].any() and inside[-3:].any()  # the second ellipse wraps
