"""The adjoint of one RK stage (eif_gpu_stage_adjoint) chained over whole
runs by the torch.autograd harness (tests/adjoint_harness.py, checkpoints
at segment starts, recomputation in between), against central finite
differences of the GPU forward (the same fixed-dt steps) and of the oracle.

dt is held at a schedule taken from decide_dt on the unperturbed run, so
the derivative is that of the run at that schedule (decide_dt's min() over
advective, diffusive and particle limits is not differentiated).  The
runs are contact-free (collisions are forward-only).  Agreement is
measured relative to the largest gradient component of the same kind, so
that components near zero are judged by the gradient's scale.

Finite differences of this discrete system carry an error the adjoint
does not see: an IBM box (auxiliary.c:11-17) moves by a whole cell when a
particle crosses a cell boundary, switching a cell of weight ~1e-5 on or
off, and the objective jumps by that cell's share.  central_fd uses
h = 1e-6 unless its result disagrees with the next smaller step's (a
switch between the two perturbed runs), and then the first smaller pair
that agrees."""
import time

import numpy as np
import pytest

import cases
from oracle import Oracle

pytestmark = pytest.mark.gpu

eif = pytest.importorskip("ellipsesinflows_amd")
torch = pytest.importorskip("torch")
import adjoint_harness as H  # noqa: E402


def laminar(s, nx, ny):
    xc = s.grid()["xc"]
    uy = np.zeros((ny + 2, nx + 2))
    uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
    return uy


def setup(nx, ny, re, parts, perturb=True, ly=2.0):
    cfg = dict(itot=nx, jtot=ny, ly=ly, Re=re, extfrcy=8.0 / re)
    s = eif.Solver(**cfg)
    uy0 = laminar(s, nx, ny)
    ux0 = np.zeros(s.shapes["ux"])
    if perturb:
        ur, vr, _ = cases.random_fields(nx, ny, 3, amp=0.05)
        ux0, uy0 = ux0 + ur, uy0 + vr
    s.upload(ux0, uy0, np.zeros(s.shapes["p"]))
    s.update_boundaries()
    if parts:
        s.set_particles(**parts)
    f = s.download()
    rec = s.particles() if parts else np.zeros((0, 30))
    return s, cfg, (f["ux"], f["uy"], f["p"], rec)


def gpu_J(s, cfg, state, dts, re=None, parts=None):
    s.upload(*state[:3])
    if s.n_particles:
        s.set_particle_records(state[3])
    return H.gpu_objective(s, dts)


def gpu_J_re(cfg, state, dts, re, parts):
    s = eif.Solver(**dict(cfg, Re=re))
    s.upload(*state[:3])
    s.update_boundaries()
    if parts:
        s.set_particles(**parts)
        s.set_particle_records(state[3])
    J = gpu_J(s, dict(cfg, Re=re), state, dts, re)
    s.close()
    return J


def fd_rel(adj, fd, scale):
    return abs(adj - fd) / max(abs(scale), 1e-300)


def central_fd(J, x0, scale, hs=(1e-6, 4e-7, 1.6e-7, 6.4e-8)):
    """Central difference at the first step size of hs whose result agrees
    with the next one's to 1e-7 of scale (a box switch between the two
    perturbed runs shows as a disagreement); returns (fd, h)."""
    vals = {}

    def fd(h):
        if h not in vals:
            vals[h] = (J(x0 + h) - J(x0 - h)) / (2 * h)
        return vals[h]
    for h0, h1 in zip(hs[:-1], hs[1:]):
        if abs(fd(h0) - fd(h1)) <= 1e-7 * abs(scale):
            return fd(h0), h0
    return fd(hs[0]), hs[0]


def test_adjoint_fluid_directions_and_re():
    """No particles: J is the wall shear; dJ/dRe and directional
    derivatives along smooth perturbations of the initial ux and uy."""
    s, cfg, st = setup(32, 64, 50.0, None)
    dts = [s.step() for _ in range(3)]
    J, g, gre = H.objective_and_gradient(s, cfg, st, dts, 2)
    assert J == pytest.approx(gpu_J(s, cfg, st, dts), rel=1e-14)
    h = 1e-6 * cfg["Re"]
    fd = (gpu_J_re(cfg, st, dts, cfg["Re"] + h, None) - gpu_J_re(cfg, st, dts, cfg["Re"] - h, None)) / (2 * h)
    print("dJ/dRe", gre, fd)
    assert fd_rel(gre, fd, fd) < 1e-6
    d_ux, d_uy, _ = cases.random_fields(32, 64, 9)
    for d in (d_ux, d_uy):
        d[:, 0] = d[:, -1] = 0.0
        d[0] = d[-1] = 0.0
    for k, d in ((0, d_ux), (1, d_uy)):
        sp = [a.copy() for a in st]
        sm = [a.copy() for a in st]
        sp[k] += 1e-6 * d
        sm[k] -= 1e-6 * d
        fd = (gpu_J(s, cfg, sp, dts) - gpu_J(s, cfg, sm, dts)) / 2e-6
        adj = float(np.sum(g[k] * d))
        print("direction", k, adj, fd)
        assert fd_rel(adj, fd, fd) < 1e-6
    # a pressure field alone is projected away without particles
    assert np.max(np.abs(g[2])) < 1e-12
    s.close()


def test_adjoint_particles_small():
    """Two contact-free ellipses, 3 steps: every particle coordinate and
    velocity, and Re, against finite differences of the GPU forward."""
    parts = dict(den=[1.0, 1.3], a=[0.08, 0.06], b=[0.05, 0.03], x=[0.35, 0.65], y=[0.6, 1.3],
                 az=[0.4, 2.0])
    s, cfg, st = setup(64, 128, 100.0, parts)
    dts = [s.step() for _ in range(3)]
    J, g, gre = H.objective_and_gradient(s, cfg, st, dts, 2)
    gp = g[3]
    scale = np.max(np.abs(gp[:, H.PCOLS]))
    worst = 0.0
    for n in range(2):
        for col in H.PCOLS:
            sp = [a.copy() for a in st]
            sm = [a.copy() for a in st]
            sp[3][n, col] += 1e-6
            sm[3][n, col] -= 1e-6
            fd = (gpu_J(s, cfg, sp, dts) - gpu_J(s, cfg, sm, dts)) / 2e-6
            worst = max(worst, fd_rel(gp[n, col], fd, scale))
    h = 1e-6 * cfg["Re"]
    fd = (gpu_J_re(cfg, st, dts, cfg["Re"] + h, parts) - gpu_J_re(cfg, st, dts, cfg["Re"] - h, parts)) / (2 * h)
    print("particles worst", worst, "Re", gre, fd)
    assert worst < 1e-6
    assert fd_rel(gre, fd, fd) < 1e-6
    s.close()


def test_c4_adjoint_checkpointed_against_gpu_and_oracle_fd():
    """c4: 256x512, ly 2, Re 200, extfrcy 8/Re, 16 ellipses a .04 b .02 at
    least .05 apart, laminar start; J over 200 steps with checkpoints every
    20.  Gradient with respect to every initial centre and orientation and
    Re; checked against central differences (h = 1e-6) of the GPU forward
    for the centre and angle of three ellipses and Re, and of the oracle
    (same dt schedule) for one centre coordinate and Re."""
    parts = cases.place_ellipses(16, 2.0, (0.04, 0.04), (0.5, 0.5), 0.05, 4)
    s, cfg, st = setup(256, 512, 200.0, parts, perturb=False)
    t0 = time.perf_counter()
    dts = [s.step() for _ in range(200)]
    s.synchronize()
    t_fwd = time.perf_counter() - t0
    assert s.colliding_pairs() == []
    t0 = time.perf_counter()
    J, g, gre = H.objective_and_gradient(s, cfg, st, dts, 20)
    t_adj = time.perf_counter() - t0
    print("c4: J %.12e, forward %.3f s, gradient %.3f s, ratio %.2f" % (J, t_fwd, t_adj, t_adj / t_fwd))
    gp = g[3]
    scale = np.max(np.abs(gp[:, [3, 4]]))
    scale_az = np.max(np.abs(gp[:, 5]))
    worst = 0.0
    for n in (0, 7, 15):
        for col, sc in ((3, scale), (4, scale), (5, scale_az)):
            def J_of(v, n=n, col=col):
                sv = [a.copy() for a in st]
                sv[3][n, col] = v
                return gpu_J(s, cfg, sv, dts)
            fd, h_used = central_fd(J_of, st[3][n, col], sc)
            e = fd_rel(gp[n, col], fd, sc)
            print("particle", n, "col", col, gp[n, col], fd, "h", h_used, "err", e)
            worst = max(worst, e)
    h = 1e-6 * cfg["Re"]
    fd_re = (gpu_J_re(cfg, st, dts, cfg["Re"] + h, parts) - gpu_J_re(cfg, st, dts, cfg["Re"] - h, parts)) / (2 * h)
    print("Re", gre, fd_re)
    assert worst < 1e-6
    assert fd_rel(gre, fd_re, fd_re) < 1e-6

    # the oracle with the same schedule
    def oracle_J(rec, re):
        o = Oracle(threads=16, **dict(cfg, Re=re))
        o.upload(*st[:3])
        o.update_boundaries()
        o.set_particles(**parts)
        o.set_particle_records(rec)
        grid = {"xc": o.grid()["xc"], "dxc": o.grid()["dxc"], "dy": 2.0 / 512}

        def step(dt):
            o.set_dt(dt)
            o.integrate()
        J_ = H.forward_objective(step, o.download, o.particles, grid, dts, re, cfg)
        o.close()
        return J_
    Jo = oracle_J(st[3], cfg["Re"])
    print("oracle J", Jo, "gpu J", J)
    assert Jo == pytest.approx(J, rel=1e-10)
    rp, rm = st[3].copy(), st[3].copy()
    rp[7, 3] += 1e-6
    rm[7, 3] -= 1e-6
    fd_o = (oracle_J(rp, cfg["Re"]) - oracle_J(rm, cfg["Re"])) / 2e-6
    fd_o_re = (oracle_J(st[3], cfg["Re"] + h) - oracle_J(st[3], cfg["Re"] - h)) / (2 * h)
    print("oracle fd x7", fd_o, "adjoint", gp[7, 3], "oracle fd Re", fd_o_re, "adjoint", gre)
    assert fd_rel(gp[7, 3], fd_o, scale) < 1e-6
    assert fd_rel(gre, fd_o_re, fd_o_re) < 1e-6
    s.close()
