"""Whole-step parity of the gfx950 path with the oracle on the workloads
the benchmark configurations name, at their real shapes and horizons, and
on the reference's own validation cases.

  validation   case 1 (Feng 1994) and case 3 (Aidun 1998) setups of
               docs/source/examples, against the oracle and the literature
  c1           512x1024, 64 ellipses, 1000 steps against an oracle
               trajectory stored by tests/golden/make_trajectories.py:
               dt and sub-iteration counts every step, colliding pairs every
               step, positions every 10 steps
  c4 forward   256x512, Re 200, 16 contact-free ellipses, 200 steps
  c3 shapes    each grid of the ensemble with its largest particle count
  dense c2     2048x4096 at ~30 % area fraction with initial contacts:
               colliding-pair sets every step
"""
import os

import numpy as np
import pytest

import cases
from oracle import Oracle
from test_validation import aidun_case, feng_case, follow, start

pytestmark = pytest.mark.gpu

eif = pytest.importorskip("ellipsesinflows_amd")

HERE = os.path.dirname(os.path.abspath(__file__))
KEYS = ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def rel_mod_mean(a, b):
    a = a[1:-1, 1:-1]
    b = b[1:-1, 1:-1]
    return rel(a - a.mean(), b - b.mean())


def pos_rel(g, o):
    rg, ro = g[:, 3:6], o[:, 3:6]
    return float(np.max(np.abs(rg - ro) / np.maximum(np.abs(ro), 1e-300)))


def laminar(s, nx, ny):
    xc = s.grid()["xc"]
    uy = np.zeros((ny + 2, nx + 2))
    uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
    return uy


def gpu_stepper(g, dt_log):
    t = [0.0]

    def step():
        dt = g.step()
        dt_log.append(dt)
        t[0] += dt
        return t[0]
    return step


@pytest.mark.parametrize("case,tmax", [("feng", 20.0), ("aidun", 1.0)])
def test_validation_cases_against_oracle_and_literature(case, tmax):
    """The GPU run of the reference's validation setups matches the oracle
    (same dt every step, positions within 1e-9 relative, the tracked
    quantity within 1e-9 absolute) and the literature within the oracle's
    own CPU bars (test_validation.py)."""
    cfg, parts, data, col = feng_case() if case == "feng" else aidun_case()
    g = eif.Solver(**cfg)
    o = Oracle(threads=8, **cfg)
    start(g, cfg, parts)
    start(o, cfg, parts)
    dts = []
    tg = follow(g, data[1:] if case == "feng" else data, col, tmax, gpu_stepper(g, dts))
    for k, dt in enumerate(dts):
        assert o.step() == dt, k
    assert pos_rel(g.particles(), o.particles()) < 1e-9
    assert abs(g.particles()[0, col] - o.particles()[0, col]) < 1e-9
    for (t_, v, _), row in zip(tg, data[1:] if case == "feng" else data):
        ref = row[1:].mean()
        assert abs(v - ref) < (0.015 if case == "feng" else 0.03), (t_, v, row)
    print(case, "steps", len(dts), "final", g.particles()[0, [3, 4, 5, col]])
    g.close()


def test_c1_1000_steps_against_stored_oracle_trajectory():
    """c1 (512x1024, ly 2, Re 500, 64 ellipses) for 1000 steps: every dt
    equal, every step's sub-iteration counts and colliding-pair set equal,
    positions within 1e-9 relative every 10 steps, fluid norms within 1e-10
    every 100 steps."""
    from golden.make_trajectories import c1_setup
    ref = np.load(os.path.join(HERE, "golden", "trajectories", "c1_1000.npz"), allow_pickle=False)
    parts, kw = c1_setup()
    g = eif.Solver(**kw)
    g.upload(uy=laminar(g, kw["itot"], kw["jtot"]))
    g.update_boundaries()
    g.set_particles(**parts)
    pairs_ref = {}
    for k, n0, n1 in ref["pairs"]:
        pairs_ref.setdefault(int(k), []).append((int(n0), int(n1)))
    worst = 0.0
    for k in range(1000):
        dt = g.step()
        assert dt == ref["dt"][k], k
        got = g.colliding_pairs()
        assert got == pairs_ref.get(k, []), (k, got, pairs_ref.get(k))
        if k % 10 == 9:
            rg = g.particles()[:, 3:6]
            ro = ref["pos"][k // 10]
            worst = max(worst, float(np.max(np.abs(rg - ro) / np.maximum(np.abs(ro), 1e-300))))
        if k % 100 == 99:
            f = g.download()
            p = f["p"][1:-1, 1:-1]
            nrm = [np.linalg.norm(f["ux"]), np.linalg.norm(f["uy"]), np.linalg.norm(p - p.mean())]
            np.testing.assert_allclose(nrm, ref["norms"][k // 100], rtol=1e-10)
    print("c1 1000 steps: worst position difference", worst, "pairs", len(ref["pairs"]))
    assert worst < 1e-9
    g.close()


def test_c4_forward_200_steps():
    """c4's forward workload (256x512, ly 2, Re 200, extfrcy 8/Re, 16
    ellipses a .04, b .02 placed at least .05 apart) for its 200 steps:
    fields within 1e-11, positions within 1e-9 relative."""
    parts = cases.place_ellipses(16, 2.0, (0.04, 0.04), (0.5, 0.5), 0.05, 4)
    kw = dict(itot=256, jtot=512, ly=2.0, Re=200.0, extfrcy=8.0 / 200.0)
    g = eif.Solver(**kw)
    o = Oracle(threads=8, **kw)
    for s in (g, o):
        s.upload(uy=laminar(s, 256, 512))
        s.update_boundaries()
        s.set_particles(**parts)
    for k in range(200):
        assert g.step() == o.step(), k
        if k % 50 == 49:
            a, b = g.download(), o.download()
            assert rel(a["ux"], b["ux"]) < 1e-11 and rel(a["uy"], b["uy"]) < 1e-11, k
            assert rel_mod_mean(a["p"], b["p"]) < 1e-10, k
    assert pos_rel(g.particles(), o.particles()) < 1e-9
    assert o.colliding_pairs() == [] == g.colliding_pairs()
    g.close()


@pytest.mark.parametrize("nx,ny,npart,seed", [(256, 512, 8, 1), (512, 1024, 128, 2),
                                             (1024, 2048, 256, 3)])
def test_c3_case_shapes(nx, ny, npart, seed):
    """The ensemble's grids with the larger particle counts it draws (a ~
    U[.01, .04], b/a ~ U[.4, 1]), a few steps each against the oracle."""
    from ellipsesinflows_amd import initparticles
    rng = np.random.default_rng(seed)
    re = float(np.exp(rng.uniform(np.log(100.0), np.log(1000.0))))
    parts = initparticles.random_ellipses(npart, 2.0, (0.01, 0.04), (0.4, 1.0), 0.0, seed)
    kw = dict(itot=nx, jtot=ny, ly=2.0, Re=re, extfrcy=8.0 / re)
    g = eif.Solver(**kw)
    o = Oracle(threads=8, **kw)
    for s in (g, o):
        s.upload(uy=laminar(s, nx, ny))
        s.update_boundaries()
        s.set_particles(**{k: parts[k] for k in KEYS})
    for k in range(3):
        assert g.step() == o.step(), k
        a, b = g.download(), o.download()
        assert rel(a["ux"], b["ux"]) < 1e-12 and rel(a["uy"], b["uy"]) < 1e-12, k
        assert rel_mod_mean(a["p"], b["p"]) < 1e-11, k
        assert g.colliding_pairs() == o.colliding_pairs(), k
    assert pos_rel(g.particles(), o.particles()) < 1e-9
    g.close()


def test_dense_c2_variant_pair_sets_every_step():
    """The collision-heavy regime the c2 description names: 1024 ellipses,
    a ~ U[.013, .019] (b/a ~ U[.4, 1], den ~ U[1, 1.5]) on 2048x4096, ~30 %
    area fraction, bounding circles allowed to overlap by up to .004 so that
    contacts exist from the start.  Two steps in the bench's mode
    (partitioned y-solve) against the oracle: the colliding-pair set of
    every step is identical, velocities within 1e-12, positions within
    1e-9 relative."""
    from ellipsesinflows_amd import initparticles
    nx, ny, ly, re = 2048, 4096, 2.0, 1000.0
    parts = initparticles.random_ellipses(1024, ly, (0.013, 0.019), (0.4, 1.0), -0.004, 7,
                                          den_range=(1.0, 1.5))
    phi = initparticles.area_fraction(parts, ly)
    assert 0.27 < phi < 0.33, phi
    kw = dict(itot=nx, jtot=ny, ly=ly, Re=re, extfrcy=8.0 / re)
    g = eif.Solver(poisson_mode=1, **kw)
    o = Oracle(threads=16, **kw)
    for s in (g, o):
        s.upload(uy=laminar(s, nx, ny))
        s.update_boundaries()
        s.set_particles(**{k: parts[k] for k in KEYS})
    for k in range(2):
        dg, do = g.step(), o.step()
        assert dg == pytest.approx(do, rel=1e-12), k
        pg, po = g.colliding_pairs(), o.colliding_pairs()
        print("step", k, "colliding pairs", len(po))
        assert len(po) > 50
        assert pg == po, k
        a, b = g.download(), o.download()
        assert rel(a["ux"], b["ux"]) < 1e-12 and rel(a["uy"], b["uy"]) < 1e-12, k
    assert pos_rel(g.particles(), o.particles()) < 1e-9
    g.close()
