"""Whole-step hipGraph mode (eif_gpu_graph_enable + eif_gpu_run_steps)
against the eager path through the C-ABI: the same kernels in the same
order per stream, so fields and particles must be bit-identical -- also
when the captured sub-iterations are too few and the device stops the step
for the host to resume it, and across replays with eager work and
synchronisations in between (the graph holds no memset nodes)."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

eif = pytest.importorskip("ellipsesinflows_amd")

KEYS = ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")


def _setup(nx, ny, parts, mode=1, re=400.0):
    g = eif.Solver(poisson_mode=mode, itot=nx, jtot=ny, ly=2.0, Re=re, extfrcy=8.0 / re)
    xc = g.grid()["xc"]
    uy = np.zeros((ny + 2, nx + 2))
    uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
    g.upload(uy=uy)
    g.update_boundaries()
    g.set_particles(**{k: parts[k] for k in KEYS if k in parts})
    return g


def _same(a, b, what):
    fa, fb = a.download(), b.download()
    for k in ("ux", "uy", "p"):
        assert np.array_equal(fa[k], fb[k]), (what, k, np.max(np.abs(fa[k] - fb[k])))
    assert np.array_equal(a.particles(), b.particles()), what


def test_graph_steps_equal_eager_steps():
    parts = cases.place_ellipses(16, 2.0, (0.03, 0.05), (0.4, 1.0), 0.01, 5, den_range=(1.0, 1.5))
    print("graph vs eager: 256x512, 16 ellipses, 3 x 8 steps", flush=True)
    e = _setup(256, 512, parts)
    g = _setup(256, 512, parts)
    g.graph(True)
    for rep in range(3):
        de, se = e.run_steps(8)
        dg, sg = g.run_steps(8)
        assert np.array_equal(de, dg), rep
        assert np.array_equal(se, sg), rep
        _same(e, g, "replay %d" % rep)
        # eager work on the graph's context between replays
        g.divergence()
        g.synchronize()
    launches, resumes, captures, k = g.graph_stats()
    print("launches", launches, "resumes", resumes, "captures", captures, "K", k)
    assert launches >= 20 and captures >= 1
    e.close()
    g.close()


def test_graph_stops_and_resumes_when_a_loop_needs_more_sub_iterations():
    """One captured sub-iteration per stage: every stage's loop needs more,
    so every step stops on the device and the host resumes it; results
    still equal the eager path bit for bit."""
    parts = cases.contact_pairs(4, 2.0, 21, overlap=(0.9, 0.98))
    print("graph resume: 128x256, 8 ellipses in contact, K = 1", flush=True)
    e = _setup(128, 256, parts, mode=0, re=200.0)
    g = _setup(128, 256, parts, mode=0, re=200.0)
    g.graph(True)
    de, se = e.run_steps(1)
    dg, sg = g.run_steps(1)
    g.graph_set_subiterations((1, 1, 1))
    de, se = e.run_steps(6)
    dg, sg = g.run_steps(6)
    assert np.array_equal(de, dg) and np.array_equal(se, sg)
    _same(e, g, "resumed")
    launches, resumes, captures, k = g.graph_stats()
    print("launches", launches, "resumes", resumes, "captures", captures, "K", k, "counts", se[-1])
    assert resumes >= 1
    assert e.colliding_pairs() == g.colliding_pairs()
    e.close()
    g.close()


def test_graph_c2_workload_steps_equal_eager():
    """The bench's workload at its real shape: 4 graph steps equal 4 eager
    steps bit for bit."""
    from ellipsesinflows_amd import initparticles
    print("graph vs eager: 2048x4096, 1024 ellipses, 4 steps", flush=True)
    parts = initparticles.random_ellipses(1024, 2.0, (0.008, 0.012), (0.4, 1.0), 0.0, 2,
                                          den_range=(1.0, 1.5))
    e = _setup(2048, 4096, parts, re=1000.0)
    de, se = e.run_steps(4)
    fe, pe = e.download(), e.particles()
    e.close()
    g = _setup(2048, 4096, parts, re=1000.0)
    g.graph(True)
    dg, sg = g.run_steps(4)
    assert np.array_equal(de, dg) and np.array_equal(se, sg)
    fg = g.download()
    for k in ("ux", "uy", "p"):
        assert np.array_equal(fe[k], fg[k]), k
    assert np.array_equal(pe, g.particles())
    g.close()
