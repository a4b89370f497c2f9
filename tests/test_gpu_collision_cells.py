"""The collision broad phase (eif_gpu_set_collision_search): the cell list
against the test of all pairs that it replaces (the pair-matrix loop of
collision.c:320-332 with the bounding-circle exit of collision.c:172-189).

Both find the same partners of every particle in the same ascending order,
so steps are bit-identical: fields, particle records and the colliding
pairs.  Cases: the dense collision regime (1024 ellipses at ~30 % area
fraction, contacts from the start), pairs across the periodic boundary and
against both walls, a box too short for three cells across y (all pairs
then), graph steps and batched ensemble steps, and 6000 particles, where
the default takes the cell list."""
import time

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

eif = pytest.importorskip("ellipsesinflows_amd")
from ellipsesinflows_amd import initparticles  # noqa: E402

KEYS = ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")


def _make(kw, parts, search, graph=False, mode=eif.POISSON_PARTITION):
    g = eif.Solver(poisson_mode=mode, **kw)
    nx, ny = kw["itot"], kw["jtot"]
    xc = g.grid()["xc"]
    uy = np.zeros((ny + 2, nx + 2))
    uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
    g.upload(uy=uy)
    g.update_boundaries()
    g.set_particles(**{k: parts[k] for k in KEYS if k in parts})
    g.set_collision_search(search)
    if graph:
        g.graph(True)
    return g


def _same(a, b, what):
    fa, fb = a.download(), b.download()
    for k in ("ux", "uy", "p"):
        assert np.array_equal(fa[k], fb[k]), (what, k)
    assert np.array_equal(a.particles(), b.particles()), what


def _dense(seed=7):
    return initparticles.random_ellipses(1024, 2.0, (0.013, 0.019), (0.4, 1.0), -0.004, seed,
                                         den_range=(1.0, 1.5))


def test_cells_equal_all_pairs_dense():
    kw = dict(itot=512, jtot=1024, ly=2.0, Re=1000.0, extfrcy=8.0 / 1000.0)
    parts = _dense()
    allp = _make(kw, parts, eif.COLLISION_ALL_PAIRS)
    cell = _make(kw, parts, eif.COLLISION_CELLS)
    print("dense: 1024 ellipses, all pairs vs cell list", flush=True)
    seen = 0
    for k in range(3):
        assert allp.step() == cell.step(), k
        pa, pc = allp.colliding_pairs(), cell.colliding_pairs()
        print("step", k, "colliding pairs", len(pa), flush=True)
        seen += len(pa)
        assert pa == pc, k
        _same(allp, cell, k)
    assert seen > 20, seen
    allp.close()
    cell.close()


@pytest.mark.parametrize("ly,jtot,parts", [
    (1.0, 128, cases.touching_pairs(1.0)),
    (1.0, 128, cases.contact_pairs(12, 1.0, 5)),
    # r > .035: two cells across y, all pairs are tested
    (0.1875, 24, cases.contact_pairs(1, 0.1875, 6, a_range=(0.035, 0.04))),
])
def test_cells_equal_all_pairs_periodic_and_walls(ly, jtot, parts):
    kw = dict(itot=128, jtot=jtot, ly=ly, Re=150.0, extfrcy=0.1)
    allp = _make(kw, parts, eif.COLLISION_ALL_PAIRS, mode=eif.POISSON_THOMAS)
    cell = _make(kw, parts, eif.COLLISION_CELLS, mode=eif.POISSON_THOMAS)
    print("contacts, ly", ly, "particles", len(parts["a"]), flush=True)
    seen = 0
    for k in range(6):
        assert allp.step() == cell.step(), k
        pa = allp.colliding_pairs()
        seen += len(pa)
        assert pa == cell.colliding_pairs(), k
    assert seen > 0 or ly < 1.0
    _same(allp, cell, "end")
    allp.close()
    cell.close()


def test_cells_in_graph_steps_and_batched_ensemble():
    kw = dict(itot=256, jtot=512, ly=2.0, Re=500.0, extfrcy=8.0 / 500.0)
    parts = [_dense(3), _dense(4)]
    ref = [_make(kw, p, eif.COLLISION_ALL_PAIRS) for p in parts]
    for g in ref:
        g.run_steps(4)
    print("graph steps with the cell list", flush=True)
    gr = _make(kw, parts[0], eif.COLLISION_CELLS, graph=True)
    gr.run_steps(4)
    _same(ref[0], gr, "graph")
    gr.close()
    print("batched ensemble with the cell list", flush=True)
    ens = [_make(kw, p, eif.COLLISION_CELLS) for p in parts]
    eif.ensemble_run_steps(ens, 4, eif.ENSEMBLE_BATCHED)
    for k in range(2):
        _same(ref[k], ens[k], ("ensemble", k))
    for g in ref + ens:
        g.close()


def test_records_of_other_sizes_resize_the_cells_in_graph_steps():
    """Particle records uploaded with 5 % larger semi-axes change the cells
    (their size follows the widest reach) and drop the captured graphs:
    graph steps with the cell list then equal eager steps testing all
    pairs."""
    kw = dict(itot=512, jtot=1024, ly=2.0, Re=1000.0, extfrcy=8.0 / 1000.0)
    parts = _dense(5)
    out = []
    for search, graph in ((eif.COLLISION_ALL_PAIRS, False), (eif.COLLISION_CELLS, True)):
        g = _make(kw, parts, search, graph=graph)
        g.run_steps(2)
        rec = g.particles()
        rec[:, 1:3] *= 1.05  # a, b
        g.set_particle_records(rec)
        g.run_steps(2)
        out.append(g)
    print("resized records: pairs", len(out[0].colliding_pairs()), flush=True)
    assert out[0].colliding_pairs() == out[1].colliding_pairs()
    _same(out[0], out[1], "resized")
    for g in out:
        g.close()


def test_default_takes_cells_for_many_particles():
    """6000 small ellipses: the default (cell list from 4096 particles)
    against all pairs, two steps, with each mode's time per step."""
    kw = dict(itot=512, jtot=1024, ly=2.0, Re=500.0, extfrcy=8.0 / 500.0)
    parts = initparticles.random_ellipses(6000, 2.0, (0.004, 0.006), (0.5, 1.0), -0.001, 9)
    auto = _make(kw, parts, eif.COLLISION_AUTO)
    allp = _make(kw, parts, eif.COLLISION_ALL_PAIRS)
    times = {}
    for name, g in (("auto", auto), ("all pairs", allp)):
        g.step()
        t0 = time.perf_counter()
        g.step()
        g.particles()
        times[name] = time.perf_counter() - t0
    print("6000 particles, s per step:", times, flush=True)
    assert auto.colliding_pairs() == allp.colliding_pairs()
    _same(auto, allp, "6000")
    auto.close()
    allp.close()


def test_set_collision_search_rejects_bad_modes():
    g = eif.Solver(itot=64, jtot=64, ly=1.0, Re=100.0)
    for mode in (-1, 3):
        with pytest.raises(eif.EIFError):
            g.set_collision_search(mode)
    g.set_collision_search(eif.COLLISION_CELLS)
    g.close()
