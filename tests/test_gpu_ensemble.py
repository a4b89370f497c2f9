"""Ensembles of independent cases (eif_gpu_ensemble_run_steps): the
concurrent paths (each case's steps as captured graphs on its own stream;
all cases' step graphs in one graph; batched launches that run each kernel
of a step once for all cases from per-case argument records) against each
case run alone through the plain path, bit for bit, on cases of uneven
grids and particle counts drawn like the benchmark's ensemble."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

eif = pytest.importorskip("ellipsesinflows_amd")
from ellipsesinflows_amd import initparticles  # noqa: E402

KEYS = ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")


def _make(cfg, parts):
    g = eif.Solver(poisson_mode=eif.POISSON_PARTITION, **cfg)
    nx, ny = cfg["itot"], cfg["jtot"]
    xc = g.grid()["xc"]
    uy = np.zeros((ny + 2, nx + 2))
    uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
    g.upload(uy=uy)
    g.update_boundaries()
    g.set_particles(**{k: parts[k] for k in KEYS})
    return g


def test_concurrent_ensemble_equals_each_case_alone():
    cases = initparticles.ensemble_cases(32, seed=11)
    # one of each grid, with small and large particle counts
    pick = []
    for nx, _ in initparticles.ENSEMBLE_GRIDS:
        ks = [k for k, (c, p) in enumerate(cases) if c["itot"] == nx]
        pick += ks[:2]
    sel = [cases[k] for k in pick]
    print("ensemble of", [(c["itot"], len(p["a"])) for c, p in sel], flush=True)
    steps = 6
    alone = []
    for cfg, parts in sel:
        g = _make(cfg, parts)
        g.run_steps(steps)
        alone.append((g.download(), g.particles()))
        g.close()
    solvers = [_make(cfg, parts) for cfg, parts in sel]
    eif.ensemble_run_steps(solvers, steps, concurrent=True)
    for k, g in enumerate(solvers):
        f = g.download()
        for name in ("ux", "uy", "p"):
            assert np.array_equal(f[name], alone[k][0][name]), (k, name)
        assert np.array_equal(g.particles(), alone[k][1]), k
        launches, resumes, captures, kk = g.graph_stats()
        assert launches >= steps - 1
    # the sequential form, one merged graph launch per ensemble step, and
    # batched launches (each kernel once for all cases) give the same again
    for g in solvers:
        g.close()
    for mode in (0, 2, 3):
        solvers = [_make(cfg, parts) for cfg, parts in sel]
        eif.ensemble_run_steps(solvers, steps, concurrent=mode)
        for k, g in enumerate(solvers):
            f = g.download()
            for name in ("ux", "uy", "p"):
                assert np.array_equal(f[name], alone[k][0][name]), (mode, k, name)
            assert np.array_equal(g.particles(), alone[k][1]), (mode, k)
            g.close()


def test_batched_ensemble_longer_run_and_rebuilds():
    """Batched launches over more steps than a graph is kept for: the
    sub-iteration bound of each stage follows the cases (a case whose loop
    needs more stops at the gate and is resumed alone; bounds that fall
    rebuild the batch), and a second call continues from the first."""
    cases = initparticles.ensemble_cases(32, seed=5)
    sel = [c for c in cases if c[0]["itot"] <= 512][:5]
    print("ensemble of", [(c["itot"], len(p["a"])) for c, p in sel], flush=True)
    alone = []
    for cfg, parts in sel:
        g = _make(cfg, parts)
        g.run_steps(30)
        alone.append((g.download(), g.particles()))
        g.close()
    solvers = [_make(cfg, parts) for cfg, parts in sel]
    eif.ensemble_run_steps(solvers, 12, concurrent=3)
    # one sub-iteration per stage: every case whose loops need more stops at
    # a gate (with the pipelined step behind it already launched) and is
    # resumed alone until its bounds have grown again
    for g in solvers:
        g.graph_set_subiterations((1, 1, 1))
    eif.ensemble_run_steps(solvers, 18, concurrent=3)
    assert sum(g.graph_stats()[1] for g in solvers) > 0
    for k, g in enumerate(solvers):
        f = g.download()
        for name in ("ux", "uy", "p"):
            assert np.array_equal(f[name], alone[k][0][name]), (k, name)
        assert np.array_equal(g.particles(), alone[k][1]), k
        g.close()


def test_ensemble_rejects_repeated_context_and_bad_mode():
    cases = initparticles.ensemble_cases(2, seed=1)
    g = _make(*cases[0])
    with pytest.raises(eif.EIFError):
        eif.ensemble_run_steps([g, g], 1, concurrent=3)
    with pytest.raises(eif.EIFError):
        eif.ensemble_run_steps([g], 1, concurrent=4)
    g.close()


def _small(nx, ny, n_parts, seed):
    cfg = dict(itot=nx, jtot=ny, ly=2.0, Re=200.0, extfrcy=8.0 / 200.0)
    parts = initparticles.random_ellipses(n_parts, 2.0, (0.03, 0.05), (0.5, 1.0), 0.02, seed) \
        if n_parts else None
    return cfg, parts


def _make_any(cfg, parts):
    g = eif.Solver(poisson_mode=eif.POISSON_PARTITION, **cfg)
    nx, ny = cfg["itot"], cfg["jtot"]
    xc = g.grid()["xc"]
    uy = np.zeros((ny + 2, nx + 2))
    uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
    g.upload(uy=uy)
    g.update_boundaries()
    if parts is not None:
        g.set_particles(**{k: parts[k] for k in KEYS})
    return g


def test_batched_ensemble_beyond_one_batch_with_fluid_only_cases():
    """66 small cases: two batches per ensemble step (64 + 2 cases); every
    fourth case has no particles, so its launch sequence has no IBM kernels
    and the cases' sequences do not line up position by position.  Each
    case still equals its own run bit for bit."""
    sel = [_small(64, 128, 0 if k % 4 == 0 else 2 + k % 3, 100 + k) for k in range(66)]
    print("66 cases, particle counts", [0 if p is None else len(p["a"]) for _, p in sel],
          flush=True)
    steps = 4
    alone = []
    for cfg, parts in sel:
        g = _make_any(cfg, parts)
        g.run_steps(steps)
        alone.append((g.download(), g.particles() if parts is not None else None))
        g.close()
    solvers = [_make_any(cfg, parts) for cfg, parts in sel]
    eif.ensemble_run_steps(solvers, steps, concurrent=3)
    for k, g in enumerate(solvers):
        f = g.download()
        for name in ("ux", "uy", "p"):
            assert np.array_equal(f[name], alone[k][0][name]), (k, name)
        if alone[k][1] is not None:
            assert np.array_equal(g.particles(), alone[k][1]), k
    for g in solvers:
        g.close()
