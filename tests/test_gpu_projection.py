"""The fused DCT-III + projection + pressure update (eif_gpu_set_projection
with a strip height of 2, 4 or 8 rows) against the oracle and against the
unfused kernels (compute_potential.c:118-137, correct_velocity.c:16-41,
update_pressure.c:14-23).

In Thomas mode the fused kernel transforms the same row pairs (2q, 2q+1)
as the unfused DCT-III kernel, so whole steps equal the oracle's bit for
bit, with dt from the fused kernel's partial minima.  Row counts cover an
odd ny (a last pair of one row), strips that do not divide ny and the
periodic wrap of the first strip; row lengths 64 to 512.  In partition mode
the fused kernel pairs consecutive rows, which changes rounding only."""
import numpy as np
import pytest

import cases
from oracle import Oracle

pytestmark = pytest.mark.gpu

eif = pytest.importorskip("ellipsesinflows_amd")


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("nx,ny,strip", [(64, 48, 2), (64, 33, 4), (128, 36, 8), (256, 7, 4),
                                         (512, 64, 8), (64, 3, 2)])
def test_fused_projection_steps_bit_exact_thomas(nx, ny, strip):
    kw = dict(itot=nx, jtot=ny, ly=1.0, Re=150.0, extfrcy=0.2, uy_wall_xm=0.1, uy_wall_xp=-0.2)
    g = eif.Solver(poisson_mode=0, **kw)
    o = Oracle(**kw)
    g.set_projection(strip)
    ux, uy, p = cases.random_fields(nx, ny, nx + 7 * ny, amp=0.3)
    for s in (g, o):
        s.upload(ux, uy, p)
        s.update_boundaries()
    print("fused projection, Thomas", nx, ny, "strip", strip, flush=True)
    for k in range(4):
        dg, do = g.step(), o.step()
        assert dg == do, k
        a, b = g.download(), o.download()
        for f in ("ux", "uy", "p"):
            assert np.array_equal(a[f], b[f]), (k, f)
    g.close()


def test_fused_projection_with_ellipses_graph_and_eager():
    """Partition mode with ellipses: fused steps (graph steps and eager
    steps give the same bits) against the unfused kernels within rounding."""
    parts = cases.place_ellipses(12, 2.0, (0.03, 0.05), (0.5, 0.8), 0.03, 5)
    cfg = dict(itot=256, jtot=512, ly=2.0, Re=300.0, extfrcy=8.0 / 300.0)
    runs = {}
    for name, strip, graph in (("unfused", 0, False), ("fused", 4, False),
                               ("fused_graph", 4, True)):
        g = eif.Solver(poisson_mode=eif.POISSON_PARTITION, **cfg)
        xc = g.grid()["xc"]
        uy = np.zeros((512 + 2, 256 + 2))
        uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
        g.upload(uy=uy)
        g.update_boundaries()
        g.set_particles(**parts)
        g.set_projection(strip)
        if graph:
            g.graph(True)
        g.run_steps(8)
        runs[name] = (g.download(), g.particles())
        g.close()
    f, fg = runs["fused"], runs["fused_graph"]
    for k in ("ux", "uy", "p"):
        assert np.array_equal(f[0][k], fg[0][k]), k
    assert np.array_equal(f[1], fg[1])
    u = runs["unfused"]
    for k in ("ux", "uy"):
        assert rel(f[0][k], u[0][k]) < 1e-12, k
    pf, pu = f[0]["p"][1:-1, 1:-1], u[0]["p"][1:-1, 1:-1]
    assert rel(pf - pf.mean(), pu - pu.mean()) < 1e-10
    assert np.max(np.abs(f[1][:, 3:6] - u[1][:, 3:6])) < 1e-12


def test_set_projection_after_graph_capture_takes_effect():
    """Graph steps captured with the unfused kernels, then the fused
    projection set: the next graph steps use it (captured graphs are
    dropped), so they equal eager steps that switch at the same step.  In
    partition mode the two forms round differently, so steps that kept the
    old graph would differ."""
    parts = cases.place_ellipses(6, 2.0, (0.03, 0.05), (0.5, 0.8), 0.03, 9)
    cfg = dict(itot=256, jtot=512, ly=2.0, Re=300.0, extfrcy=8.0 / 300.0)
    runs = []
    for graph in (False, True):
        g = eif.Solver(poisson_mode=eif.POISSON_PARTITION, **cfg)
        xc = g.grid()["xc"]
        uy = np.zeros((512 + 2, 256 + 2))
        uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
        g.upload(uy=uy)
        g.update_boundaries()
        g.set_particles(**parts)
        if graph:
            g.graph(True)
        g.run_steps(3)
        g.set_projection(4)
        g.run_steps(3)
        runs.append((g.download(), g.particles()))
        g.close()
    (a, pa), (b, pb) = runs
    for k in ("ux", "uy", "p"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(pa, pb)


def test_set_projection_rejects_bad_arguments():
    g = eif.Solver(itot=48, jtot=16, ly=1.0, Re=100.0)
    with pytest.raises(eif.EIFError):
        g.set_projection(4)  # nx not a power of two
    g.set_projection(0)
    g.close()
    g = eif.Solver(itot=64, jtot=16, ly=1.0, Re=100.0)
    with pytest.raises(eif.EIFError):
        g.set_projection(3)
    g.close()
