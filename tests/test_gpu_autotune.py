"""The autotuner (eif_gpu_autotune): the launch shapes it picks change no
result (predictor and projection rows) or only the rounding of the
partitioned y-solve (its chunk length, and the row pairing of the fused
projection), and a Thomas-mode context stays bit-identical to an untuned
one."""
import re

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

eif = pytest.importorskip("ellipsesinflows_amd")

KEYS = ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")


def _make(mode, parts):
    g = eif.Solver(poisson_mode=mode, itot=512, jtot=1024, ly=2.0, Re=500.0, extfrcy=8.0 / 500.0)
    xc = g.grid()["xc"]
    uy = np.zeros((1026, 514))
    uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
    g.upload(uy=uy)
    g.update_boundaries()
    g.set_particles(**parts)
    return g


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("mode", [eif.POISSON_THOMAS, eif.POISSON_PARTITION])
def test_autotuned_context_matches_untuned(mode):
    parts = cases.place_ellipses(32, 2.0, (0.03, 0.05), (0.5, 0.5), 0.01, 1)
    ref = _make(mode, parts)
    tuned = _make(mode, parts)
    rep = tuned.autotune()
    print("autotune:", rep)
    assert re.search(r"predictor_rows:.* -> (8|16|32)", rep)
    assert re.search(r"projection_rows:.* -> (4|8|16)", rep)
    assert re.search(r"projection:.* -> (unfused|2|4|8);", rep)
    assert re.search(r"fft_values_per_thread:.* -> (4|8|16)", rep)
    for _ in range(3):
        assert ref.step() == pytest.approx(tuned.step(), rel=1e-13)
    a, b = ref.download(), tuned.download()
    # the fused projection pairs rows as the Thomas-mode transform does:
    # in partition mode it changes rounding
    same_split = mode == eif.POISSON_THOMAS or ("partition 64" in rep and
                                                "-> unfused;" in rep)
    for k in ("ux", "uy"):
        if same_split:
            assert np.array_equal(a[k], b[k]), k
        else:
            assert _rel(b[k], a[k]) < 1e-12, k
    if same_split:
        assert np.array_equal(ref.particles(), tuned.particles())
    else:
        assert np.max(np.abs(ref.particles()[:, 3:6] - tuned.particles()[:, 3:6])) < 1e-12
    ref.close()
    tuned.close()


def test_tuned_shapes_copied_to_a_case_of_the_same_grid():
    """eif_gpu_tune_copy: a context given another's tuned shapes runs
    exactly as the tuned one; a different grid is refused."""
    parts = cases.place_ellipses(16, 2.0, (0.03, 0.05), (0.5, 0.5), 0.01, 2)
    src = _make(eif.POISSON_PARTITION, parts)
    dst = _make(eif.POISSON_PARTITION, parts)
    rep = src.autotune()
    print("autotune:", rep)
    dst.tune_copy(src)
    for _ in range(3):
        assert src.step() == dst.step()
    a, b = src.download(), dst.download()
    for k in ("ux", "uy", "p"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(src.particles(), dst.particles())
    other = eif.Solver(itot=256, jtot=512, ly=2.0, Re=500.0)
    with pytest.raises(eif.EIFError):
        other.tune_copy(src)
    for g in (src, dst, other):
        g.close()
