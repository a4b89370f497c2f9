"""Pins the oracle against the reference's OWN compiled sources
(oracle/_ref, built from NaokiHori/EllipsesInFlows src/ellipse.c,
src/suspensions/{auxiliary,collision,reset,init}.c, src/param/init.c,
src/fileio.c, src/simple_npyio.c).  Skipped where the reference's sources
are not available; the restatement is compared bit for bit."""
import ctypes
import os
import tempfile

import numpy as np
import pytest

import cases
from oracle import Oracle, oracle_lib, ref_lib, NREC

_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)

REF = ref_lib()
pytestmark = pytest.mark.skipif(REF is None, reason="reference sources not buildable here")


def _bits(x):
    return np.asarray(x, dtype=np.float64).view(np.uint64)


def test_record_layout_matches_particle_t():
    assert REF.ref_record_size() == NREC


def test_find_normal_t_bit_exact():
    o = oracle_lib()
    rng = np.random.default_rng(11)
    pts = []
    for _ in range(4000):
        a = rng.uniform(0.005, 0.2)
        b = a * rng.uniform(0.25, 1.0)
        x, y = rng.uniform(-2.5 * a, 2.5 * a, 2)
        pts.append((a, b, x, y))
    # axes, origin neighbourhood, circles, boundary points
    for a, b in ((0.15, 0.075), (0.1, 0.1), (0.012, 0.005)):
        for x, y in ((0.0, 0.03), (0.04, 0.0), (-0.02, 0.0), (0.0, -0.01), (1e-9, 1e-9),
                     (a, 0.0), (0.0, b), (a * np.cos(0.3), b * np.sin(0.3))):
            pts.append((a, b, x, y))
    got = np.array([o.orc_find_normal_t(*p) for p in pts])
    ref = np.array([REF.ref_find_normal_t(*p) for p in pts])
    assert np.array_equal(_bits(got), _bits(ref))


def test_weights_and_boxes_bit_exact():
    o = oracle_lib()
    rng = np.random.default_rng(12)
    for _ in range(1500):
        a = rng.uniform(0.01, 0.1)
        b = a * rng.uniform(0.3, 1.0)
        px, py = rng.uniform(0.2, 0.8, 2)
        az = rng.uniform(0, 2 * np.pi)
        h = 1.0 / rng.choice([64, 128, 256])
        x, y = px + rng.uniform(-2 * a, 2 * a), py + rng.uniform(-2 * a, 2 * a)
        args = (h, a, b, px, py, az, x, y)
        assert _bits(o.orc_s_weight(*args)) == _bits(REF.ref_s_weight(*args))
        assert _bits(o.orc_v_weight(*args)) == _bits(REF.ref_v_weight(*args))
        m1, m2, r1, r2 = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        o.orc_decide_loop_size(1, 256, h, max(a, b), px, ctypes.byref(m1), ctypes.byref(m2))
        REF.ref_decide_loop_size(1, 256, h, max(a, b), px, ctypes.byref(r1), ctypes.byref(r2))
        assert (m1.value, m2.value) == (r1.value, r2.value)


def test_mass_and_inertia_formula():
    # A23: V = pi a b, m = rho V, I = m (a^2 + b^2) / 4 (auxiliary.c:63-75)
    rng = np.random.default_rng(13)
    for _ in range(200):
        den, a, b = rng.uniform(0.5, 2), rng.uniform(0.01, 0.2), rng.uniform(0.01, 0.2)
        m = REF.ref_mass(den, a, b)
        assert m == den * (3.1415926535897932 * a * b)
        assert REF.ref_moment_of_inertia(den, a, b) == 0.25 * m * (a * a + b * b)


def _records(parts, rng=None, with_increments=False):
    n = len(parts["a"])
    rec = np.zeros((n, NREC))
    for k, name in enumerate(("den", "a", "b", "x", "y", "az")):
        rec[:, k] = parts[name]
    if with_increments:
        rec[:, 6:9] = rng.uniform(-1e-3, 1e-3, size=(n, 3))
        rec[:, 9:12] = rng.uniform(-0.1, 0.1, size=(n, 3))
    return rec


@pytest.mark.parametrize("cnstep", [0, 1])
def test_collision_bit_exact(cnstep):
    rng = np.random.default_rng(14)
    for trial in range(4):
        if trial == 0:
            parts = cases.touching_pairs(1.0)
            ly = 1.0
        else:
            ly = 1.0
            parts = cases.contact_pairs(8, ly, seed=100 + trial)
        rec = _records(parts, rng, with_increments=trial > 1)
        ref = rec.copy()
        REF.ref_collision(len(rec), 1.0, ly, cnstep, ref.ctypes.data_as(_PD))
        o = Oracle(16, 16, ly=ly)
        o.set_particles(**{k: parts[k] for k in ("den", "a", "b", "x", "y", "az")})
        o.set_particle_records(rec)
        o.compute_collision_force(cnstep)
        got = o.particles()
        assert np.array_equal(_bits(got), _bits(ref)), trial
        if trial == 0:
            assert o.colliding_pairs()[:2] == [(0, 1), (2, 3)]


@pytest.mark.parametrize("itot,jtot,ly", [(128, 128, 1.0), (48, 96, 2.0), (2048, 4096, 2.0),
                                          (256, 512, 2.0), (37, 11, 0.7)])
def test_grid_and_rk_coefficients_bit_exact(itot, jtot, ly):
    arr = [np.zeros(n) for n in (itot + 1, itot + 2, itot, itot + 1, jtot + 1, jtot)]
    rk = np.zeros(9)
    dxdy = np.zeros(2)
    REF.ref_param_grid(itot, jtot, ly, *[a.ctypes.data_as(_PD) for a in arr],
                       rk.ctypes.data_as(_PD), dxdy.ctypes.data_as(_PD))
    g = Oracle(itot, jtot, ly=ly).grid()
    for a, k in zip(arr, ("xf", "xc", "dxf", "dxc", "yf", "yc")):
        assert np.array_equal(_bits(a), _bits(g[k])), k
    assert list(rk) == [32 / 60, 25 / 60, 45 / 60, 0 / 60, -17 / 60, -25 / 60,
                        32 / 60, 25 / 60 - 17 / 60, 45 / 60 - 25 / 60]


def test_reset_matches():
    rng = np.random.default_rng(15)
    rec = rng.standard_normal((5, NREC))
    ref = rec.copy()
    REF.ref_reset(5, ref.ctypes.data_as(_PD))
    expect = rec.copy()
    expect[:, 6:9] = 0.0
    expect[:, 12:15] = 0.0
    assert np.array_equal(ref, expect)


def test_reference_initparticles_fixture():
    """The reference ships initparticles/*.npy (one ellipse a=.15, b=.10 at
    (.4, 0), copied to tests/golden/initparticles); its own loader
    (src/suspensions/init.c) and numpy agree."""
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "initparticles")
    buf = np.zeros((4, NREC))
    n = REF.ref_load_particles(d.encode(), 4, buf.ctypes.data_as(_PD))
    assert n == 1
    assert buf[0, 1] == 0.15 and buf[0, 2] == 0.10 and buf[0, 3] == 0.4 and buf[0, 4] == 0.0
    assert np.load(os.path.join(d, "particle_as.npy"))[0] == 0.15
