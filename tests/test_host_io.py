"""Host side of the driver (ellipsesinflows_amd/host): npy save/restart
format, environment parameters, schedules and log formats.

Pinned against the reference's own fileio.c / simple_npyio.c / param/init.c
(built into oracle/_ref from the reference sources) where that library is
present, and against numpy's independent npy implementation and literal
expected text otherwise.  CPU only.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import ref_lib
from ellipsesinflows_amd import savefile
from ellipsesinflows_amd.savefile import Param, host_library

_PD = ctypes.POINTER(ctypes.c_double)

PARAM_VARS = ("dirname_restart", "timemax", "wtimemax", "log_rate", "log_after", "save_rate",
              "save_after", "stat_rate", "stat_after", "itot", "jtot", "ly", "safefactor_adv",
              "safefactor_dif", "safefactor_par", "Re", "Fr", "extfrcy")


@pytest.fixture(scope="module")
def hlib():
    return host_library()


@pytest.fixture(scope="module")
def rlib():
    lib = ref_lib()
    if lib is None:
        pytest.skip("reference sources not built (oracle/_ref)")
    lib.ref_param_values.argtypes = [_PD]
    lib.ref_w_field.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, _PD]
    lib.ref_r_field.argtypes = lib.ref_w_field.argtypes
    return lib


def _b(s):
    return os.fsencode(str(s))


def _p(a):
    return a.ctypes.data_as(_PD)


def _bytes(path):
    with open(path, "rb") as f:
        return f.read()


# -- npy format ---------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 5, 9, 10, 99, 100, 4097, 123456])
def test_npy_1d_bytes_match_reference_writer(hlib, rlib, tmp_path, n):
    data = np.random.default_rng(n).standard_normal(n)
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    assert hlib.eifh_w_1d(_b(tmp_path / "a"), b"x", n, _p(data)) == 0
    rlib.ref_w_1d(_b(tmp_path / "b"), b"x", n, _p(data))
    assert _bytes(tmp_path / "a" / "x.npy") == _bytes(tmp_path / "b" / "x.npy")


def test_npy_0d_bytes_match_reference_writer(hlib, rlib, tmp_path):
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    for v in (0, 7, -3, 2 ** 31 - 1):
        hlib.eifh_w_0d_int(_b(tmp_path / "a"), b"i", v)
        rlib.ref_w_0d_int(_b(tmp_path / "b"), b"i", v)
        assert _bytes(tmp_path / "a" / "i.npy") == _bytes(tmp_path / "b" / "i.npy")
    for v in (0.0, 1.0 / 3.0, -2.5e-300, 1e300):
        hlib.eifh_w_0d_double(_b(tmp_path / "a"), b"d", v)
        rlib.ref_w_0d_double(_b(tmp_path / "b"), b"d", v)
        assert _bytes(tmp_path / "a" / "d.npy") == _bytes(tmp_path / "b" / "d.npy")


@pytest.mark.parametrize("itot,jtot", [(4, 3), (16, 8), (33, 17)])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_field_files_match_reference_fileio(hlib, rlib, tmp_path, itot, jtot, kind):
    """src/fileio.c:283-365 writes rows 1..jtot of the halo-padded array
    as (jtot, itot+1) or (jtot, itot+2); both writers give the same bytes
    and each reader reads the other's file back exactly."""
    nc = itot + 1 if kind == 0 else itot + 2
    data = np.random.default_rng(itot * 7 + kind).standard_normal((jtot + 2, nc))
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    assert hlib.eifh_w_field(_b(tmp_path / "a"), b"f", jtot, nc, _p(data)) == 0
    rlib.ref_w_field(_b(tmp_path / "b"), b"f", kind, itot, jtot, _p(data))
    assert _bytes(tmp_path / "a" / "f.npy") == _bytes(tmp_path / "b" / "f.npy")
    back_ref = np.zeros_like(data)
    assert rlib.ref_r_field(_b(tmp_path / "a"), b"f", kind, itot, jtot, _p(back_ref)) == 0
    back_ours = np.zeros_like(data)
    assert hlib.eifh_r_field(_b(tmp_path / "b"), b"f", jtot, nc, _p(back_ours)) == 0
    np.testing.assert_array_equal(back_ref[1:jtot + 1], data[1:jtot + 1])
    np.testing.assert_array_equal(back_ours[1:jtot + 1], data[1:jtot + 1])
    assert not back_ours[0].any() and not back_ours[jtot + 1].any()


@pytest.mark.parametrize("shape", [(), (1,), (5,), (1000,), (3, 7), (64, 65)])
def test_npy_header_layout_and_numpy_roundtrip(hlib, tmp_path, shape):
    """The header is 64-byte aligned, ends in '\\n', and numpy's own reader
    returns the data."""
    data = np.ascontiguousarray(np.arange(int(np.prod(shape)), dtype=np.float64).reshape(shape) * 0.5 - 3.0)
    path = tmp_path / "x.npy"
    sh = (ctypes.c_size_t * max(len(shape), 1))(*shape)
    assert hlib.eifh_npy_write(_b(path), b"'float64'", len(shape), sh,
                               data.ctypes.data_as(ctypes.c_void_p), data.nbytes) == 0
    raw = _bytes(path)
    hlen = int.from_bytes(raw[8:10], "little")
    assert raw[:8] == b"\x93NUMPY\x01\x00"
    assert (10 + hlen) % 64 == 0 and raw[10 + hlen - 1:10 + hlen] == b"\n"
    dict_text = raw[10:10 + hlen].decode().rstrip()
    expect_shape = "(" + "".join("%d," % s for s in shape) + ")"
    assert dict_text == "{'descr':'float64','fortran_order':False,'shape':%s}" % expect_shape
    back = np.load(path, allow_pickle=False)
    assert back.shape == shape
    np.testing.assert_array_equal(back, data)


def test_reader_accepts_numpy_written_files(hlib, tmp_path):
    """np.save writes '<f8' / '<i8' with spaces in the dictionary."""
    a = np.random.default_rng(1).standard_normal((6, 9))
    np.save(tmp_path / "a.npy", a)
    out = np.zeros_like(a)
    sh = (ctypes.c_size_t * 2)(6, 9)
    assert hlib.eifh_npy_read_doubles(_b(tmp_path / "a.npy"), 2, sh, _p(out)) == 0
    np.testing.assert_array_equal(out, a)
    bad = (ctypes.c_size_t * 2)(9, 6)
    assert hlib.eifh_npy_read_doubles(_b(tmp_path / "a.npy"), 2, bad, _p(out)) != 0
    for dtype in (np.int32, np.int64):
        np.save(tmp_path / "n.npy", np.array(17, dtype=dtype))
        v = ctypes.c_int()
        assert hlib.eifh_npy_read_int(_b(tmp_path / "n.npy"), ctypes.byref(v)) == 0
        assert v.value == 17
    np.save(tmp_path / "t.npy", np.array(0.125))
    d = ctypes.c_double()
    assert hlib.eifh_npy_read_double(_b(tmp_path / "t.npy"), ctypes.byref(d)) == 0
    assert d.value == 0.125
    (tmp_path / "junk.npy").write_bytes(b"not an npy file at all")
    assert hlib.eifh_npy_read_double(_b(tmp_path / "junk.npy"), ctypes.byref(d)) != 0


# -- save / restart directory ----------------------------------------------------

def _particles(n, seed):
    rng = np.random.default_rng(seed)
    return {"den": rng.uniform(1, 2, n), "a": rng.uniform(.02, .04, n),
            "b": rng.uniform(.01, .02, n), "x": rng.uniform(.2, .8, n),
            "y": rng.uniform(0, 1, n), "az": rng.uniform(0, np.pi, n),
            "ux": rng.standard_normal(n), "uy": rng.standard_normal(n),
            "vz": rng.standard_normal(n)}


def test_save_directory_roundtrip(tmp_path):
    itot, jtot = 12, 10
    rng = np.random.default_rng(3)
    ux = rng.standard_normal((jtot + 2, itot + 1))
    uy = rng.standard_normal((jtot + 2, itot + 2))
    p = rng.standard_normal((jtot + 2, itot + 2))
    parts = _particles(4, 5)
    d = tmp_path / "step0000000042"
    savefile.write_save(d, itot=itot, jtot=jtot, ly=2.0, Re=500.0, ux=ux, uy=uy, p=p,
                        particles=parts, step=42, time=1.25)
    names = sorted(os.listdir(d))
    assert names == sorted(["step.npy", "itot.npy", "jtot.npy", "time.npy", "lx.npy", "ly.npy",
                            "Re.npy", "xf.npy", "xc.npy", "yf.npy", "yc.npy", "ux.npy", "uy.npy",
                            "p.npy", "n_particles.npy"] +
                           [f + ".npy" for f in savefile.PARTICLE_FILES])
    back = savefile.read_save(d, itot, jtot)
    assert (back["step"], back["itot"], back["jtot"]) == (42, itot, jtot)
    assert (back["time"], back["lx"], back["ly"], back["Re"]) == (1.25, 1.0, 2.0, 500.0)
    for k, a in (("ux", ux), ("uy", uy), ("p", p)):
        np.testing.assert_array_equal(back[k][1:jtot + 1], a[1:jtot + 1])
    for k in savefile.PARTICLE_KEYS:
        np.testing.assert_array_equal(back["particles"][k], parts[k])
    # numpy reads every file; shapes are the reference's
    assert np.load(d / "ux.npy").shape == (jtot, itot + 1)
    assert np.load(d / "p.npy").shape == (jtot, itot + 2)
    yf = np.load(d / "yf.npy")
    np.testing.assert_allclose(yf, np.arange(jtot + 1) * (2.0 / jtot), rtol=0, atol=1e-15)
    np.testing.assert_allclose(np.load(d / "yc.npy"), 0.5 * (yf[:-1] + yf[1:]), rtol=0, atol=0)


def test_save_directory_read_by_reference(rlib, tmp_path):
    """The reference's own loaders read a directory written here:
    suspensions_init (particles), fileio_r_*_like_parallel (fields) and
    param_init (step and time of a restart)."""
    itot, jtot = 16, 8
    rng = np.random.default_rng(9)
    ux = rng.standard_normal((jtot + 2, itot + 1))
    uy = rng.standard_normal((jtot + 2, itot + 2))
    p = rng.standard_normal((jtot + 2, itot + 2))
    parts = _particles(3, 11)
    d = tmp_path / "save"
    savefile.write_save(d, itot=itot, jtot=jtot, ly=1.0, Re=100.0, ux=ux, uy=uy, p=p,
                        particles=parts, step=7, time=0.375)
    rec = np.zeros((8, 30))
    n = rlib.ref_load_particles(_b(d), 8, _p(rec))
    assert n == 3
    for key, m in zip(savefile.PARTICLE_KEYS, savefile.PARTICLE_MEMBERS):
        np.testing.assert_array_equal(rec[:3, m], parts[key])
    for kind, name, a in ((0, b"ux", ux), (1, b"uy", uy), (2, b"p", p)):
        back = np.zeros_like(a)
        assert rlib.ref_r_field(_b(d), name, kind, itot, jtot, _p(back)) == 0
        np.testing.assert_array_equal(back[1:jtot + 1], a[1:jtot + 1])


# -- parameters --------------------------------------------------------------------

def _ours(hlib):
    par = Param()
    assert hlib.eifh_param_load(ctypes.byref(par), 0) == 0
    return par


def _as_ref_vector(par):
    return np.array([par.timemax, par.wtimemax, par.log.rate, par.log.after, par.log.next,
                     par.save.rate, par.save.after, par.save.next, par.stat.rate, par.stat.after,
                     par.stat.next, par.itot, par.jtot, par.lx, par.ly, par.safefactors[0],
                     par.safefactors[1], par.safefactors[2], par.Re, par.Fr, par.extfrcy,
                     par.step, par.time, par.load_flow_field])


ENV_CASES = [
    {},
    {"timemax": "2.5e1", "log_rate": "1.e-1", "log_after": "0.35", "save_rate": "5",
     "save_after": "10", "stat_rate": "0.2", "stat_after": "1e-3", "itot": "96", "jtot": "64",
     "ly": "2.0", "safefactor_adv": "0.5", "safefactor_dif": "0.4", "safefactor_par": "0.9",
     "Re": "1000", "Fr": "2.0", "extfrcy": "8e-3", "wtimemax": "30"},
    {"itot": "12abc", "jtot": " 24", "Re": "1e-300", "log_rate": "0.007", "ly": "junk"},
]


@pytest.mark.parametrize("env", ENV_CASES)
def test_param_env_matches_reference(hlib, rlib, monkeypatch, env):
    for k in PARAM_VARS:
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    ref = np.zeros(24)
    rlib.ref_param_values(_p(ref))
    np.testing.assert_array_equal(_as_ref_vector(_ours(hlib)), ref)


def test_param_restart_matches_reference(hlib, rlib, monkeypatch, tmp_path):
    """With dirname_restart, step and time come from the directory and the
    schedules are recomputed from that time (src/param/init.c:225-238)."""
    for k in PARAM_VARS:
        monkeypatch.delenv(k, raising=False)
    d = tmp_path / "r"
    d.mkdir()
    rlib.ref_w_0d_int(_b(d), b"step", 1234)
    rlib.ref_w_0d_double(_b(d), b"time", 20.48)
    monkeypatch.setenv("dirname_restart", str(d))
    monkeypatch.setenv("stat_rate", "0.1")
    monkeypatch.setenv("stat_after", "13.80")
    monkeypatch.setenv("log_rate", "0.1")
    monkeypatch.setenv("log_after", "20.21")
    monkeypatch.setenv("itot", "8")
    monkeypatch.setenv("jtot", "8")
    ref = np.zeros(24)
    rlib.ref_param_values(_p(ref))
    ours = _as_ref_vector(_ours(hlib))
    np.testing.assert_array_equal(ours, ref)
    assert ours[21] == 1234 and ours[22] == 20.48 and ours[23] == 1


def test_param_int_clamp(hlib, monkeypatch):
    """strtol results outside the int range are clamped (src/param/init.c:41-43);
    the reference itself would then fail to allocate, so only ours is run."""
    for k in PARAM_VARS:
        monkeypatch.delenv(k, raising=False)
    monkeypatch.setenv("itot", "99999999999")
    monkeypatch.setenv("jtot", "-99999999999")
    par = _ours(hlib)
    assert (par.itot, par.jtot) == (2 ** 31 - 1, -2 ** 31)


def test_next_schedule_examples(hlib):
    """The two worked examples of src/param/init.c:194-205."""
    assert abs(hlib.eifh_next_schedule(0.1, 20.21, 13.48) - 20.30) < 1e-12
    assert abs(hlib.eifh_next_schedule(0.1, 13.80, 20.48) - 20.50) < 1e-12
    # an exact multiple is rounded up to the next one
    assert hlib.eifh_next_schedule(1.0, 0.0, 3.0) == 4.0
    assert hlib.eifh_next_schedule(1.0, 0.0, 0.0) == 1.0


def test_param_defaults_and_options(hlib, monkeypatch):
    for k in PARAM_VARS + ("uy_wall_xm", "uy_wall_xp", "poisson_ysolve"):
        monkeypatch.delenv(k, raising=False)
    par = _ours(hlib)
    assert (par.itot, par.jtot, par.ly, par.Re, par.extfrcy) == (32, 32, 1.0, 2.334e3, 2.337e-4)
    assert par.Fr == np.finfo(float).max
    assert tuple(par.safefactors) == (0.75, 0.75, 0.95)
    assert (par.uy_wall_xm, par.uy_wall_xp, par.poisson_ysolve) == (0.0, 0.0, 1)
    monkeypatch.setenv("poisson_ysolve", "thomas")
    monkeypatch.setenv("uy_wall_xm", "-0.5")
    par = _ours(hlib)
    assert (par.uy_wall_xm, par.poisson_ysolve) == (-0.5, 0)


# -- logging -----------------------------------------------------------------------

@pytest.mark.parametrize("rate,digits", [(1.0, 1), (0.5, 1), (0.1, 2), (0.05, 2), (0.01, 3),
                                         (1e-3, 4), (2.0, 1), (1e-6, 7)])
def test_log_digits(hlib, rate, digits):
    """Digits of (int)(1/rate) (src/logging.c:12-19)."""
    assert hlib.eifh_log_digits(rate) == digits


def test_logging_lines(hlib, tmp_path):
    par = Param()
    par.step, par.time, par.dt = 3, 0.5, 1.0e-3
    par.log.rate = 0.1
    rec = np.zeros((2, 30))
    rec[0, [3, 4, 5, 9, 10, 11]] = [0.5, 0.25, 1.0, -1e-3, 2.0, 0.0]
    rec[1, [3, 4, 5, 9, 10, 11]] = [0.75, 1.5, 3.0, 0.0, -0.125, 1e-9]
    d = tmp_path / "log"
    d.mkdir()
    moms = np.array([1.0 / 3.0, -2.0])
    ener = np.array([0.125, 2.0 ** -60])
    assert hlib.eifh_logging(_b(d), ctypes.byref(par), 2.5e-14, -3.0e-17, _p(moms), _p(ener),
                             2, _p(rec)) == 0
    assert (d / "progress.dat").read_text() == "step        3, time  0.50, dt 1.00e-03\n"
    assert (d / "divergence.dat").read_text() == " 0.50  2.5e-14 -3.0e-17\n"
    assert (d / "momentum.dat").read_text() == \
        " 0.50  3.333333333333333e-01 -2.000000000000000e+00\n"
    assert (d / "energy.dat").read_text() == \
        " 0.50  1.250000000000000e-01  8.673617379884035e-19\n"
    assert (d / "particle0000000000.dat").read_text() == \
        " 0.50  5.0000000e-01  2.5000000e-01  1.0000000e+00 -1.0000000e-03  2.0000000e+00  0.0000000e+00\n"
    assert (d / "particle0000000001.dat").read_text() == \
        " 0.50  7.5000000e-01  1.5000000e+00  3.0000000e+00  0.0000000e+00 -1.2500000e-01  1.0000000e-09\n"
    # later steps append
    par.step, par.time = 4, 0.6
    hlib.eifh_logging(_b(d), ctypes.byref(par), 0.0, 0.0, _p(moms), _p(ener), 2, _p(rec))
    assert len((d / "progress.dat").read_text().splitlines()) == 2


# -- the driver binary -----------------------------------------------------------------

@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present: covered by GPU tests")
def test_driver_reports_missing_device(tmp_path):
    from ellipsesinflows_amd import build
    if not os.path.exists(build.EXE):
        pytest.skip("driver not built")
    from ellipsesinflows_amd import initparticles
    initparticles.write(tmp_path / "initparticles", _particles(2, 1))
    env = dict(os.environ, itot="16", jtot="16", timemax="0.01")
    r = subprocess.run([build.EXE], cwd=tmp_path, env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 1
    assert "no HIP device" in r.stderr
    assert "------- parameters are loaded -------" in r.stdout
    assert "  itot: 16\n" in r.stdout and "  Re: 2.334e+03 (default)\n" in r.stdout
