"""Save/restart directories in the reference's format, through the host
library (``libeif_host.so``, ellipsesinflows_amd/host/save.c).

A directory ``output/save/stepNNNNNNNNNN`` (src/save.c:13-122) holds one
npy file per dataset: 0-d ``step``, ``itot``, ``jtot``, ``n_particles``
(int32) and ``time``, ``lx``, ``ly``, ``Re`` (float64); 1-d ``xf``, ``xc``,
``yf``, ``yc`` and the nine ``particle_*`` arrays; 2-d ``ux`` (jtot, itot+1)
and ``uy``, ``p`` (jtot, itot+2) without their y-halo rows.  A directory
written here is a valid ``dirname_restart`` for the reference and for
``eif_gpu_main``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOSTLIB = os.path.join(HERE, "libeif_host.so")
NREC = 30
PARTICLE_FILES = ("particle_dens", "particle_as", "particle_bs", "particle_xs", "particle_ys",
                  "particle_azs", "particle_uxs", "particle_uys", "particle_vzs")
PARTICLE_KEYS = ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")
# particle_t member offsets of the saved quantities (include/suspensions.h:8-31)
PARTICLE_MEMBERS = (0, 1, 2, 3, 4, 5, 9, 10, 11)

_PD = ctypes.POINTER(ctypes.c_double)
_lib = None


class Schedule(ctypes.Structure):
    _fields_ = [("rate", ctypes.c_double), ("after", ctypes.c_double), ("next", ctypes.c_double)]


class Param(ctypes.Structure):
    """eifh_param (ellipsesinflows_amd/host/eif_host.h)."""
    _fields_ = [
        ("load_flow_field", ctypes.c_int), ("dirname_restart", ctypes.c_char * 4096),
        ("itot", ctypes.c_int), ("jtot", ctypes.c_int), ("lx", ctypes.c_double),
        ("ly", ctypes.c_double), ("safefactors", ctypes.c_double * 3), ("Re", ctypes.c_double),
        ("Fr", ctypes.c_double), ("extfrcy", ctypes.c_double), ("time", ctypes.c_double),
        ("dt", ctypes.c_double), ("step", ctypes.c_int), ("timemax", ctypes.c_double),
        ("wtimemax", ctypes.c_double), ("log", Schedule), ("save", Schedule), ("stat", Schedule),
        ("uy_wall_xm", ctypes.c_double), ("uy_wall_xp", ctypes.c_double),
        ("poisson_ysolve", ctypes.c_int),
    ]


def host_library():
    global _lib
    if _lib is None:
        if not os.path.exists(HOSTLIB):
            raise OSError("%s is not built (run ellipsesinflows_amd.build.build_host())" % HOSTLIB)
        lib = ctypes.CDLL(HOSTLIB)
        lib.eifh_next_schedule.restype = ctypes.c_double
        lib.eifh_next_schedule.argtypes = [ctypes.c_double] * 3
        lib.eifh_npy_write.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_size_t), ctypes.c_void_p,
                                       ctypes.c_size_t]
        lib.eifh_npy_read_doubles.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_size_t), _PD]
        lib.eifh_npy_read_int.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        lib.eifh_npy_read_double.argtypes = [ctypes.c_char_p, _PD]
        lib.eifh_w_0d_int.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        lib.eifh_w_0d_double.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double]
        lib.eifh_w_1d.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t, _PD]
        lib.eifh_w_field.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                     ctypes.c_int, _PD]
        lib.eifh_r_field.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                     ctypes.c_int, _PD]
        lib.eifh_save.argtypes = [ctypes.c_char_p, ctypes.POINTER(Param), _PD, _PD, _PD, _PD,
                                  _PD, ctypes.c_int, _PD]
        lib.eifh_load_fields.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _PD, _PD,
                                         _PD]
        lib.eifh_load_n_particles.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        lib.eifh_load_particles.argtypes = [ctypes.c_char_p, ctypes.c_int, _PD]
        lib.eifh_param_load.argtypes = [ctypes.POINTER(Param), ctypes.c_int]
        lib.eifh_logging.argtypes = [ctypes.c_char_p, ctypes.POINTER(Param), ctypes.c_double,
                                     ctypes.c_double, _PD, _PD, ctypes.c_int, _PD]
        lib.eifh_log_digits.argtypes = [ctypes.c_double]
        lib.eifh_mkdirs.argtypes = [ctypes.c_char_p]
        lib.eifh_stat_output.argtypes = [ctypes.c_char_p, ctypes.POINTER(Param), _PD, _PD,
                                         ctypes.c_int, _PD, _PD, _PD, _PD, _PD]
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(_PD)


def _b(s):
    return os.fsencode(s)


def grid_x(itot):
    """xf[itot+1], xc[itot+2] of src/param/init.c:125-148 (one process)."""
    dx = 1.0 / itot
    xf = np.array([1.0 * (i - 1) * dx for i in range(1, itot + 2)])
    xf[0], xf[itot] = 0.0, 1.0
    xc = np.empty(itot + 2)
    xc[0] = xf[0]
    xc[1:itot + 1] = 0.5 * (xf[:-1] + xf[1:])
    xc[itot + 1] = xf[itot]
    return xf, xc


def write_save(dirname, *, itot, jtot, ly, Re, ux, uy, p, particles, step=0, time=0.0):
    """Write a restart directory.  ux, uy, p are halo-padded arrays
    ((jtot+2) x (itot+1) / (itot+2)) or interior rows (jtot x ...);
    particles is a dict of the nine arrays (den, a, b, x, y, az, ux, uy, vz;
    missing velocities are zero) or an (n, 30) record array."""
    lib = host_library()
    par = Param()
    par.itot, par.jtot, par.lx, par.ly, par.Re = itot, jtot, 1.0, ly, Re
    par.step, par.time = int(step), float(time)
    padded = []
    for a, nc in ((ux, itot + 1), (uy, itot + 2), (p, itot + 2)):
        a = np.asarray(a, dtype=np.float64)
        if a.shape == (jtot, nc):
            full = np.zeros((jtot + 2, nc))
            full[1:jtot + 1] = a
            a = full
        if a.shape != (jtot + 2, nc):
            raise ValueError("field of shape %s, expected (%d, %d)" % (a.shape, jtot + 2, nc))
        padded.append(np.ascontiguousarray(a))
    if isinstance(particles, dict):
        n = len(particles["a"])
        rec = np.zeros((n, NREC))
        for key, m in zip(PARTICLE_KEYS, PARTICLE_MEMBERS):
            v = particles.get(key)
            if v is not None:
                rec[:, m] = v
    else:
        rec = np.ascontiguousarray(particles, dtype=np.float64).reshape(-1, NREC)
        n = rec.shape[0]
    xf, xc = grid_x(itot)
    if lib.eifh_save(_b(dirname), ctypes.byref(par), _p(xf), _p(xc), *[_p(a) for a in padded],
                     int(n), _p(rec if n else np.zeros((1, NREC)))):
        raise OSError("writing %s failed" % dirname)


def read_save(dirname, itot, jtot):
    """Read a restart directory: fields as halo-padded arrays (halo rows
    zero), scalars, and the particle arrays under "particles"."""
    lib = host_library()
    out = {"ux": np.zeros((jtot + 2, itot + 1)), "uy": np.zeros((jtot + 2, itot + 2)),
           "p": np.zeros((jtot + 2, itot + 2))}
    if lib.eifh_load_fields(_b(dirname), itot, jtot, _p(out["ux"]), _p(out["uy"]), _p(out["p"])):
        raise OSError("reading fields of %s failed" % dirname)
    n = ctypes.c_int()
    if lib.eifh_load_n_particles(_b(dirname), ctypes.byref(n)):
        raise OSError("reading n_particles of %s failed" % dirname)
    soa = np.zeros((9, max(n.value, 1)))
    if lib.eifh_load_particles(_b(dirname), n.value, _p(soa)):
        raise OSError("reading particles of %s failed" % dirname)
    out["particles"] = {key: soa[k, :n.value].copy() for k, key in enumerate(PARTICLE_KEYS)}
    out["n_particles"] = n.value
    iv = ctypes.c_int()
    dv = ctypes.c_double()
    for name in ("step", "itot", "jtot"):
        if lib.eifh_npy_read_int(_b(os.path.join(dirname, name + ".npy")), ctypes.byref(iv)):
            raise OSError("reading %s of %s failed" % (name, dirname))
        out[name] = iv.value
    for name in ("time", "lx", "ly", "Re"):
        if lib.eifh_npy_read_double(_b(os.path.join(dirname, name + ".npy")), ctypes.byref(dv)):
            raise OSError("reading %s of %s failed" % (name, dirname))
        out[name] = dv.value
    return out
