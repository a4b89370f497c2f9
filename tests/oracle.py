"""ctypes view of oracle/liboracle.so (the CPU restatement, test checker
only) and of oracle/_ref/libeifref.so (the reference's own compiled
sources, where they build), mirroring ellipsesinflows_amd.Solver so the
parity tests call both the same way."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libeifref.so")

_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p
NREC = 30

FIELDS = {"ux": 0, "uy": 1, "p": 2, "psi": 3, "srcuxa": 4, "srcuxb": 5, "srcuxg": 6,
          "srcuya": 7, "srcuyb": 8, "srcuyg": 9, "dux": 10, "duy": 11}
GRIDS = {"xf": 0, "xc": 1, "dxf": 2, "dxc": 3, "yf": 4, "yc": 5}

_olib = None


def ensure_built():
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, stdout=subprocess.DEVNULL)


def oracle_lib():
    global _olib
    if _olib is None:
        ensure_built()
        lib = ctypes.CDLL(ORACLE_LIB)
        lib.orc_create.restype = _VP
        lib.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D, _D, _PD]
        lib.orc_destroy.argtypes = [_VP]
        lib.orc_field.restype = _PD
        lib.orc_field.argtypes = [_VP, ctypes.c_int]
        lib.orc_field_len.restype = ctypes.c_long
        lib.orc_field_len.argtypes = [_VP, ctypes.c_int]
        lib.orc_grid.restype = _PD
        lib.orc_grid.argtypes = [_VP, ctypes.c_int]
        lib.orc_get_dt.restype = _D
        lib.orc_get_dt.argtypes = [_VP]
        lib.orc_set_dt.argtypes = [_VP, _D]
        lib.orc_get_time.restype = _D
        for name in ("orc_find_normal_t", "orc_s_weight", "orc_v_weight"):
            getattr(lib, name).restype = _D
        lib.orc_find_normal_t.argtypes = [_D] * 4
        lib.orc_s_weight.argtypes = [_D] * 8
        lib.orc_v_weight.argtypes = [_D] * 8
        lib.orc_increment_particles.argtypes = [_VP, ctypes.c_int, _PD]
        lib.orc_collide_pair.argtypes = [_D, ctypes.c_int, _PD, _PD]
        lib.orc_collide_wall.argtypes = [_D, ctypes.c_int, _PD]
        lib.orc_tdm_periodic.argtypes = [ctypes.c_int, _PD, _PD, _PD, _PD]
        lib.orc_dct2.argtypes = [ctypes.c_int, _PD]
        lib.orc_dct3.argtypes = [ctypes.c_int, _PD]
        lib.orc_dct2_pair.argtypes = [ctypes.c_int, _PD, _PD]
        lib.orc_dct3_pair.argtypes = [ctypes.c_int, _PD, _PD]
        lib.orc_set_threads.argtypes = [ctypes.c_int]
        lib.orc_decide_loop_size.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, _PI, _PI]
        _olib = lib
    return _olib


def _ptr(a):
    return a.ctypes.data_as(_PD)


class Oracle:
    """CPU restatement of the reference time step, one process."""

    def __init__(self, itot, jtot, ly=1.0, Re=2.334e3, Fr=np.finfo(float).max,
                 extfrcy=2.337e-4, uy_wall_xm=0.0, uy_wall_xp=0.0,
                 safefactors=(0.75, 0.75, 0.95), threads=None):
        self.lib = oracle_lib()
        if threads:
            self.lib.orc_set_threads(int(threads))
        sf = (_D * 3)(*safefactors)
        self.h = _VP(self.lib.orc_create(int(itot), int(jtot), ly, Re, Fr, extfrcy,
                                         uy_wall_xm, uy_wall_xp, sf))
        if not self.h.value:
            raise ValueError("bad oracle configuration")
        self.itot, self.jtot, self.ly = int(itot), int(jtot), float(ly)
        self.n_particles = 0

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shapes(self):
        nx, ny = self.itot, self.jtot
        return {"ux": (ny + 2, nx + 1), "uy": (ny + 2, nx + 2), "p": (ny + 2, nx + 2),
                "psi": (ny + 2, nx + 2), "srcuxa": (ny, nx - 1), "srcuxb": (ny, nx - 1),
                "srcuxg": (ny, nx - 1), "srcuya": (ny, nx), "srcuyb": (ny, nx),
                "srcuyg": (ny, nx), "dux": (ny + 2, nx + 2), "duy": (ny + 2, nx + 2)}

    def view(self, name):
        k = FIELDS[name]
        n = self.lib.orc_field_len(self.h, k)
        return np.ctypeslib.as_array(self.lib.orc_field(self.h, k), shape=(n,)).reshape(self.shapes[name])

    def field(self, name):
        return self.view(name).copy()

    def grid(self):
        nx, ny = self.itot, self.jtot
        lens = {"xf": nx + 1, "xc": nx + 2, "dxf": nx, "dxc": nx + 1, "yf": ny + 1, "yc": ny}
        return {k: np.ctypeslib.as_array(self.lib.orc_grid(self.h, v), shape=(lens[k],)).copy()
                for k, v in GRIDS.items()}

    def upload(self, ux=None, uy=None, p=None):
        for name, a in (("ux", ux), ("uy", uy), ("p", p)):
            if a is not None:
                self.view(name)[...] = np.asarray(a, dtype=np.float64).reshape(self.shapes[name])

    def download(self):
        return {k: self.field(k) for k in ("ux", "uy", "p")}

    def set_field(self, name, arr):
        self.view(name)[...] = np.asarray(arr, dtype=np.float64).reshape(self.shapes[name])

    def update_boundaries(self):
        for k, f in (("ux", self.lib.orc_update_boundaries_ux), ("uy", self.lib.orc_update_boundaries_uy),
                     ("p", self.lib.orc_update_boundaries_p)):
            f.argtypes = [_VP, _PD]
            f(self.h, self.lib.orc_field(self.h, FIELDS[k]))

    def decide_dt(self):
        self.lib.orc_decide_dt(self.h)
        return self.lib.orc_get_dt(self.h)

    def set_dt(self, dt):
        self.lib.orc_set_dt(self.h, float(dt))

    def get_dt(self):
        return self.lib.orc_get_dt(self.h)

    def update_velocity(self, rk):
        self.lib.orc_update_velocity(self.h, int(rk))

    def compute_potential(self, rk):
        self.lib.orc_compute_potential(self.h, int(rk))

    def correct_velocity(self, rk):
        self.lib.orc_correct_velocity(self.h, int(rk))

    def update_pressure(self):
        self.lib.orc_update_pressure(self.h)

    def divergence(self):
        a, b = _D(), _D()
        self.lib.orc_divergence(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def momentum(self):
        o = np.zeros(2)
        self.lib.orc_momentum(self.h, _ptr(o))
        return o

    def energy(self):
        o = np.zeros(2)
        self.lib.orc_energy(self.h, _ptr(o))
        return o

    def set_particles(self, den, a, b, x, y, az, ux=None, uy=None, vz=None):
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (den, a, b, x, y, az)]
        n = arrs[0].size
        for v in (ux, uy, vz):
            arrs.append(np.zeros(n) if v is None else np.ascontiguousarray(v, dtype=np.float64))
        self.lib.orc_set_particles.argtypes = [_VP, ctypes.c_int] + [_PD] * 9
        self.lib.orc_set_particles(self.h, n, *[_ptr(v) for v in arrs])
        self.n_particles = n

    def particles(self):
        out = np.zeros((self.n_particles, NREC))
        if self.n_particles:
            self.lib.orc_get_particle_records.argtypes = [_VP, _PD]
            self.lib.orc_get_particle_records(self.h, _ptr(out))
        return out

    def set_particle_records(self, rec):
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        self.lib.orc_set_particle_records.argtypes = [_VP, _PD]
        self.lib.orc_set_particle_records(self.h, _ptr(rec))

    def reset_particle_increments(self):
        self.lib.orc_reset_particle_increments(self.h)

    def compute_inertia(self, cnstep):
        self.lib.orc_compute_inertia(self.h, int(cnstep))

    def exchange_momentum(self):
        self.lib.orc_exchange_momentum(self.h)

    def update_momentum_field(self):
        self.lib.orc_update_momentum_field(self.h)

    def compute_collision_force(self, cnstep):
        self.lib.orc_compute_collision_force(self.h, int(cnstep))

    def increment_particles(self, rk):
        r = _D()
        self.lib.orc_increment_particles(self.h, int(rk), ctypes.byref(r))
        return r.value

    def update_particles(self):
        self.lib.orc_update_particles(self.h)

    def colliding_pairs(self):
        n = self.lib.orc_n_colliding(self.h)
        buf = np.zeros(2 * max(n, 1), dtype=np.int32)
        self.lib.orc_get_colliding.argtypes = [_VP, _PI]
        self.lib.orc_get_colliding(self.h, buf.ctypes.data_as(_PI))
        pairs = [tuple(int(v) for v in buf[2 * k:2 * k + 2]) for k in range(n)]
        pp = sorted(p for p in pairs if p[1] >= 0)
        walls = sorted([p for p in pairs if p[1] < 0], key=lambda p: (-p[1], p[0]))
        return pp + walls

    def integrate(self):
        self.lib.orc_integrate(self.h)
        sub = (ctypes.c_int * 3)()
        self.lib.orc_last_substeps(self.h, sub)
        return list(sub)

    def step(self):
        self.lib.orc_step(self.h)
        return self.lib.orc_get_dt(self.h)

    def last_substeps(self):
        sub = (ctypes.c_int * 3)()
        self.lib.orc_last_substeps(self.h, sub)
        return list(sub)

    def time(self):
        return self.lib.orc_get_time(self.h)

    def statistics_collect(self):
        if not hasattr(self, "_stats"):
            self._stats = {k: np.zeros(self.shapes["ux" if k[:2] == "ux" else "uy"])
                           for k in ("ux1", "ux2", "uy1", "uy2", "phi")}
            self._stats["num"] = 0
        st = self._stats
        self.lib.orc_statistics_collect(self.h, *[_ptr(st[k]) for k in ("ux1", "ux2", "uy1",
                                                                       "uy2", "phi")])
        st["num"] += 1

    def statistics(self):
        return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self._stats.items()}


def ref_lib():
    """The reference's own sources built by `make -C oracle ref`, or None."""
    if not os.path.exists(REF_LIB):
        if not os.path.isdir("/root/reference"):
            return None
        r = subprocess.run(["make", "-C", ORACLE_DIR, "ref"], stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
        if r.returncode != 0 or not os.path.exists(REF_LIB):
            return None
    lib = ctypes.CDLL(REF_LIB)
    for name in ("ref_find_normal_t", "ref_compute_ex", "ref_compute_ey", "ref_curvature"):
        getattr(lib, name).restype = _D
        getattr(lib, name).argtypes = [_D] * (4 if name == "ref_find_normal_t" else 3)
    for name in ("ref_s_weight", "ref_v_weight"):
        getattr(lib, name).restype = _D
        getattr(lib, name).argtypes = [_D] * 8
    for name in ("ref_mass", "ref_moment_of_inertia"):
        getattr(lib, name).restype = _D
        getattr(lib, name).argtypes = [_D] * 3
    lib.ref_decide_loop_size.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, _PI, _PI]
    lib.ref_collision.argtypes = [ctypes.c_int, _D, _D, ctypes.c_int, _PD]
    lib.ref_reset.argtypes = [ctypes.c_int, _PD]
    lib.ref_param_grid.argtypes = [ctypes.c_int, ctypes.c_int, _D] + [_PD] * 8
    lib.ref_w_1d.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _PD]
    lib.ref_w_0d_int.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    lib.ref_w_0d_double.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _D]
    lib.ref_r_1d.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _PD]
    lib.ref_r_0d_int.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _PI]
    lib.ref_load_particles.argtypes = [ctypes.c_char_p, ctypes.c_int, _PD]
    return lib
