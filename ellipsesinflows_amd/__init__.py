"""MI355X (gfx950) EllipsesInFlows time-step path.

Python view of the C-ABI in ``include/eif_gpu.h`` (``libeif_gpu.so``).  The
functions keep the reference's names and argument meaning
(NaokiHori/EllipsesInFlows include/fluid.h:27-38, include/suspensions.h:44-63,
include/param.h:55-58, src/main.c:13-76); arrays are numpy float64 in the
reference's halo-padded layouts.

The compute path is the HIP library only: importing works without a GPU, but
creating a :class:`Solver` raises if the library is missing or no device is
visible.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

__all__ = ["Solver", "EIFError", "load_library", "PARTICLE_FIELDS", "library_path"]

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libeif_gpu.so")

# member order of particle_t (include/suspensions.h:8-31)
PARTICLE_FIELDS = (
    "den", "a", "b", "x", "y", "az", "dx", "dy", "daz", "ux", "uy", "vz",
    "dux", "duy", "dvz", "fux", "fuy", "tvz", "iux0", "iux1", "iuy0", "iuy1",
    "ivz0", "ivz1", "cfx0", "cfx1", "cfy0", "cfy1", "ctz0", "ctz1",
)
NREC = len(PARTICLE_FIELDS)

POISSON_THOMAS = 0
POISSON_PARTITION = 1

FIELD_PSI, FIELD_SRCUXA, FIELD_SRCUYA, FIELD_DUX, FIELD_DUY = range(5)

# every symbol declared in include/eif_gpu.h
EXPORTED = (
    "fluid_gpu_create", "fluid_gpu_destroy", "fluid_gpu_get_grid", "fluid_gpu_upload",
    "fluid_gpu_download", "fluid_gpu_download_field", "fluid_gpu_upload_field",
    "fluid_gpu_update_boundaries_ux", "fluid_gpu_update_boundaries_uy",
    "fluid_gpu_update_boundaries_p", "fluid_gpu_decide_dt", "fluid_gpu_set_dt",
    "fluid_gpu_get_dt", "fluid_gpu_update_velocity", "poisson_gpu_compute_potential",
    "poisson_gpu_partition_layout",
    "fluid_gpu_correct_velocity", "fluid_gpu_update_pressure", "fluid_gpu_divergence",
    "fluid_gpu_momentum", "fluid_gpu_energy", "fluid_gpu_diagnostics_noalloc",
    "suspensions_gpu_create",
    "suspensions_gpu_n_particles", "suspensions_gpu_download", "suspensions_gpu_upload",
    "suspensions_gpu_reset_particle_increments", "suspensions_gpu_compute_inertia",
    "suspensions_gpu_exchange_momentum", "suspensions_gpu_update_momentum_fleid",
    "suspensions_gpu_compute_collision_force", "suspensions_gpu_increment_particles",
    "suspensions_gpu_update_particles", "suspensions_gpu_colliding_pairs",
    "eif_gpu_integrate", "eif_gpu_step", "eif_gpu_synchronize", "eif_gpu_last_error",
    "eif_gpu_run_steps", "eif_gpu_graph_enable", "eif_gpu_graph_stats",
    "eif_gpu_graph_set_subiterations", "eif_gpu_ensemble_run_steps", "eif_gpu_set_projection",
    "eif_gpu_set_collision_search",
    "eif_gpu_tune_copy",
    "eif_gpu_tape_begin", "eif_gpu_step_fixed_dt", "eif_gpu_stage_adjoint",
    "eif_gpu_adjoint_get", "eif_gpu_adjoint_add", "fluid_gpu_wall_shear",
    "eif_gpu_adjoint_add_wall_shear", "eif_gpu_autotune",
    "eif_gpu_ibm_stats", "eif_gpu_ibm_reference_iterations", "statistics_gpu_collect", "statistics_gpu_reset",
    "statistics_gpu_download",
    "eif_gpu_profile_enable", "eif_gpu_profile_select", "eif_gpu_profile_reset",
    "eif_gpu_profile_query", "eif_gpu_profile_sample",
    "eif_gpu_profile_names",
)


class EIFError(RuntimeError):
    pass


def partition_layout(jtot):
    """(chunk rows, chunks) of the partitioned y-solve for jtot, or (0, 0)
    where the Thomas sweep is used instead (host only, no device)."""
    lib = load_library()
    L, P = ctypes.c_int(), ctypes.c_int()
    lib.poisson_gpu_partition_layout(int(jtot), ctypes.byref(L), ctypes.byref(P))
    return L.value, P.value


ENSEMBLE_SEQUENTIAL, ENSEMBLE_STREAMS, ENSEMBLE_ONE_GRAPH, ENSEMBLE_BATCHED = 0, 1, 2, 3
COLLISION_AUTO, COLLISION_ALL_PAIRS, COLLISION_CELLS = 0, 1, 2


def ensemble_run_steps(solvers, nsteps, concurrent=True):
    """Advance independent Solvers by nsteps each: one after the other
    through the plain path (0/False), each step of every case one graph
    launch on its own stream (1/True), all cases' step graphs nested in one
    graph (2), or batched launches, each kernel of a step launched once for
    all cases (3)."""
    lib = load_library()
    arr = (_VP * max(len(solvers), 1))(*[s.h for s in solvers])
    rc = lib.eif_gpu_ensemble_run_steps(arr, len(solvers), int(nsteps), int(concurrent))
    if rc != 0:
        raise EIFError("eif_gpu error %d: %s" % (rc, lib.eif_gpu_last_error().decode()))


class _Config(ctypes.Structure):
    _fields_ = [
        ("itot", ctypes.c_int), ("jtot", ctypes.c_int), ("ly", ctypes.c_double),
        ("Re", ctypes.c_double), ("Fr", ctypes.c_double), ("extfrcy", ctypes.c_double),
        ("uy_wall_xm", ctypes.c_double), ("uy_wall_xp", ctypes.c_double),
        ("safefactors", ctypes.c_double * 3), ("poisson_mode", ctypes.c_int),
    ]


_lib = None
_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p


def load_library(path=None):
    """Load libeif_gpu.so (no device is touched)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or library_path
    if not os.path.exists(p):
        raise EIFError("libeif_gpu.so not built (%s); run __graft_entry__.build()" % p)
    lib = ctypes.CDLL(p)
    lib.eif_gpu_last_error.restype = ctypes.c_char_p
    lib.eif_gpu_profile_names.restype = ctypes.c_char_p
    lib.fluid_gpu_create.argtypes = [ctypes.POINTER(_Config), ctypes.POINTER(_VP)]
    lib.suspensions_gpu_n_particles.argtypes = [_VP]
    for name in EXPORTED:
        f = getattr(lib, name)
        if name not in ("eif_gpu_last_error", "eif_gpu_profile_names"):
            f.restype = ctypes.c_int
    lib.eif_gpu_profile_query.argtypes = [_VP, ctypes.c_char_p,
                                          ctypes.POINTER(ctypes.c_long), _PD]
    if path is None:
        _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(_PD)


class Solver:
    """One simulation context on the GPU (fluid + suspensions)."""

    def __init__(self, itot, jtot, ly=1.0, Re=2.334e3, Fr=np.finfo(float).max,
                 extfrcy=2.337e-4, uy_wall_xm=0.0, uy_wall_xp=0.0,
                 safefactors=(0.75, 0.75, 0.95), poisson_mode=POISSON_THOMAS):
        self.lib = load_library()
        cfg = _Config()
        cfg.itot, cfg.jtot, cfg.ly = int(itot), int(jtot), float(ly)
        cfg.Re, cfg.Fr, cfg.extfrcy = float(Re), float(Fr), float(extfrcy)
        cfg.uy_wall_xm, cfg.uy_wall_xp = float(uy_wall_xm), float(uy_wall_xp)
        for k in range(3):
            cfg.safefactors[k] = float(safefactors[k])
        cfg.poisson_mode = int(poisson_mode)
        h = _VP()
        self._check(self.lib.fluid_gpu_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.itot, self.jtot, self.ly = int(itot), int(jtot), float(ly)
        self.n_particles = 0

    # -- helpers ----------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise EIFError("eif_gpu error %d: %s" % (rc, self.lib.eif_gpu_last_error().decode()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.fluid_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shapes(self):
        nx, ny = self.itot, self.jtot
        return {"ux": (ny + 2, nx + 1), "uy": (ny + 2, nx + 2), "p": (ny + 2, nx + 2),
                "psi": (ny + 2, nx + 2), "srcuxa": (ny, nx - 1), "srcuya": (ny, nx),
                "dux": (ny + 2, nx + 2), "duy": (ny + 2, nx + 2)}

    # -- grid and fields --------------------------------------------------
    def grid(self):
        nx, ny = self.itot, self.jtot
        arrs = [np.zeros(n) for n in (nx + 1, nx + 2, nx, nx + 1, ny + 1, ny)]
        self._check(self.lib.fluid_gpu_get_grid(self.h, *[_ptr(a) for a in arrs]))
        return dict(zip(("xf", "xc", "dxf", "dxc", "yf", "yc"), arrs))

    def upload(self, ux=None, uy=None, p=None):
        arrs = []
        for name, a in (("ux", ux), ("uy", uy), ("p", p)):
            if a is None:
                arrs.append(None)
                continue
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.size != int(np.prod(self.shapes[name])):
                raise ValueError("%s has %d values, expected shape %s" % (name, a.size, self.shapes[name]))
            arrs.append(a)
        self._check(self.lib.fluid_gpu_upload(self.h, *[None if a is None else _ptr(a) for a in arrs]))

    def download(self):
        out = {k: np.empty(self.shapes[k]) for k in ("ux", "uy", "p")}
        self._check(self.lib.fluid_gpu_download(self.h, _ptr(out["ux"]), _ptr(out["uy"]), _ptr(out["p"])))
        return out

    def field(self, which):
        sel = {"psi": FIELD_PSI, "srcuxa": FIELD_SRCUXA, "srcuya": FIELD_SRCUYA,
               "dux": FIELD_DUX, "duy": FIELD_DUY}[which]
        out = np.empty(self.shapes[which])
        self._check(self.lib.fluid_gpu_download_field(self.h, sel, _ptr(out)))
        return out

    def set_field(self, which, arr):
        sel = {"psi": FIELD_PSI, "srcuxa": FIELD_SRCUXA, "srcuya": FIELD_SRCUYA}[which]
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != int(np.prod(self.shapes[which])):
            raise ValueError("bad size for %s" % which)
        self._check(self.lib.fluid_gpu_upload_field(self.h, sel, _ptr(a)))

    # -- the reference's functions -----------------------------------------
    def update_boundaries(self):
        self._check(self.lib.fluid_gpu_update_boundaries_ux(self.h))
        self._check(self.lib.fluid_gpu_update_boundaries_uy(self.h))
        self._check(self.lib.fluid_gpu_update_boundaries_p(self.h))

    def decide_dt(self):
        dt = _D()
        self._check(self.lib.fluid_gpu_decide_dt(self.h, ctypes.byref(dt)))
        return dt.value

    def set_dt(self, dt):
        self._check(self.lib.fluid_gpu_set_dt(self.h, _D(dt)))

    def get_dt(self):
        dt = _D()
        self._check(self.lib.fluid_gpu_get_dt(self.h, ctypes.byref(dt)))
        return dt.value

    def update_velocity(self, rk):
        self._check(self.lib.fluid_gpu_update_velocity(self.h, int(rk)))

    def compute_potential(self, rk):
        self._check(self.lib.poisson_gpu_compute_potential(self.h, int(rk)))

    def correct_velocity(self, rk):
        self._check(self.lib.fluid_gpu_correct_velocity(self.h, int(rk)))

    def update_pressure(self):
        self._check(self.lib.fluid_gpu_update_pressure(self.h))

    def divergence(self):
        a, b = _D(), _D()
        self._check(self.lib.fluid_gpu_divergence(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def momentum(self):
        o = np.zeros(2)
        self._check(self.lib.fluid_gpu_momentum(self.h, _ptr(o)))
        return o

    def energy(self):
        o = np.zeros(2)
        self._check(self.lib.fluid_gpu_energy(self.h, _ptr(o)))
        return o

    def diagnostics_noalloc(self, which, out_dev_ptr):
        """Queue divergence (0), momentum (1) or energy (2) into 2 doubles of
        device memory at out_dev_ptr (an integer address); asynchronous."""
        self._check(self.lib.fluid_gpu_diagnostics_noalloc(self.h, int(which),
                                                           _VP(int(out_dev_ptr))))

    def set_particles(self, den, a, b, x, y, az, ux=None, uy=None, vz=None):
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (den, a, b, x, y, az)]
        n = arrs[0].size
        for v in (ux, uy, vz):
            arrs.append(np.zeros(n) if v is None else np.ascontiguousarray(v, dtype=np.float64))
        if any(v.size != n for v in arrs):
            raise ValueError("particle arrays differ in length")
        self._check(self.lib.suspensions_gpu_create(self.h, int(n), *[_ptr(v) for v in arrs]))
        self.n_particles = n

    def particles(self):
        out = np.zeros((self.n_particles, NREC))
        if self.n_particles:
            self._check(self.lib.suspensions_gpu_download(self.h, _ptr(out)))
        return out

    def set_particle_records(self, rec):
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        if rec.shape != (self.n_particles, NREC):
            raise ValueError("expected (%d, %d) records" % (self.n_particles, NREC))
        self._check(self.lib.suspensions_gpu_upload(self.h, _ptr(rec)))

    def reset_particle_increments(self):
        self._check(self.lib.suspensions_gpu_reset_particle_increments(self.h))

    def compute_inertia(self, cnstep):
        self._check(self.lib.suspensions_gpu_compute_inertia(self.h, int(cnstep)))

    def exchange_momentum(self):
        self._check(self.lib.suspensions_gpu_exchange_momentum(self.h))

    def update_momentum_field(self):
        self._check(self.lib.suspensions_gpu_update_momentum_fleid(self.h))

    def compute_collision_force(self, cnstep):
        self._check(self.lib.suspensions_gpu_compute_collision_force(self.h, int(cnstep)))

    def increment_particles(self, rk):
        r = _D()
        self._check(self.lib.suspensions_gpu_increment_particles(self.h, int(rk), ctypes.byref(r)))
        return r.value

    def update_particles(self):
        self._check(self.lib.suspensions_gpu_update_particles(self.h))

    def colliding_pairs(self):
        n = ctypes.c_int()
        self._check(self.lib.suspensions_gpu_colliding_pairs(self.h, 0, None, ctypes.byref(n)))
        buf = np.zeros(2 * max(n.value, 1), dtype=np.int32)
        self._check(self.lib.suspensions_gpu_colliding_pairs(
            self.h, n.value, buf.ctypes.data_as(_PI), ctypes.byref(n)))
        return [tuple(int(v) for v in buf[2 * k:2 * k + 2]) for k in range(n.value)]

    def integrate(self):
        sub = (ctypes.c_int * 3)()
        self._check(self.lib.eif_gpu_integrate(self.h, sub))
        return list(sub)

    def step(self):
        dt = _D()
        self._check(self.lib.eif_gpu_step(self.h, ctypes.byref(dt)))
        return dt.value

    def run_steps(self, nsteps):
        """nsteps time steps; returns (dts, substeps[nsteps, 3])."""
        dts = np.zeros(max(int(nsteps), 1))
        sub = np.zeros((max(int(nsteps), 1), 3), dtype=np.int32)
        self._check(self.lib.eif_gpu_run_steps(self.h, int(nsteps), _ptr(dts),
                                               sub.ctypes.data_as(_PI)))
        return dts[:nsteps], sub[:nsteps]

    # -- adjoint of one RK stage ---------------------------------------------
    ADJ_UX, ADJ_UY, ADJ_P, ADJ_PARTICLES, ADJ_FORCE, ADJ_RE = range(6)

    def autotune(self, allow_ysolve_switch=False):
        """Time the candidate launch shapes and keep the fastest; returns
        the report string."""
        buf = ctypes.create_string_buffer(1024)
        self._check(self.lib.eif_gpu_autotune(self.h, int(bool(allow_ysolve_switch)), buf, 1024))
        return buf.value.decode()

    def wall_shear(self):
        out = _D()
        self._check(self.lib.fluid_gpu_wall_shear(self.h, ctypes.byref(out)))
        return out.value

    def tape_begin(self):
        self._check(self.lib.eif_gpu_tape_begin(self.h))

    def step_fixed_dt(self, dt, record=False):
        self._check(self.lib.eif_gpu_step_fixed_dt(self.h, _D(dt), int(bool(record))))

    def stage_adjoint(self):
        self._check(self.lib.eif_gpu_stage_adjoint(self.h))

    def _adj_shape(self, which):
        if which == self.ADJ_UX:
            return self.shapes["ux"]
        if which in (self.ADJ_UY, self.ADJ_P):
            return self.shapes["uy"]
        if which == self.ADJ_PARTICLES:
            return (self.n_particles, 6)
        if which == self.ADJ_FORCE:
            return (self.n_particles, 3)
        return (1,)

    def adjoint_add_wall_shear(self, weight):
        self._check(self.lib.eif_gpu_adjoint_add_wall_shear(self.h, _D(weight)))

    def adjoint_get(self, which):
        out = np.zeros(self._adj_shape(which))
        self._check(self.lib.eif_gpu_adjoint_get(self.h, int(which), _ptr(out)))
        return out

    def adjoint_add(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self._adj_shape(which))
        self._check(self.lib.eif_gpu_adjoint_add(self.h, int(which), _ptr(a)))

    def graph(self, enable=True):
        self._check(self.lib.eif_gpu_graph_enable(self.h, int(bool(enable))))

    def tune_copy(self, src):
        """Apply the launch shapes autotune() chose for src, a Solver of the
        same grid, to this one."""
        self._check(self.lib.eif_gpu_tune_copy(self.h, src.h))

    def set_projection(self, strip):
        """0: DCT-III and projection as two kernels (psi stored); 2, 4, 8:
        one fused kernel over strips of that many rows (psi not stored)."""
        self._check(self.lib.eif_gpu_set_projection(self.h, int(strip)))

    def set_collision_search(self, mode):
        """Collision broad phase: COLLISION_AUTO (cell list from 4096
        particles), COLLISION_ALL_PAIRS or COLLISION_CELLS; the same results."""
        self._check(self.lib.eif_gpu_set_collision_search(self.h, int(mode)))

    def graph_set_subiterations(self, k):
        arr = (ctypes.c_int * 3)(*[int(v) for v in k])
        self._check(self.lib.eif_gpu_graph_set_subiterations(self.h, arr))

    def graph_stats(self):
        """(graph launches, resumed stops, captures, K per stage)."""
        out = (ctypes.c_long * 4)()
        self._check(self.lib.eif_gpu_graph_stats(self.h, out))
        k = int(out[3])
        return int(out[0]), int(out[1]), int(out[2]), (k // 10000, k // 100 % 100, k % 100)

    def ibm_stats(self, reset=False):
        """(cells visited, cells evaluated, FP64 iterations, IBM calls, FP32 iterations)."""
        out = (ctypes.c_ulonglong * 5)()
        self._check(self.lib.eif_gpu_ibm_stats(self.h, out, int(bool(reset))))
        return tuple(int(v) for v in out)

    def ibm_reference_iterations(self):
        """(cells, iterations) of the reference's FP64 find_normal_t over the
        current IBM boxes."""
        out = (ctypes.c_ulonglong * 2)()
        self._check(self.lib.eif_gpu_ibm_reference_iterations(self.h, out))
        return int(out[0]), int(out[1])

    def synchronize(self):
        self._check(self.lib.eif_gpu_synchronize(self.h))

    # -- statistics (src/statistics/collect.c, output.c) -------------------------
    def statistics_collect(self):
        self._check(self.lib.statistics_gpu_collect(self.h))

    def statistics_reset(self):
        self._check(self.lib.statistics_gpu_reset(self.h))

    def statistics(self):
        """Accumulators ux1, ux2 (ux layout), uy1, uy2, phi (uy layout) and num."""
        out = {k: np.empty(self.shapes["ux" if k[:2] == "ux" else "uy"])
               for k in ("ux1", "ux2", "uy1", "uy2", "phi")}
        num = ctypes.c_int()
        self._check(self.lib.statistics_gpu_download(
            self.h, *[_ptr(out[k]) for k in ("ux1", "ux2", "uy1", "uy2", "phi")], ctypes.byref(num)))
        out["num"] = num.value
        return out

    # -- profiling ------------------------------------------------------------
    def profile(self, enable=True):
        self._check(self.lib.eif_gpu_profile_enable(self.h, int(bool(enable))))

    def profile_select(self, name=None):
        self._check(self.lib.eif_gpu_profile_select(self.h, None if not name else name.encode()))

    def profile_sample(self, every):
        self._check(self.lib.eif_gpu_profile_sample(self.h, int(every)))

    def profile_reset(self):
        self._check(self.lib.eif_gpu_profile_reset(self.h))

    def profile_query(self, name):
        cnt = ctypes.c_long()
        tot = _D()
        self._check(self.lib.eif_gpu_profile_query(self.h, name.encode(), ctypes.byref(cnt), ctypes.byref(tot)))
        return cnt.value, tot.value

    def profile_names(self):
        return self.lib.eif_gpu_profile_names().decode().split(",")
