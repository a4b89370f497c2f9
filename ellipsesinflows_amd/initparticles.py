"""Seeded initial-particle generation and the initparticles .npy format.

The reference reads ten files from ``initparticles/`` (or a restart
directory): ``n_particles.npy`` (int32, 0-d) and ``particle_{dens,as,bs,xs,
ys,azs,uxs,uys,vzs}.npy`` (float64, 1-d) -- src/suspensions/init.c:37-62,
written by initparticles/main.py:28-48.  ``random_ellipses`` places
ellipses by seeded rejection sampling the way
initparticles/RandomGenerator/src/main.c:175-252 does (sizes and angles
first, then centres tested against the bounding circles of the ellipses
already placed, periodic in y), using a cell list instead of the O(N^2)
scan.
"""
import os

import numpy as np

NPY_NAMES = ("dens", "as", "bs", "xs", "ys", "azs", "uxs", "uys", "vzs")
KEYS = ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")


def random_ellipses(n, ly, a_range, ba_range, gap, seed, den_range=(1.0, 1.0), lx=1.0,
                    max_tries=2_000_000):
    """n ellipses with a ~ U(a_range), b/a ~ U(ba_range), az ~ U[0, pi),
    density ~ U(den_range), zero velocities; centres x in [r+gap, lx-r-gap],
    y in [0, ly), no two bounding circles closer than gap."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(*a_range, size=n)
    b = a * rng.uniform(*ba_range, size=n)
    den = rng.uniform(*den_range, size=n)
    az = rng.uniform(0.0, np.pi, size=n)
    r = np.maximum(a, b)
    cell = 2.0 * r.max() + gap
    ncy = max(1, int(ly // cell))
    hy = ly / ncy
    grid = {}
    xs = np.zeros(n)
    ys = np.zeros(n)
    tries = 0
    for k in range(n):
        while True:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("could not place %d ellipses" % n)
            x = rng.uniform(r[k] + gap, lx - r[k] - gap)
            y = rng.uniform(0.0, ly)
            cx, cy = int(x // cell), int(y // hy) % ncy
            ok = True
            for dx in (-1, 0, 1):
                for dyc in (-1, 0, 1):
                    for m in grid.get((cx + dx, (cy + dyc) % ncy), ()):
                        ddy = ys[m] - y
                        ddy -= ly * np.round(ddy / ly)
                        if np.hypot(xs[m] - x, ddy) < r[m] + r[k] + gap:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                break
        xs[k], ys[k] = x, y
        grid.setdefault((cx, cy), []).append(k)
    zeros = np.zeros(n)
    return dict(den=den, a=a, b=b, x=xs, y=ys, az=az, ux=zeros.copy(), uy=zeros.copy(),
                vz=zeros.copy())


ENSEMBLE_GRIDS = ((256, 512), (512, 1024), (1024, 2048))
ENSEMBLE_COUNTS = (8, 32, 128, 256)


def ensemble_cases(n_cases=32, seed=0, ly=2.0):
    """Seeded draws of an ensemble of independent cases: per case a grid
    from ENSEMBLE_GRIDS (lx = 1 walls, ly periodic), Re ~ log-U[100, 1000]
    with the Poiseuille forcing extfrcy = 8/Re, a particle count from
    ENSEMBLE_COUNTS, a ~ U[.01, .04], b/a ~ U[.4, 1], orientation
    ~ U[0, pi), density 1, rejection-placed centres.  Returns a list of
    (config, particles) pairs; config holds the eif_config fields."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_cases):
        nx, ny = ENSEMBLE_GRIDS[rng.integers(len(ENSEMBLE_GRIDS))]
        re = float(np.exp(rng.uniform(np.log(100.0), np.log(1000.0))))
        npart = int(ENSEMBLE_COUNTS[rng.integers(len(ENSEMBLE_COUNTS))])
        parts = random_ellipses(npart, ly, (0.01, 0.04), (0.4, 1.0), 0.0,
                                int(rng.integers(1 << 31)))
        cfg = dict(itot=int(nx), jtot=int(ny), ly=float(ly), Re=re, extfrcy=8.0 / re)
        out.append((cfg, parts))
    return out


def area_fraction(parts, ly, lx=1.0):
    return float(np.sum(np.pi * np.asarray(parts["a"]) * np.asarray(parts["b"])) / (lx * ly))


def write(dirname, parts):
    """Write the ten initparticles files (numpy .npy v1.0)."""
    os.makedirs(dirname, exist_ok=True)
    n = len(parts["a"])
    np.save(os.path.join(dirname, "n_particles.npy"), np.array(n, dtype=np.int32))
    for name, key in zip(NPY_NAMES, KEYS):
        v = parts.get(key)
        v = np.zeros(n) if v is None else np.asarray(v, dtype=np.float64)
        np.save(os.path.join(dirname, "particle_%s.npy" % name), v)


def read(dirname):
    """Read the ten initparticles files (loaders that execute nothing)."""
    n = int(np.load(os.path.join(dirname, "n_particles.npy"), allow_pickle=False))
    out = {}
    for name, key in zip(NPY_NAMES, KEYS):
        out[key] = np.load(os.path.join(dirname, "particle_%s.npy" % name), allow_pickle=False)[:n]
    return out
