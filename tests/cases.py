"""Input builders shared by the tests: the same seeded inputs are given to
the GPU path and to the oracle."""
import numpy as np


def c0_config():
    """Single ellipse rotating in simple shear: 128x128, ly=1, Re=100, walls
    moving at uy=-0.5/+0.5, initial uy=-0.5+x (the reference's case-3 boundary
    and initial-condition files,
    docs/source/examples/case3/config/src_fluid_boundary_conditions.c:38-41 and
    src_fluid_init.c:59-66), one ellipse a=0.15, b=0.075 at (0.5, 0.5)."""
    cfg = dict(itot=128, jtot=128, ly=1.0, Re=100.0, extfrcy=0.0,
               uy_wall_xm=-0.5, uy_wall_xp=0.5)
    parts = dict(den=[1.0], a=[0.15], b=[0.075], x=[0.5], y=[0.5], az=[0.0])
    return cfg, parts


def couette_fields(itot, jtot, xc):
    """ux = 0, uy = -0.5 + XC(i) in the reference layouts."""
    ux = np.zeros((jtot + 2, itot + 1))
    uy = np.zeros((jtot + 2, itot + 2))
    uy[1:jtot + 1, 1:itot + 1] = -0.5 + xc[None, 1:itot + 1]
    p = np.zeros((jtot + 2, itot + 2))
    return ux, uy, p


def random_fields(itot, jtot, seed, amp=1.0, smooth=True):
    """Smooth random velocity/pressure fields (interior only; boundary values
    are applied by update_boundaries on both sides)."""
    rng = np.random.default_rng(seed)
    ux = np.zeros((jtot + 2, itot + 1))
    uy = np.zeros((jtot + 2, itot + 2))
    p = np.zeros((jtot + 2, itot + 2))
    x = (np.arange(itot + 2) - 0.5) / itot
    y = (np.arange(jtot + 2) - 0.5) / jtot
    X, Y = np.meshgrid(x, y)
    if smooth:
        f = np.zeros_like(X)
        for _ in range(6):
            kx, ky = rng.integers(1, 5, size=2)
            f += rng.standard_normal() * np.sin(np.pi * kx * X) * np.cos(2 * np.pi * ky * Y + rng.uniform(0, 6.28))
        g = np.zeros_like(X)
        for _ in range(6):
            kx, ky = rng.integers(1, 5, size=2)
            g += rng.standard_normal() * np.sin(np.pi * kx * X + 0.3) * np.sin(2 * np.pi * ky * Y + rng.uniform(0, 6.28))
    else:
        f = rng.standard_normal(X.shape)
        g = rng.standard_normal(X.shape)
    ux[:, :] = amp * f[:, :itot + 1]
    uy[:, :] = amp * g
    p[:, :] = rng.standard_normal(p.shape) * 0.1 + 0.2 * f
    return ux, uy, p


def place_ellipses(n, ly, a_range, ba_range, gap, seed, den_range=(1.0, 1.0), lx=1.0,
                   max_tries=200000):
    """Seeded rejection placement of non-overlapping ellipses using their
    bounding circles (r = max(a, b)) plus a minimum gap; positions in
    [r, lx-r] x [0, ly), orientations U[0, pi)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(*a_range, size=n)
    b = a * rng.uniform(*ba_range, size=n)
    den = rng.uniform(*den_range, size=n)
    az = rng.uniform(0.0, np.pi, size=n)
    r = np.maximum(a, b)
    xs = np.zeros(n)
    ys = np.zeros(n)
    placed = 0
    tries = 0
    while placed < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place %d ellipses" % n)
        x = rng.uniform(r[placed] + gap, lx - r[placed] - gap)
        y = rng.uniform(0.0, ly)
        if placed:
            dx = xs[:placed] - x
            dy = ys[:placed] - y
            dy = dy - ly * np.round(dy / ly)
            if np.any(np.hypot(dx, dy) < r[:placed] + r[placed] + gap):
                continue
        xs[placed] = x
        ys[placed] = y
        placed += 1
    return dict(den=den, a=a, b=b, x=xs, y=ys, az=az)


def touching_pairs(ly=1.0):
    """A handful of ellipses with overlapping and near-touching pairs, one
    across the periodic boundary and two near the walls."""
    return dict(
        den=[1.0, 1.2, 1.0, 1.5, 1.0, 1.1, 1.0],
        a=[0.06, 0.05, 0.05, 0.04, 0.05, 0.05, 0.03],
        b=[0.03, 0.04, 0.025, 0.04, 0.02, 0.03, 0.03],
        x=[0.30, 0.39, 0.70, 0.75, 0.04, 0.955, 0.5],
        y=[0.20, 0.25, 0.02, ly - 0.03, 0.60, 0.40, 0.8],
        az=[0.3, 1.2, 0.1, 2.0, 1.5707963267948966, 0.7, 0.0],
    )


def contact_pairs(n_pairs, ly, seed, overlap=(0.85, 1.02), a_range=(0.02, 0.04), ba=(0.4, 1.0)):
    """Pairs of ellipses at controlled, shallow contact (centre distance
    between 0.85 and 1.02 of the sum of the bounding radii), each pair in
    its own cell of a grid of slots so that pairs do not interact; one pair
    straddles the periodic boundary and one touches each wall."""
    rng = np.random.default_rng(seed)
    out = {k: [] for k in ("den", "a", "b", "x", "y", "az")}
    nslot_y = max(1, n_pairs // 2)
    for k in range(n_pairs):
        a0, a1 = rng.uniform(*a_range, size=2)
        b0, b1 = a0 * rng.uniform(*ba), a1 * rng.uniform(*ba)
        r0, r1 = max(a0, b0), max(a1, b1)
        d = (r0 + r1) * rng.uniform(*overlap)
        th = rng.uniform(0, 2 * np.pi)
        cx = 0.25 + 0.5 * (k % 2)
        cy = (k // 2 + 0.5) * ly / nslot_y
        if k == 0:
            cy = 0.0  # across the periodic boundary
        x0, y0 = cx - 0.5 * d * np.cos(th), cy - 0.5 * d * np.sin(th)
        x1, y1 = cx + 0.5 * d * np.cos(th), cy + 0.5 * d * np.sin(th)
        if k == 1:
            x0 = min(a0, b0) * rng.uniform(0.8, 0.95)  # overlapping the wall x = 0
        for name, v in (("den", rng.uniform(1, 1.5)), ("a", a0), ("b", b0), ("x", x0),
                        ("y", y0 % ly), ("az", rng.uniform(0, np.pi))):
            out[name].append(v)
        for name, v in (("den", rng.uniform(1, 1.5)), ("a", a1), ("b", b1), ("x", x1),
                        ("y", y1 % ly), ("az", rng.uniform(0, np.pi))):
            out[name].append(v)
    return {k: np.array(v) for k, v in out.items()}
