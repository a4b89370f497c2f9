"""Writes long oracle trajectories as compact fixtures for the GPU parity
tests (the oracle needs minutes for these horizons; the GPU seconds).

    python tests/golden/make_trajectories.py [c1]

c1: the dilute Poiseuille suspension at its real shape (512x1024, ly=2,
Re=500, extfrcy=8/Re, 64 ellipses a~U[.03,.05], b/a=.5, gap .01, laminar
initial profile) for 1000 steps.  Stored: dt and the sub-iteration counts
of every step, the colliding pairs of every step's last collision call
(src/suspensions/collision.c:232,292), particle x, y, az every 10 steps,
and the fluid's norms every 100 steps.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import cases  # noqa: E402
from oracle import Oracle  # noqa: E402

OUT = os.path.join(HERE, "trajectories")


def c1_setup(threads=None):
    parts = cases.place_ellipses(64, 2.0, (0.03, 0.05), (0.5, 0.5), 0.01, 1)
    kw = dict(itot=512, jtot=1024, ly=2.0, Re=500.0, extfrcy=8.0 / 500.0)
    return parts, kw


def laminar_uy(xc, nx, ny):
    uy = np.zeros((ny + 2, nx + 2))
    uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
    return uy


def norms(f):
    p = f["p"][1:-1, 1:-1]
    return [np.linalg.norm(f["ux"]), np.linalg.norm(f["uy"]), np.linalg.norm(p - p.mean())]


def make_c1(nsteps=1000):
    parts, kw = c1_setup()
    o = Oracle(threads=os.cpu_count(), **kw)
    o.upload(uy=laminar_uy(o.grid()["xc"], kw["itot"], kw["jtot"]))
    o.update_boundaries()
    o.set_particles(**parts)
    dts, subs, pairs, pos, fl = [], [], [], [], []
    t0 = time.time()
    for k in range(nsteps):
        dts.append(o.step())
        subs.append(o.last_substeps())
        for n0, n1 in o.colliding_pairs():
            pairs.append((k, n0, n1))
        if k % 10 == 9:
            pos.append(o.particles()[:, 3:6].copy())
        if k % 100 == 99:
            fl.append(norms(o.download()))
            print("c1 step %d  %.0fs  pairs so far %d" % (k + 1, time.time() - t0, len(pairs)),
                  flush=True)
    os.makedirs(OUT, exist_ok=True)
    np.savez_compressed(os.path.join(OUT, "c1_1000.npz"), dt=np.array(dts),
                        substeps=np.array(subs, dtype=np.int32),
                        pairs=np.array(pairs, dtype=np.int32).reshape(-1, 3),
                        pos=np.array(pos), norms=np.array(fl))


if __name__ == "__main__":
    which = sys.argv[1:] or ["c1"]
    if "c1" in which:
        make_c1()
