"""Benchmark: FP64 timesteps/s of the EllipsesInFlows time step on one MI355X.

Workload (the 2048x4096 + 1024-ellipse configuration): lx=1 walls, ly=2
periodic, 2048x4096 grid, Re=1000, Poiseuille forcing extfrcy=8/Re, 1024
ellipses a~U[.008,.012], b/a~U[.4,1], az~U[0,pi), density~U[1,1.5], seeded
rejection placement; initial fluid = laminar profile uy = 4x(1-x).  A step
is the reference's main-loop body (src/main.c:71-76): param_decide_dt plus
integrate() with its three RK stages and Crank-Nicolson sub-iterations.
Inputs are resident on the GPU before the timed region.

Prints one JSON line (other output goes to stderr).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import ellipsesinflows_amd as eif  # noqa: E402
from ellipsesinflows_amd import initparticles  # noqa: E402

METRIC = "FP64 timesteps/s, 2048x4096 grid + 1024 ellipses; batched vs plain; adjoint/fwd"
NX, NY, LY, RE = 2048, 4096, 2.0, 1000.0
NP = 1024
SEED = 2

# Algorithmic FP64 operations of one IBM cell evaluation in the reference
# (src/suspensions/auxiliary.c:19-59, src/ellipse.c:30-65,
# src/suspensions/inertia.c:55-60), every +,-,*,/,sqrt,fabs,fmin,fmax,
# pow and libm call counted as one operation: 63 outside the Newton loop,
# 51 per Newton iteration.
FLOPS_CELL = 63
FLOPS_ITER = 51
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12
# the dominant launch: a Crank-Nicolson sub-iteration's IBM inertia
# integrals, with its collision forces in the same launch (k_inertia_collide)
DOMINANT = "ibm_substep"
# one launch in PROFILE_EVERY is timed by HIP events that the launch itself
# stamps on the kernel's stream; 13 is prime to the usual 3-6 such launches
# of a stage, so the timed launches rotate through them
PROFILE_EVERY = 13


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_inputs():
    parts = initparticles.random_ellipses(NP, LY, (0.008, 0.012), (0.4, 1.0), 0.0, SEED,
                                          den_range=(1.0, 1.5))
    return parts


def init_fields(xc):
    ux = np.zeros((NY + 2, NX + 1))
    uy = np.zeros((NY + 2, NX + 2))
    p = np.zeros((NY + 2, NX + 2))
    uy[1:NY + 1, 1:NX + 1] = 4.0 * xc[None, 1:NX + 1] * (1.0 - xc[None, 1:NX + 1])
    return ux, uy, p


def box_cells(rec):
    """IBM box cells per call (auxiliary.c:11-17 boxes of the 3 images)."""
    dx, dy = 1.0 / NX, LY / NY
    tot = 0
    for q in rec:
        r = max(q[1], q[2])
        px = q[3] + q[6]
        wi = min(int(np.ceil((px + r) / dx) + 3), NX) - max(int(np.floor((px - r) / dx) - 3), 1) + 1
        for per in (-1, 0, 1):
            py = q[4] + q[7] + LY * per
            hj = min(int(np.ceil((py + r) / dy) + 3), NY) - max(int(np.floor((py - r) / dy) - 3), 1) + 1
            if wi > 0 and hj > 0:
                tot += wi * hj
    return tot


def traffic_from_profiles(kernel):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        v = t.get(kernel)
        return None if v is None else float(v)
    except (OSError, ValueError):
        return None


def cpu_baseline(parts, xc):
    """The oracle (C restatement, OpenMP) timed on the host cores for ONE step
    of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import Oracle  # test infrastructure; cpu_baseline leg only
    o = Oracle(NX, NY, ly=LY, Re=RE, extfrcy=8.0 / RE)
    cores = o.lib.orc_set_threads(0)
    ux, uy, p = init_fields(xc)
    o.upload(ux, uy, p)
    o.update_boundaries()
    o.set_particles(**{k: parts[k] for k in ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")})
    t0 = time.perf_counter()
    o.step()
    el = time.perf_counter() - t0
    o.close()
    return {"value": 1.0 / el, "unit": "timesteps/s", "cores": int(cores), "kind": "port",
            "sample": "1 full time step (decide_dt + 3 RK stages with CN sub-iterations) of the "
                      "same 2048x4096 + 1024-ellipse workload from the same initial state, "
                      "oracle/eif_oracle.c with OpenMP over %d threads, %.1f s" % (cores, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--poisson", choices=("partition", "thomas"), default="partition")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="before the workload, copy a buffer of this size once on the GPU "
                         "(a kernel of known traffic for scaling FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.calibration_copy_mib > 0:
        import torch
        n = args.calibration_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: %d MiB read + %d MiB written" % (args.calibration_copy_mib,
                                                                args.calibration_copy_mib))
        del src, dst
    parts = workload_inputs()
    mode = eif.POISSON_PARTITION if args.poisson == "partition" else eif.POISSON_THOMAS
    g = eif.Solver(itot=NX, jtot=NY, ly=LY, Re=RE, extfrcy=8.0 / RE, poisson_mode=mode)
    xc = g.grid()["xc"]
    ux, uy, p = init_fields(xc)
    g.upload(ux, uy, p)
    g.update_boundaries()
    g.set_particles(**{k: parts[k] for k in ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")})
    log("area fraction %.3f" % initparticles.area_fraction(parts, LY))
    for _ in range(args.warmup):
        g.step()
    g.synchronize()
    cells0 = box_cells(g.particles())
    g.ibm_stats(reset=True)
    g.profile_select(DOMINANT)
    g.profile_sample(PROFILE_EVERY)
    g.profile(True)
    g.profile_reset()
    t0 = time.perf_counter()
    dts = []
    for _ in range(args.steps):
        dts.append(g.step())
    g.synchronize()
    el = time.perf_counter() - t0
    g.profile(False)
    ncall, kms = g.profile_query(DOMINANT)
    visited, evaluated, it64, calls, it32 = g.ibm_stats()
    cells1 = box_cells(g.particles())
    # the reference's own FP64 iteration count on the current boxes (measured
    # outside the timed region), for the algorithmic operation count
    ref_cells, ref_iters = g.ibm_reference_iterations()
    dmax, _ = g.divergence()
    g.close()

    steps_per_s = args.steps / el
    ms = 1e3 * el / args.steps
    mean_iters = ref_iters / max(ref_cells, 1)
    # inertia and exchange share the counters; per-launch cells from the boxes
    cells = 0.5 * (cells0 + cells1)
    avg_s = kms * 1e-3 / max(ncall, 1)
    # no timed launch (every stage's loop closed at its first sub-iteration):
    # no figure rather than a division by zero
    achieved = cells * (FLOPS_CELL + FLOPS_ITER * mean_iters) / avg_s if avg_s > 0 else None
    roofline = {"bound": "mfma", "achieved": achieved / 1e12 if achieved else None,
                "peak": PEAK_FP64 / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64 if achieved else None,
                "traffic": traffic_from_profiles(DOMINANT),
                "kernel": "k_inertia_collide (sub-iteration IBM inertia integrals, FP64 "
                          "vector ALU; its collision forces in the same launch are not counted)",
                "avg_launch_ms": kms / max(ncall, 1), "launches_timed": ncall,
                "algorithmic_flops_per_cell": FLOPS_CELL + FLOPS_ITER * mean_iters,
                "cells_per_launch": cells}
    out = {
        "metric": METRIC, "value": steps_per_s, "unit": "timesteps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded ellipse placement, laminar initial profile)",
        "config": {"workload": "2048x4096 grid, ly=2, Re=1000, extfrcy=8/Re, 1024 ellipses "
                               "a~U[.008,.012] b/a~U[.4,1] rho~U[1,1.5]; plain single case",
                   "poisson_y_solve": args.poisson, "parallelism": "single GPU"},
        "cell_updates_per_s": NX * NY * steps_per_s,
        "ibm": {"cells_per_step": visited / args.steps,
                "cells_evaluated_per_step": evaluated / args.steps,
                "reference_newton_iterations_per_cell": mean_iters,
                "fp64_iterations_per_evaluated_cell": it64 / max(evaluated, 1),
                "fp32_iterations_per_evaluated_cell": it32 / max(evaluated, 1),
                "cell_evals_per_s": visited / el, "calls_per_step": calls / args.steps},
        "final_max_divergence": dmax,
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(parts, xc)
        except Exception as e:  # reported, never fatal for the GPU figure
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
