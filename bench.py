"""Benchmark: FP64 timesteps/s of the EllipsesInFlows time step on one MI355X.

Workload (the 2048x4096 + 1024-ellipse configuration): lx=1 walls, ly=2
periodic, 2048x4096 grid, Re=1000, Poiseuille forcing extfrcy=8/Re, 1024
ellipses a~U[.008,.012], b/a~U[.4,1], az~U[0,pi), density~U[1,1.5], seeded
rejection placement; initial fluid = laminar profile uy = 4x(1-x).  A step
is the reference's main-loop body (src/main.c:71-76): param_decide_dt plus
integrate() with its three RK stages and Crank-Nicolson sub-iterations.
Inputs are resident on the GPU before the timed region.  By default the
timed steps run as captured whole-step hipGraphs (--mode graph); the
dominant kernel's launch time and the fluid kernels' HBM fractions come
from HIP events in an eager segment of the same steps right after it.

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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(parts, xc, warm=1, timed=4):
    """The oracle (C restatement, OpenMP) timed on the host cores: `warm`
    untimed steps, then `timed` steps of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import Oracle  # test infrastructure; cpu_baseline leg only
    o = Oracle(NX, NY, ly=LY, Re=RE, extfrcy=8.0 / RE)
    cores = o.lib.orc_set_threads(0)
    ux, uy, p = init_fields(xc)
    o.upload(ux, uy, p)
    o.update_boundaries()
    o.set_particles(**{k: parts[k] for k in ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")})
    for _ in range(warm):
        o.step()
    t0 = time.perf_counter()
    for _ in range(timed):
        o.step()
    el = time.perf_counter() - t0
    o.close()
    return {"value": timed / el, "unit": "timesteps/s", "cores": int(cores), "kind": "port",
            "cpu": cpu_model(),
            "sample": "%d warm-up + %d timed full time steps (decide_dt + 3 RK stages with CN "
                      "sub-iterations) of the same 2048x4096 + 1024-ellipse workload from the "
                      "same initial state, oracle/eif_oracle.c with OpenMP over %d threads, "
                      "%.1f s timed" % (warm, timed, cores, el)}


# algorithmic HBM bytes per launch of the fluid kernels, in units of
# N = nx*ny values of 8 B (compulsory traffic, each field once): the predictor
# reads ux, uy, p and (stages 1-2) the kept previous velocity, and writes
# the new velocity: 5N at stage 0, 7N after, 19/3 N on average; the fused
# DCT-III + projection reads the solved coefficients, ux, uy and p and
# writes ux, uy and p (7N; psi is never stored)
FLUID_BYTES_N = {"predictor": 19.0 / 3.0, "rhs_dct2": 3.0, "ysolve_partition": 2.0,
                 "thomas": 2.0, "dct3_psi": 2.0, "correct_pressure": 7.0,
                 "dct3_project": 7.0}


ENSEMBLE_SEED = 3


def _ensemble_solvers(cases, tune=True):
    """One context per case; with tune, one case per grid shape is
    autotuned and its shapes applied to the others of that grid."""
    out = []
    tuned = {}
    for cfg, parts in cases:
        g = eif.Solver(poisson_mode=eif.POISSON_PARTITION, **cfg)
        nx, ny = cfg["itot"], cfg["jtot"]
        xc = g.grid()["xc"]
        uy = np.zeros((ny + 2, nx + 2))
        uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
        g.upload(uy=uy)
        g.update_boundaries()
        g.set_particles(**{k: parts[k] for k in ("den", "a", "b", "x", "y", "az", "ux", "uy",
                                                 "vz")})
        if tune:
            key = (nx, ny)
            if key not in tuned:
                tuned[key] = g
                g.autotune()
            else:
                g.tune_copy(tuned[key])
        out.append(g)
    return out


def ensemble_bench(nsteps, warm=3):
    """The 32-case ensemble (grids 256x512..1024x2048, Re ~ logU[100, 1000],
    8-256 ellipses): nsteps steps of every case through the plain path one
    case after the other, then the same from the same state with each case's
    step graphs on their own streams, all step graphs in one graph, and
    batched launches (each kernel of a step once for all 32 cases);
    case-steps/s of each and batched over sequential."""
    cases = initparticles.ensemble_cases(32, seed=ENSEMBLE_SEED)
    res = {}
    for concurrent in (0, 1, 2, 3):
        solvers = _ensemble_solvers(cases)
        eif.ensemble_run_steps(solvers, warm, concurrent=concurrent)
        for g in solvers:
            g.synchronize()
        t0 = time.perf_counter()
        eif.ensemble_run_steps(solvers, nsteps, concurrent=concurrent)
        for g in solvers:
            g.synchronize()
        el = time.perf_counter() - t0
        res[("sequential", "concurrent", "merged", "batched")[concurrent]] = len(cases) * nsteps / el
        for g in solvers:
            g.close()
    grids = {}
    for cfg, parts in cases:
        key = "%dx%d" % (cfg["itot"], cfg["jtot"])
        grids[key] = grids.get(key, 0) + 1
    return {"workload": "32 independent cases, seed %d: grids %s, 8-256 ellipses each, "
                        "Re~logU[100,1000]; launch shapes autotuned on one case per grid "
                        "and applied to the others, the same in every mode"
                        % (ENSEMBLE_SEED, grids),
            "steps_timed_per_case": nsteps,
            "case_steps_per_s_sequential": res["sequential"],
            "case_steps_per_s_graphs_on_streams": res["concurrent"],
            "case_steps_per_s_graphs_in_one_graph": res["merged"],
            "case_steps_per_s_batched": res["batched"],
            "cases_per_s_sequential_1000_steps": res["sequential"] / 1000.0,
            "cases_per_s_batched_1000_steps": res["batched"] / 1000.0,
            "batched_over_sequential": res["batched"] / res["sequential"]}


def adjoint_bench(nsteps=200, every=20):
    """The differentiation workload (256x512, ly 2, Re 200, extfrcy 8/Re, 16
    ellipses a .04 b .02 placed at least .05 apart, laminar start): nsteps
    forward steps (decide_dt + integrate) timed, then the gradient of the
    time-averaged wall shear + drag over the same steps (dt held at that
    schedule) by checkpoints every `every` steps: a forward pass keeping the
    checkpoints, then per segment in reverse the recorded recomputation and
    the stage adjoints (eif_gpu_stage_adjoint)."""
    rng = np.random.default_rng(4)
    parts = initparticles.random_ellipses(16, 2.0, (0.04, 0.04), (0.5, 0.5), 0.05, 4)
    cfg = dict(itot=256, jtot=512, ly=2.0, Re=200.0, extfrcy=8.0 / 200.0)
    g = eif.Solver(**cfg)
    xc = g.grid()["xc"]
    uy = np.zeros((514, 258))
    uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
    g.upload(uy=uy)
    g.update_boundaries()
    g.set_particles(**{k: parts[k] for k in ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")})
    f0, r0 = g.download(), g.particles()
    g.synchronize()
    t0 = time.perf_counter()
    dts = [g.step() for _ in range(nsteps)]
    g.synchronize()
    t_fwd = time.perf_counter() - t0
    mass = r0[:, 0] * (3.1415926535897932 * r0[:, 1] * r0[:, 2])
    w = 1.0 / nsteps
    t0 = time.perf_counter()
    # forward pass keeping the checkpoints
    g.upload(f0["ux"], f0["uy"], f0["p"])
    g.set_particle_records(r0)
    ckpt = []
    J = 0.0
    for k, dt in enumerate(dts):
        if k % every == 0:
            f = g.download()
            ckpt.append((f["ux"], f["uy"], f["p"], g.particles()))
        g.step_fixed_dt(dt)
        J += w * (g.wall_shear() - float(np.sum(mass * g.particles()[:, 16])))
    # reverse, segment by segment
    g.tape_begin()
    pb = np.zeros((len(r0), 6))
    first = True
    for seg in reversed(range(len(ckpt))):
        ux, uyc, p, rec = ckpt[seg]
        g.upload(ux, uyc, p)
        g.set_particle_records(rec)
        carry = None if first else (g.adjoint_get(g.ADJ_UX), g.adjoint_get(g.ADJ_UY),
                                    g.adjoint_get(g.ADJ_P), g.adjoint_get(g.ADJ_PARTICLES),
                                    g.adjoint_get(g.ADJ_RE))
        g.tape_begin()
        seg_dts = dts[seg * every:(seg + 1) * every]
        for dt in seg_dts:
            g.step_fixed_dt(dt, record=True)
        if carry is not None:
            for which, v in zip((g.ADJ_UX, g.ADJ_UY, g.ADJ_P, g.ADJ_PARTICLES, g.ADJ_RE), carry):
                g.adjoint_add(which, v)
        first = False
        fseed = np.zeros((len(r0), 3))
        fseed[:, 1] = -mass * w
        for _ in seg_dts:
            g.adjoint_add_wall_shear(w)
            g.adjoint_add(g.ADJ_FORCE, fseed)
            for _ in range(3):
                g.stage_adjoint()
    pb = g.adjoint_get(g.ADJ_PARTICLES)
    g.synchronize()
    t_adj = time.perf_counter() - t0
    g.close()
    return {"workload": "256x512 grid, ly=2, Re=200, 16 ellipses a=.04 b=.02 gap .05; J = "
                        "time-averaged wall shear + drag over %d steps, checkpoints every %d"
                        % (nsteps, every),
            "forward_ms_per_step": 1e3 * t_fwd / nsteps,
            "gradient_ms_per_step": 1e3 * t_adj / nsteps,
            "adjoint_over_forward": t_adj / t_fwd,
            "objective": J, "grad_norm_positions": float(np.linalg.norm(pb[:, :2]))}


# the collision broad phase's sweep: particle counts at ~17-30 % area
# fraction, each with a grid that resolves its ellipses by 2.5-8 cells
COLLISION_SWEEP = (
    ("1024 ellipses, a .013-.019", 512, 1024, 1024, (0.013, 0.019), -0.004),
    ("4096 ellipses, a .006-.009", 1024, 2048, 4096, (0.006, 0.009), -0.001),
    ("16384 ellipses, a .003-.0045", 1024, 2048, 16384, (0.003, 0.0045), -0.0005),
)


def collision_sweep(nsteps=3):
    """Time per step, and the average collision launch (with the cell-list
    build) and sub-iteration launch, of eager steps with the broad
    phase testing all pairs and with the cell list, on the same cases; the
    two runs' particle records must be identical."""
    out = []
    for name, nx, ny, npart, a_range, gap in COLLISION_SWEEP:
        parts = initparticles.random_ellipses(npart, 2.0, a_range, (0.5, 1.0), gap, npart)
        res = {}
        for search in (eif.COLLISION_ALL_PAIRS, eif.COLLISION_CELLS):
            g = eif.Solver(itot=nx, jtot=ny, ly=2.0, Re=500.0, extfrcy=8.0 / 500.0,
                           poisson_mode=eif.POISSON_PARTITION)
            xc = g.grid()["xc"]
            uy = np.zeros((ny + 2, nx + 2))
            uy[1:-1, 1:-1] = 4.0 * xc[None, 1:-1] * (1 - xc[None, 1:-1])
            g.upload(uy=uy)
            g.update_boundaries()
            g.set_particles(**{k: parts[k] for k in ("den", "a", "b", "x", "y", "az", "ux", "uy", "vz")})
            g.set_collision_search(search)
            g.step()
            # timed eager steps, then eager steps with each launch between
            # events
            g.synchronize()
            t0 = time.perf_counter()
            g.run_steps(nsteps)
            g.synchronize()
            r = {"ms_per_step": 1e3 * (time.perf_counter() - t0) / nsteps}
            g.profile_select(None)
            g.profile_sample(1)
            g.profile(True)
            g.profile_reset()
            g.run_steps(nsteps)
            g.synchronize()
            g.profile(False)
            for k in ("collision", "ibm_substep"):
                cnt, tms = g.profile_query(k)
                r[k + "_avg_launch_ms"] = tms / max(cnt, 1)
            res[search] = (r, g.particles())
            g.close()
        a, c = res[eif.COLLISION_ALL_PAIRS], res[eif.COLLISION_CELLS]
        if not np.array_equal(a[1], c[1]):
            raise RuntimeError("cell list and all pairs differ: " + name)
        row = {"case": name, "grid": [nx, ny], "particles": npart, "all_pairs": a[0],
               "cell_list": c[0], "speedup": a[0]["ms_per_step"] / c[0]["ms_per_step"]}
        log("collision sweep:", json.dumps(row))
        out.append(row)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--poisson", choices=("partition", "thomas"), default="partition")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch shapes")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint/forward cost measurement")
    ap.add_argument("--ensemble-steps", type=int, default=20,
                    help="steps per case of the 32-case ensemble comparison (0: skip)")
    ap.add_argument("--mode", choices=("graph", "eager"), default="graph",
                    help="timed steps as captured whole-step graphs or eager launches")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="before the workload, copy a buffer of this size once on the GPU "
                         "(a kernel of known traffic for scaling FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--collision-sweep", action="store_true",
                    help="only compare the collision broad phases (all pairs, cell list) "
                         "at 1024, 4096 and 16384 particles")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.collision_sweep:
        print(json.dumps({"collision_sweep": collision_sweep()}))
        return

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
    tuning = g.autotune() if not args.no_autotune else None
    if tuning:
        log("autotune: " + tuning)
    if args.mode == "graph":
        g.graph(True)
    g.run_steps(args.warmup)
    g.synchronize()
    cells0 = box_cells(g.particles())
    g.ibm_stats(reset=True)
    t0 = time.perf_counter()
    g.run_steps(args.steps)
    g.synchronize()
    el = time.perf_counter() - t0
    visited, evaluated, it64, calls, it32 = g.ibm_stats()
    graph_stats = g.graph_stats() if args.mode == "graph" else None
    # kernel times: HIP events in an eager segment right after the timed
    # steps (a captured graph carries no host-recorded events): the dominant
    # kernel by launch-stamped events on its own stream, one launch in 13;
    # then the fluid kernels, each between events on the main stream
    seg = max(4, min(args.steps, 12))
    g.graph(False)
    g.profile_select(DOMINANT)
    g.profile_sample(PROFILE_EVERY)
    g.profile(True)
    g.profile_reset()
    g.run_steps(seg)
    g.synchronize()
    g.profile(False)
    ncall, kms = g.profile_query(DOMINANT)
    cells1 = box_cells(g.particles())
    fluid = {}
    g.profile_select(None)
    g.profile_sample(1)
    g.profile(True)
    g.profile_reset()
    g.run_steps(4)
    g.synchronize()
    g.profile(False)
    for name, nvals in FLUID_BYTES_N.items():
        cnt, tms = g.profile_query(name)
        if cnt:
            avg = tms * 1e-3 / cnt
            nbytes = nvals * NX * NY * 8.0
            fluid[name] = {"avg_launch_ms": 1e3 * avg, "algorithmic_bytes": nbytes,
                           "achieved_GBps": nbytes / avg / 1e9,
                           "frac_hbm": nbytes / avg / PEAK_HBM}
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
    # FP64 vector ALU: MI355X's FP64 vector and matrix peaks are both 78.6
    # TFLOP/s; the contract's bound field names "mfma" for that ceiling
    roofline = {"bound": "mfma", "achieved": achieved / 1e12 if achieved else None,
                "peak": PEAK_FP64 / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64 if achieved else None,
                "traffic": traffic_from_profiles(DOMINANT),
                "kernel": "k_inertia_collide (sub-iteration IBM inertia integrals, FP64 "
                          "vector ALU; its collision forces in the same launch are not counted)",
                "avg_launch_ms": kms / max(ncall, 1), "launches_timed": ncall,
                "algorithmic_flops_per_cell": FLOPS_CELL + FLOPS_ITER * mean_iters,
                "cells_per_launch": cells, "compute_unit": "FP64 vector ALU",
                "timing": "eager segment of %d steps after the timed steps" % seg}
    # the FP64 work the kernel actually issues (rocprofv3 FP64 instruction
    # counters, profiles/ibm_substep_valu_mix.md) over the same launch time
    executed = traffic_from_profiles(DOMINANT + "_fp64_lane_flops")
    if executed and avg_s > 0:
        roofline["executed_fp64_tflops"] = executed / avg_s / 1e12
        roofline["executed_frac"] = executed / avg_s / PEAK_FP64
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
        "fluid_kernels_hbm": fluid,
        "step_mode": args.mode,
        "autotune": tuning,
    }
    if graph_stats:
        out["graph"] = {"launches": graph_stats[0], "resumed_stops": graph_stats[1],
                        "captures": graph_stats[2], "subiterations_captured": graph_stats[3]}
    if args.ensemble_steps > 0:
        out["ensemble"] = ensemble_bench(args.ensemble_steps)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench()
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(parts, xc)
        except Exception as e:  # reported, never fatal for the GPU figure
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
