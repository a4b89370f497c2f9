"""Test harness for the adjoint of one RK stage (eif_gpu_stage_adjoint):
a torch.autograd.Function over a segment of time steps, with the state at
each segment's start as the checkpoint (saved by autograd) and the steps in
between recomputed, recorded and reversed in its backward.  torch is used
here only as the tape of segments; every derivative comes from the HIP
adjoint kernels through the C-ABI.

Objective (time-averaged over the N steps of a run, dt fixed at a given
schedule): the wall shear stress summed over both walls,
    tau(u) = (1/Re) sum_j dy [ (UY(1,j) - uy_wall_xm) / DXC(1)
                             + (UY(nx,j) - uy_wall_xp) / DXC(nx+1) ],
plus the summed particle drag -m fuy of each step's last exchange
(src/suspensions/exchange.c:73-76).
"""
import numpy as np
import torch

import ellipsesinflows_amd as eif

# particle_t members the adjoint carries: x, y, az, ux, uy, vz
PCOLS = [3, 4, 5, 9, 10, 11]


def particle_mass(rec):
    return rec[:, 0] * (3.1415926535897932 * rec[:, 1] * rec[:, 2])


def shear(grid, uy, re, wxm=0.0, wxp=0.0):
    nx = uy.shape[1] - 2
    dy = grid["dy"]
    dxc = grid["dxc"]
    inner = uy[1:-1]
    return (1.0 / re) * np.sum(dy * ((inner[:, 1] - wxm) / dxc[0] + (inner[:, nx] - wxp) / dxc[nx]))


def shear_seed(grid, shape, re):
    """d tau / d uy (reference layout)."""
    nx = shape[1] - 2
    s = np.zeros(shape)
    s[1:-1, 1] = grid["dy"] / (re * grid["dxc"][0])
    s[1:-1, nx] = grid["dy"] / (re * grid["dxc"][nx])
    return s


def drag(rec):
    return float(np.sum(-particle_mass(rec) * rec[:, 16]))


def step_objective(grid, f, rec, re, cfg):
    return shear(grid, f["uy"], re, cfg.get("uy_wall_xm", 0.0), cfg.get("uy_wall_xp", 0.0)) + drag(rec)


def gpu_step_objective(s):
    """the step's objective on the device: wall shear (fluid_gpu_wall_shear)
    plus the drag of the records' last exchange"""
    return s.wall_shear() + (drag(s.particles()) if s.n_particles else 0.0)


def gpu_objective(s, dts):
    """J of a run of fixed-dt steps from the solver's current state."""
    J = 0.0
    for dt in dts:
        s.step_fixed_dt(dt)
        J += gpu_step_objective(s) / len(dts)
    return J


def grid_info(solver):
    g = solver.grid()
    g["dy"] = solver.ly / solver.jtot
    return g


class Segment(torch.autograd.Function):
    """K time steps: (ux, uy, p, records, Re) -> (ux, uy, p, records, J_K)."""

    @staticmethod
    def forward(ctx, ux, uy, p, prt, re, run):
        s, dts, w, cfg = run["solver"], run["dts"], run["weight"], run["cfg"]
        grid = run["grid"]
        s.upload(ux.numpy(), uy.numpy(), p.numpy())
        if s.n_particles:
            s.set_particle_records(prt.numpy())
        J = 0.0
        for dt in dts:
            s.step_fixed_dt(dt)
            J += w * gpu_step_objective(s)
        ctx.save_for_backward(ux, uy, p, prt, re)
        ctx.run = run
        f = s.download()
        return (torch.from_numpy(f["ux"]), torch.from_numpy(f["uy"]), torch.from_numpy(f["p"]),
                torch.from_numpy(s.particles()), torch.tensor(J, dtype=torch.float64))

    @staticmethod
    def backward(ctx, gux, guy, gp, gprt, gJ):
        ux, uy, p, prt, re = ctx.saved_tensors
        run = ctx.run
        s, dts, w, grid = run["solver"], run["dts"], run["weight"], run["grid"]
        re_v = float(re)
        # recompute the segment from its checkpoint, recording
        s.upload(ux.numpy(), uy.numpy(), p.numpy())
        if s.n_particles:
            s.set_particle_records(prt.numpy())
        s.tape_begin()
        shears = []
        for dt in dts:
            s.step_fixed_dt(dt, record=True)
            shears.append(s.wall_shear())
        gJ = float(gJ)
        s.adjoint_add(s.ADJ_UX, gux.numpy())
        s.adjoint_add(s.ADJ_UY, guy.numpy())
        s.adjoint_add(s.ADJ_P, gp.numpy())
        if s.n_particles:
            s.adjoint_add(s.ADJ_PARTICLES, gprt.numpy()[:, PCOLS])
        mass = particle_mass(prt.numpy()) if s.n_particles else None
        re_bar = 0.0
        for n in reversed(range(len(dts))):
            # the objective of step n: its final velocity and its last exchange
            s.adjoint_add_wall_shear(w * gJ)
            re_bar += -shears[n] / re_v * (w * gJ)
            if s.n_particles:
                fseed = np.zeros((s.n_particles, 3))
                fseed[:, 1] = -mass * (w * gJ)
                s.adjoint_add(s.ADJ_FORCE, fseed)
            for _ in range(3):
                s.stage_adjoint()
        g_ux = torch.from_numpy(s.adjoint_get(s.ADJ_UX))
        g_uy = torch.from_numpy(s.adjoint_get(s.ADJ_UY))
        g_p = torch.from_numpy(s.adjoint_get(s.ADJ_P))
        g_prt = torch.zeros_like(prt)
        if s.n_particles:
            g_prt[:, PCOLS] = torch.from_numpy(s.adjoint_get(s.ADJ_PARTICLES))
        re_bar += float(s.adjoint_get(s.ADJ_RE)[0])
        return g_ux, g_uy, g_p, g_prt, torch.tensor(re_bar, dtype=torch.float64), None


def objective_and_gradient(solver, cfg, state, dts, every):
    """J over len(dts) steps from state = (ux, uy, p, records) and its
    gradient with respect to the initial state and Re, with checkpoints
    every `every` steps."""
    grid = grid_info(solver)
    ts = [torch.tensor(np.ascontiguousarray(a, dtype=np.float64), requires_grad=True)
          for a in state]
    re_t = torch.tensor(float(cfg["Re"]), dtype=torch.float64, requires_grad=True)
    w = 1.0 / len(dts)
    cur = ts
    J = torch.zeros((), dtype=torch.float64)
    for k0 in range(0, len(dts), every):
        run = {"solver": solver, "dts": list(dts[k0:k0 + every]), "weight": w, "cfg": cfg,
               "grid": grid}
        out = Segment.apply(*cur, re_t, run)
        cur = list(out[:4])
        J = J + out[4]
    J.backward()
    return float(J.detach()), [t.grad.numpy() for t in ts], float(re_t.grad)


def forward_objective(stepper, download, particles, grid, dts, re, cfg):
    """J of a run through any stepper (GPU fixed-dt steps, or the oracle)."""
    J = 0.0
    for dt in dts:
        stepper(dt)
        J += step_objective(grid, download(), particles(), re, cfg) / len(dts)
    return J
