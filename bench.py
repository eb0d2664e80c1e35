#!/usr/bin/env python3
"""Throughput of the flow-past-a-cylinder time step on one MI355X.

Workload (the metric's configuration, one GPU): uniform 512 x 256 x 64 cells over
20 x 10 x 0.5 (D = 1), cylinder at (5, 5) with R = 0.5 (sharp area-fraction mask),
Re = 200, dt = 0.001, inflow U = 1, convective outflow with flux correction, slip
top/bottom walls, periodic z, FP64, pressure by BiCGSTAB on the SPE operator at
tol 1e-10; initial field u = 1 +- 1e-3 seeded noise.  W warm-up steps (from the
initial field; step 1 has no pressure solve, as in the reference driver), then K
timed steps.  value = Mcell-steps/s = N*K / wall / 1e6 with all fields resident on
the device.

The JSON line also carries:
  roofline      -- the dominant kernel's algorithmic bytes per launch over its mean
                   launch duration, from HIP events on the library's stream during the
                   timed steps, against the 8 TB/s HBM peak;
  cpu_baseline  -- the reference's own CPU code (oracle/_ref, built from the reference
                   headers) timed on this host on a bounded sample of the same grid;
  stream_ceiling -- what plain streaming kernels (copy, and a two-read-one-write triad
                   of phase B's shape) reach at the same array size with the caches
                   flushed: the attainable rate the phases are read against;
  mixed         -- the mixed-precision mode and the same algorithm in FP64 (FP64
                   multigrid) over the FP64 run's steps, with field errors against both;
                   mixed.sustained: both multigrid modes timed over 20 steps after 100;
  rbsor, adjoint -- the red-black SOR leg and the drag adjoint (their own workloads).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# algorithmic HBM bytes per interior cell and launch (FP64; each array counted once,
# ghosts and 1-D metric tables excluded)
BYTES_PER_CELL = {
    "bicg_A": 32,    # r r, q, r0;  w Ap  (p = r + beta q is formed in registers)
    "bicg_B": 24,    # r r, Ap;         w As
    "bicg_C": 72,    # r r, Ap, q, As, r0, x;  w x, r, q  (p formed again from r, q)
    "bicg_init": 48, # r b, x;  w r0, r, px1
    "convdiff": 48,  # r u, v, w;  w u*, v*, w*
    "rhs": 32,       # r u*, v*, w*;  w b
    "project": 96,   # r u*, v*, w*, X, eta;  w u, v, w, p, f(3)
    "sor_half": 16,  # colour-split half sweep: r p (other colour), p, rhs (own);  w p (own)
}

# the BiCGSTAB iteration (phases A + B + C) is the FP64 step's dominant work: 128 B/cell
# per iteration (32 + 24 + 72)
ITERATION_PHASES = ("bicg_A", "bicg_B", "bicg_C")

MIXED_INNER_TOL = 0.0  # the library's per-mode default: 1e-5 in both multigrid modes

# multigrid V-cycles from captured HIP graphs (cfdmx_config.mg_graphs); --no-mg-graphs
# launches them kernel by kernel (the same bits and, at these sizes, the same time)
MG_GRAPHS = 1

# the same phases on float Krylov vectors (mixed-precision inner solve); the FP64
# residual/update passes of the refinement are counted under bicg_init
BYTES_PER_CELL_FP32 = {
    "convdiff": 48, "rhs": 32, "project": 96,
}

WORKLOADS = {
    "cylinder512": dict(nx=512, ny=256, nz=64, re=200.0, dt=0.001, tol=1e-10),
    "cylinder256": dict(nx=256, ny=128, nz=32, re=100.0, dt=0.002, tol=1e-10),
    "cylinder128": dict(nx=128, ny=64, nz=4, re=100.0, dt=0.002, tol=1e-8),
}


def make_solver(wl, precond, solver, seed, maxit=5000, precision=0, smag_cs=0.0,
                inner_tol=None):
    import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg
    from cfd_nssolver_forhpc_flowpassingcylinder_amd.api import (PRECOND_JACOBI, PRECOND_NONE,
                                                                  SOLVER_BICGSTAB, SOLVER_RBSOR)
    slip_num = [[0, 0, 0], [1, 1, 1], [1, 0, 1], [1, 0, 1], [1, 1, 1], [1, 1, 1]]
    slip_dir = [[1, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]]
    cfg = pkg.default_config(
        nx=wl["nx"], ny=wl["ny"], nz=wl["nz"], lx=20.0, ly=10.0, lz=0.5, re=wl["re"],
        dt=wl["dt"], tol=wl["tol"], maxit=maxit,
        precond=PRECOND_JACOBI if precond == "jacobi" else PRECOND_NONE,
        solver=SOLVER_RBSOR if solver == "rbsor" else SOLVER_BICGSTAB, sor_omega=1.8,
        bc_num=slip_num, bc_dir=slip_dir, z_periodic=1, outflow_convective=1,
        convective_speed=1.0, flux_correction=1, skip_first_pressure=1,
        precision=precision, inner_tol=MIXED_INNER_TOL if inner_tol is None else inner_tol,
        mg_graphs=MG_GRAPHS, smagorinsky_cs=smag_cs)
    s = pkg.FlowSolver(cfg)
    s.init_flow(1.0, 1e-3, seed)
    s.set_eta_cylinder_z(5.0, 5.0, 0.5, 100)
    return s


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                      "launches": 3}), flush=True)


def stream_ceiling(n_cells, launches=60):
    """Streaming rate that HBM gives this array size when nothing is reused from the
    caches: torch's elementwise copy (16 B/cell) and z = x - 0.5 y (24 B/cell, the
    shape of BiCGSTAB phase B: two arrays read, one written), rotating over nine arrays
    of n_cells doubles (more than the 256 MB memory-side cache) so that each launch reads
    arrays the previous two launches did not touch."""
    import torch
    bufs = [torch.zeros(n_cells, dtype=torch.float64, device="cuda") for _ in range(9)]
    out = {}
    for name, nbytes in (("copy", 16), ("triad", 24)):
        def launch(q):
            x, y, z = bufs[(3 * q) % 9], bufs[(3 * q + 1) % 9], bufs[(3 * q + 2) % 9]
            if name == "copy":
                z.copy_(x)
            else:
                torch.add(x, y, alpha=-0.5, out=z)
        for q in range(6):
            launch(q)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for q in range(launches):
            launch(q)
        b.record()
        torch.cuda.synchronize()
        us = a.elapsed_time(b) * 1e3 / launches
        out[name] = {"us": round(us, 2), "GB/s": round(nbytes * n_cells / us / 1e3, 1),
                     "bytes_per_cell": nbytes}
    del bufs
    torch.cuda.empty_cache()
    return out


def stream_ceiling_child(n_cells):
    """stream_ceiling in a child process: torch brings a HIP runtime of its own, which
    does not share a process with the library's."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--stream-ceiling",
                        str(n_cells)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        return {"error": r.stderr.strip().splitlines()[-1:] or ["failed"]}
    return json.loads(r.stdout.strip().splitlines()[-1])


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary for this
    workload (FETCH_SIZE + WRITE_SIZE per dispatch, calibrated on a copy kernel), or
    None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("#traffic "):
                rec = json.loads(line[len("#traffic "):])
                if rec.get("kernel") == kernel:
                    return rec.get("bytes_per_launch")
    return None


def box_cores():
    """Threads for the reference's OpenMP run: the host cores this process may use, capped
    by OMP_NUM_THREADS when that is set (the GPU box gives the process of its one GPU a
    16-core share and sets OMP_NUM_THREADS to it)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        cores = min(cores, int(env))
    return max(1, cores)


def _ref_timings(wl, threads, iters_per_step, whole_step):
    """The reference's stages (oracle/_ref) at the workload's grid on `threads` OpenMP
    threads: one pass of each explicit stage, the BiCGSTAB iteration cost from solves
    capped at 2 and 5 iterations, and (whole_step) one whole loop body of runSeqOmp
    (ref_step) with its solve capped at 2 iterations."""
    import numpy as np
    import oracle_py as op
    nx, ny, nz = wl["nx"], wl["ny"], wl["nz"]
    r = op.RefHarness(nx, ny, nz, 20.0, 10.0, 0.5, sml=0, re=wl["re"], dt=wl["dt"],
                      precond=0, tol=1e-300, threads=threads)
    u0 = r.field("u0")
    u0[:] = 1.0 + 1e-3 * (2.0 * np.random.default_rng(1).random(u0.size) - 1.0)
    r.call("eta")
    r.call("start")
    ph = {}

    def timed(name, *calls):
        t = time.perf_counter()
        for c in calls:
            c()
        ph[name] = ph.get(name, 0.0) + time.perf_counter() - t

    timed("conv_diff", lambda: r.call("convdiff"))
    timed("rhs", lambda: r.call("rhs"))
    timed("projection_ib", lambda: r.call("transform"), lambda: r.call("update"))
    timed("copy_vel", lambda: r.call("copyvel"))
    timed("bc", lambda: r.call("bc_main"), lambda: r.call("bc_copy"))
    timed("cd_cl", r.cdcl)
    sol = {}
    for k in (1, 4):  # the solver stops after k + 1 iterations
        r.set_maxit(k)
        t = time.perf_counter()
        r.solve()
        sol[k] = time.perf_counter() - t
    t_iter = (sol[4] - sol[1]) / 3.0
    out = {"threads": threads,
           "phase_s": {k: round(v, 4) for k, v in ph.items()},
           "bicgstab_iteration_s": round(t_iter, 5)}
    out["phase_s"]["poisson"] = round(iters_per_step * t_iter + ph["rhs"], 4)
    out["explicit_s"] = round(sum(ph.values()), 4)
    if whole_step:
        r.set_maxit(1)
        t = time.perf_counter()
        r.step(2)  # loop 2: the pressure solve runs
        out["whole_step_2_iterations_s"] = round(time.perf_counter() - t, 4)
        out["step_s"] = out["whole_step_2_iterations_s"] + (iters_per_step - 2.0) * t_iter
    else:
        out["step_s"] = out["explicit_s"] + iters_per_step * t_iter
    del r
    return out


def _ref_whole_steps_c1(threads, nsteps, budget_s):
    """Whole, uncapped loop bodies of runSeqOmp (ref_step) at the CPU-runnable grid,
    128x64x4, Re 100, dt 0.002, BiCGSTAB on the Jacobi rows at tol 1e-8: the reference's
    own FLOWpassing boundary table (it has no convective, slip or periodic faces)."""
    import numpy as np
    import oracle_py as op
    r = op.RefHarness(128, 64, 4, 20.0, 10.0, 0.5, sml=0, re=100.0, dt=0.002, precond=1,
                      tol=1e-8, threads=threads)
    u0 = r.field("u0")
    u0[:] = 1.0 + 1e-3 * (2.0 * np.random.default_rng(1).random(u0.size) - 1.0)
    r.call("eta")
    r.call("start")
    iters, t0 = [], time.perf_counter()
    for q in range(nsteps):
        iters.append(r.step(q + 1)["iters"])
        if time.perf_counter() - t0 > budget_s:
            break
    wall = time.perf_counter() - t0
    n = len(iters)
    return {"value": round(128 * 64 * 4 * n / wall / 1e6, 4), "unit": "Mcell-steps/s",
            "steps": n, "iters_per_step": iters, "wall_s": round(wall, 3), "threads": threads,
            "note": ("whole reference steps, measured (no model): the first steps of the "
                     "128x64x4 case with the reference's own FLOWpassing table, so its "
                     "iteration counts are the reference's, not the GPU run's")}


def cpu_baseline(wl, iters_per_step, budget_s):
    """The reference's CPU path (oracle/_ref) on this host at the workload's grid, on all
    the cores this process may use and on one thread.  A whole step at the GPU's
    iteration count is far beyond the time budget (minutes per step), so the per-step
    cost is a measured whole loop body with its solve capped at 2 iterations plus the
    remaining iterations at the measured per-iteration cost; the per-phase breakdown
    comes from timing each stage.  Whole uncapped steps are measured at 128x64x4."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as op

    if not op.ref_available():
        return None
    cores = box_cores()
    nx, ny, nz = wl["nx"], wl["ny"], wl["nz"]
    n = nx * ny * nz
    main = _ref_timings(wl, cores, iters_per_step, True)
    one = _ref_timings(wl, 1, iters_per_step, False)
    c1 = _ref_whole_steps_c1(cores, 5, budget_s)
    return {
        "value": n / main["step_s"] / 1e6,
        "unit": "Mcell-steps/s",
        "cores": cores,
        "kind": "reference",
        "sample": (f"reference CPU path (oracle/_ref, OpenMP {cores} threads) at "
                   f"{nx}x{ny}x{nz}: one whole runSeqOmp loop body with its BiCGSTAB solve "
                   f"capped at 2 iterations ({main['whole_step_2_iterations_s']:.2f} s) plus "
                   f"{iters_per_step - 2:.1f} more iterations at the measured "
                   f"{main['bicgstab_iteration_s']:.4f} s each (from solves capped at 2 and 5 "
                   "iterations); the GPU's mean iterations per timed step"),
        "per_phase_s": main["phase_s"],
        "one_thread": {"value": n / one["step_s"] / 1e6, "unit": "Mcell-steps/s",
                       "per_phase_s": one["phase_s"],
                       "bicgstab_iteration_s": one["bicgstab_iteration_s"],
                       "note": "each explicit stage once plus the iterations at their "
                               "measured cost (no whole step at one thread)"},
        "whole_steps_128x64x4": c1,
        "cores_note": ("all cores this process may use, capped by OMP_NUM_THREADS (16 on the "
                       "GPU box: the CPU share of its one GPU)"),
    }


def run_mode(wl, args, precision):
    """W warm-up steps, K timed steps (wall clock), then profiled steps (HIP events)."""
    tag = {0: "FP64", 1: "mixed", 2: "FP64-MG"}[precision]
    s = make_solver(wl, args.precond, args.solver, args.seed, args.maxit, precision)
    warm = []
    for q in range(args.warmup):
        warm.append(s.step(1))
        print(f"{tag} warmup step {q + 1}: {warm[-1]['iters']} iterations", file=sys.stderr,
              flush=True)
    s.synchronize()
    infos = []
    t0 = time.perf_counter()
    for q in range(args.steps):
        infos.append(s.step(1))
        print(f"{tag} timed step {q + 1}: {infos[-1]['iters']} iterations, residual "
              f"{infos[-1]['err']:.3e}", file=sys.stderr, flush=True)
    s.synchronize()
    wall = time.perf_counter() - t0
    # per-kernel device time from HIP events on the library's stream, over further
    # steps of the same run (events are kept out of the wall-clock region)
    s.profile(True)
    for q in range(args.profile_steps):
        s.step(1)
        print(f"{tag} profiled step {q + 1}", file=sys.stderr, flush=True)
    s.synchronize()
    prof = s.profile_read()
    s.profile(False)
    return s, warm, infos, wall, prof


# gradients of the time-averaged drag: 256 x 128 x 32, Re 100, 50 steps, checkpoints
# every 10 steps, 128-point inflow profile 1 +- 0.01 (seeded), smoothed mask of width
# 1.5 cells
ADJOINT_CASE = dict(nx=256, ny=128, nz=32, re=100.0, dt=0.002, tol=1e-10, nsteps=50, K=10)


SLIP_NUM = [[0, 0, 0], [1, 1, 1], [1, 0, 1], [1, 0, 1], [1, 1, 1], [1, 1, 1]]
SLIP_DIR = [[1, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]]

# pressure tolerance of the gradient checks: at tol 1e-10 the solve's own error in the
# smooth pressure modes (up to the residual times the condition number) moves J by more
# than the h * dJ/dR that central differences with h = 1e-6 resolve
GRAD_CHECK_TOL = 1e-13

# the smoothed mask's sine ramp is 1.5 cells wide in all: half-width 0.75 cells on
# either side of the circle
MASK_HALF_WIDTH_CELLS = 0.75


def case_solver(a, seed, precision=0, tol=None, inflow_amp=0.0):
    """A solver of case a (cylinder BCs, smoothed mask at (5, 5), R = 0.5, its ramp 1.5 cells
    wide)
    and its inflow profile (1 + inflow_amp * seeded noise)."""
    import numpy as np
    import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg
    cfg = pkg.default_config(
        nx=a["nx"], ny=a["ny"], nz=a["nz"], lx=20.0, ly=10.0, lz=0.5, re=a["re"], dt=a["dt"],
        tol=tol if tol else a["tol"], maxit=5000, bc_num=SLIP_NUM, bc_dir=SLIP_DIR,
        z_periodic=1, outflow_convective=1, convective_speed=1.0, flux_correction=1,
        skip_first_pressure=1, precision=precision, inner_tol=MIXED_INNER_TOL,
        mg_graphs=MG_GRAPHS)
    s = pkg.FlowSolver(cfg)
    prof = None
    if inflow_amp:
        rng = np.random.default_rng(seed)
        prof = np.ones(a["ny"] + 4)
        prof[2:-2] += inflow_amp * (2.0 * rng.random(a["ny"]) - 1.0)
        s.set_inflow(prof)
    s.init_flow(1.0, 1e-3, seed)
    s.set_eta_smoothed(5.0, 5.0, 0.5, MASK_HALF_WIDTH_CELLS * 20.0 / a["nx"])
    return s, prof


def gradient_check(a, seed, nsteps, K, params, inflow_amp=0.0, h=1e-6, extra_h=()):
    """Adjoint gradients against central differences (step h) of the same objective, both
    with the FP64 multigrid pressure solve at tol GRAD_CHECK_TOL.  params: list of
    ("R" | "cx" | "cy" | inflow index j).  extra_h: further steps for the mask parameters,
    reported with the one-sided differences (J(p+h) - J)/h and (J - J(p-h))/h, which
    differ by more than O(h) when a QUICK upwind switch (a transport velocity crossing
    zero, where the update is not differentiable) lies between p - h and p + h.
    rel_diff is |adjoint - FD| over the largest entry of the same gradient (the mask's
    three, or the inflow profile's); rel_diff_own is over the entry's own size."""
    import numpy as np
    s, prof = case_solver(a, seed, precision=2, tol=GRAD_CHECK_TOL, inflow_amp=inflow_amp)
    eps = MASK_HALF_WIDTH_CELLS * 20.0 / a["nx"]
    J, gi, gm, info = s.adjoint_drag(nsteps, K)
    mask = [5.0, 5.0, 0.5]
    out = {"tol": GRAD_CHECK_TOL, "h": h, "nsteps": nsteps, "checkpoint_every": K, "J": J}
    for q in params:
        if isinstance(q, str):
            m = {"cx": 0, "cy": 1, "R": 2}[q]

            def pm(hh):
                vals = []
                for sgn in (1.0, -1.0):
                    mm = list(mask)
                    mm[m] += sgn * hh
                    s.set_eta_smoothed(mm[0], mm[1], mm[2], eps)
                    vals.append(s.drag_objective(nsteps))
                s.set_eta_smoothed(mask[0], mask[1], mask[2], eps)
                return vals

            vals = pm(h)
            fd, adj, key = (vals[0] - vals[1]) / (2 * h), gm[m], q
            # relative to the largest mask-gradient entry, as the inflow entries are to
            # theirs: dJ/dcy is near zero by the up-down symmetry of the case, so its own
            # size would measure the differences' noise, not the adjoint
            scale = float(np.max(np.abs(gm)))
            steps = {}
            for hh in (h,) + tuple(extra_h):
                v2 = vals if hh == h else pm(hh)
                steps[f"{hh:g}"] = {"central": (v2[0] - v2[1]) / (2 * hh),
                                    "forward": (v2[0] - J) / hh, "backward": (J - v2[1]) / hh}
                print(f"gradient check {q} h {hh:g}: {steps[f'{hh:g}']}", file=sys.stderr,
                      flush=True)
            if extra_h:
                out[f"{q}_by_step"] = steps
        else:
            vals = []
            for sgn in (1.0, -1.0):
                pp = prof.copy()
                pp[q] += sgn * h
                s.set_inflow(pp)
                vals.append(s.drag_objective(nsteps))
            s.set_inflow(prof)
            fd, adj, key = (vals[0] - vals[1]) / (2 * h), gi[q], f"u_in[{q}]"
            scale = float(np.max(np.abs(gi)))  # relative to the largest profile entry
        out[key] = {"adjoint": adj, "central_fd": fd, "rel_diff": abs(adj - fd) / scale,
                    "rel_diff_own": abs(adj - fd) / max(abs(fd), 1e-300)}
        print(f"gradient check {key}: adjoint {adj:.10e} FD {fd:.10e}", file=sys.stderr,
              flush=True)
    s.close()
    out["max_rel_diff"] = max(v["rel_diff"] for v in out.values()
                              if isinstance(v, dict) and "rel_diff" in v)
    return out


def adjoint_run(seed, fd=True):
    import numpy as np
    a = ADJOINT_CASE
    s, prof = case_solver(a, seed, inflow_amp=0.01)
    t = time.perf_counter()
    J0 = s.drag_objective(a["nsteps"])
    plain = time.perf_counter() - t
    print(f"adjoint case: plain forward {plain:.2f} s", file=sys.stderr, flush=True)
    J, gi, gm, info = s.adjoint_drag(a["nsteps"], a["K"])
    print(f"adjoint case: checkpointed forward {info['forward_ms'] / 1e3:.2f} s, reverse "
          f"{info['reverse_ms'] / 1e3:.2f} s", file=sys.stderr, flush=True)
    s.close()
    n = a["nsteps"]
    out = {
        "config": (f"{a['nx']}x{a['ny']}x{a['nz']}, Re={a['re']:g}, dt={a['dt']:g}, "
                   f"{n} steps, checkpoint every {a['K']}, tol {a['tol']:g}, FP64; objective "
                   "mean Cd; parameters: 128-point inflow profile, cylinder (cx, cy, R) of a "
                   "smoothed mask whose ramp is 1.5 cells wide (half-width 0.75 cells)"),
        "J": J, "J_plain_forward": J0,
        "grad_mask": {"cx": gm[0], "cy": gm[1], "R": gm[2]},
        "grad_inflow_l2": float(np.linalg.norm(gi)),
        "forward_ms_per_step": round(plain * 1e3 / n, 3),
        "checkpointed_forward_ms_per_step": round(info["forward_ms"] / n, 3),
        "reverse_ms_per_step": round(info["reverse_ms"] / n, 3),
        "adjoint_step_cost": round(info["reverse_ms"] / (plain * 1e3), 3),
        "adjoint_step_cost_note": ("wall(recompute + adjoint step) / wall(forward step); "
                                   "the reverse pass re-runs every checkpointed segment "
                                   "but the last, which the forward pass tapes, with the "
                                   "pressure solutions the forward pass kept in place of "
                                   "the solves"),
        "gradient_cost": round((info["forward_ms"] + info["reverse_ms"]) / (plain * 1e3), 3),
        "gradient_cost_note": ("wall(checkpointed forward + reverse pass) / wall(plain "
                               "forward): J and both gradients in plain forward runs"),
        "forward_iters": info["forward_iters"], "adjoint_iters": info["adjoint_iters"],
    }
    # the same adjoint with the multigrid pressure solves in FP64 (forward, recompute and
    # pressure adjoint), solved to the same tolerance
    s, _ = case_solver(a, seed, precision=2, inflow_amp=0.01)
    t = time.perf_counter()
    s.drag_objective(n)
    plain_mg = time.perf_counter() - t
    Jm, gim, gmm, infom = s.adjoint_drag(n, a["K"])
    s.close()
    out["fp64_multigrid"] = {
        "forward_ms_per_step": round(plain_mg * 1e3 / n, 3),
        "reverse_ms_per_step": round(infom["reverse_ms"] / n, 3),
        "adjoint_step_cost": round(infom["reverse_ms"] / (plain_mg * 1e3), 3),
        "gradient_cost": round((infom["forward_ms"] + infom["reverse_ms"]) / (plain_mg * 1e3),
                               3),
        "rel_diff_J": abs(Jm - J) / abs(J),
        "rel_diff_grad_mask": float(np.max(np.abs(gmm - gm)) / np.max(np.abs(gm))),
        "rel_diff_grad_inflow": float(np.max(np.abs(gim - gi)) / np.max(np.abs(gi))),
    }
    if fd:
        out["gradient_check"] = gradient_check(a, seed, n, a["K"],
                                               ["R", "cx", "cy", 2 + 20, 2 + 64, 2 + 100],
                                               inflow_amp=0.01)
    return out


# 1024 x 512 x 128 (67M cells), Re 3900, dt 0.0005, 100 steps: mixed-precision forward
# and the drag adjoint w.r.t. the cylinder radius with checkpoints every 20 steps; the
# smoothed mask (1.5 cells) makes the radius differentiable, and the forward uses the
# same mask so that both solve one problem
LARGE_CASE = dict(nx=1024, ny=512, nz=128, re=3900.0, dt=0.0005, tol=1e-10, nsteps=100, K=20)


def hbm_used_bytes():
    """Device memory in use (total - free) from hipMemGetInfo."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        return None, None
    return total.value - free.value, total.value


def large_run(seed, nsteps=None, fd=True):
    import numpy as np
    import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg
    a = dict(LARGE_CASE)
    if nsteps:
        a["nsteps"] = nsteps
    n = a["nsteps"]
    n_cells = a["nx"] * a["ny"] * a["nz"]
    shp = (a["nx"] + 4, a["ny"] + 4, a["nz"] + 4)
    used0, total = hbm_used_bytes()

    def make(precision):
        return case_solver(a, seed, precision=precision)[0]

    def forward(sv, tag):
        sv.synchronize()
        t = time.perf_counter()
        infos = []
        for q in range(n):
            infos.append(sv.step(1))
            if (q + 1) % 10 == 0:
                print(f"large {tag}: step {q + 1}, {infos[-1]['iters']} inner iterations, "
                      f"residual {infos[-1]['err']:.2e}, cd {infos[-1]['cd']:.6f}",
                      file=sys.stderr, flush=True)
        sv.synchronize()
        wall = time.perf_counter() - t
        u = sv.download_velocity(0)
        pr = sv.pressure().reshape(shp)[2:-2, 2:-2, 2:-2]
        return wall, infos, u, pr

    out = {"config": (f"{a['nx']}x{a['ny']}x{a['nz']} uniform over 20x10x0.5, Re={a['re']:g}, "
                      f"dt={a['dt']:g}, {n} steps, smoothed mask (ramp 1.5 cells wide) at (5, 5) R=0.5, "
                      f"convective outflow, slip y walls, periodic z, tol {a['tol']:g}; "
                      f"adjoint of mean Cd w.r.t. R, checkpoint every {a['K']}"),
           "hbm_total_bytes": total}
    # FP64 multigrid: the converged FP64 solution of the same steps
    sg = make(2)
    gwall, ginfos, gu, gp = forward(sg, "FP64-MG")
    sg.close()
    sm = make(1)
    mwall, minfos, mu, mp = forward(sm, "mixed")

    def rel(x, y):
        return float(np.linalg.norm(x - y) / np.linalg.norm(y))

    unorm = float(np.linalg.norm(gu[0]))
    out["mixed_forward"] = {
        "value": round(n_cells * n / mwall / 1e6, 3), "unit": "Mcell-steps/s",
        "ms_per_step": round(mwall / n * 1e3, 3),
        "mean_inner_iters": sum(i["iters"] for i in minfos) / n,
        "errors_vs_fp64_multigrid": {
            "u": rel(mu[0], gu[0]), "v": float(np.linalg.norm(mu[1] - gu[1])) / unorm,
            "w": float(np.linalg.norm(mu[2] - gu[2])) / unorm,
            "p": rel(mp - mp.mean(), gp - gp.mean()),
            "cd_last": abs(minfos[-1]["cd"] - ginfos[-1]["cd"]) / abs(ginfos[-1]["cd"])},
        "cd_last": minfos[-1]["cd"], "cl_last": minfos[-1]["cl"],
    }
    out["fp64_multigrid_forward"] = {
        "value": round(n_cells * n / gwall / 1e6, 3), "unit": "Mcell-steps/s",
        "ms_per_step": round(gwall / n * 1e3, 3),
        "mean_inner_iters": sum(i["iters"] for i in ginfos) / n}
    del gu, gp, mu, mp
    # adjoint from the initial field (a fresh context of the same configuration)
    sm.close()
    sa = make(1)
    used_pre, _ = hbm_used_bytes()
    J, gi, gm, info = sa.adjoint_drag(n, a["K"], mask=True, inflow=False)
    used_adj, _ = hbm_used_bytes()
    print(f"large adjoint: J {J:.10f}, dJ/dR {gm[2]:.8e}, forward {info['forward_ms']:.0f} "
          f"ms, reverse {info['reverse_ms']:.0f} ms", file=sys.stderr, flush=True)
    out["adjoint"] = {
        "J": J, "dJ_dR": gm[2], "dJ_dcx": gm[0], "dJ_dcy": gm[1],
        "checkpointed_forward_ms_per_step": round(info["forward_ms"] / n, 3),
        "reverse_ms_per_step": round(info["reverse_ms"] / n, 3),
        "adjoint_step_cost": round(info["reverse_ms"] / (mwall * 1e3), 3),
        "gradient_cost": round((info["forward_ms"] + info["reverse_ms"]) / (mwall * 1e3), 3),
        "forward_iters": info["forward_iters"], "adjoint_iters": info["adjoint_iters"],
        "hbm_used_bytes": used_adj, "hbm_used_before_adjoint_bytes": used_pre,
        "hbm_used_by_process_at_start_bytes": used0,
    }
    sa.close()
    if fd:
        out["gradient_check"] = gradient_check(a, seed, n, a["K"], ["R"], extra_h=(1e-7, 1e-8))
    return out


def rbsor_run(wl, args):
    """The metric's grid with the second pressure solver: red-black SOR on SorPipeLine's
    coefficients (omega 1.8, tol 1e-10 on max |dp| per sweep), sweeps per step capped at
    args.sor_maxit (the reference's p_sor_iter_max is 5000).  W warm-up and K timed steps
    (wall clock), then one profiled step for the half-sweep kernel's rate."""
    s = make_solver(wl, "none", "rbsor", args.seed, args.sor_maxit, 0)
    for q in range(args.warmup):
        s.step(1)
    s.synchronize()
    infos = []
    t = time.perf_counter()
    for q in range(args.steps):
        infos.append(s.step(1))
        print(f"RB-SOR timed step {q + 1}: {infos[-1]['iters'] + 1} sweeps, max |dp| "
              f"{infos[-1]['err']:.3e}", file=sys.stderr, flush=True)
    s.synchronize()
    wall = time.perf_counter() - t
    s.profile(True)
    s.step(1)
    s.synchronize()
    prof = s.profile_read()
    s.profile(False)
    s.close()
    n_cells = wl["nx"] * wl["ny"] * wl["nz"]
    ms, launches = prof["sor_half"]
    avg_s = ms / 1e3 / launches
    achieved = BYTES_PER_CELL["sor_half"] * n_cells / avg_s / 1e9
    sweeps = [i["iters"] + 1 for i in infos]
    return {
        "value": round(n_cells * args.steps / wall / 1e6, 4), "unit": "Mcell-steps/s",
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "config": (f"{wl['nx']}x{wl['ny']}x{wl['nz']}, Re={wl['re']:g}, dt={wl['dt']:g}, "
                   f"red-black SOR omega 1.8, tol {wl['tol']:g} (max |dp| per sweep), at most "
                   f"{args.sor_maxit} sweeps per step, FP64, colour-split storage"),
        "sweep_cap": args.sor_maxit,
        "sweeps_per_step": sweeps, "max_dp_per_step": [i["err"] for i in infos],
        "steps_reaching_tol": sum(1 for i in infos if i["err"] < wl["tol"]),
        "ms_per_sweep": round(wall / sum(sweeps) * 1e3, 4),
        "roofline": {"bound": "hbm", "kernel": "k_sor_rb2 (one colour's half sweep)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.workload, "sor_half"),
                     "avg_launch_us": round(avg_s * 1e6, 2),
                     "bytes_per_launch": BYTES_PER_CELL["sor_half"] * n_cells,
                     "bytes_note": ("16 B per cell of the grid per half sweep: the other "
                                    "colour read, own p and right-hand side read, own p "
                                    "written (32 B/cell per full sweep)")},
        "kernels": kernel_table(prof, n_cells, BYTES_PER_CELL),
    }


def sustained_run(wl, args, warm=100, steps=20):
    """Both multigrid modes past the start-up: `warm` untimed steps from the seeded field,
    then `steps` timed ones, each mode at its default inner tolerance.  The mixed line's
    own window (the FP64 run's warm-up and timed steps, for the field errors) lies in the
    first steps, where the pressure still changes fast and every solve needs more inner
    iterations."""
    out = {"warmup_steps": warm, "steps": steps}
    for name, precision in (("mixed", 1), ("fp64_multigrid", 2)):
        sv = make_solver(wl, "none", "bicgstab", args.seed, 5000, precision)
        for _ in range(warm):
            sv.step(1)
        sv.synchronize()
        its = []
        t = time.perf_counter()
        for _ in range(steps):
            its.append(sv.step(1)["iters"])
        sv.synchronize()
        wall = time.perf_counter() - t
        sv.close()
        out[name] = {"value": round(wl["nx"] * wl["ny"] * wl["nz"] * steps / wall / 1e6, 4),
                     "unit": "Mcell-steps/s", "ms_per_step": round(wall / steps * 1e3, 3),
                     "mean_inner_iters": sum(its) / steps}
        print(f"sustained {name}: {out[name]['ms_per_step']} ms per step", file=sys.stderr,
              flush=True)
    out["speedup_vs_fp64_multigrid"] = round(out["mixed"]["value"]
                                             / out["fp64_multigrid"]["value"], 3)
    return out


def c3_run(wl, args):
    """The mixed-precision configuration at its full length: 512x256x64, Re 200, dt 0.001,
    args.c3_steps (500) steps from the seeded field, FP32 inner BiCGSTAB to 1e-5 with FP64
    refinement to 1e-10, against the same steps in the FP64 multigrid mode (the same
    algorithm in FP64, converged to tol) and, unless --c3-no-fp64, the FP64 path (the
    reference's unpreconditioned BiCGSTAB, 5000-iteration cap): the Cd/Cl series and the
    final fields."""
    import numpy as np
    n = args.c3_steps
    shp = (wl["nx"] + 4, wl["ny"] + 4, wl["nz"] + 4)
    # the configuration's inner tolerance 1e-5 (also both modes' default)
    modes = [("mixed", 1, 1e-5), ("fp64_multigrid", 2, 1e-5)]
    if not args.c3_no_fp64:
        modes.append(("fp64", 0, None))
    runs = {}
    for name, precision, itol in modes:
        sv = make_solver(wl, "none", "bicgstab", args.seed, 5000, precision, inner_tol=itol)
        sv.synchronize()
        cd, cl, its, errs = [], [], [], []
        t = time.perf_counter()
        for q in range(n):
            info = sv.step(1)
            cd.append(info["cd"])
            cl.append(info["cl"])
            its.append(info["iters"])
            errs.append(info["err"])
            if (q + 1) % 50 == 0:
                print(f"C3 {name}: step {q + 1}, {info['iters']} iterations, residual "
                      f"{info['err']:.2e}, cd {info['cd']:.8f}", file=sys.stderr, flush=True)
        sv.synchronize()
        wall = time.perf_counter() - t
        u = sv.download_velocity(0)
        p = sv.pressure().reshape(shp)[2:-2, 2:-2, 2:-2]
        runs[name] = dict(wall=wall, cd=np.array(cd), cl=np.array(cl), its=its, errs=errs,
                          u=u, p=p - p.mean(), unconverged=info["unconverged"])
        sv.close()
    n_cells = wl["nx"] * wl["ny"] * wl["nz"]

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    out = {"config": (f"{wl['nx']}x{wl['ny']}x{wl['nz']}, Re={wl['re']:g}, dt={wl['dt']:g}, "
                      f"{n} steps from the seeded field; mixed: FP32 inner BiCGSTAB with a "
                      "multigrid preconditioner to 1e-5, FP64 refinement to 1e-10, FP64 dots "
                      "and drag, the correction accumulated in FP64; every multigrid solve "
                      "starts from the extrapolation of the last two")}
    for name, r in runs.items():
        out[name] = {"value": round(n_cells * n / r["wall"] / 1e6, 4), "unit": "Mcell-steps/s",
                     "ms_per_step": round(r["wall"] / n * 1e3, 3),
                     "mean_iters": sum(r["its"]) / n,
                     "steps_above_tol": sum(1 for e in r["errs"][1:] if not e < wl["tol"]),
                     "cd_last": float(r["cd"][-1]), "cl_last": float(r["cl"][-1])}
    for mk in ("mixed",):
        m = runs[mk]
        for ref in ("fp64_multigrid", "fp64"):
            if ref not in runs:
                continue
            r = runs[ref]
            unorm = float(np.linalg.norm(r["u"][0]))
            out[mk][f"errors_vs_{ref}"] = {
                "u": rel(m["u"][0], r["u"][0]),
                "v": float(np.linalg.norm(m["u"][1] - r["u"][1])) / unorm,
                "w": float(np.linalg.norm(m["u"][2] - r["u"][2])) / unorm,
                "p": rel(m["p"], r["p"]),
                "cd_series_max_rel": float(np.max(np.abs(m["cd"] - r["cd"])
                                                  / np.abs(r["cd"]))),
                "cl_series_max_abs_over_max_cl": float(np.max(np.abs(m["cl"] - r["cl"]))
                                                       / np.max(np.abs(r["cl"])))}
        out[mk]["speedup_vs_fp64_multigrid"] = round(
            runs["fp64_multigrid"]["wall"] / m["wall"], 3)
        if "fp64" in runs:
            out[mk]["speedup_vs_fp64"] = round(runs["fp64"]["wall"] / m["wall"], 3)
    out["errors_note"] = ("relative L2 of the final fields (v and w relative to |u|, p with "
                          "its mean removed) and the largest Cd / Cl difference over all "
                          "steps; fp64 stops at its 5000-iteration cap short of tol on most "
                          "steps, so its distance from the converged solution is in its "
                          "column")
    return out


def plot3d_run(wl, args):
    """Plot3D q output of the 512x256x64 mixed-precision run: the staging kernel and the
    copy to pinned host memory overlap the next steps; the file is written at the flush."""
    import tempfile
    s = make_solver(wl, args.precond, args.solver, args.seed, args.maxit, 1)
    for _ in range(3):
        s.step(1)
    s.synchronize()
    n = 3
    t = time.perf_counter()
    for _ in range(n):
        s.step(1)
    s.synchronize()
    plain = (time.perf_counter() - t) / n
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "P3D01.q")
        t = time.perf_counter()
        s.plot3d_q(path, 0.0)
        t_stage = time.perf_counter() - t
        t = time.perf_counter()
        for _ in range(n):
            s.step(1)
        s.synchronize()
        with_out = (time.perf_counter() - t) / n
        t = time.perf_counter()
        s.plot3d_flush()
        t_flush = time.perf_counter() - t
        size = os.path.getsize(path)
        t = time.perf_counter()
        s.plot3d_x(os.path.join(d, "P3D.x"))
        t_x = time.perf_counter() - t
    s.close()
    return {"q_file_bytes": size, "step_ms": round(plain * 1e3, 3),
            "step_ms_with_q_staged": round(with_out * 1e3, 3),
            "stage_call_ms": round(t_stage * 1e3, 3), "flush_ms": round(t_flush * 1e3, 3),
            "x_file_ms": round(t_x * 1e3, 3),
            "note": ("q output started before 3 mixed-precision steps (staging kernel on the "
                     "step stream, device-to-host copy on a second stream); step_ms_with_q_staged "
                     "times those steps alone, stage_call_ms the call that starts the output "
                     "(the first call also allocates the staging buffers); flush = wait + "
                     "file write to a temporary directory")}


def les_run(wl, args):
    """The mixed-precision step at the metric's grid with the Smagorinsky model off and on
    (Cs 0.18, the reference's constant): step time, and the eddy viscosity against 1/Re."""
    import numpy as np
    out = {}
    for name, cs in (("off", 0.0), ("on", 0.18)):
        s = make_solver(wl, args.precond, args.solver, args.seed, args.maxit, 1, smag_cs=cs)
        for _ in range(3):
            s.step(1)
        s.synchronize()
        n = 5
        t = time.perf_counter()
        for _ in range(n):
            s.step(1)
        s.synchronize()
        out[name] = {"step_ms": round((time.perf_counter() - t) / n * 1e3, 3)}
        if cs > 0:
            vis = s.viseff()
            out[name]["viseff_max_over_nu"] = float(vis.max() * wl["re"])
            out[name]["viseff_mean_over_nu"] = float(vis.mean() * wl["re"])
        s.close()
    out["note"] = ("mixed precision, 5 timed steps after 3; 'on' adds the eddy-viscosity "
                   "kernel and the per-cell viscosity in conv-diff")
    return out


def kernel_table(prof, n_cells, bytes_per_cell):
    return {k: {"ms_total": round(v[0], 3), "launches": v[1],
                "avg_us": round(v[0] * 1e3 / v[1], 2) if v[1] else None,
                "GBps": (round(bytes_per_cell[k] * n_cells / (v[0] / 1e3 / v[1]) / 1e9, 1)
                         if v[1] and k in bytes_per_cell else None)}
            for k, v in prof.items() if v[1] > 0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cylinder512", choices=sorted(WORKLOADS))
    ap.add_argument("--precond", default="none", choices=["none", "jacobi"])
    ap.add_argument("--solver", default="bicgstab", choices=["bicgstab", "rbsor"])
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the drag-adjoint run (256x128x32, 50 steps)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision run (FP32 inner BiCGSTAB, FP64 refinement)")
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--profile-steps", type=int, default=1)
    ap.add_argument("--maxit", type=int, default=5000,
                    help="BiCGSTAB iteration cap (5000, the reference default; lower only for "
                         "counter-collection passes, which need few launches)")
    ap.add_argument("--sor-maxit", type=int, default=5000,
                    help="red-black SOR sweeps per step at most (the reference's "
                         "p_sor_iter_max, 5000)")
    ap.add_argument("--c3-steps", type=int, default=500,
                    help="steps of --part c3 (the configuration's 500)")
    ap.add_argument("--c3-no-fp64", action="store_true",
                    help="--part c3: leave out the FP64 BiCGSTAB run (about 1 s per step)")
    ap.add_argument("--no-rbsor", action="store_true",
                    help="skip the red-black SOR run at the metric's grid")
    ap.add_argument("--part", default="all",
                    choices=["all", "mixed", "fp64mg", "adjoint", "large", "plot3d", "les",
                             "rbsor", "c3"],
                    help="run only the mixed-precision or only the adjoint measurement and "
                         "print its JSON object (for profiling one part)")
    ap.add_argument("--large-steps", type=int, default=0,
                    help="steps of --part large (default: the configuration's 100)")
    ap.add_argument("--no-fd", action="store_true",
                    help="--part large / adjoint: skip the finite-difference gradient checks")
    ap.add_argument("--no-mg-graphs", action="store_true",
                    help="launch the multigrid V-cycles kernel by kernel instead of "
                         "replaying captured HIP graphs (for profiler runs)")
    ap.add_argument("--no-stream-ceiling", action="store_true",
                    help="skip the streaming-ceiling child process (under a profiler, whose "
                         "preloaded library would also wrap the child's own HIP runtime)")
    ap.add_argument("--stream-ceiling", type=int, default=0, metavar="CELLS",
                    help="only measure the cache-flushed copy and triad rates over arrays "
                         "of CELLS doubles, print them as JSON and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (the "
                         "calibration kernel for HBM byte counters) and exit")
    args = ap.parse_args()
    global MG_GRAPHS
    MG_GRAPHS = 0 if args.no_mg_graphs else 1

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.stream_ceiling:
        print(json.dumps(stream_ceiling(args.stream_ceiling)), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not run "
              "(multi-GPU x-slab sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    wl = WORKLOADS[args.workload]
    n_cells = wl["nx"] * wl["ny"] * wl["nz"]
    if args.part == "adjoint":
        print(json.dumps({"adjoint": adjoint_run(args.seed, not args.no_fd)}), flush=True)
        return
    if args.part == "plot3d":
        print(json.dumps({"plot3d": plot3d_run(wl, args)}), flush=True)
        return
    if args.part == "les":
        print(json.dumps({"les": les_run(wl, args)}), flush=True)
        return
    if args.part == "rbsor":
        print(json.dumps({"rbsor": rbsor_run(wl, args)}), flush=True)
        return
    if args.part == "c3":
        print(json.dumps({"c3": c3_run(wl, args)}), flush=True)
        return
    if args.part == "large":
        print(json.dumps({"large": large_run(args.seed, args.large_steps or None,
                                             not args.no_fd)}), flush=True)
        return
    if args.part in ("mixed", "fp64mg"):
        sm, mwarm, minfos, mwall, mprof = run_mode(wl, args, 1 if args.part == "mixed" else 2)
        print(json.dumps({args.part: {
            "value": round(n_cells * args.steps / mwall / 1e6, 4), "unit": "Mcell-steps/s",
            "iters_per_step": [i["iters"] for i in minfos],
            "residual_per_step": [i["err"] for i in minfos],
            "kernels": kernel_table(mprof, n_cells, BYTES_PER_CELL_FP32)}}), flush=True)
        sm.close()
        return
    s, warm, infos, wall, prof = run_mode(wl, args, 0)
    value = n_cells * args.steps / wall / 1e6
    iters = [i["iters"] for i in infos]
    mean_iters = sum(iters) / max(1, len(iters))

    # the dominant work by device time during the profiled steps: with BiCGSTAB the
    # iteration, its three phase kernels together (128 B/cell per iteration); otherwise
    # the kernel class with the most device time
    cands = {k: v for k, v in prof.items() if k in BYTES_PER_CELL and v[1] > 0}
    if all(k in cands for k in ITERATION_PHASES):
        iters_run = min(cands[k][1] for k in ITERATION_PHASES)
        avg_s = sum(cands[k][0] for k in ITERATION_PHASES) / 1e3 / iters_run
        alg_bytes = sum(BYTES_PER_CELL[k] for k in ITERATION_PHASES) * n_cells
        dom = "bicgstab_iteration (k_march<0> + k_march<1> + k_march_C)"
        traffic = [measured_traffic(args.workload, k) for k in ITERATION_PHASES]
        traffic = sum(traffic) if all(t is not None for t in traffic) else None
    else:
        k0 = max(cands, key=lambda k: cands[k][0])
        avg_s = cands[k0][0] / 1e3 / cands[k0][1]
        alg_bytes = BYTES_PER_CELL[k0] * n_cells
        dom, traffic = k0, measured_traffic(args.workload, k0)
    achieved = alg_bytes / avg_s / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "avg_launch_us": round(avg_s * 1e6, 2),
                "bytes_per_launch": alg_bytes}
    kernels = kernel_table(prof, n_cells, BYTES_PER_CELL)
    ceiling = None if args.no_stream_ceiling else stream_ceiling_child(n_cells)

    mixed = None
    if not args.no_mixed and args.solver == "bicgstab":
        import numpy as np
        shp = (wl["nx"] + 4, wl["ny"] + 4, wl["nz"] + 4)

        def fields(sv):
            return sv.download_velocity(0), sv.pressure().reshape(shp)[2:-2, 2:-2, 2:-2]

        def rel(a, b):
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))

        def errors(got, ref):
            # v and w relative to |u| (they are small perturbations); pressure is defined
            # up to a constant: interior fields with the mean removed
            unorm = float(np.linalg.norm(ref[0][0]))
            return {"u": rel(got[0][0], ref[0][0]),
                    "v": float(np.linalg.norm(got[0][1] - ref[0][1])) / unorm,
                    "w": float(np.linalg.norm(got[0][2] - ref[0][2])) / unorm,
                    "p": rel(got[1] - got[1].mean(), ref[1] - ref[1].mean())}

        ref64 = fields(s)
        s.close()
        s = None
        # the same algorithm with the inner solve in FP64 (multigrid-preconditioned
        # BiCGSTAB, FP64 refinement): the converged FP64 solution and the precision-only
        # comparison
        sg, gwarm, ginfos, gwall, gprof = run_mode(wl, args, 2)
        gvalue = n_cells * args.steps / gwall / 1e6
        refmg = fields(sg)
        sg.close()
        sm, mwarm, minfos, mwall, mprof = run_mode(wl, args, 1)
        mvalue = n_cells * args.steps / mwall / 1e6
        got = fields(sm)
        sm.close()
        mixed = {
            "value": round(mvalue, 4), "unit": "Mcell-steps/s",
            "ms_per_step": round(mwall / args.steps * 1e3, 3),
            "speedup_vs_fp64": round(mvalue / value, 3),
            "speedup_vs_fp64_multigrid": round(mvalue / gvalue, 3),
            "speedup_note": ("speedup_vs_fp64 is against the FP64 path, whose unpreconditioned "
                             "BiCGSTAB stops at its 5000-iteration cap short of tol on most "
                             "steps: it is mostly algorithmic (multigrid against thousands of "
                             "iterations).  speedup_vs_fp64_multigrid is the precision-only "
                             "figure: the same algorithm in FP64, converged to tol, each mode "
                             "at its own default inner tolerance"),
            "error_gate": {"bound": 1e-6, "against": "fp64_multigrid (converged to tol)",
                           "passed": None},
            "inner_tol": MIXED_INNER_TOL or {"mixed": 1e-5, "fp64_multigrid": 1e-5},
            "tol": wl["tol"],
            "iters_per_step": [i["iters"] for i in minfos],
            "residual_per_step": [i["err"] for i in minfos],
            "errors_vs_fp64": errors(got, ref64),
            "errors_vs_fp64_multigrid": errors(got, refmg),
            "errors_note": ("relative L2 after the same warm-up + timed + profiled steps from "
                            "the same initial field; v and w relative to |u|.  fp64: the "
                            "library's FP64 path (the reference's unpreconditioned BiCGSTAB, "
                            "bitwise equal to the FP64 oracle; it stops at the 5000-iteration "
                            "cap short of tol); fp64_multigrid: the mixed mode's algorithm in "
                            "FP64, converged to tol"),
            "kernels": kernel_table(mprof, n_cells, BYTES_PER_CELL_FP32),
            "sustained": sustained_run(wl, args),
            "fp64_multigrid": {
                "value": round(gvalue, 4), "unit": "Mcell-steps/s",
                "ms_per_step": round(gwall / args.steps * 1e3, 3),
                "iters_per_step": [i["iters"] for i in ginfos],
                "residual_per_step": [i["err"] for i in ginfos],
                "errors_vs_fp64": errors(refmg, ref64),
                "kernels": kernel_table(gprof, n_cells, BYTES_PER_CELL),
            },
        }

    if mixed is not None:
        e = mixed["errors_vs_fp64_multigrid"]
        mixed["error_gate"]["passed"] = bool(max(e["u"], e["v"], e["w"]) <= 1e-6)

    rbsor = None
    if not args.no_rbsor and args.solver == "bicgstab":
        if s is not None:
            s.close()
            s = None
        rbsor = rbsor_run(wl, args)

    adjoint = None
    if not args.no_adjoint and args.workload == "cylinder512":
        if s is not None:
            s.close()
            s = None
        adjoint = adjoint_run(args.seed)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, mean_iters, args.cpu_budget)

    line = {
        "metric": "Mcell-steps/s, FP64 vs mixed @≤1e-6 rel err; % HBM peak; adjoint step cost",
        "value": round(value, 4),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"flow past a cylinder, {wl['nx']}x{wl['ny']}x{wl['nz']} uniform "
                         f"cells over 20x10x0.5, Re={wl['re']:g}, dt={wl['dt']:g}, inflow U=1, "
                         "convective outflow, slip y walls, periodic z, "
                         f"{args.solver} ({args.precond}) tol {wl['tol']:g}, FP64"),
            "grid": [wl["nx"], wl["ny"], wl["nz"]],
            "solver": args.solver, "precond": args.precond, "tol": wl["tol"],
            "maxit": args.maxit,
            "iters_per_step": iters, "mean_iters": mean_iters,
            "residual_per_step": [i["err"] for i in infos],
            "warmup_iters": [i["iters"] for i in warm],
            "cd_last": infos[-1]["cd"], "cl_last": infos[-1]["cl"],
        },
        "roofline": roofline,
        "profile_steps": args.profile_steps,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "stream_ceiling": ceiling,
        "mixed": mixed,
        "rbsor": rbsor,
        "adjoint": adjoint,
    }
    print(json.dumps(line), flush=True)
    if s is not None:
        s.close()


if __name__ == "__main__":
    main()
