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
                   flushed: the attainable rate the phases are read against.
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
    "bicg_A": 40,    # r r, q, r0;  w p, Ap
    "bicg_B": 24,    # r r, Ap;         w As
    "bicg_C": 72,    # r r, Ap, p, As, r0, x;  w x, r, q
    "bicg_init": 48, # r b, x;  w r0, r, px1
    "convdiff": 48,  # r u, v, w;  w u*, v*, w*
    "rhs": 32,       # r u*, v*, w*;  w b
    "project": 96,   # r u*, v*, w*, X, eta;  w u, v, w, p, f(3)
    "sor_half": 24,  # r p (all), rhs (half);  w p (half)
}

MIXED_INNER_TOL = 0.0  # the library's per-mode default: 3e-3 mixed, 1e-5 FP64 multigrid

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


def make_solver(wl, precond, solver, seed, maxit=5000, precision=0, smag_cs=0.0):
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
        precision=precision, inner_tol=MIXED_INNER_TOL, mg_graphs=MG_GRAPHS,
        smagorinsky_cs=smag_cs)
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


def cpu_baseline(wl, iters_per_step, budget_s):
    """The reference's CPU path (oracle/_ref) on this host: one pass of every explicit
    stage at the full grid, plus a BiCGSTAB solve capped at a few iterations; the
    per-step cost is t_explicit + iters_per_step * t_iteration."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_py as op

    if not op.ref_available():
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    nx, ny, nz = wl["nx"], wl["ny"], wl["nz"]
    r = op.RefHarness(nx, ny, nz, 20.0, 10.0, 0.5, sml=0, re=wl["re"], dt=wl["dt"],
                      precond=0, tol=1e-300, threads=cores)
    u0 = r.field("u0")
    rng = np.random.default_rng(1)
    u0[:] = 1.0 + 1e-3 * (2.0 * rng.random(u0.size) - 1.0)
    r.call("eta")
    r.call("start")
    t = time.perf_counter()
    r.call("convdiff")
    r.call("rhs")
    t_rhs = time.perf_counter()
    r.call("transform")
    r.call("update")
    r.call("copyvel")
    r.call("bc_main")
    r.call("bc_copy")
    r.cdcl()
    t_explicit = time.perf_counter() - t
    # capped solves: k+1 iterations each (the solver breaks after the iteration limit)
    timings = {}
    for k in (1, 4):
        r.set_maxit(k)
        t = time.perf_counter()
        r.solve()
        timings[k] = time.perf_counter() - t
        if time.perf_counter() - t_rhs > budget_s:
            break
    if 4 in timings:
        t_iter = (timings[4] - timings[1]) / 3.0
    else:
        t_iter = timings[1] / 2.0
    step_s = t_explicit + iters_per_step * t_iter
    n = nx * ny * nz
    return {
        "value": n / step_s / 1e6,
        "unit": "Mcell-steps/s",
        "cores": cores,
        "kind": "reference",
        "sample": (f"reference CPU path (oracle/_ref, OpenMP {cores} threads) at "
                   f"{nx}x{ny}x{nz}: one pass of conv-diff, RHS, transform, projection, "
                   f"copy, BCs and Cd/Cl ({t_explicit:.2f} s) plus BiCGSTAB solves capped at "
                   f"2 and 5 iterations ({t_iter:.3f} s per iteration); per-step cost = "
                   f"explicit + {iters_per_step:.1f} iterations (the GPU's mean per timed step)"),
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


def case_solver(a, seed, precision=0, tol=None, inflow_amp=0.0):
    """A solver of case a (cylinder BCs, smoothed mask 1.5 cells wide at (5, 5), R = 0.5)
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
    s.set_eta_smoothed(5.0, 5.0, 0.5, 1.5 * 20.0 / a["nx"])
    return s, prof


def gradient_check(a, seed, nsteps, K, params, inflow_amp=0.0, h=1e-6, extra_h=()):
    """Adjoint gradients against central differences (step h) of the same objective, both
    with the FP64 multigrid pressure solve at tol GRAD_CHECK_TOL.  params: list of
    ("R" | "cx" | "cy" | inflow index j).  extra_h: further steps for the mask parameters,
    reported with the one-sided differences (J(p+h) - J)/h and (J - J(p-h))/h, which
    differ by more than O(h) when a QUICK upwind switch (a transport velocity crossing
    zero, where the update is not differentiable) lies between p - h and p + h."""
    import numpy as np
    s, prof = case_solver(a, seed, precision=2, tol=GRAD_CHECK_TOL, inflow_amp=inflow_amp)
    eps = 1.5 * 20.0 / a["nx"]
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
            scale = abs(fd)
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
        out[key] = {"adjoint": adj, "central_fd": fd, "rel_diff": abs(adj - fd) / scale}
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
                   "smoothed mask 1.5 cells wide"),
        "J": J, "J_plain_forward": J0,
        "grad_mask": {"cx": gm[0], "cy": gm[1], "R": gm[2]},
        "grad_inflow_l2": float(np.linalg.norm(gi)),
        "forward_ms_per_step": round(plain * 1e3 / n, 3),
        "checkpointed_forward_ms_per_step": round(info["forward_ms"] / n, 3),
        "reverse_ms_per_step": round(info["reverse_ms"] / n, 3),
        "adjoint_step_cost": round(info["reverse_ms"] / (plain * 1e3), 3),
        "adjoint_step_cost_note": ("wall(recompute + adjoint step) / wall(forward step); "
                                   "the reverse pass re-runs each checkpointed segment"),
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
                      f"dt={a['dt']:g}, {n} steps, smoothed mask (1.5 cells) at (5, 5) R=0.5, "
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
        "forward_iters": info["forward_iters"], "adjoint_iters": info["adjoint_iters"],
        "hbm_used_bytes": used_adj, "hbm_used_before_adjoint_bytes": used_pre,
        "hbm_used_by_process_at_start_bytes": used0,
    }
    sa.close()
    if fd:
        out["gradient_check"] = gradient_check(a, seed, n, a["K"], ["R"], extra_h=(1e-7, 1e-8))
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
    ap.add_argument("--part", default="all",
                    choices=["all", "mixed", "fp64mg", "adjoint", "large", "plot3d", "les"],
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

    # dominant kernel by device time during the profiled steps
    cands = {k: v for k, v in prof.items() if k in BYTES_PER_CELL and v[1] > 0}
    dom = max(cands, key=lambda k: cands[k][0])
    ms, launches = cands[dom]
    avg_s = ms / 1e3 / launches
    alg_bytes = BYTES_PER_CELL[dom] * n_cells
    achieved = alg_bytes / avg_s / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(args.workload, dom),
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
            "inner_tol": MIXED_INNER_TOL or {"mixed": 3e-3, "fp64_multigrid": 1e-5},
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
            "fp64_multigrid": {
                "value": round(gvalue, 4), "unit": "Mcell-steps/s",
                "ms_per_step": round(gwall / args.steps * 1e3, 3),
                "iters_per_step": [i["iters"] for i in ginfos],
                "residual_per_step": [i["err"] for i in ginfos],
                "errors_vs_fp64": errors(refmg, ref64),
                "kernels": kernel_table(gprof, n_cells, BYTES_PER_CELL),
            },
        }

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
        "adjoint": adjoint,
    }
    print(json.dumps(line), flush=True)
    if s is not None:
        s.close()


if __name__ == "__main__":
    main()
