"""Shared set-up for the GPU parity tests: one HIP context and one oracle run on
identical inputs."""
import numpy as np

import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg
import oracle_py as op

# boundary-condition presets: (num, dir, z_periodic, convective, flux_correction)
BC_PRESETS = {
    "flowpassing": (op.FLOWPASSING_NUM, op.FLOWPASSING_DIR, 0, 0, 0),
    "slip": (op.SLIP_NUM, op.SLIP_DIR, 0, 0, 0),
    # inflow U=1, convective outflow with flux correction, slip y walls, periodic z
    "cylinder": (op.SLIP_NUM, op.SLIP_DIR, 1, 1, 1),
    # the same without the flux correction: the convective update alone sets the outlet
    "convective": (op.SLIP_NUM, op.SLIP_DIR, 1, 1, 0),
}


def make_gpu(nxc, nyc, nzc, *, lengths=(20.0, 10.0, 0.5), kind=0, re=100.0, dt=0.002,
             solver=0, precond=0, tol=1e-6, maxit=5000, omega=1.8, bc="flowpassing",
             seed=20261202, amp=1e-3, cyl=(5.0, 5.0, 0.5), xcd_remap=1, smoothed_eps=None,
             skip_first=1, inflow=None, precision=0, inner_tol=1e-5, mg_graphs=1,
             smag_cs=0.0, steady_check=0, mg_extrapolate=1):
    """The device context of a case, with its seeded initial field and mask."""
    num, dirr, zper, conv, flux = BC_PRESETS[bc]
    cfg = pkg.default_config(nx=nxc, ny=nyc, nz=nzc, lx=lengths[0], ly=lengths[1],
                             lz=lengths[2], grid_kind=kind, re=re, dt=dt, solver=solver,
                             precond=precond, tol=tol, maxit=maxit, sor_omega=omega,
                             bc_num=num, bc_dir=dirr, z_periodic=zper, outflow_convective=conv,
                             convective_speed=1.0, flux_correction=flux,
                             skip_first_pressure=skip_first, xcd_remap=xcd_remap,
                             precision=precision, inner_tol=inner_tol, mg_graphs=mg_graphs,
                             smagorinsky_cs=smag_cs, steady_check=steady_check,
                             mg_extrapolate=mg_extrapolate)
    gpu = pkg.FlowSolver(cfg)
    if inflow is not None:
        gpu.set_inflow(inflow)
    gpu.init_flow(1.0, amp, seed)
    if cyl is not None:
        if smoothed_eps is None:
            gpu.set_eta_cylinder_z(*cyl, 100)
        else:
            gpu.set_eta_smoothed(*cyl, smoothed_eps)
    return gpu


def make_pair(nxc, nyc, nzc, *, lengths=(20.0, 10.0, 0.5), kind=0, re=100.0, dt=0.002,
              solver=0, precond=0, tol=1e-6, maxit=5000, omega=1.8, bc="flowpassing",
              seed=20261202, amp=1e-3, cyl=(5.0, 5.0, 0.5), xcd_remap=1, smoothed_eps=None,
              skip_first=1, inflow=None, precision=0, inner_tol=1e-5, mg_graphs=1,
              smag_cs=0.0, steady_check=0, mg_extrapolate=1):
    """The device context of a case and the oracle's run of the same case."""
    num, dirr, zper, conv, flux = BC_PRESETS[bc]
    gpu = make_gpu(nxc, nyc, nzc, lengths=lengths, kind=kind, re=re, dt=dt, solver=solver,
                   precond=precond, tol=tol, maxit=maxit, omega=omega, bc=bc, seed=seed,
                   amp=amp, cyl=cyl, xcd_remap=xcd_remap, smoothed_eps=smoothed_eps,
                   skip_first=skip_first, inflow=inflow, precision=precision,
                   inner_tol=inner_tol, mg_graphs=mg_graphs, smag_cs=smag_cs,
                   steady_check=steady_check, mg_extrapolate=mg_extrapolate)
    o = op.Oracle()
    g = o.grid(nxc, nyc, nzc, *lengths, kind=kind)
    obc = op.make_bc(num, dirr, inflow=inflow, z_periodic=zper, outflow_convective=conv,
                     conv_speed=1.0, flux_correction=flux)
    osolver = {0: op.SOLVER_BICGSTAB, 1: op.SOLVER_RBSOR}[solver]
    sim = op.Sim(o, g, obc, dt, re, solver=osolver, precond=precond, tol=tol, maxit=maxit,
                 omega=omega, reduction=op.RED_BLOCKED, skip_first=skip_first, swap=1,
                 smag_cs=smag_cs)
    u, v, w = o.init_velocity(g, 1.0, amp, seed)
    sim.field("u0")[:] = u
    sim.field("v0")[:] = v
    sim.field("w0")[:] = w
    if cyl is not None:
        if smoothed_eps is None:
            sim.field("eta")[:] = o.eta_cylinder_z(g, *cyl, 100)
        else:
            sim.field("eta")[:] = o.eta_smoothed(g, *cyl, smoothed_eps)
    sim.start()
    return gpu, sim, o, g


def fields_equal(gpu, sim):
    """Compare every device field with the oracle's; returns the list of mismatches."""
    bad = []
    for which, names in ((0, ("u0", "v0", "w0")), (1, ("u1", "v1", "w1")),
                         (2, ("u3", "v3", "w3"))):
        got = gpu.download_velocity(which)
        for a, n in zip(got, names):
            if not np.array_equal(a, sim.field(n), equal_nan=True):
                bad.append(n)
    for name, got in (("p", gpu.pressure()), ("f", gpu.force()), ("x", gpu.xresult())):
        if not np.array_equal(got, sim.field(name), equal_nan=True):
            bad.append(name)
    return bad


def max_rel(a, b):
    d = np.max(np.abs(a - b))
    s = np.max(np.abs(b))
    return d / s if s > 0 else d
