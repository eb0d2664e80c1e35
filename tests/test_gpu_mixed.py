"""Multigrid pressure solves: a BiCGSTAB preconditioned by a multigrid V-cycle inside
FP64 iterative refinement of the pressure residual, with the inner solve in FP32 (mixed
precision: float vectors, FP64 sums) or in FP64 (the same algorithm in FP64).  They have
no counterpart in the reference, so they are checked against the FP64 oracle to the
tolerances the modes promise: the pressure system's FP64 residual, recomputed on the CPU
with the oracle's operator, is below the requested tolerance, and velocities, pressure
and Cd/Cl after several time steps agree with the FP64 oracle run to 1e-6 relative L2
(mixed) or 1e-8 (FP64 multigrid; both solve to tol 1e-10, so what is left is the
difference between two solutions within that tolerance).  The multigrid preconditioner
is checked by what it must deliver: a few inner iterations per solve, independent of
the grid size."""
import numpy as np
import pytest

import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg
import oracle_py as op
from cfd_nssolver_forhpc_flowpassingcylinder_amd.api import PRECISION_FP64_MG, PRECISION_MIXED
from gpu_helpers import make_pair

pytestmark = pytest.mark.gpu

# velocity / pressure / Cd, Cl error bound of each multigrid mode
REL_TOL = {PRECISION_MIXED: 1e-6, PRECISION_FP64_MG: 1e-8}
PRECISIONS = [PRECISION_MIXED, PRECISION_FP64_MG]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("n", [(48, 24, 8), (64, 32, 16), (128, 64, 32)])
def test_mixed_solve_reaches_fp64_residual(n, precision):
    tol, precond = 1e-10, 0
    gpu, sim, o, g = make_pair(*n, bc="cylinder", precond=precond, tol=tol,
                               precision=precision)
    gpu.convdiff()
    o.convdiff(g, 0.002, 1.0 / 100.0, sim.field("u0"), sim.field("v0"), sim.field("w0"),
               sim.field("u1"), sim.field("v1"), sim.field("w1"), z_periodic=1)
    gpu.pressure_rhs()
    b = o.rhs(g, 0.002, sim.field("u1"), sim.field("v1"), sim.field("w1"), precond, 1)
    assert np.array_equal(gpu.rhs(), b)  # the RHS stays on the FP64 path
    for rep in range(2):  # cold start, then a warm start from the converged x
        it, res = gpu.solve_bicgstab()
        x = gpu.xresult()
        r = b - o.apply_A(g, x, precond, 1)
        rel = np.linalg.norm(r) / np.linalg.norm(b)
        print(f"{n} solve {rep}: {it} inner iterations, residual {rel:.3e}")
        assert rel < tol, (rep, it, rel)
        assert it <= 40, it  # multigrid-preconditioned: a few iterations per outer pass
        assert np.isclose(res, rel, rtol=1e-6, atol=0.0), (res, rel)
    # the FP64 oracle's own solution of the same system satisfies the same bound; the
    # two solutions differ by at most a multiple of the tolerance in the operator norm
    xo = np.zeros(g.N)
    op.BiCG(o, g).solve(b, xo, tol, precond, 1, op.RED_BLOCKED)
    d = o.apply_A(g, x - xo, precond, 1)
    assert np.linalg.norm(d) / np.linalg.norm(b) < 2.5 * tol
    gpu.close()


# consistent pressure systems only (flux-corrected outlet, periodic z): at tol 1e-10 an
# inconsistent system ends at the iteration limit in either precision
# inner_tol 0: the mode's default (1e-5 in both modes)
@pytest.mark.parametrize("inner_tol", [1e-5, 3e-3, 0.0])
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("kind,n,precond,steps", [(0, (48, 24, 8), 0, 6),
                                                  (0, (64, 32, 16), 0, 4)])
def test_mixed_time_steps_against_fp64_oracle(kind, n, precond, steps, precision, inner_tol):
    tol = 1e-10
    gpu, sim, o, g = make_pair(*n, bc="cylinder", kind=kind, precond=precond, tol=tol,
                               precision=precision, inner_tol=inner_tol)
    for q in range(steps):
        info = gpu.step(1)
        out = sim.step()
        assert info["iters"] <= 5000 and info["err"] < tol, (q, info)
        print(f"step {q + 1}: mixed {info['iters']} inner iterations (res {info['err']:.2e}); "
              f"FP64 oracle {out['iters']} iterations")
    errs = {}
    for name, got in zip(("u", "v", "w"), gpu.download_velocity(0)):
        errs[name] = rel_l2(got, sim.field(name + "0"))
    # pressure is defined up to a constant (singular Neumann/periodic system): compare
    # the interior fields with their means removed
    sl = g.interior()
    pg = gpu.pressure().reshape(g.shape())[sl]
    po = sim.field("p").reshape(g.shape())[sl]
    errs["p"] = rel_l2(pg - pg.mean(), po - po.mean())
    cd, cl, ocd, ocl = info["cd"], info["cl"], out["cd"], out["cl"]
    errs["cd"] = abs(cd - ocd) / abs(ocd)
    print("relative errors vs FP64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()})
    bound = REL_TOL[precision]
    for k in ("u", "p", "cd"):
        assert errs[k] <= bound, (k, errs)
    # v, w and Cl are O(1e-3) perturbations here; bound them against |u|
    for name, got in zip(("v", "w"), gpu.download_velocity(0)[1:]):
        ref = sim.field(name + "0")
        assert np.linalg.norm(got - ref) / np.linalg.norm(sim.field("u0")) <= bound
    assert abs(cl - ocl) <= bound * abs(ocd)
    gpu.close()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_mixed_configuration_limits(precision):
    import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg
    from cfd_nssolver_forhpc_flowpassingcylinder_amd.api import CfdmxError
    # multigrid levels are uniform SPE Laplacians: no Jacobi rows, no SML grid
    for kw in (dict(precond=1), dict(grid_kind=1, nx=112, ny=104)):
        cfg = dict(nx=48, ny=24, nz=8, precision=precision)
        cfg.update(kw)
        with pytest.raises(CfdmxError):
            pkg.FlowSolver(pkg.default_config(**cfg))
    # a hierarchy that cannot reach 512 cells is refused at creation
    with pytest.raises(CfdmxError, match="multigrid"):
        pkg.FlowSolver(pkg.default_config(nx=45, ny=27, nz=7, precision=precision))
    # inner_tol: 0 (the default) takes the mode's own value; outside [0, 1) is refused
    for bad in (-1e-3, 1.0):
        with pytest.raises(CfdmxError, match="inner_tol"):
            pkg.FlowSolver(pkg.default_config(nx=48, ny=24, nz=8, precision=precision,
                                              inner_tol=bad))
    pkg.FlowSolver(pkg.default_config(nx=48, ny=24, nz=8, precision=precision,
                                      inner_tol=0.0)).close()


def test_odd_nz_levels_take_the_per_cell_kernels():
    """nz = 5 (lz chosen so that all spacings are equal): level 48x24x5 has odd nz, so the
    V-cycle runs its one-thread-per-cell smoother and prolongation (coarsest level
    12x6x5, solved exactly); the solve must still reach the FP64 tolerance."""
    tol = 1e-10
    gpu, sim, o, g = make_pair(48, 24, 5, lengths=(20.0, 10.0, 20.0 * 5 / 48), bc="cylinder",
                               tol=tol, precision=PRECISION_MIXED)
    gpu.convdiff()
    gpu.pressure_rhs()
    o.convdiff(g, 0.002, 1.0 / 100.0, sim.field("u0"), sim.field("v0"), sim.field("w0"),
               sim.field("u1"), sim.field("v1"), sim.field("w1"), z_periodic=1)
    b = o.rhs(g, 0.002, sim.field("u1"), sim.field("v1"), sim.field("w1"), 0, 1)
    it, res = gpu.solve_bicgstab()
    r = b - o.apply_A(g, gpu.xresult(), 0, 1)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    print(f"48x24x5: {it} inner iterations, residual {rel:.3e}")
    assert rel < tol and it <= 60, (it, rel)
    gpu.close()


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("n,lz", [((48, 24, 5), 20.0 * 5 / 48), ((96, 48, 6), 1.25),
                                  ((64, 32, 8), 2.5)])
def test_multigrid_runs_are_deterministic(precision, n, lz):
    """Two runs from the same state give the same bits.  With periodic z and odd nz the
    red-black colouring joins two cells of one colour across the wrap (k = 0 and
    k = nz - 1); the smoother updates the last plane in a launch of its own so that no
    cell is read while it is written.  48x24x5 has odd nz on every level, 96x48x6 from
    its second level on; 64x32x8 has even nz down to one plane and is stored by colour."""
    runs = []
    for _ in range(2):
        gpu, *_ = make_pair(*n, lengths=(20.0, 10.0, lz), bc="cylinder", tol=1e-9,
                            precision=precision)
        infos = [gpu.step(1) for _ in range(3)]
        runs.append((infos, gpu.download_velocity(0), gpu.pressure(), gpu.xresult()))
        gpu.close()
    (ia, ua, pa, xa), (ib, ub, pb, xb) = runs
    assert ia == ib
    for a, b in zip(ua, ub):
        assert np.array_equal(a, b)
    assert np.array_equal(pa, pb) and np.array_equal(xa, xb)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_extrapolated_start_reaches_the_same_solution(precision):
    """mg_extrapolate = 1 (the default) starts each pressure solve from 2 x_n - x_(n-1)
    from the fourth step on (the first has no solve, the next two give the history).  The
    solves still end at tol, so the fields agree with the runs started from x_n to the
    solve's accuracy.  (Fewer iterations is a property of the developed flow, measured at
    512x256x64; in these first steps from the seeded field the counts are alike.)"""
    runs = []
    for ex in (0, 1):
        gpu, sim, o, g = make_pair(64, 32, 16, bc="cylinder", tol=1e-10, dt=0.001,
                                   precision=precision, mg_extrapolate=ex)
        infos = [gpu.step(1) for _ in range(12)]
        assert all(i["err"] < 1e-10 for i in infos[1:]), infos
        p = gpu.pressure().reshape(g.shape())[2:-2, 2:-2, 2:-2]
        runs.append((sum(i["iters"] for i in infos), gpu.download_velocity(0), p - p.mean()))
        gpu.close()
    (ia, ua, pa), (ib, ub, pb) = runs
    assert abs(ib - ia) <= 0.25 * ia, (ia, ib)
    for a, b in zip(ua, ub):
        assert np.linalg.norm(b - a) <= 1e-10 * np.linalg.norm(ua[0])
    assert np.linalg.norm(pb - pa) <= 1e-7 * np.linalg.norm(pa)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("n,lz", [((64, 32, 8), 2.5), ((48, 24, 5), 20.0 * 5 / 48)])
def test_graph_replay_matches_launched_cycles(precision, n, lz):
    """The V-cycle replayed from its captured HIP graph (the default) gives the bits of the
    cycle launched kernel by kernel.  Each step replays both graphs (one per level-0
    binding) once per inner iteration, with the Krylov kernels and a host
    synchronisation between the replays, and every step is compared."""
    runs = []
    for graphs in (0, 1):
        gpu, *_ = make_pair(*n, lengths=(20.0, 10.0, lz), bc="cylinder", tol=1e-9,
                            precision=precision, mg_graphs=graphs)
        steps = []
        for _ in range(4):
            info = gpu.step(1)
            steps.append((info, gpu.download_velocity(0), gpu.xresult()))
        runs.append(steps)
        gpu.close()
    for (ia, ua, xa), (ib, ub, xb) in zip(*runs):
        assert ia == ib
        for a, b in zip(ua, ub):
            assert np.array_equal(a, b)
        assert np.array_equal(xa, xb)


@pytest.mark.parametrize("precision,max_inner", [(PRECISION_FP64_MG, 2), (PRECISION_MIXED, 6)])
@pytest.mark.parametrize("n,lengths", [((16, 8, 4), (20.0, 10.0, 5.0)),
                                       ((9, 7, 5), (9.0, 7.0, 5.0))])
def test_single_level_is_the_exact_solve(precision, max_inner, n, lengths):
    """A grid of at most 512 cells is its own coarsest level, which the V-cycle solves
    exactly by fast diagonalisation (cosine modes in x and y, Fourier modes in periodic z,
    even and odd sizes).  The preconditioned BiCGSTAB then needs one iteration per
    refinement pass, and the FP64 residual, recomputed with the oracle's operator, meets
    the tolerance."""
    tol = 1e-10
    gpu, sim, o, g = make_pair(*n, lengths=lengths, bc="cylinder", tol=tol, precision=precision)
    gpu.convdiff()
    gpu.pressure_rhs()
    o.convdiff(g, 0.002, 1.0 / 100.0, sim.field("u0"), sim.field("v0"), sim.field("w0"),
               sim.field("u1"), sim.field("v1"), sim.field("w1"), z_periodic=1)
    b = o.rhs(g, 0.002, sim.field("u1"), sim.field("v1"), sim.field("w1"), 0, 1)
    it, res = gpu.solve_bicgstab()
    r = b - o.apply_A(g, gpu.xresult(), 0, 1)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    print(f"{n}: {it} inner iterations, residual {rel:.3e}")
    assert rel < tol and it <= max_inner, (it, rel)
    gpu.close()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_refinement_that_cannot_converge_is_reported(precision):
    """Without the flux correction the singular system is inconsistent: the FP64
    refinement cannot reach tol 1e-10, ends (pass limit, residual rise or maxit) with
    err >= tol, and the step report counts the solve as unconverged; the reported err
    is the FP64 residual of the x the solve returns (recomputed on the CPU)."""
    gpu, sim, o, g = make_pair(48, 24, 8, bc="convective", tol=1e-10, precision=precision,
                               maxit=200)
    gpu.step(1)
    info = gpu.step(1)
    print(info)
    assert info["err"] >= 1e-10 and info["unconverged"] == 1
    b = gpu.rhs()
    r = b - o.apply_A(g, gpu.xresult(), 0, 1)
    assert np.isclose(np.linalg.norm(r) / np.linalg.norm(b), info["err"], rtol=1e-6)
    gpu.close()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_a_pass_that_raises_the_residual_is_dropped(precision):
    """A digest case with walled z and no outlet flux correction: the system is
    inconsistent, the inner solves stall, and a refinement pass can raise the FP64 residual.
    Such a pass is dropped, not applied.  Every solve ends with an x whose residual is no
    higher than its start's. The err it reports is that x's own, formed again here with the
    library's operator. A solve whose first pass is dropped returns its start bit for bit.
    Without the extrapolated start, each solve starts from the solution before it."""
    from cfd_nssolver_forhpc_flowpassingcylinder_amd.api import PRECOND_NONE, SOLVER_BICGSTAB
    slip_num = [[0, 0, 0], [1, 1, 1], [1, 0, 1], [1, 0, 1], [1, 1, 1], [1, 1, 1]]
    slip_dir = [[1, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]]
    cfg = pkg.default_config(nx=64, ny=32, nz=16, lx=20.0, ly=10.0, lz=0.5, re=200.0,
                             dt=0.001, tol=1e-10, maxit=5000, precond=PRECOND_NONE,
                             solver=SOLVER_BICGSTAB, bc_num=slip_num, bc_dir=slip_dir,
                             z_periodic=0, outflow_convective=1, convective_speed=1.0,
                             flux_correction=0, skip_first_pressure=1, precision=precision,
                             inner_tol=3e-3 if precision == 1 else 1e-5, mg_extrapolate=0)
    s = pkg.FlowSolver(cfg)
    s.init_flow(1.0, 1e-3, 11)
    s.set_eta_cylinder_z(5.0, 5.0, 0.5, 100)

    def rel(b, x):
        return np.linalg.norm(b - s.apply_operator(x)) / np.linalg.norm(b)

    solved = 0
    for n in range(5):
        x0 = s.xresult()
        info = s.step(1)
        if not info["solved"]:
            continue
        solved += 1
        b, x1 = s.rhs(), s.xresult()
        r0, r1 = rel(b, x0), rel(b, x1)
        print(f"step {n + 1}: start {r0:.3e}, end {r1:.3e}, reported {info['err']:.3e}, "
              f"{info['iters']} inner iterations")
        assert np.isclose(r1, info["err"], rtol=1e-9)
        assert r1 <= r0 * (1.0 + 1e-9)
        if np.isclose(r1, r0, rtol=1e-12):
            assert np.array_equal(x1, x0)
    assert solved == 4
    s.close()


def test_multigrid_fields_match_the_committed_digests():
    """Both multigrid modes on six grids (every V-cycle kernel variant) give the fields
    whose digests tests/golden/multigrid_digests.json holds, bit for bit."""
    import json
    import os
    import sys
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    import make_multigrid_digests as mk
    ref = json.load(open(os.path.join(here, "multigrid_digests.json")))
    assert ref["steps"] == mk.STEPS and ref["seed"] == mk.SEED
    bad = []
    for c in mk.CASES:
        for precision in (PRECISION_MIXED, PRECISION_FP64_MG):
            k = mk.key(*c, precision)
            got = mk.run_case(*c, precision)
            print(k, got)
            if got != ref["digests"][k]:
                bad.append(k)
    assert not bad, bad
