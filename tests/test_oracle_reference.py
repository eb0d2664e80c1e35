"""Pins the C oracle to the reference itself.

oracle/_ref/libcfdref.so is the reference's own ConvectionDifussion,
BC_staggered_main/copy, createBMatrix_omp, SPE_matrix / CSR_matrix,
bicgstabRe2, Pressure_transform_X_result, SorPipeLine_seq, update_UandF_omp,
DFIB_CylinderZ and virtualF_Int, compiled from the reference headers by
oracle/Makefile (one OpenMP thread, so math::dot is a plain index-order sum).
The oracle in serial-reduction mode with copyVel semantics must agree bit for bit.
"""
import numpy as np
import pytest

import oracle_py as op

pytestmark = pytest.mark.skipif(not op.ref_available(),
                                reason="reference harness not built (oracle/_ref)")

GRIDS = [
    dict(n=(32, 16, 4), L=(20.0, 10.0, 0.5), kind=0),
    dict(n=(112, 104, 4), L=(17.0, 12.0, 3.0), kind=1),  # the reference's SML generator
]


@pytest.fixture(scope="module")
def oracle():
    return op.Oracle()


@pytest.mark.parametrize("gd", GRIDS, ids=["uniform", "sml"])
def test_grid_metrics(oracle, gd):
    g = oracle.grid(*gd["n"], *gd["L"], kind=gd["kind"])
    r = op.RefHarness(*gd["n"], *gd["L"], sml=gd["kind"])
    for name, arr in r.metrics().items():
        assert np.array_equal(getattr(g, name), arr), name


def _fields(g, seed):
    rng = np.random.default_rng(seed)
    u = 1.0 + 0.2 * rng.standard_normal(g.G)
    v = 0.2 * rng.standard_normal(g.G)
    w = 0.2 * rng.standard_normal(g.G)
    return u, v, w


@pytest.mark.parametrize("gd", GRIDS, ids=["uniform", "sml"])
def test_stages(oracle, gd):
    n, L, kind = gd["n"], gd["L"], gd["kind"]
    g = oracle.grid(*n, *L, kind=kind)
    r = op.RefHarness(*n, *L, sml=kind, re=100.0, dt=0.002)
    u, v, w = _fields(g, 1)
    # ConvectionDifussion
    for nm, a in (("u0", u), ("v0", v), ("w0", w)):
        r.field(nm)[:] = a
    r.call("convdiff")
    us, vs, ws = r.field("u1").copy(), r.field("v1").copy(), r.field("w1").copy()
    ou, ov, ow = r.field("u1").copy(), r.field("v1").copy(), r.field("w1").copy()
    oracle.convdiff(g, 0.002, 1.0 / 100.0, u, v, w, ou, ov, ow)
    assert np.array_equal(ou, us) and np.array_equal(ov, vs) and np.array_equal(ow, ws)
    # BC_staggered_main / BC_staggered_copy with the FLOWpassing table
    p = np.random.default_rng(2).standard_normal(g.G)
    r.field("p")[:] = p
    r.call("bc_main")
    bu, bv, bw, bp = u.copy(), v.copy(), w.copy(), p.copy()
    oracle.bc_main(g, op.make_bc(), bu, bv, bw, bp)
    assert np.array_equal(r.field("u0"), bu) and np.array_equal(r.field("v0"), bv)
    assert np.array_equal(r.field("w0"), bw) and np.array_equal(r.field("p"), bp)
    r.call("bc_copy")
    cu, cv, cw = us.copy(), vs.copy(), ws.copy()
    oracle.bc_copy(g, bu, bv, bw, cu, cv, cw)
    assert np.array_equal(r.field("u1"), cu) and np.array_equal(r.field("v1"), cv)
    assert np.array_equal(r.field("w1"), cw)
    # createBMatrix_omp
    r.call("rhs")
    assert np.array_equal(r.field("b"), oracle.rhs(g, 0.002, cu, cv, cw))
    # update_UandF_omp with DFIB_CylinderZ's mask
    r.call("eta")
    eta = oracle.eta_cylinder_z(g, 6.5, L[1] / 2.0, 0.5, 100)
    assert np.array_equal(r.field("eta"), eta)
    r.call("update")
    t3 = [r.field(n).copy() for n in ("u3", "v3", "w3")]
    o3 = [np.full(g.G, 1.0), np.zeros(g.G), np.zeros(g.G)]  # T3's reference initial state
    f = np.zeros(3 * g.G)
    oracle.update_uf(g, 0.002, bp, eta, cu, cv, cw, *o3, f)
    for a, b in zip(t3, o3):
        assert np.array_equal(a, b)
    assert np.array_equal(r.field("f"), f)
    # CD_CL
    assert r.cdcl() == oracle.cd_cl(g, f, op.RED_SERIAL)


@pytest.mark.parametrize("precond", [op.PRECOND_NONE, op.PRECOND_JACOBI], ids=["spe", "csr"])
def test_operators(oracle, precond):
    g = oracle.grid(12, 10, 6, 17.0, 12.0, 3.0)
    r = op.RefHarness(12, 10, 6, 17.0, 12.0, 3.0, precond=precond)
    x = np.random.default_rng(4).standard_normal(g.N)
    ref = r.apply_A(x)
    assert np.array_equal(oracle.apply_A(g, x, precond), ref)
    if precond == op.PRECOND_NONE:
        # matrix-free rows equal SPE_matrix::multiply_omp over the stored table
        assert np.array_equal(oracle.spe_apply_vals(g, oracle.spe_vals(g), x), ref)


def test_spe_y_swap_on_sml_grid(oracle):
    """The SPE table stores -yp(j) in the -y slot and -ym(j) in the +y slot
    (SPE_sparseMatrix.hpp:79,83); on the SML grid these differ."""
    g = oracle.grid(112, 104, 4, 17.0, 12.0, 3.0, kind=op.GRID_SML)
    r = op.RefHarness(112, 104, 4, 17.0, 12.0, 3.0, sml=1)
    x = np.random.default_rng(5).standard_normal(g.N)
    ref = r.apply_A(x)
    assert np.array_equal(oracle.apply_A(g, x), ref)
    A = oracle.spe_vals(g).reshape(-1, 7)
    assert not np.array_equal(A[:, 1], A[:, 5])


def test_sor_lexicographic(oracle):
    g = oracle.grid(16, 12, 4)
    r = op.RefHarness(16, 12, 4)
    u, v, w = _fields(g, 6)
    p = np.random.default_rng(7).standard_normal(g.G) * 0.01
    for nm, a in (("u1", u), ("v1", v), ("w1", w), ("p", p)):
        r.field(nm)[:] = a
    it_r, err_r = r.sor_seq(1.8, 1e-6, 200)
    po = p.copy()
    it_o, err_o = oracle.sor(g, 0.002, u, v, w, po, omega=1.8, tol=1e-6, itmax=200, redblack=0)
    assert (it_o, err_o) == (it_r, err_r)
    assert np.array_equal(po, r.field("p"))


CASES = [
    dict(n=(48, 24, 4), L=(20.0, 10.0, 0.5), kind=0, precond=0, tol=1e-3, steps=4),
    dict(n=(48, 24, 4), L=(20.0, 10.0, 0.5), kind=0, precond=1, tol=1e-3, steps=4),
    # inconsistent Neumann system at a tight tolerance: ends at the 5000-iteration
    # limit every step and goes through the NaN/Inf reset (bicgstabRe2.hpp:133-153)
    dict(n=(24, 12, 4), L=(20.0, 10.0, 0.5), kind=0, precond=1, tol=1e-9, steps=3),
    dict(n=(112, 104, 4), L=(17.0, 12.0, 3.0), kind=1, precond=0, tol=1e-3, steps=3),
]


@pytest.mark.parametrize("c", CASES, ids=["spe", "jacobi", "jacobi-maxit", "sml"])
def test_time_loop(oracle, c):
    """runSeqOmp's loop body (RUN_CPU.cpp:257-436) against the oracle: every field,
    the iteration count, the residual and Cd/Cl, step by step."""
    n, L, kind, precond, tol = c["n"], c["L"], c["kind"], c["precond"], c["tol"]
    g = oracle.grid(*n, *L, kind=kind)
    r = op.RefHarness(*n, *L, sml=kind, re=100.0, dt=0.002, precond=precond, tol=tol)
    sim = op.Sim(oracle, g, op.make_bc(), 0.002, 100.0, precond=precond, tol=tol,
                 reduction=op.RED_SERIAL, swap=0)
    u, v, w = oracle.init_velocity(g, 1.0, 1e-3, 11)
    for nm, a in (("u0", u), ("v0", v), ("w0", w)):
        sim.field(nm)[:] = a
        r.field(nm)[:] = a
    r.call("eta")
    sim.field("eta")[:] = r.field("eta")
    sim.start()
    r.call("start")
    names = ("u0", "v0", "w0", "u1", "v1", "w1", "u3", "v3", "w3", "p", "f", "x")
    for s in range(1, c["steps"] + 1):
        a, b = sim.step(), r.step(s)
        for key in ("iters", "err", "cd", "cl", "time"):
            assert a[key] == b[key] or (np.isnan(a[key]) and np.isnan(b[key])), (s, key)
        bad = [nm for nm in names
               if not np.array_equal(sim.field(nm), r.field(nm), equal_nan=True)]
        assert bad == [], (s, bad)


@pytest.mark.parametrize("gd", GRIDS, ids=["uniform", "sml"])
def test_steady_state_check(oracle, gd):
    """CheckSteadyStatebyMaxVal_omp (CheckSteadystate.hpp:247-307): the signed maximum
    of T0 - T3 per component over the interior cells, starting at 0.  The reference
    prints it; the harness reads its line back at 17 digits, so the values compare
    exactly."""
    n, L, kind = gd["n"], gd["L"], gd["kind"]
    g = oracle.grid(*n, *L, kind=kind)
    r = op.RefHarness(*n, *L, sml=kind)
    rng = np.random.default_rng(4)
    fields = [1.0 + 0.1 * rng.standard_normal(g.G) for _ in range(6)]
    # make w one-sided (T0 below T3 everywhere) so that its maximum is the starting 0
    fields[5] = fields[2] + 1.0
    for nm, a in zip(("u0", "v0", "w0", "u3", "v3", "w3"), fields):
        r.field(nm)[:] = a
    want = r.maxdiff()
    got = oracle.maxdiff(g, *fields)
    assert got == want, (got, want)
    assert want[2] == 0.0 and want[0] > 0.0 and want[1] > 0.0
