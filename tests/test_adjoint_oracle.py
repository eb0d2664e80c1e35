"""The CPU reference adjoint of the time-averaged drag (oracle/cfd_adjoint.c).

It has no counterpart in the reference, so it is pinned by mathematics against the
forward oracle (itself pinned bit for bit to the reference): each stage transpose
passes the dot-product test <l, J d> = <J^T l, d> with J d from central differences
of the forward stage, and the whole gradient matches central finite differences of
the objective J = mean_n Cd^n."""
import numpy as np
import pytest

import oracle_py as op


@pytest.fixture(scope="module")
def o():
    return op.Oracle()


def _smooth_field(g, rng, base, amp):
    """A smooth random field (no sign changes of the transport velocities near 0 for
    u ~ 1), so that central differences do not cross an upwind switch."""
    x = np.arange(g.nx)[:, None, None] / g.nx
    y = np.arange(g.ny)[None, :, None] / g.ny
    z = np.arange(g.nz)[None, None, :] / g.nz
    f = base + amp * (np.sin(2 * np.pi * (x + rng.random())) * np.cos(2 * np.pi * (y + rng.random()))
                      + 0.5 * np.sin(2 * np.pi * (z + rng.random())))
    return np.ascontiguousarray(f.reshape(-1))


@pytest.mark.parametrize("zp", [0, 1])
def test_convdiff_transpose_dot_product(o, zp):
    g = o.grid(14, 10, 8)
    rng = np.random.default_rng(3)
    U = [_smooth_field(g, rng, 1.0, 0.3), _smooth_field(g, rng, 0.2, 0.2),
         _smooth_field(g, rng, -0.1, 0.2)]
    D = [rng.standard_normal(g.G) for _ in range(3)]
    Lout = [rng.standard_normal(g.G) for _ in range(3)]
    dt, nu, h = 0.01, 0.01, 1e-6

    def cd(Uin):
        outs = [np.zeros(g.G) for _ in range(3)]
        o.convdiff(g, dt, nu, *Uin, *outs, z_periodic=zp)
        return outs

    # only the computed faces are outputs of the update
    mask = [np.zeros(g.G, bool) for _ in range(3)]
    ones = [np.ones(g.G) for _ in range(3)]
    probe = [np.full(g.G, np.nan) for _ in range(3)]
    o.convdiff(g, dt, nu, *ones, *probe, z_periodic=zp)
    for c in range(3):
        mask[c] = ~np.isnan(probe[c])
        if zp:  # the periodic images written after the update are not update outputs
            m = mask[c].reshape(g.nx, g.ny, g.nz)
            m[:, :, :2] = False
            m[:, :, -2:] = False
        Lout[c] = np.where(mask[c], Lout[c], 0.0)
    plus = cd([U[c] + h * D[c] for c in range(3)])
    minus = cd([U[c] - h * D[c] for c in range(3)])
    lhs = sum(np.dot(Lout[c], (plus[c] - minus[c]) / (2 * h)) for c in range(3))
    LU = o.adj_convdiff(g, dt, nu, *U, *Lout, z_periodic=zp)
    rhs = sum(np.dot(LU[c], D[c]) for c in range(3))
    assert abs(lhs - rhs) <= 1e-7 * abs(rhs), (lhs, rhs)


@pytest.mark.parametrize("preset,advance", [("cylinder", 1), ("cylinder", 0), ("walls", 1)])
def test_bc_transpose_dot_product(o, preset, advance):
    g = o.grid(12, 8, 6)
    rng = np.random.default_rng(5)
    inflow = 1.0 + 0.01 * rng.standard_normal(g.ny)
    if preset == "cylinder":
        bc = op.make_bc(op.SLIP_NUM, op.SLIP_DIR, inflow=inflow, z_periodic=1,
                        outflow_convective=1, conv_speed=1.0, flux_correction=1)
    else:
        bc = op.make_bc(op.FLOWPASSING_NUM, op.FLOWPASSING_DIR, inflow=inflow)
    dt = 0.01
    U = [rng.standard_normal(g.G) for _ in range(3)]
    D = [rng.standard_normal(g.G) for _ in range(3)]
    dinf = rng.standard_normal(g.ny)
    L = [rng.standard_normal(g.G) for _ in range(3)]

    def bcmap(Uin, inf):
        out = [a.copy() for a in Uin]
        inflow[:] = inf
        o.bc_main(g, bc, *out, np.zeros(g.G), advance=advance, dt=dt)
        return out

    base_inf = inflow.copy()
    # the map is affine: the difference of two evaluations is the linear part
    a = bcmap([U[c] + D[c] for c in range(3)], base_inf + dinf)
    b = bcmap(U, base_inf)
    lhs = sum(np.dot(L[c], a[c] - b[c]) for c in range(3))
    LL = [x.copy() for x in L]
    li = o.adj_bc_main(g, bc, *LL, advance=advance, dt=dt)
    rhs = sum(np.dot(LL[c], D[c]) for c in range(3)) + np.dot(li, dinf)
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(rhs)), (lhs, rhs)


def _adjoint_case(o, nxc=24, nyc=12, nzc=4, nsteps=4, skip_first=1):
    g = o.grid(nxc, nyc, nzc)
    rng = np.random.default_rng(7)
    inflow = np.ones(g.ny)
    inflow[2:-2] += 0.01 * (2 * rng.random(nyc) - 1)
    bc = op.make_bc(op.SLIP_NUM, op.SLIP_DIR, inflow=inflow, z_periodic=1,
                    outflow_convective=1, conv_speed=1.0, flux_correction=1)
    prm = op.OrcParams(0.002, 1.0 / 100.0, op.SOLVER_BICGSTAB, op.PRECOND_NONE, 1e-13, 5000,
                       1.8, op.RED_SERIAL, skip_first, 1)
    u, v, w = o.init_velocity(g, 1.0, 1e-3, 20261202)
    dx = 20.0 / nxc
    mask = (5.0, 5.0, 1.0, 1.5 * dx)
    return g, bc, prm, inflow, (u, v, w), mask, nsteps


@pytest.mark.parametrize("skip_first", [1, 0])
def test_adjoint_gradient_matches_finite_differences(o, skip_first):
    g, bc, prm, inflow, U, mask, nsteps = _adjoint_case(o, skip_first=skip_first)
    J, gi, gm = o.adjoint_drag(g, bc, prm, *U, mask, nsteps)
    assert J == o.drag_objective(g, bc, prm, *U, mask, nsteps)
    h = 1e-6
    for q, name in enumerate(("cx", "cy", "R")):
        mp, mm = list(mask), list(mask)
        mp[q] += h
        mm[q] -= h
        fd = (o.drag_objective(g, bc, prm, *U, tuple(mp), nsteps) -
              o.drag_objective(g, bc, prm, *U, tuple(mm), nsteps)) / (2 * h)
        print(f"d J / d {name}: adjoint {gm[q]:.10e}  central FD {fd:.10e}")
        assert abs(gm[q] - fd) <= 1e-5 * max(abs(fd), 1e-3 * abs(J)), (name, gm[q], fd)
    base = inflow.copy()
    for j in (2, 5, g.ny // 2, g.ny - 3):
        inflow[:] = base
        inflow[j] += h
        jp = o.drag_objective(g, bc, prm, *U, mask, nsteps)
        inflow[:] = base
        inflow[j] -= h
        jm = o.drag_objective(g, bc, prm, *U, mask, nsteps)
        inflow[:] = base
        fd = (jp - jm) / (2 * h)
        print(f"d J / d u_in[{j}]: adjoint {gi[j]:.10e}  central FD {fd:.10e}")
        assert abs(gi[j] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.max(np.abs(gi))), (j, gi[j], fd)
    # ghost entries of the profile are not read by the forward map
    assert gi[0] == gi[1] == gi[-1] == gi[-2] == 0.0
