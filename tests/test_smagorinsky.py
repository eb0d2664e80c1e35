"""Smagorinsky LES (SmagorinskyModel.hpp:6-79; the eddy viscosity enters the diffusion at
Convection_Difussion.hpp:251-255,334-338,414-418).  The reference's model does not compile
(SmagorinskyModel.hpp:61) and is switched off in its build, so it cannot be run: the
oracle restates its formulas and is pinned here by closed forms (a linear shear, the
viscosity's scaling, conv-diff with a constant eddy viscosity equal to conv-diff at the
summed viscosity), and the GPU is checked against the oracle bit for bit."""
import numpy as np
import pytest

import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg
import oracle_py as op
from cfd_nssolver_forhpc_flowpassingcylinder_amd.api import CfdmxError
from gpu_helpers import fields_equal, make_pair

CS = 0.18  # SmagorinskyModel.hpp:38


@pytest.fixture(scope="module")
def o():
    return op.Oracle()


def _shear(g, a, axis_from, comp):
    """velocity component comp = a * coordinate along axis_from (of that component's own
    positions), zero elsewhere"""
    nx, ny, nz = g.shape()
    f = [np.zeros((nx, ny, nz)) for _ in range(3)]
    n = (nx, ny, nz)[axis_from]
    D = (g.Dx, g.Dy, g.Dz)[axis_from]
    x = np.concatenate([[0.0], np.cumsum(D[:-1])])
    if axis_from != comp:
        x = x + 0.5 * D  # cell centres along the transverse axis
    shape = [1, 1, 1]
    shape[axis_from] = n
    f[comp] = np.broadcast_to(a * x.reshape(shape), (nx, ny, nz)).copy()
    return [np.ascontiguousarray(c.ravel()) for c in f]


@pytest.mark.parametrize("comp,axis_from", [(0, 1), (0, 2), (1, 0), (2, 1)])
def test_linear_shear_closed_form(o, comp, axis_from):
    """A uniform shear du_c/dx_a = a has |temp| = a^2 (one cross term), so Viseff =
    (Cs delta)^2 |a| in every cell, with delta the cube root of the cell volume."""
    g = o.grid(12, 10, 8)
    a = 0.7
    u, v, w = _shear(g, a, axis_from, comp)
    vis = o.smagorinsky(g, CS, u, v, w)
    delta = (g.Dx[2] * g.Dy[2] * g.Dz[2]) ** (1.0 / 3.0)
    assert np.allclose(vis, (CS * delta) ** 2 * a, rtol=1e-12, atol=0)


def test_normal_strain_weights(o):
    """u = a x alone gives temp = 2 a^2 (the factor 2 sits on the u term only, as written);
    v = a y alone gives temp = a^2; the viscosity scales with Cs^2."""
    g = o.grid(12, 10, 8)
    a = 0.3
    delta = (g.Dx[2] * g.Dy[2] * g.Dz[2]) ** (1.0 / 3.0)
    u, v, w = _shear(g, a, 0, 0)
    assert np.allclose(o.smagorinsky(g, CS, u, v, w), (CS * delta) ** 2 * np.sqrt(2) * a,
                       rtol=1e-12)
    u, v, w = _shear(g, a, 1, 1)
    assert np.allclose(o.smagorinsky(g, CS, u, v, w), (CS * delta) ** 2 * a, rtol=1e-12)
    assert np.allclose(o.smagorinsky(g, 2 * CS, u, v, w), 4 * o.smagorinsky(g, CS, u, v, w),
                       rtol=1e-12)


def test_convdiff_with_constant_eddy_viscosity(o):
    """conv-diff with Viseff = 0 is conv-diff; with a constant Viseff = c it is conv-diff at
    nu + c, bit for bit (the reference adds Viseff to nu before multiplying by dt)."""
    g = o.grid(16, 12, 6)
    rng = np.random.default_rng(3)
    u, v, w = (1.0 + 0.1 * rng.standard_normal(g.G) for _ in range(3))
    nu, c, dt = 0.01, 0.0037, 0.002
    outs = []
    for nu_, vis in ((nu, np.zeros(g.N)), (nu, np.full(g.N, c))):
        us, vs, ws = np.zeros(g.G), np.zeros(g.G), np.zeros(g.G)
        o.convdiff_les(g, dt, nu_, vis, u, v, w, us, vs, ws, z_periodic=1)
        outs.append((us, vs, ws))
    for nu_, k in ((nu, 0), (nu + c, 1)):
        us, vs, ws = np.zeros(g.G), np.zeros(g.G), np.zeros(g.G)
        o.convdiff(g, dt, nu_, u, v, w, us, vs, ws, z_periodic=1)
        for x, y in zip((us, vs, ws), outs[k]):
            assert np.array_equal(x, y)


# ---------------------------------------------------------------- GPU
def _random_fields(g, seed):
    rng = np.random.default_rng(seed)
    return [1.0 * (c == 0) + 0.2 * rng.standard_normal(g.G) for c in range(3)]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,L", [(0, (40, 24, 8), (20.0, 10.0, 0.5)),
                                      (1, (112, 104, 6), (17.0, 12.0, 3.0))])
def test_gpu_eddy_viscosity_and_convdiff_bitwise(o, kind, n, L):
    cfg = pkg.default_config(nx=n[0], ny=n[1], nz=n[2], lx=L[0], ly=L[1], lz=L[2],
                             grid_kind=kind, re=3900.0, dt=0.001, z_periodic=1,
                             smagorinsky_cs=CS)
    gpu = pkg.FlowSolver(cfg)
    g = o.grid(*n, *L, kind=kind)
    u, v, w = _random_fields(g, 11)
    gpu.upload_velocity(0, u, v, w)
    gpu.smagorinsky()
    vis = o.smagorinsky(g, CS, u, v, w)
    assert np.array_equal(gpu.viseff(), vis)
    z = np.zeros(g.G)
    gpu.upload_velocity(1, z, z, z)
    gpu.convdiff()
    us, vs, ws = np.zeros(g.G), np.zeros(g.G), np.zeros(g.G)
    o.convdiff_les(g, 0.001, 1.0 / 3900.0, vis, u, v, w, us, vs, ws, z_periodic=1)
    a, b, c = gpu.download_velocity(1)
    assert np.array_equal(a, us) and np.array_equal(b, vs) and np.array_equal(c, ws)
    gpu.close()


@pytest.mark.gpu
@pytest.mark.parametrize("solver", [0, 1])
def test_gpu_time_steps_with_smagorinsky_bitwise(solver):
    """Whole steps with the model on (eddy viscosity from T0, then conv-diff, pressure,
    projection), Re 3900, against the oracle's time loop."""
    gpu, sim, o, g = make_pair(40, 20, 6, bc="cylinder", re=3900.0, dt=0.001, tol=1e-8,
                               solver=solver, smag_cs=CS)
    for s in range(4):
        a = gpu.step(1)
        b = sim.step()
        for key in ("iters", "err", "cd", "cl", "cfl", "time"):
            assert a[key] == b[key], (s, key, a, b)
        assert fields_equal(gpu, sim) == [], s
    gpu.close()


@pytest.mark.gpu
def test_gpu_smagorinsky_switches():
    cfg = pkg.default_config(nx=16, ny=8, nz=4)
    gpu = pkg.FlowSolver(cfg)
    with pytest.raises(CfdmxError, match="Smagorinsky model is off"):
        gpu.smagorinsky()
    gpu.close()
    with pytest.raises(CfdmxError, match="smagorinsky_cs"):
        pkg.FlowSolver(pkg.default_config(nx=16, ny=8, nz=4, smagorinsky_cs=-0.1))
    gpu = pkg.FlowSolver(pkg.default_config(nx=16, ny=8, nz=4, smagorinsky_cs=CS))
    gpu.init_flow(1.0, 1e-3, 1)
    gpu.set_eta_smoothed(5.0, 5.0, 0.5, 2.0)
    with pytest.raises(CfdmxError, match="does not differentiate"):
        gpu.adjoint_drag(2, 1)
    gpu.close()
