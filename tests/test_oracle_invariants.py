"""Pins for the oracle from mathematics (closed forms, invariants, brute force),
covering what the reference cannot pin: the device's canonical reduction order,
periodic z, the convective outflow with flux correction, the red-black SOR
ordering, the smoothed mask and the seeded initial field."""
import numpy as np
import pytest

import oracle_py as op


@pytest.fixture(scope="module")
def o():
    return op.Oracle()


# ---------------------------------------------------------------- reductions
def test_blocked_sum_exact_on_integers(o):
    rng = np.random.default_rng(0)
    for n in (1, 255, 1024, 1025, 70001, 300000):
        v = rng.integers(-2**20, 2**20, n).astype(np.float64)
        assert o.blocked_sum(v) == float(v.sum()) == o.serial_sum(v)


def test_blocked_sum_order_brute_force(o):
    """The canonical order, written out independently: per block of 1024 values,
    thread t sums v[b*1024 + c*256 + t] for c = 0..3, waves fold pairwise by
    halves, blocks combine (w0+w1)+(w2+w3); partials fold the same way."""
    rng = np.random.default_rng(1)
    v = rng.standard_normal(5000) * 10.0 ** rng.integers(-8, 8, 5000)

    def fold(a):
        a = list(a)
        for w in range(4):
            s = 32
            while s >= 1:
                for l in range(s):
                    a[64 * w + l] = a[64 * w + l] + a[64 * w + l + s]
                s //= 2
        return (a[0] + a[64]) + (a[128] + a[192])

    parts = []
    for b in range((len(v) + 1023) // 1024):
        acc = []
        for t in range(256):
            s = 0.0
            for c in range(4):
                m = b * 1024 + c * 256 + t
                if m < len(v):
                    s += v[m]
            acc.append(s)
        parts.append(fold(acc))
    acc = []
    for t in range(256):
        s = 0.0
        for m in range(t, len(parts), 256):
            s += parts[m]
        acc.append(s)
    assert o.blocked_sum(v) == fold(acc)


def test_march_sum_order_brute_force(o):
    """March order of the plane-marching kernels, written out independently: block
    b = chunk*ntile + tile, thread t sums plane by plane (8 planes per chunk, downwards
    in even chunks, upwards in odd ones) the plane cells q = 1024*tile + 256c + t;
    block and partial folds as above."""
    g = o.grid(20, 13, 7)  # cyz = 91 < 1024: one tile; 3 chunks (8, 8, 4 planes)
    g2 = o.grid(9, 40, 60)  # cyz = 2400: three tiles, the last partial; 2 chunks
    rng = np.random.default_rng(2)

    def fold(a):
        a = list(a)
        for w in range(4):
            s = 32
            while s >= 1:
                for l in range(s):
                    a[64 * w + l] = a[64 * w + l] + a[64 * w + l + s]
                s //= 2
        return (a[0] + a[64]) + (a[128] + a[192])

    for gg in (g, g2):
        v = rng.standard_normal(gg.N) * 10.0 ** rng.integers(-6, 6, gg.N)
        cyz = gg.nyc * gg.nzc
        ntile, nchunk = -(-cyz // 1024), -(-gg.nxc // 8)
        parts = []
        for b in range(ntile * nchunk):
            tile, chunk = b % ntile, b // ntile
            acc = []
            for t in range(256):
                s = 0.0
                planes = list(range(chunk * 8, min(chunk * 8 + 8, gg.nxc)))
                for i in (planes if chunk % 2 else planes[::-1]):
                    for c in range(4):
                        q = tile * 1024 + c * 256 + t
                        if q < cyz:
                            s += v[i * cyz + q]
                acc.append(s)
            parts.append(fold(acc))
        acc = []
        for t in range(256):
            s = 0.0
            for m in range(t, len(parts), 256):
                s += parts[m]
            acc.append(s)
        assert o.march_sum(gg, v) == fold(acc)


# ---------------------------------------------------------------- grid
def test_uniform_grid_closed_form(o):
    g = o.grid(40, 20, 8)
    dx = 20.0 / 40
    assert np.all(g.Dx == dx) and np.all(g.Dxs == dx)
    assert g.X[2] == 0.0 and abs(g.X[g.nx - 2] - 20.0) < 1e-12
    assert np.all(np.diff(g.X) > 0)


def test_sml_grid_lengths(o):
    g = o.grid(112, 104, 6, 17.0, 12.0, 3.0, kind=op.GRID_SML)
    assert abs(g.X[g.nx - 2] - 17.0) < 1e-5 and abs(g.Y[g.ny - 2] - 12.0) < 1e-5  # checkgC
    assert np.all(np.diff(g.X) > 0) and np.all(np.diff(g.Y) > 0)
    fine = 0.3 * 17.0 / 100
    assert np.sum(np.isclose(g.Dx[2:-2], fine)) == 100


# ---------------------------------------------------------------- conv-diff
def test_free_stream_preserved(o):
    """u = 1, v = w = 0 with the FLOWpassing table: QUICK and diffusion of a constant
    vanish, so u* = u exactly."""
    g = o.grid(24, 12, 6)
    u, v, w = np.ones(g.G), np.zeros(g.G), np.zeros(g.G)
    p = np.zeros(g.G)
    o.bc_main(g, op.make_bc(), u, v, w, p)
    us, vs, ws = u.copy(), v.copy(), w.copy()
    o.convdiff(g, 0.002, 0.01, u, v, w, us, vs, ws)
    assert np.array_equal(us, u) and np.all(vs == 0) and np.all(ws == 0)


def test_quick_uniform_leonard_form(o):
    """On a uniform grid with a linear-in-x u >0 and v = w = 0, conv-diff of u equals
    -dt d(u^2)/dx with QUICK face values 3/4 c + 3/8 p - 1/8 n (Leonard), exact for
    linear profiles: u_f = face midpoint value, so the result is -dt*(uf+^2-uf-^2)/dx."""
    g = o.grid(16, 8, 4)
    nx, ny, nz = g.shape()
    i = np.arange(nx)[:, None, None] * np.ones((1, ny, nz))
    a, c0 = 0.01, 1.0
    u = (c0 + a * i).ravel()  # u at faces i
    v, w = np.zeros(g.G), np.zeros(g.G)
    us = u.copy()
    dt = 0.003
    o.convdiff(g, dt, 0.0, u, v, w, us, v.copy(), w.copy())
    us = us.reshape(nx, ny, nz)
    ii = np.arange(2, nx - 3)
    # QUICK of a linear field is exact: face values at i+1/2 are the linear midpoints
    up = c0 + a * (ii + 0.5)
    un = c0 + a * (ii - 0.5)
    expect = (c0 + a * ii) - dt * (up ** 2 - un ** 2) / g.Dx[0]
    got = us[2:nx - 3, 5, 3]
    assert np.allclose(got, expect, rtol=0, atol=1e-13)


def test_diffusion_of_quadratic(o):
    """v = w = 0, u = q*(y - y0)^2 constant in x: the QUICK flux differences cancel and
    the diffusion is exact: u* = u + nu*dt*2q."""
    g = o.grid(12, 16, 4)
    nx, ny, nz = g.shape()
    yc = (g.Y[:-1] + g.Y[1:]) / 2
    q, nu, dt = 0.3, 0.05, 0.01
    u = (q * (yc - 5.0) ** 2)[None, :, None] * np.ones((nx, ny, nz))
    u = u.ravel().copy()
    v, w = np.zeros(g.G), np.zeros(g.G)
    us = u.copy()
    o.convdiff(g, dt, nu, u, v, w, us, v.copy(), w.copy())
    d = (us - u).reshape(nx, ny, nz)[2:nx - 3, 4:ny - 4, 2:nz - 2]
    assert np.allclose(d, nu * dt * 2 * q, rtol=1e-9, atol=0)


def test_periodic_z_translation(o):
    """Periodic z: rolling the field by one cell in z commutes with conv-diff and with
    the pressure operator, bit for bit."""
    g = o.grid(16, 10, 6)
    nx, ny, nz = g.shape()
    rng = np.random.default_rng(3)
    f = [rng.standard_normal((nx, ny, nz)) * 0.1 + (1.0 if c == 0 else 0.0) for c in range(3)]
    bc = op.make_bc(op.SLIP_NUM, op.SLIP_DIR, z_periodic=1)
    p = np.zeros(g.G)
    u, v, w = (a.ravel().copy() for a in f)
    o.bc_main(g, bc, u, v, w, p)

    def roll(a):
        a = a.reshape(nx, ny, nz)
        inner = np.roll(a[:, :, 2:nz - 2], 1, axis=2)
        out = a.copy()
        out[:, :, 2:nz - 2] = inner
        return out.ravel()

    ur, vr, wr = roll(u), roll(v), roll(w)
    o.bc_main(g, bc, ur, vr, wr, p.copy())
    outs = [np.zeros(g.G) for _ in range(6)]
    o.convdiff(g, 0.002, 0.01, u, v, w, *outs[:3], z_periodic=1)
    o.convdiff(g, 0.002, 0.01, ur, vr, wr, *outs[3:], z_periodic=1)
    sl = (slice(2, nx - 3), slice(2, ny - 3), slice(2, nz - 2))
    for a, b in zip(outs[:3], outs[3:]):
        assert np.array_equal(roll(a).reshape(nx, ny, nz)[sl], b.reshape(nx, ny, nz)[sl])
    x = rng.standard_normal(g.N).reshape(g.nxc, g.nyc, g.nzc)
    Ax = o.apply_A(g, x.ravel(), 0, 1).reshape(x.shape)
    Axr = o.apply_A(g, np.roll(x, 1, axis=2).ravel(), 0, 1).reshape(x.shape)
    assert np.array_equal(np.roll(Ax, 1, axis=2), Axr)


# ---------------------------------------------------------------- operator
@pytest.mark.parametrize("zper", [0, 1])
def test_operator_constant_null_space_and_symmetry(o, zper):
    g = o.grid(12, 8, 6)
    Ax = o.apply_A(g, np.ones(g.N), 0, zper)
    diag = o.spe_vals(g)[3::7].max()
    assert np.max(np.abs(Ax)) <= 1e-12 * diag
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal(g.N), rng.standard_normal(g.N)
    lhs, rhs = np.dot(o.apply_A(g, x, 0, zper), y), np.dot(x, o.apply_A(g, y, 0, zper))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("zper", [0, 1])
def test_operator_eigenpairs(o, zper):
    """Neumann cosine modes cos(pi m (i+1/2)/n) (periodic: cos(2 pi m k/n)) are
    eigenvectors with eigenvalue sum over axes of (2 - 2 cos(pi m / n)) / h^2."""
    g = o.grid(16, 12, 8)
    n = (g.nxc, g.nyc, g.nzc)
    h = (g.Dx[2], g.Dy[2], g.Dz[2])
    m = (3, 2, 2)
    vecs, lam = [], 0.0
    for ax in range(3):
        idx = np.arange(n[ax])
        if ax == 2 and zper:
            vecs.append(np.cos(2 * np.pi * m[ax] * idx / n[ax]))
            lam += (2 - 2 * np.cos(2 * np.pi * m[ax] / n[ax])) / h[ax] ** 2
        else:
            vecs.append(np.cos(np.pi * m[ax] * (idx + 0.5) / n[ax]))
            lam += (2 - 2 * np.cos(np.pi * m[ax] / n[ax])) / h[ax] ** 2
    v = np.einsum("i,j,k->ijk", *vecs).ravel()
    Av = o.apply_A(g, v, 0, zper)
    assert np.allclose(Av, lam * v, rtol=0, atol=1e-9 * lam)


def _axis_chains(o, g):
    """The x and y parts of the SPE operator as dense 1-D chains, read from the rows at
    an interior (i, k) / (j, k): the operator is a Kronecker sum of per-axis chains."""
    V = o.spe_vals(g).reshape(g.nxc, g.nyc, g.nzc, 7)
    i, j, k = g.nxc // 2, g.nyc // 2, 1
    Ay = np.zeros((g.nyc, g.nyc))
    for jj in range(g.nyc):
        r = V[i, jj, k]
        Ay[jj, jj] = r[3] + r[0] + r[6] + r[2] + r[4]
        if jj > 0:
            Ay[jj, jj - 1] = r[1]
        if jj < g.nyc - 1:
            Ay[jj, jj + 1] = r[5]
    Ax = np.zeros((g.nxc, g.nxc))
    for ii in range(g.nxc):
        r = V[ii, j, k]
        Ax[ii, ii] = r[3] + r[1] + r[5] + r[2] + r[4]
        if ii > 0:
            Ax[ii, ii - 1] = r[0]
        if ii < g.nxc - 1:
            Ax[ii, ii + 1] = r[6]
    return Ax, Ay


def _left_null(A):
    w, vl = np.linalg.eig(A.T)
    m = np.argmin(np.abs(w))
    v = np.real(vl[:, m])
    return v / v.sum()


def test_sml_y_slot_swap_moves_the_left_null_vector(o):
    """Why the SML grid has no adjoint or multigrid mode.  The SPE rows sum to zero on
    either grid, so A 1 = 0.  The discrete divergence b of a flux-balanced field is
    orthogonal to the cell volumes, so A p = b is consistent only if the volumes span
    A's left null space.  That holds per axis where the chain is the flux form divided
    by the widths (Dx on either grid, Dy on the uniform grid), but the y-slot swap of
    SPE_sparseMatrix.hpp:79,83 breaks it on the stretched y of the SML grid: the left
    null vector leaves the widths by O(1), and the chain gains a second near-null
    direction.  Tight-tolerance pressure solves there stall (the reference's own
    BiCGSTAB stops at its iteration cap near 1e-6 relative residual), so a time step is
    not a smooth function of its inputs and an adjoint has no finite-difference pin."""
    for kind, n, L in ((1, (112, 104, 4), (17.0, 12.0, 3.0)), (0, (40, 24, 4), (20.0, 10.0, 0.5))):
        g = o.grid(*n, *L, kind=kind)
        Ax, Ay = _axis_chains(o, g)
        for A, D in ((Ax, g.Dx[2:2 + g.nxc]), (Ay, g.Dy[2:2 + g.nyc])):
            scale = np.abs(np.diag(A)).max()
            assert np.abs(A.sum(axis=1)).max() <= 1e-12 * scale
        lx, ly = _left_null(Ax), _left_null(Ay)
        wx = g.Dx[2:2 + g.nxc] / g.Dx[2:2 + g.nxc].sum()
        wy = g.Dy[2:2 + g.nyc] / g.Dy[2:2 + g.nyc].sum()
        assert np.abs(lx - wx).max() <= 1e-9 * wx.max()
        sv = np.linalg.svd(Ay, compute_uv=False)
        if kind == 0:
            assert np.abs(ly - wy).max() <= 1e-9 * wy.max()
            assert sv[-2] > 1e-2
        else:
            assert np.abs(ly - wy).max() >= 0.5 * wy.max()
            assert sv[-2] < 1e-4 * sv[0]


def test_jacobi_operator_is_row_scaled_spe(o):
    """CSR + Jacobi rows are the SPE rows (without the y-slot swap, equal on a uniform
    grid) divided by the diagonal."""
    g = o.grid(10, 8, 6)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(g.N)
    diag = o.spe_vals(g)[3::7]
    assert np.allclose(o.apply_A(g, x, 1), o.apply_A(g, x, 0) / diag, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("precond,zper", [(0, 0), (1, 0), (0, 1)])
def test_bicgstab_matches_dense_solve(o, precond, zper):
    g = o.grid(8, 6, 4)
    A = np.stack([o.apply_A(g, e, precond, zper) for e in np.eye(g.N)], axis=1)
    rng = np.random.default_rng(6)
    b = A @ rng.standard_normal(g.N)  # consistent
    x = np.zeros(g.N)
    bic = op.BiCG(o, g)
    it, res = bic.solve(b, x, 1e-12, precond, zper, op.RED_BLOCKED)
    assert res < 1e-12
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-11
    xd = np.linalg.lstsq(A, b, rcond=None)[0]
    assert np.allclose(x - x.mean(), xd - xd.mean(), rtol=0, atol=1e-8 * np.abs(xd).max())


# ---------------------------------------------------------------- SOR
def test_red_black_sor_reaches_the_lexicographic_fixed_point(o):
    """Red-black and lexicographic SOR (the latter pinned to SorPipeLine_seq) solve the
    same equations with the same stale-ghost boundary values."""
    g = o.grid(16, 12, 6)
    rng = np.random.default_rng(7)
    u, v, w = 1 + 0.1 * rng.standard_normal(g.G), 0.1 * rng.standard_normal(g.G), \
        0.1 * rng.standard_normal(g.G)
    p0 = 0.01 * rng.standard_normal(g.G)
    pa, pb = p0.copy(), p0.copy()
    ia, ea = o.sor(g, 0.002, u, v, w, pa, omega=1.5, tol=1e-13, itmax=20000, redblack=0)
    ib, eb = o.sor(g, 0.002, u, v, w, pb, omega=1.5, tol=1e-13, itmax=20000, redblack=1)
    assert ea <= 1e-13 and eb <= 1e-13
    inner = np.zeros((g.nx, g.ny, g.nz), bool)
    inner[2:-2, 2:-2, 2:-2] = True
    d = np.abs(pa - pb)[inner.ravel()].max()
    assert d < 1e-9 * np.abs(pa).max()


# ---------------------------------------------------------------- outflow, projection
def test_flux_correction_balances_outlet(o):
    g = o.grid(20, 12, 6)
    rng = np.random.default_rng(8)
    u = 1 + 0.2 * rng.standard_normal(g.G)
    v, w, p = 0.1 * rng.standard_normal(g.G), 0.1 * rng.standard_normal(g.G), np.zeros(g.G)
    bc = op.make_bc(op.SLIP_NUM, op.SLIP_DIR, z_periodic=1, outflow_convective=1,
                    flux_correction=1)
    o.bc_main(g, bc, u, v, w, p, advance=1, dt=0.002)
    U = u.reshape(g.shape())
    A = np.outer(g.Dy, g.Dz)[2:-2, 2:-2]
    qin = np.sum(U[1, 2:-2, 2:-2] * A)
    qout = np.sum(U[g.nx - 3, 2:-2, 2:-2] * A)
    assert abs(qin - qout) <= 1e-13 * abs(qin)
    # slip walls: zero normal velocity on the y faces
    V = v.reshape(g.shape())
    assert np.all(V[2:-2, 1, 2:-2] == 0) and np.all(V[2:-2, g.ny - 3, 2:-2] == 0)


def test_cylinder_bc_rhs_is_consistent(o):
    """With inflow, flux-corrected outflow, slip walls and periodic z the divergence
    of u* sums to zero (the singular Neumann system is consistent)."""
    g = o.grid(24, 12, 6)
    bc = op.make_bc(op.SLIP_NUM, op.SLIP_DIR, z_periodic=1, outflow_convective=1,
                    flux_correction=1)
    sim = op.Sim(o, g, bc, 0.002, 100.0, tol=1e-10, reduction=op.RED_BLOCKED, swap=1)
    u, v, w = o.init_velocity(g, 1.0, 1e-2, 3)
    sim.field("u0")[:] = u
    sim.field("eta")[:] = o.eta_cylinder_z(g, 5.0, 5.0, 1.5, 50)
    sim.start()
    for _ in range(3):
        sim.step()
    b = o.rhs(g, 0.002, sim.field("u1"), sim.field("v1"), sim.field("w1"), 0, 1)
    assert abs(b.sum()) <= 1e-9 * np.abs(b).sum()


def test_projection_makes_velocity_solenoidal_without_body(o):
    """eta = 0: u = u* - dt grad p with A p = -div(u*)/dt gives div u = 0 to the solve
    tolerance (identity D(u* - dt G p) = D u* - dt L p)."""
    g = o.grid(24, 12, 6)
    bc = op.make_bc(op.SLIP_NUM, op.SLIP_DIR, z_periodic=1, outflow_convective=1,
                    flux_correction=1)
    sim = op.Sim(o, g, bc, 0.002, 100.0, tol=1e-12, reduction=op.RED_BLOCKED, swap=1,
                 skip_first=0)
    u, v, w = o.init_velocity(g, 1.0, 5e-2, 5)
    sim.field("u0")[:] = u
    sim.start()
    sim.step()
    # divergence of the projected field over interior cells; the last x cell is left
    # out because the boundary condition then advances the outlet face
    U, V, W = (sim.field(n).reshape(g.shape()) for n in ("u0", "v0", "w0"))
    div = ((U[2:-3, 2:-2, 2:-2] - U[1:-4, 2:-2, 2:-2]) / g.Dx[2]
           + (V[2:-3, 2:-2, 2:-2] - V[2:-3, 1:-3, 2:-2]) / g.Dy[2]
           + (W[2:-3, 2:-2, 2:-2] - W[2:-3, 2:-2, 1:-3]) / g.Dz[2])
    assert np.abs(div).max() < 1e-8
    f = sim.field("f")
    assert np.all(f == 0)  # no body: no forcing


# ---------------------------------------------------------------- mask, drag, misc
def test_sharp_mask_area_and_symmetry(o):
    g = o.grid(160, 80, 4)
    eta = o.eta_cylinder_z(g, 5.0, 5.0, 0.5, 100).reshape(g.shape())
    plane = eta[2:-2, 2:-2, 2]
    area = plane.sum() * g.Dx[2] * g.Dy[2]
    assert abs(area - np.pi * 0.25) < 2e-3
    assert np.all((plane >= 0) & (plane <= 1))
    # mirror symmetry about the centre (5, 5): x cell i <-> 79 - i, y cell j <-> 79 - j
    # (sub-sample midpoints are rounded, so allow one sample of difference)
    assert np.allclose(plane[:80, :], plane[79::-1, :], rtol=0, atol=1.01e-4)
    assert np.allclose(plane, plane[:, ::-1], rtol=0, atol=1.01e-4)
    assert np.array_equal(eta[:, :, 2], eta[:, :, g.nz - 3])


def test_smoothed_mask(o):
    g = o.grid(160, 80, 4)
    eps = 1.5 * g.Dx[2] / 2
    eta = o.eta_smoothed(g, 5.0, 5.0, 0.5, eps).reshape(g.shape())[2:-2, 2:-2, 2]
    # area = pi R^2 + O(eps^2)
    assert abs(eta.sum() * g.Dx[2] * g.Dy[2] - np.pi * 0.25) < 2.0 * eps ** 2
    xc = (g.X[2:-3] + g.X[3:-2]) / 2
    yc = (g.Y[2:-3] + g.Y[3:-2]) / 2
    r = np.sqrt((xc[:, None] - 5.0) ** 2 + (yc[None, :] - 5.0) ** 2)
    assert np.all(eta[r < 0.5 - eps] == 1.0) and np.all(eta[r > 0.5 + eps] == 0.0)
    # monotone non-increasing in the distance, 1/2 on the circle
    order = np.argsort(r.ravel())
    assert np.all(np.diff(eta.ravel()[order]) <= 1e-15)
    band = np.abs(r - 0.5) < 0.2 * eps
    assert np.all(np.abs(eta[band] - 0.5) < 0.2)


def test_drag_closed_form(o):
    g = o.grid(20, 10, 4)
    f = np.zeros(3 * g.G)
    f[:g.G] = 0.25
    f[g.G:2 * g.G] = -0.5
    cd, cl = o.cd_cl(g, f, op.RED_BLOCKED)
    assert np.isclose(cd, -2 * 0.25 * 20.0 * 10.0, rtol=1e-13)
    assert np.isclose(cl, 2 * 0.5 * 20.0 * 10.0, rtol=1e-13)


def test_cfl_closed_form(o):
    g = o.grid(20, 10, 4)
    u, v, w = 1.7 * np.ones(g.G), np.zeros(g.G), np.zeros(g.G)
    assert np.isclose(o.cfl(g, 0.003, u, v, w), 1.7 * 0.003 / g.Dx[2], rtol=1e-15)


def test_initial_field_splitmix64(o):
    """splitmix64(0) = 0xE220A8397B1DCDAF (the generator's published first output),
    so with seed 0 the interior cell 0 gets u0 + amp*(2U-1) for that word."""
    g = o.grid(8, 6, 4)
    u, v, w = o.init_velocity(g, 1.0, 1e-3, 0)
    U = (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert u.reshape(g.shape())[2, 2, 2] == 1.0 + 1e-3 * (2.0 * U - 1.0)
    inner = u.reshape(g.shape())[2:-2, 2:-2, 2:-2]
    assert np.all(np.abs(inner - 1.0) <= 1e-3) and np.all(v == 0) and np.all(w == 0)
    assert abs(inner.mean() - 1.0) < 2e-4


def test_steady_check_over_the_projected_faces(o):
    """The per-step steady-state check covers the faces the projection writes: u for
    i < nx-3, v for j < ny-3, w for k < nz-3 (every k with periodic z); a brute-force
    maximum over those index ranges (starting at 0, as the reference's loop does)."""
    g = o.grid(12, 10, 6)
    rng = np.random.default_rng(9)
    f = [rng.standard_normal(g.G) for _ in range(6)]
    shape = g.shape()
    nx, ny, nz = g.nx, g.ny, g.nz
    for zper in (0, 1):
        d = [(f[c] - f[c + 3]).reshape(shape) for c in range(3)]
        want = (max(0.0, d[0][2:nx - 3, 2:-2, 2:-2].max()),
                max(0.0, d[1][2:-2, 2:ny - 3, 2:-2].max()),
                max(0.0, d[2][2:-2, 2:-2, 2:(nz - 2 if zper else nz - 3)].max()))
        assert o.maxdiff_faces(g, *f, z_periodic=zper) == want
    # a change only on a skipped face does not count
    u3 = f[0].copy().reshape(shape)
    u0 = u3.copy()
    u0[nx - 3, 4, 4] += 5.0
    z = np.zeros(g.G)
    assert o.maxdiff_faces(g, u0.ravel(), z, z, u3.ravel(), z, z) == (0.0, 0.0, 0.0)
    assert o.maxdiff(g, u0.ravel(), z, z, u3.ravel(), z, z)[0] == 5.0
