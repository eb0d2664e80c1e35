"""CPU pins of the two oracle pieces that the reference cannot pin: the convective
outflow update and one red-black SOR sweep.

Neither has a reference counterpart (the reference's outlet is the zero-gradient row
of FLOWpassing.hpp:6-40, and its SOR is lexicographic, SorPipeLine.seq.hpp:80-118), so
each is checked against what the mathematics fixes:

- convective outflow, the first-order upwind form of du/dt + Uc du/dx = 0 at the outlet
  face: a linear profile advances by exactly -Uc dt du/dx; at Courant number 1 the
  face takes its upwind neighbour's value exactly; an advected pulse leaves through the
  face without overshoot and without a residue;
- one red-black SOR sweep, against a brute-force Gauss-Seidel sweep assembled from the
  finite-volume Poisson equation (face area over centre distance, taken from the node
  coordinates, not from the oracle's metric arrays), and, with omega = 1, against the
  residual of the reference-pinned SPE operator: every black row that touches no
  boundary is solved exactly after the sweep.

Each check is chosen so that a plausible mistake fails it: a sign, the upwind index,
the spacing index, the colour order, the place of omega, a missing periodic wrap.
"""
import numpy as np
import pytest

import oracle_py as op


@pytest.fixture(scope="module")
def o():
    return op.Oracle()


def _stretch_x(g, ratio=1.07):
    """Overwrite the oracle grid's x metrics with a geometric stretching (the C side
    reads the same numpy buffers), so that neighbouring x spacings all differ."""
    n = g.nx
    d = 0.1 * ratio ** np.arange(n)
    X = np.concatenate([[0.0], np.cumsum(d)]) - 2 * d[0]
    g.X[:] = X
    g.Dx[:] = X[1:] - X[:-1]
    g.Dxs[:-1] = (X[2:] - X[:-2]) / 2
    g.Dxs[-1] = g.Dxs[-2]
    return g


def _conv_bc(speed):
    return op.make_bc(op.SLIP_NUM, op.SLIP_DIR, z_periodic=1, outflow_convective=1,
                      conv_speed=speed, flux_correction=0)


# ---------------------------------------------------------------- convective outflow
@pytest.mark.parametrize("stretched", [False, True])
def test_convective_outflow_advances_a_linear_profile_exactly(o, stretched):
    """u = a(j,k) + s x on the faces near the outlet: one boundary step gives
    u_b - Uc dt s on the outlet face and its two ghosts (first-order upwind is exact on
    a linear profile); the interior is not touched."""
    g = o.grid(24, 10, 6)
    if stretched:
        _stretch_x(g)
    rng = np.random.default_rng(11)
    shape = g.shape()
    xf = g.X[1:g.nx + 1]  # u[i] sits on the +x face of cell i, at X[i+1]
    a = 1.0 + 0.1 * rng.standard_normal((1, shape[1], shape[2]))
    s = 0.37
    u = (a + s * xf[:, None, None]).ravel().copy()
    v = 0.1 * rng.standard_normal(g.G)
    w = 0.1 * rng.standard_normal(g.G)
    p = rng.standard_normal(g.G)
    u_in = u.copy()
    speed, dt = 0.7, 0.013
    o.bc_main(g, _conv_bc(speed), u, v, w, p, advance=1, dt=dt)
    U, U0 = u.reshape(shape), u_in.reshape(shape)
    nx = g.nx
    j, k = slice(2, shape[1] - 2), slice(2, shape[2] - 2)
    want = U0[nx - 3, j, k] - speed * dt * s
    assert np.allclose(U[nx - 3, j, k], want, rtol=2e-15, atol=0)
    # the expected move is far above rounding, so a doubled, missing or reversed step fails
    assert np.abs(U[nx - 3, j, k] - U0[nx - 3, j, k]).min() > 1e3 * 2e-15
    assert np.array_equal(U[nx - 2, j, k], U[nx - 3, j, k])
    assert np.array_equal(U[nx - 1, j, k], U[nx - 3, j, k])
    assert np.array_equal(U[2:nx - 3, j, k], U0[2:nx - 3, j, k])
    # without advance the outlet takes the table's zero-gradient row
    u2 = u_in.copy()
    o.bc_main(g, _conv_bc(speed), u2, v.copy(), w.copy(), p.copy(), advance=0, dt=dt)
    U2 = u2.reshape(shape)
    assert np.array_equal(U2[nx - 3, j, k], U0[nx - 4, j, k])


def test_convective_outflow_at_courant_one_is_an_exact_shift(o):
    """Uc dt = Dx of the outlet face: the face takes the value of its upwind neighbour,
    bit for bit (values within a factor 2, so the differences are exact)."""
    g = _stretch_x(o.grid(20, 8, 4))
    rng = np.random.default_rng(12)
    u = 1.0 + 0.9 * rng.random(g.G)
    v, w, p = np.zeros(g.G), np.zeros(g.G), np.zeros(g.G)
    shape, nx = g.shape(), g.nx
    u_old = u.reshape(shape)[nx - 4].copy()
    o.bc_main(g, _conv_bc(1.0), u, v, w, p, advance=1, dt=float(g.Dx[nx - 3]))
    U = u.reshape(shape)
    assert np.array_equal(U[nx - 3, 2:-2, 2:-2], u_old[2:-2, 2:-2])


def test_convective_outflow_lets_a_pulse_leave(o):
    """A pulse advected at Uc crosses the outlet: with the interior faces set to the
    exact solution each step, the outlet face follows it to first order (error bounded
    by dx * max|u'|), never leaves the range of the profile (upwind at Courant number
    0.5 is a convex combination) and returns to the free stream once the pulse is
    gone: nothing is reflected or left behind."""
    g = o.grid(40, 6, 4)
    shape, nx = g.shape(), g.nx
    dx = float(g.Dx[nx - 3])
    speed = 1.0
    dt = 0.5 * dx / speed
    xb = g.X[nx - 2]
    x0, sig, amp = xb - 6 * dx, 2.0 * dx, 0.3

    def prof(x, t):
        return 1.0 + amp * np.exp(-((x - x0 - speed * t) / sig) ** 2)

    xf = g.X[1:nx + 1]
    bc = _conv_bc(speed)
    u = np.broadcast_to(prof(xf, 0.0)[:, None, None], shape).ravel().copy()
    v, w, p = np.zeros(g.G), np.zeros(g.G), np.zeros(g.G)
    errs, vals = [], []
    for n in range(1, 90):
        t = n * dt
        # interior faces at the exact solution of time t; the boundary step then moves
        # the outlet face from t - dt to t using the upwind face at time t - dt
        U = u.reshape(shape)
        keep = U[nx - 3].copy()
        U[:nx - 3] = prof(xf[:nx - 3], t - dt)[:, None, None]
        U[nx - 3] = keep
        o.bc_main(g, bc, u, v, w, p, advance=1, dt=dt)
        ub = u.reshape(shape)[nx - 3, 2, 2]
        vals.append(ub)
        errs.append(abs(ub - prof(xb, t)))
    slope = amp * np.sqrt(2 / np.e) / sig  # max |u'| of the Gaussian
    assert max(errs) < 1.5 * dx * slope
    assert max(errs) > 1e-4  # first order, not exact: the test sees the scheme
    assert min(vals) >= 1.0 - 1e-15 and max(vals) <= 1.0 + amp + 1e-15
    assert abs(vals[-1] - 1.0) < 1e-9


# ---------------------------------------------------------------- red-black SOR
def _sor_sweep_brute_force(g, dt, omega, us, vs, ws, p, zper):
    """One red-black Gauss-Seidel/SOR sweep of the finite-volume Poisson equation
    sum_f (A_f / d_f)(p_nb - p_c) = (net outflow of u* through the faces) / dt,
    cell by cell.  Face areas and centre distances come from the node coordinates;
    p outside the interior (the ghosts) keeps its value during the sweep."""
    X, Y, Z = g.X, g.Y, g.Z
    nx, ny, nz = g.nx, g.ny, g.nz
    P = p.reshape(g.shape()).copy()
    Us, Vs, Ws = (a.reshape(g.shape()) for a in (us, vs, ws))

    def cen(C, i):
        return 0.5 * (C[i] + C[i + 1])

    for colour in (0, 1):
        for i in range(2, nx - 2):
            for j in range(2, ny - 2):
                for k in range(2, nz - 2):
                    if (i + j + k) % 2 != colour:
                        continue
                    ax = (Y[j + 1] - Y[j]) * (Z[k + 1] - Z[k])
                    ay = (X[i + 1] - X[i]) * (Z[k + 1] - Z[k])
                    az = (X[i + 1] - X[i]) * (Y[j + 1] - Y[j])
                    nbs = [((i + 1, j, k), ax / (cen(X, i + 1) - cen(X, i))),
                           ((i - 1, j, k), ax / (cen(X, i) - cen(X, i - 1))),
                           ((i, j + 1, k), ay / (cen(Y, j + 1) - cen(Y, j))),
                           ((i, j - 1, k), ay / (cen(Y, j) - cen(Y, j - 1))),
                           ((i, j, k + 1), az / (cen(Z, k + 1) - cen(Z, k))),
                           ((i, j, k - 1), az / (cen(Z, k) - cen(Z, k - 1)))]
                    if zper:  # the z neighbours of the end planes are across the wrap
                        nzc = g.nzc
                        nbs[4] = ((i, j, k + 1 - nzc if k == nz - 3 else k + 1), nbs[4][1])
                        nbs[5] = ((i, j, k - 1 + nzc if k == 2 else k - 1), nbs[5][1])
                    outflow = ((Us[i, j, k] - Us[i - 1, j, k]) * ax
                               + (Vs[i, j, k] - Vs[i, j - 1, k]) * ay
                               + (Ws[i, j, k] - Ws[i, j, k - 1]) * az)
                    wsum = sum(wt for _, wt in nbs)
                    gs = (sum(wt * P[c] for c, wt in nbs) - outflow / dt) / wsum
                    P[i, j, k] += omega * (gs - P[i, j, k])
    return P.ravel()


@pytest.mark.parametrize("kind,n,L,zper,omega", [
    (0, (7, 6, 4), (3.5, 3.0, 1.0), 0, 1.8),
    (0, (6, 5, 4), (3.0, 2.5, 2.0), 1, 1.3),
    (1, (112, 104, 4), (17.0, 12.0, 3.0), 0, 1.8),
])
def test_red_black_sor_one_sweep_equals_brute_force(o, kind, n, L, zper, omega):
    g = o.grid(*n, *L, kind=kind)
    rng = np.random.default_rng(21)
    us, vs, ws = (rng.standard_normal(g.G) for _ in range(3))
    p = rng.standard_normal(g.G)
    dt = 0.01
    want = None
    if kind == 0:
        want = _sor_sweep_brute_force(g, dt, omega, us, vs, ws, p, zper)
    got = p.copy()
    it, err = o.sor(g, dt, us, vs, ws, got, omega=omega, tol=0.0, itmax=1, redblack=1,
                    z_periodic=zper)
    assert it == 0  # one sweep performed; the count, like the reference's, is one fewer
    if want is None:
        # SML grid (112 x 104 cells: too many for the whole cell loop): the spacings
        # vary along x and y, so a swapped Dxs[i]/Dxs[i-1] or Dys[j]/Dys[j-1] shows.
        # Three 4 x 4 column blocks (a corner and two where the spacings change) are
        # checked cell by cell; one sweep reads only neighbours, so each block's values
        # follow from the old values around it.
        for (i0, j0) in ((2, 2), (40, 4), (60, 50)):
            sub = _subgrid_sweep(g, dt, omega, us, vs, ws, p, i0, j0)
            for (i, j, k), val in sub.items():
                assert np.isclose(got.reshape(g.shape())[i, j, k], val, rtol=1e-12,
                                  atol=1e-12), (i, j, k)
        return
    inner = np.zeros(g.shape(), bool)
    inner[2:-2, 2:-2, 2:-2] = True
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert np.array_equal(got[~inner.ravel()], p[~inner.ravel()])  # ghosts untouched


def _subgrid_sweep(g, dt, omega, us, vs, ws, p, i0, j0, w=4):
    """Brute-force values of one red-black sweep on the cells (i0..i0+w-1, j0..j0+w-1,
    all k): the red cells there need the old black values around them, and the black
    cells need the red values around them after the red half sweep, which the red
    cells one ring further out provide.  So the red half sweep is done on a box one
    cell wider, from the old values."""
    X, Y, Z = g.X, g.Y, g.Z
    P0 = p.reshape(g.shape())
    Us, Vs, Ws = (a.reshape(g.shape()) for a in (us, vs, ws))

    def cen(C, i):
        return 0.5 * (C[i] + C[i + 1])

    def gs_value(P, i, j, k):
        ax = (Y[j + 1] - Y[j]) * (Z[k + 1] - Z[k])
        ay = (X[i + 1] - X[i]) * (Z[k + 1] - Z[k])
        az = (X[i + 1] - X[i]) * (Y[j + 1] - Y[j])
        nbs = [((i + 1, j, k), ax / (cen(X, i + 1) - cen(X, i))),
               ((i - 1, j, k), ax / (cen(X, i) - cen(X, i - 1))),
               ((i, j + 1, k), ay / (cen(Y, j + 1) - cen(Y, j))),
               ((i, j - 1, k), ay / (cen(Y, j) - cen(Y, j - 1))),
               ((i, j, k + 1), az / (cen(Z, k + 1) - cen(Z, k))),
               ((i, j, k - 1), az / (cen(Z, k) - cen(Z, k - 1)))]
        outflow = ((Us[i, j, k] - Us[i - 1, j, k]) * ax + (Vs[i, j, k] - Vs[i, j - 1, k]) * ay
                   + (Ws[i, j, k] - Ws[i, j, k - 1]) * az)
        wsum = sum(wt for _, wt in nbs)
        return (sum(wt * P(*c) for c, wt in nbs) - outflow / dt) / wsum

    def interior(i, j, k):
        return 2 <= i < g.nx - 2 and 2 <= j < g.ny - 2 and 2 <= k < g.nz - 2

    red = {}
    for i in range(i0 - 1, i0 + w + 1):
        for j in range(j0 - 1, j0 + w + 1):
            for k in range(2, g.nz - 2):
                if interior(i, j, k) and (i + j + k) % 2 == 0:
                    old = P0[i, j, k]
                    red[(i, j, k)] = old + omega * (gs_value(lambda a, b, c: P0[a, b, c],
                                                             i, j, k) - old)

    def after_red(a, b, c):
        return red.get((a, b, c), P0[a, b, c])

    out = {}
    for i in range(i0, i0 + w):
        for j in range(j0, j0 + w):
            for k in range(2, g.nz - 2):
                if (i + j + k) % 2 == 0:
                    out[(i, j, k)] = red[(i, j, k)]
                else:
                    old = P0[i, j, k]
                    out[(i, j, k)] = old + omega * (gs_value(after_red, i, j, k) - old)
    return out


@pytest.mark.parametrize("zper", [0, 1])
def test_red_black_sweep_solves_the_black_rows_of_the_spe_system(o, zper):
    """omega = 1 on a uniform grid: after one sweep every black row of A p = b (the
    reference-pinned SPE operator and right-hand side) that does not touch the x or y
    boundary (nor the z boundary without periodicity) is satisfied to rounding, and
    the red rows are not (the black half sweep moved their neighbours): the sweep is
    Gauss-Seidel, red first."""
    g = o.grid(10, 8, 6, 5.0, 4.0, 3.0)
    rng = np.random.default_rng(31)
    us, vs, ws = (rng.standard_normal(g.G) for _ in range(3))
    p = rng.standard_normal(g.G)
    dt = 0.02
    o.sor(g, dt, us, vs, ws, p, omega=1.0, tol=0.0, itmax=1, redblack=1, z_periodic=zper)
    b = o.rhs(g, dt, us, vs, ws, op.PRECOND_NONE, zper)
    x = p.reshape(g.shape())[2:-2, 2:-2, 2:-2].ravel().copy()
    r = (b - o.apply_A(g, x, op.PRECOND_NONE, zper)).reshape(g.nxc, g.nyc, g.nzc)
    i, j, k = np.meshgrid(np.arange(g.nxc), np.arange(g.nyc), np.arange(g.nzc),
                          indexing="ij")
    deep = (i > 0) & (i < g.nxc - 1) & (j > 0) & (j < g.nyc - 1)
    if not zper:
        deep &= (k > 0) & (k < g.nzc - 1)
    black = (i + j + k) % 2 == 1  # interior offsets 2+2+2 keep the parity
    scale = np.abs(b).max()
    assert np.abs(r[deep & black]).max() < 1e-12 * scale
    assert np.abs(r[deep & ~black]).max() > 1e-3 * scale
