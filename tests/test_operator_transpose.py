"""The transpose of the pressure operator, oracle orc_apply_AT (no reference counterpart).

It is pinned against A itself. orc_apply_A is held bit for bit to the reference's
SPE_matrix and CSR_matrix products (SPE_sparseMatrix.hpp:92-106, CSR_sparseMatrix.hpp:
231-245) in test_oracle_reference.py. Its entries are read back exactly by probing with
the indicator vectors of cell classes that no row couples twice: (i mod 3, j mod 3,
k mod 3), or k itself with a periodic z wrap. Row m of A e_c then holds the single entry
A[m][n] of the class member n in its stencil. A^T y is then formed from those entries.

- With walled z, summing over ascending rows is the slot order. A^T y must equal
  orc_apply_AT bit for bit.
- Across a periodic wrap the orders differ, and the bar there is the sum's rounding.
- The dot-product identity <y, A x> = <A^T y, x> is checked on every case.
- On the uniform grid with walled z the SPE operator is symmetric, so A^T y is A y to the
  bit.
- On the SML grid it is not: the widths are stretched, and the y-slot swap
  (SPE_sparseMatrix.hpp:79,83) puts the +y coefficient in the -y slot.
"""
import numpy as np
import pytest

import oracle_py as op

CASES = [
    (op.GRID_UNIFORM, (13, 11, 5), (20.0, 10.0, 0.5)),
    (op.GRID_SML, (112, 104, 4), (17.0, 12.0, 3.0)),
]


def _neighbours(g, zper):
    """Index of each cell's -x, -y, -z, +z, +y, +x neighbour (z wrapped when periodic)
    and whether it exists."""
    i, j, k = np.meshgrid(np.arange(g.nxc), np.arange(g.nyc), np.arange(g.nzc), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    sx, sy = g.nyc * g.nzc, g.nzc
    n = np.arange(g.N)
    km = np.where(k == 0, g.nzc - 1, k - 1)
    kp = np.where(k == g.nzc - 1, 0, k + 1)
    nb = [
        (n - sx, i > 0),
        (n - sy, j > 0),
        (n + (km - k), (k > 0) | bool(zper)),
        (n + (kp - k), (k < g.nzc - 1) | bool(zper)),
        (n + sy, j < g.nyc - 1),
        (n + sx, i < g.nxc - 1),
    ]
    return (i, j, k), [(np.where(ok, m, n), ok) for m, ok in nb]


def _colour(g, zper, ijk):
    i, j, k = ijk
    kc = k if zper else k % 3
    nk = g.nzc if zper else 3
    return (i % 3) * 3 * nk + (j % 3) * nk + kc, 9 * nk


def _probed_transpose(o, g, precond, zper, y):
    """A^T y from the probed entries of A, summed over the rows in ascending order."""
    ijk, nb = _neighbours(g, zper)
    col, ncol = _colour(g, zper, ijk)
    probes = []
    for c in range(ncol):
        probes.append(o.apply_A(g, (col == c).astype(np.float64), precond, zper))
    probes = np.array(probes)
    n = np.arange(g.N)
    # the rows m holding column n, with A[m][n] = (A e_colour(n))[m]
    rows = [(m, ok) for m, ok in nb] + [(n, np.ones(g.N, bool))]
    terms = [(m, ok, probes[col, m]) for m, ok in rows]
    # ascending order of the row index, per column
    order = np.argsort(np.stack([np.where(ok, m, -1) for m, ok, _ in terms]), axis=0,
                       kind="stable")
    s = np.zeros(g.N)
    for slot in range(len(terms)):
        pick = order[slot]
        m = np.choose(pick, [t[0] for t in terms])
        ok = np.choose(pick, [t[1] for t in terms])
        a = np.choose(pick, [t[2] for t in terms])
        s = s + np.where(ok, a * y[m], 0.0)
    mag = np.zeros(g.N)
    for m, ok, a in terms:
        mag = mag + np.where(ok, np.abs(a * y[m]), 0.0)
    return s, mag


@pytest.fixture(scope="module")
def o():
    return op.Oracle()


@pytest.mark.parametrize("kind,n,L", CASES)
@pytest.mark.parametrize("precond", [0, 1])
@pytest.mark.parametrize("zper", [0, 1])
def test_transpose_equals_the_probed_matrix_transposed(o, kind, n, L, precond, zper):
    g = o.grid(*n, *L, kind=kind)
    rng = np.random.default_rng(11)
    y = rng.standard_normal(g.N)
    at = o.apply_AT(g, y, precond, zper)
    ref, mag = _probed_transpose(o, g, precond, zper, y)
    if zper:
        # the wrap planes sum in another order: two orders of seven terms differ by at
        # most 12 u of the terms' magnitudes (u = 2^-53); every other plane to the bit
        k = np.arange(g.N) % g.nzc
        wrap = (k == 0) | (k == g.nzc - 1)
        assert np.all(np.abs(at - ref)[wrap] <= 12 * 2.0**-53 * mag[wrap])
        assert np.array_equal(at[~wrap], ref[~wrap])
    else:
        assert np.array_equal(at, ref)


@pytest.mark.parametrize("kind,n,L", CASES)
@pytest.mark.parametrize("precond", [0, 1])
@pytest.mark.parametrize("zper", [0, 1])
def test_dot_product_identity(o, kind, n, L, precond, zper):
    g = o.grid(*n, *L, kind=kind)
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal(g.N), rng.standard_normal(g.N)
    ax, aty = o.apply_A(g, x, precond, zper), o.apply_AT(g, y, precond, zper)
    lhs, rhs = float(np.dot(y, ax)), float(np.dot(aty, x))
    scale = float(np.dot(np.abs(y), np.abs(ax))) + float(np.dot(np.abs(aty), np.abs(x)))
    assert abs(lhs - rhs) <= 1e-13 * scale


def test_symmetric_on_the_uniform_grid_and_not_on_sml(o):
    rng = np.random.default_rng(3)
    kind, n, L = CASES[0]
    g = o.grid(*n, *L, kind=kind)
    y = rng.standard_normal(g.N)
    assert np.array_equal(o.apply_AT(g, y, 0, 0), o.apply_A(g, y, 0, 0))
    kind, n, L = CASES[1]
    g = o.grid(*n, *L, kind=kind)
    y = rng.standard_normal(g.N)
    d = np.abs(o.apply_AT(g, y, 0, 0) - o.apply_A(g, y, 0, 0))
    assert d.max() > 1e-3 * np.abs(o.apply_A(g, y, 0, 0)).max()


@pytest.mark.parametrize("nz", [1, 2, 3])
@pytest.mark.parametrize("precond", [0, 1])
def test_short_periodic_z(o, nz, precond):
    """With periodic z and nz <= 2 a row couples to one cell through both z slots (to itself
    at nz = 1), so the probing above cannot separate the entries; the dot-product identity
    still pins the transpose there, on the uniform and (nz >= 2) the SML grid.  With one z
    cell the SML grid's ghost widths (init_gC_dx, structureGrid.hpp:433-486) come out
    zero, so its operator is undefined there, A included."""
    rng = np.random.default_rng(17 + nz)
    cases = [(op.GRID_UNIFORM, (9, 7, nz), (20.0, 10.0, 0.5))]
    if nz >= 2:
        cases.append((op.GRID_SML, (112, 104, nz), (17.0, 12.0, 3.0)))
    for kind, n, L in cases:
        g = o.grid(*n, *L, kind=kind)
        x, y = rng.standard_normal(g.N), rng.standard_normal(g.N)
        ax, aty = o.apply_A(g, x, precond, 1), o.apply_AT(g, y, precond, 1)
        lhs, rhs = float(np.dot(y, ax)), float(np.dot(aty, x))
        scale = float(np.dot(np.abs(y), np.abs(ax))) + float(np.dot(np.abs(aty), np.abs(x)))
        assert abs(lhs - rhs) <= 1e-13 * scale
