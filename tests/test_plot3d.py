"""Plot3D output (OutputPLOT3D_Qfile, qfile.hpp:6-108; OutputPlot3D_Xfile,
xfileWrite.hpp:6-64).  CPU: the reference's own writers (oracle/_ref) produce exactly the
bytes of the layout restated here in numpy, which pins the format.  GPU: cfdmx_plot3d_q
and cfdmx_plot3d_x write the same bytes as the reference's writers for the same state."""
import os

import numpy as np
import pytest

import oracle_py as op

NXC, NYC, NZC = 16, 8, 6


def run_in(tmp_path, fn):
    """The reference writes mx_out/<name> relative to the working directory and creates
    ../mx_out when it is missing: provide both, then call fn there."""
    run = tmp_path / "run"
    (run / "mx_out").mkdir(parents=True)
    (tmp_path / "mx_out").mkdir()
    cwd = os.getcwd()
    os.chdir(run)
    try:
        fn()
    finally:
        os.chdir(cwd)
    return run / "mx_out"


def q_bytes(nxc, nyc, nzc, re, time, u, v, w, p, eta):
    """The q layout: int32 1, nx, ny, nz; float32 0, 0, Re, time; p, cell-centred u, v, w
    (average of the two faces in double, then float), eta; i fastest (k, j, i loops)."""
    shp = (nxc + 4, nyc + 4, nzc + 4)
    U, V, W, P, E = (a.reshape(shp) for a in (u, v, w, p, eta))
    I, J, K = slice(2, nxc + 2), slice(2, nyc + 2), slice(2, nzc + 2)
    arrays = [P[I, J, K],
              0.5 * (U[1:nxc + 1, J, K] + U[I, J, K]),
              0.5 * (V[I, 1:nyc + 1, K] + V[I, J, K]),
              0.5 * (W[I, J, 1:nzc + 1] + W[I, J, K]),
              E[I, J, K]]
    out = np.array([1, nxc, nyc, nzc], np.int32).tobytes()
    out += np.array([0.0, 0.0, re, time], np.float32).tobytes()
    for a in arrays:
        out += a.astype(np.float32).transpose(2, 1, 0).tobytes()  # (k, j, i), i fastest
    return out


def x_bytes(nxc, nyc, nzc, X, Y, Z):
    xc = ((X[2:nxc + 2] + X[3:nxc + 3]) / 2.0).astype(np.float32)
    yc = ((Y[2:nyc + 2] + Y[3:nyc + 3]) / 2.0).astype(np.float32)
    zc = ((Z[2:nzc + 2] + Z[3:nzc + 3]) / 2.0).astype(np.float32)
    out = np.array([1, nxc, nyc, nzc], np.int32).tobytes()
    for a in np.meshgrid(xc, yc, zc, indexing="ij"):
        out += a.transpose(2, 1, 0).tobytes()
    return out


@pytest.fixture
def ref():
    if not op.ref_available():
        pytest.skip("reference harness not built")
    r = op.RefHarness(NXC, NYC, NZC, 20.0, 10.0, 0.5, sml=0, re=150.0)
    yield r


def test_reference_q_writer_matches_layout(ref, tmp_path):
    rng = np.random.default_rng(5)
    for name in ("u3", "v3", "w3", "p", "eta"):
        ref.field(name)[:] = rng.standard_normal(ref.G)
    out = run_in(tmp_path, lambda: ref.lib.ref_plot3d_q(ref.h, 7, 1.25))
    got = (out / "P3D07.q").read_bytes()
    want = q_bytes(NXC, NYC, NZC, 150.0, 1.25, *(ref.field(n) for n in ("u3", "v3", "w3", "p",
                                                                      "eta")))
    assert got == want


def test_reference_x_writer_matches_layout(ref, tmp_path):
    out = run_in(tmp_path, lambda: ref.lib.ref_plot3d_x(ref.h))
    m = ref.metrics()
    assert (out / "P3D.x").read_bytes() == x_bytes(NXC, NYC, NZC, m["X"], m["Y"], m["Z"])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [0, 1])
def test_gpu_plot3d_equals_reference_writers(tmp_path, kind):
    """After a few time steps on the GPU, the library's q and x files equal the
    reference writers' bytes for the same state (its T3 set to the GPU's T0, whose
    boundary faces it would otherwise leave at their initial values), on the uniform
    and on the SML grid."""
    if not op.ref_available():
        pytest.skip("reference harness not built")
    from gpu_helpers import make_pair
    n = (16, 8, 6) if kind == 0 else (112, 104, 4)
    gpu, sim, o, g = make_pair(*n, kind=kind, bc="flowpassing", tol=1e-6)
    for _ in range(3):
        gpu.step(1)
    gpu.plot3d_q(str(tmp_path / "gpu.q"), 0.006)
    gpu.plot3d_x(str(tmp_path / "gpu.x"))
    gpu.plot3d_flush()
    u, v, w = gpu.download_velocity(0)
    r = op.RefHarness(*n, 20.0, 10.0, 0.5, sml=kind, re=100.0)
    r.field("u3")[:] = u
    r.field("v3")[:] = v
    r.field("w3")[:] = w
    r.field("p")[:] = gpu.pressure()
    r.field("eta")[:] = gpu.eta()
    out = run_in(tmp_path / "ref", lambda: (r.lib.ref_plot3d_q(r.h, 3, 0.006),
                                            r.lib.ref_plot3d_x(r.h)))
    assert (tmp_path / "gpu.q").read_bytes() == (out / "P3D03.q").read_bytes()
    assert (tmp_path / "gpu.x").read_bytes() == (out / "P3D.x").read_bytes()
    gpu.close()
