"""FP64 checkpoint / resume (cfdmx_checkpoint_save / _load).  No reference counterpart:
the reference's only resume path reads a float32 q file back (qfileRead.hpp:6-100) and is
commented out (RUN_CPU.cpp:89).  A run resumed from a checkpoint in a fresh context must
continue bit for bit: against the oracle's uninterrupted run where the oracle covers the
configuration, and against the uninterrupted GPU run for the multigrid modes."""
import os

import numpy as np
import pytest

import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg
from cfd_nssolver_forhpc_flowpassingcylinder_amd.api import (CfdmxError, PRECISION_FP64_MG,
                                                              PRECISION_MIXED)
from gpu_helpers import fields_equal, make_pair

pytestmark = pytest.mark.gpu


def _same(a, b):
    return a == b or (np.isnan(a) and np.isnan(b))


def _fresh(gpu):
    """A new context with the configuration of `gpu` and no state of its own."""
    return pkg.FlowSolver(gpu.cfg)


@pytest.mark.parametrize("case", [
    dict(n=(40, 20, 6), bc="cylinder", tol=1e-8),
    dict(n=(40, 20, 6), bc="cylinder", tol=1e-8, solver=1, omega=1.8),
    dict(n=(112, 104, 4), kind=1, lengths=(17.0, 12.0, 3.0), cyl=(6.5, 6.0, 0.5),
         bc="flowpassing", tol=1e-3),
    dict(n=(48, 24, 8), bc="cylinder", tol=1e-9, smoothed_eps=1.5 * 20.0 / 48,
         inflow=1.0 + 0.01 * np.sin(np.arange(28) / 3.0)),
], ids=["bicgstab", "rbsor", "sml", "smoothed-inflow"])
def test_resume_continues_the_oracle_run(case, tmp_path):
    c = dict(case)
    n = c.pop("n")
    gpu, sim, o, g = make_pair(*n, **c)
    for _ in range(3):
        gpu.step(1)
        sim.step()
    path = os.fspath(tmp_path / "run.ck")
    gpu.checkpoint_save(path)
    G, N = gpu.G, gpu.N
    # the header, then 14 ghosted, 3 interior and one y-row array of FP64
    assert os.path.getsize(path) > (14 * G + 3 * N + n[1] + 4) * 8
    gpu.close()
    res = _fresh(gpu)
    res.checkpoint_load(path)
    for s in range(3):
        a = res.step(1)
        b = sim.step()
        for key in ("iters", "err", "cd", "cl", "cfl", "time"):
            assert _same(a[key], b[key]), (s, key, a, b)
        assert fields_equal(res, sim) == [], s
    res.close()


@pytest.mark.parametrize("precision", [PRECISION_MIXED, PRECISION_FP64_MG])
def test_resume_in_the_multigrid_modes(precision, tmp_path):
    gpu, *_ = make_pair(64, 32, 8, bc="cylinder", tol=1e-10, precision=precision)
    gpu.step(2)
    path = os.fspath(tmp_path / "mg.ck")
    gpu.checkpoint_save(path)
    res = _fresh(gpu)
    res.checkpoint_load(path)
    for _ in range(3):
        a, b = gpu.step(1), res.step(1)
        assert a == b
    for which in range(3):
        for x, y in zip(gpu.download_velocity(which), res.download_velocity(which)):
            assert np.array_equal(x, y)
    assert np.array_equal(gpu.pressure(), res.pressure())
    assert np.array_equal(gpu.xresult(), res.xresult())
    gpu.close()
    res.close()


def test_checkpoint_errors(tmp_path):
    gpu, *_ = make_pair(24, 12, 4, bc="cylinder")
    path = os.fspath(tmp_path / "a.ck")
    gpu.checkpoint_save(path)
    other, *_ = make_pair(24, 16, 4, bc="cylinder")
    with pytest.raises(CfdmxError, match="another size"):
        other.checkpoint_load(path)
    junk = tmp_path / "junk.ck"
    junk.write_bytes(b"\0" * 4096)
    with pytest.raises(CfdmxError, match="not a cfdmx checkpoint"):
        gpu.checkpoint_load(os.fspath(junk))
    short = tmp_path / "short.ck"
    short.write_bytes(open(path, "rb").read()[: os.path.getsize(path) // 2])
    with pytest.raises(CfdmxError, match="truncated"):
        gpu.checkpoint_load(os.fspath(short))
    with pytest.raises(CfdmxError, match="cannot open"):
        gpu.checkpoint_load(os.fspath(tmp_path / "missing.ck"))
    # refused loads change nothing: the context continues as the saved run
    res = _fresh(gpu)
    res.checkpoint_load(path)
    assert gpu.step(1) == res.step(1)
    res.close()
    gpu.close()
    other.close()
    fresh = pkg.FlowSolver(gpu.cfg)
    with pytest.raises(CfdmxError, match="before cfdmx_start"):
        fresh.checkpoint_save(os.fspath(tmp_path / "b.ck"))
    fresh.close()


@pytest.mark.parametrize("field,value,what", [
    ("precision", PRECISION_FP64_MG, "precision"), ("solver", 1, "solver"),
    ("precond", 1, "precond"), ("flux_correction", 0, "flux_correction"),
    ("smagorinsky_cs", 0.1, "smagorinsky_cs")])
def test_checkpoint_of_another_configuration_is_refused(field, value, what, tmp_path):
    """A load into a context whose precision, solver, operator, boundary conditions or
    Smagorinsky constant differ from the saved run's would not continue it: refused,
    and the context is left as it was."""
    gpu, *_ = make_pair(32, 16, 8, bc="cylinder", tol=1e-8)
    gpu.step(2)
    path = os.fspath(tmp_path / "a.ck")
    gpu.checkpoint_save(path)
    cfg = pkg.default_config()
    for name, _ in cfg._fields_:
        setattr(cfg, name, getattr(gpu.cfg, name))
    setattr(cfg, field, value)
    other = pkg.FlowSolver(cfg)
    other.init_flow(1.0, 1e-3, 5)
    before = other.download_velocity(0)
    with pytest.raises(CfdmxError, match="another " + what):
        other.checkpoint_load(path)
    for a, b in zip(other.download_velocity(0), before):
        assert np.array_equal(a, b)
    other.close()
    gpu.close()
