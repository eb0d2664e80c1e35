"""The C-ABI library loads and exports every function include/cfdmx.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cfdmx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cfdmx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = declared_functions()
    assert "cfdmx_create" in names and "cfdmx_step" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    if not os.path.exists(pkg.LIB_PATH):
        pytest.skip("libcfdmx.so not built")
    lib = ctypes.CDLL(pkg.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_library_exports_nothing_else():
    """Built with hidden visibility: the dynamic symbol table holds the header's entry
    points and no internal function."""
    import shutil
    import subprocess
    if not os.path.exists(pkg.LIB_PATH) or not shutil.which("nm"):
        pytest.skip("libcfdmx.so not built or nm missing")
    out = subprocess.run(["nm", "-D", "--defined-only", pkg.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    text_syms = sorted(l.split()[-1] for l in out.splitlines() if l.split()[1:2] == ["T"])
    assert text_syms == declared_functions()


def test_python_binding_covers_header():
    assert sorted(pkg.EXPORTS) == declared_functions()


def test_config_struct_matches_header_layout():
    """cfdmx_config_default fills the struct without a device; the ctypes layout must
    match the C one (spot-check values written by the library)."""
    if not os.path.exists(pkg.LIB_PATH):
        pytest.skip("libcfdmx.so not built")
    cfg = pkg.default_config()
    assert (cfg.nx, cfg.ny, cfg.nz) == (128, 64, 4)
    assert (cfg.lx, cfg.ly, cfg.lz) == (20.0, 10.0, 0.5)
    assert cfg.tol == 1e-3 and cfg.maxit == 5000 and cfg.sor_omega == 1.8
    assert [list(r) for r in cfg.bc_dir][0] == [1.0, 0.0, 0.0]
    assert cfg.skip_first_pressure == 1 and cfg.xcd_remap == 1
    assert cfg.inner_tol == 0.0
