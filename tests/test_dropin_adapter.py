"""The C++ drop-in adapter (include/cfdmx_reference_api.hpp) compiles against the
reference's own types, with the calls runSeqOmp makes (CPU, compile only)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = "/root/reference/inc"


@pytest.mark.skipif(not os.path.isdir(REF_INC) or shutil.which("g++") is None,
                    reason="reference headers or g++ absent")
def test_adapter_compiles_against_reference_types():
    src = os.path.join(ROOT, "tests", "dropin", "dropin_compile.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-fopenmp", "-w",
                        "-I" + REF_INC, "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.skipif(not os.path.isdir(REF_INC) or shutil.which("g++") is None,
                    reason="reference headers or g++ absent")
def test_dropin_run_program_compiles():
    """The end-to-end drop-in program (its GPU run is tests/test_dropin_run.py)."""
    src = os.path.join(ROOT, "tests", "dropin", "dropin_run.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-fopenmp", "-w",
                        "-I" + REF_INC, "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
