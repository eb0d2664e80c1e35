"""The C++ drop-in adapter run end to end on the GPU (oracle/_ref/dropin_run, built from
tests/dropin/dropin_run.cpp against the reference headers and libcfdmx.so): the
reference's own time loop, the same loop through the cfdmx_ref:: stage calls, and
cfdmx_ref::runSeqOmp, compared step by step; the program states its bounds and exits
non-zero when one is missed."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "dropin_run")

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/dropin_run not built "
                    "(needs the reference headers at build time)")
def test_dropin_adapter_runs_the_reference_loop(tmp_path):
    r = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-6000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "DROPIN OK" in r.stdout
    # runSeqOmp wrote the reference's files: the Plot3D grid and q files, the Cd/Cl series
    assert (tmp_path / "mx_out" / "P3D.x").exists()
    assert (tmp_path / "mx_out" / "P3D00.q").exists()
    text = (tmp_path / "Information" / "Time_cDcL.dat").read_text()
    assert text.count("TITLE") == 2  # the stage loop's CD_CL, then runSeqOmp's
