"""cfdmx_cd_cl_noalloc, the drag/lift reduction that allocates nothing (its block partials
live in a file-scope device array): through the C-ABI against cfdmx_cd_cl on the same
force field, bit for bit, on two calls in a row, the second on a smaller grid (fewer
reduction blocks than the first left in the shared array)."""
import numpy as np
import pytest
import torch

import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg

pytestmark = pytest.mark.gpu


def _check(n, seed):
    s = pkg.FlowSolver(nx=n[0], ny=n[1], nz=n[2])
    f = np.random.default_rng(seed).standard_normal(3 * s.G)
    s.set_force(f)
    cd, cl = s.cd_cl()
    f_dev = torch.from_numpy(f).to("cuda:0")
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    s.cd_cl_noalloc(f_dev.data_ptr(), out.data_ptr())
    s.synchronize()
    got = out.cpu().numpy()
    print(f"{n}: cd {cd!r} cl {cl!r}; no-alloc {got[0]!r} {got[1]!r}")
    assert got[0] == cd and got[1] == cl
    s.close()


def test_cd_cl_noalloc_matches_cd_cl_on_two_calls():
    _check((256, 128, 32), 1)  # 1024 reduction blocks
    _check((40, 20, 6), 2)     # 5 blocks: the first call's partials stay behind them


def test_cd_cl_noalloc_rejects_null_pointers():
    s = pkg.FlowSolver(nx=16, ny=8, nz=4)
    with pytest.raises(pkg.CfdmxError, match="null device pointer"):
        s.cd_cl_noalloc(0, 0)
    s.close()
