"""Writes multigrid_digests.json: SHA-256 digests of the fields after three time steps in
the two multigrid modes (FP32 inner solve with FP64 refinement, and the same algorithm in
FP64) on six grids that cover the kernel variants of the V-cycle: levels with nz a
multiple of 4, even nz, odd nz, periodic and non-periodic z, z-only and full coarsening.

The multigrid modes have no counterpart in the reference and no CPU restatement, so these
digests pin them against later changes of their kernels: each kernel variant keeps every
cell update's terms and order, and a run must give the same bits.  The digests were made
by this library on an MI355X (ROCm 7.2) before the two-cells-per-thread FP64 half sweep
was added, and the run with it gave the same digests, as did the run with the two
coarsest levels in one workgroup.  The mixed-mode digests were made again when the
inner solve's correction became FP64 (its arithmetic changed); the FP64 multigrid ones
stayed the same.  The FP64 multigrid digest of the walled 64x32x16 case was made again
when a refinement pass that would raise the residual came to be dropped instead of
applied: that case is inconsistent (no outlet flux correction) and has such a pass; the
eleven others kept their bits.  The inner tolerances are the ones they were made with
(INNER_TOL).  Three steps end before the extrapolated start takes effect (it needs two
earlier solves).

Usage (on a machine with the GPU): python tests/golden/make_multigrid_digests.py [out.json]
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

OUT = os.path.join(HERE, "multigrid_digests.json")

# (nx, ny, nz, periodic z, lz)
CASES = [(512, 256, 64, 1, 0.5), (64, 32, 16, 0, 0.5), (48, 24, 8, 1, 0.5),
         (96, 48, 6, 1, 1.25), (64, 32, 8, 1, 2.5), (48, 24, 5, 1, 20.0 * 5 / 48)]
STEPS = 3
SEED = 11
# the inner tolerances the digests were made with, named here so that they keep pinning
# the kernels whatever the library's per-mode defaults are
INNER_TOL = {1: 3e-3, 2: 1e-5}


def run_case(nx, ny, nz, zper, lz, precision):
    """Digest of T0 (u, v, w), p and X_result after STEPS steps."""
    import numpy as np
    import cfd_nssolver_forhpc_flowpassingcylinder_amd as pkg
    from cfd_nssolver_forhpc_flowpassingcylinder_amd.api import PRECOND_NONE, SOLVER_BICGSTAB
    slip_num = [[0, 0, 0], [1, 1, 1], [1, 0, 1], [1, 0, 1], [1, 1, 1], [1, 1, 1]]
    slip_dir = [[1, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]]
    cfg = pkg.default_config(nx=nx, ny=ny, nz=nz, lx=20.0, ly=10.0, lz=lz, re=200.0,
                             dt=0.001, tol=1e-10, maxit=5000, precond=PRECOND_NONE,
                             solver=SOLVER_BICGSTAB, bc_num=slip_num, bc_dir=slip_dir,
                             z_periodic=zper, outflow_convective=1, convective_speed=1.0,
                             flux_correction=zper, skip_first_pressure=1, precision=precision,
                             inner_tol=INNER_TOL[precision], mg_extrapolate=1)
    s = pkg.FlowSolver(cfg)
    try:
        s.init_flow(1.0, 1e-3, SEED)
        s.set_eta_cylinder_z(5.0, 5.0, 0.5, 100)
        for _ in range(STEPS):
            s.step(1)
        s.synchronize()
        h = hashlib.sha256()
        for a in list(s.download_velocity(0)) + [s.pressure(), s.xresult()]:
            h.update(np.ascontiguousarray(a, dtype="<f8").tobytes())
        return h.hexdigest()
    finally:
        s.close()


def key(nx, ny, nz, zper, lz, precision):
    return f"{nx}x{ny}x{nz} zper {zper} lz {lz:.6g} precision {precision}"


def main():
    out = {}
    for c in CASES:
        for precision in (1, 2):
            out[key(*c, precision)] = run_case(*c, precision)
            print(key(*c, precision), out[key(*c, precision)], flush=True)
    with open(sys.argv[1] if len(sys.argv) > 1 else OUT, "w") as f:
        json.dump({"steps": STEPS, "seed": SEED, "digests": out}, f, indent=1)


if __name__ == "__main__":
    main()
