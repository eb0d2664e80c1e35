"""Writes c1_128x64x4_200steps.json: the oracle's run of the CPU-runnable configuration
(128x64x4 cells over 20 x 10 x 0.5, cylinder of diameter 1 at (5, 5), uniform inflow 1,
convective outflow with flux correction, slip y walls, periodic z, Re 100, dt 0.002,
BiCGSTAB on the Jacobi-scaled rows at tol 1e-8, seeded +-1e-3 noise) over its full 200
steps, in the device's canonical reduction order.

Each step's report (iterations, residual, Cd, Cl, CFL, time) is stored exactly, as
float.hex strings, and the final fields (T0 u, v, w, p, X_result) as SHA-256 digests of
their little-endian float64 bytes, so the GPU run can be compared bit for bit without
running the oracle for 200 steps.

Usage: python tests/golden/make_c1_golden.py   (about 5 minutes on one core)
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import oracle_py as op  # noqa: E402

OUT = os.path.join(HERE, "c1_128x64x4_200steps.json")
CASE = dict(n=(128, 64, 4), lengths=(20.0, 10.0, 0.5), re=100.0, dt=0.002, tol=1e-8,
            precond=op.PRECOND_JACOBI, maxit=5000, seed=20261202, amp=1e-3,
            cylinder=(5.0, 5.0, 0.5, 100), steps=200, skip_first_pressure=1)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def c1_sim(o):
    c = CASE
    g = o.grid(*c["n"], *c["lengths"])
    bc = op.make_bc(op.SLIP_NUM, op.SLIP_DIR, z_periodic=1, outflow_convective=1,
                    conv_speed=1.0, flux_correction=1)
    sim = op.Sim(o, g, bc, c["dt"], c["re"], precond=c["precond"], tol=c["tol"],
                 maxit=c["maxit"], reduction=op.RED_BLOCKED, swap=1,
                 skip_first=c["skip_first_pressure"])
    u, v, w = o.init_velocity(g, 1.0, c["amp"], c["seed"])
    sim.field("u0")[:] = u
    sim.field("v0")[:] = v
    sim.field("w0")[:] = w
    sim.field("eta")[:] = o.eta_cylinder_z(g, *c["cylinder"])
    sim.start()
    return sim


def report(r):
    return {"iters": int(r["iters"]), "err": float(r["err"]).hex(), "cd": float(r["cd"]).hex(),
            "cl": float(r["cl"]).hex(), "cfl": float(r["cfl"]).hex(),
            "time": float(r["time"]).hex()}


def main():
    o = op.Oracle()
    sim = c1_sim(o)
    steps = []
    for s in range(CASE["steps"]):
        steps.append(report(sim.step()))
        if (s + 1) % 20 == 0:
            print(f"step {s + 1}: {steps[-1]['iters']} iterations", flush=True)
    fields = {n: digest(sim.field(n)) for n in ("u0", "v0", "w0", "p", "x")}
    with open(OUT, "w") as f:
        json.dump({"case": CASE, "steps": steps, "final_sha256": fields}, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
