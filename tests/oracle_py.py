"""ctypes front end of the CPU checkers (test infrastructure only).

`Oracle` wraps oracle/liboracle.so (the C restatement of the reference hot path);
`RefHarness` wraps oracle/_ref/libcfdref.so (the reference's own functions,
built here from its headers; absent on machines without the reference sources).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcfdref.so")

DP = C.POINTER(C.c_double)

GRID_UNIFORM, GRID_SML = 0, 1
RED_SERIAL, RED_BLOCKED = 0, 1
SOLVER_BICGSTAB, SOLVER_RBSOR, SOLVER_SOR_SEQ = 0, 1, 2
PRECOND_NONE, PRECOND_JACOBI = 0, 1


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


class OrcGrid(C.Structure):
    _fields_ = [("nxc", C.c_int), ("nyc", C.c_int), ("nzc", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double)] + \
               [(n, DP) for n in ("X", "Y", "Z", "Dx", "Dy", "Dz", "Dxs", "Dys", "Dzs")]


class OrcBC(C.Structure):
    _fields_ = [("num", (C.c_double * 3) * 6), ("dir", (C.c_double * 3) * 6),
                ("inflow", DP), ("z_periodic", C.c_int), ("outflow_convective", C.c_int),
                ("conv_speed", C.c_double), ("flux_correction", C.c_int)]


class OrcParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("nu", C.c_double), ("solver", C.c_int),
                ("precond", C.c_int), ("tol", C.c_double), ("maxit", C.c_int),
                ("omega", C.c_double), ("reduction", C.c_int),
                ("skip_first_pressure", C.c_int), ("swap_buffers", C.c_int),
                ("smag_cs", C.c_double)]


# Num/Dir table of the reference (FLOWpassing.hpp:6-40): inlet u = 1 Dirichlet with
# mirrored v, w; every other face zero-gradient.
FLOWPASSING_NUM = [[0, 0, 0], [1, 1, 1], [1, 1, 1], [1, 1, 1], [1, 1, 1], [1, 1, 1]]
FLOWPASSING_DIR = [[1, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]]
# slip top/bottom walls: normal v = 0 (Num 0, Dir 0), tangential zero-gradient
SLIP_NUM = [[0, 0, 0], [1, 1, 1], [1, 0, 1], [1, 0, 1], [1, 1, 1], [1, 1, 1]]
SLIP_DIR = FLOWPASSING_DIR


class Grid:
    """Grid metrics held in numpy arrays, plus the C view of them."""

    def __init__(self, lib, nxc, nyc, nzc, lx, ly, lz, kind=GRID_UNIFORM):
        self.nxc, self.nyc, self.nzc = nxc, nyc, nzc
        self.nx, self.ny, self.nz = nxc + 4, nyc + 4, nzc + 4
        self.lx, self.ly, self.lz = lx, ly, lz
        self.G = self.nx * self.ny * self.nz
        self.N = nxc * nyc * nzc
        self.X = np.zeros(self.nx + 1)
        self.Y = np.zeros(self.ny + 1)
        self.Z = np.zeros(self.nz + 1)
        self.Dx, self.Dxs = np.zeros(self.nx), np.zeros(self.nx)
        self.Dy, self.Dys = np.zeros(self.ny), np.zeros(self.ny)
        self.Dz, self.Dzs = np.zeros(self.nz), np.zeros(self.nz)
        self.c = OrcGrid(nxc, nyc, nzc, self.nx, self.ny, self.nz, lx, ly, lz,
                         *[ptr(getattr(self, n)) for n in
                           ("X", "Y", "Z", "Dx", "Dy", "Dz", "Dxs", "Dys", "Dzs")])
        rc = lib.orc_grid_build(C.byref(self.c), kind)
        if rc != 0:
            raise ValueError(f"orc_grid_build failed: {rc}")

    def shape(self):
        return (self.nx, self.ny, self.nz)

    def interior(self):
        return (slice(2, self.nx - 2), slice(2, self.ny - 2), slice(2, self.nz - 2))


def make_bc(num=FLOWPASSING_NUM, dir=FLOWPASSING_DIR, inflow=None, z_periodic=0,
            outflow_convective=0, conv_speed=1.0, flux_correction=0):
    bc = OrcBC()
    for f in range(6):
        for c in range(3):
            bc.num[f][c] = num[f][c]
            bc.dir[f][c] = dir[f][c]
    bc.inflow = ptr(inflow) if inflow is not None else None
    bc._inflow_keepalive = inflow
    bc.z_periodic = z_periodic
    bc.outflow_convective = outflow_convective
    bc.conv_speed = conv_speed
    bc.flux_correction = flux_correction
    return bc


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (run: make -C oracle)")
        L = self.lib = C.CDLL(path)
        L.orc_grid_build.argtypes = [C.POINTER(OrcGrid), C.c_int]
        L.orc_blocked_sum.restype = C.c_double
        L.orc_blocked_sum.argtypes = [DP, C.c_long]
        L.orc_march_sum.restype = C.c_double
        L.orc_march_sum.argtypes = [C.POINTER(OrcGrid), DP]
        L.orc_serial_sum.restype = C.c_double
        L.orc_serial_sum.argtypes = [DP, C.c_long]
        G = C.POINTER(OrcGrid)
        L.orc_convdiff.argtypes = [G, C.c_double, C.c_double, C.c_int] + [DP] * 6
        L.orc_convdiff_les.argtypes = [G, C.c_double, C.c_double, DP, C.c_int] + [DP] * 6
        L.orc_smagorinsky_coef.argtypes = [G, C.c_double, DP]
        L.orc_smagorinsky.argtypes = [G] + [DP] * 5
        L.orc_bc_main.argtypes = [G, C.POINTER(OrcBC), C.c_int, C.c_double] + [DP] * 4
        L.orc_bc_copy.argtypes = [G] + [DP] * 6
        L.orc_rhs.argtypes = [G, C.c_double, C.c_int, C.c_int] + [DP] * 4
        L.orc_apply_A.argtypes = [G, C.c_int, C.c_int, DP, DP]
        L.orc_apply_AT.argtypes = [G, C.c_int, C.c_int, DP, DP]
        L.orc_spe_build.argtypes = [G, DP]
        L.orc_spe_apply_vals.argtypes = [G, DP, DP, DP]
        L.orc_bicg_new.restype = C.c_void_p
        L.orc_bicg_new.argtypes = [C.c_long]
        L.orc_bicg_free.argtypes = [C.c_void_p]
        L.orc_bicg_solve.argtypes = [C.c_void_p, G, C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_int, DP, DP, DP]
        L.orc_transform.argtypes = [G, DP, DP]
        L.orc_sor.argtypes = [G, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                              DP, DP, DP, DP, DP]
        L.orc_update_uf.argtypes = [G, C.c_double, C.c_int, C.c_int] + [DP] * 9
        L.orc_eta_cylinder_z.argtypes = [G, C.c_double, C.c_double, C.c_double, C.c_int, DP]
        L.orc_eta_smoothed.argtypes = [G, C.c_double, C.c_double, C.c_double, C.c_double, DP]
        L.orc_cd_cl.argtypes = [G, C.c_int, DP, DP, DP]
        L.orc_cfl.restype = C.c_double
        L.orc_cfl.argtypes = [G, C.c_double, C.c_int] + [DP] * 3
        L.orc_maxdiff.argtypes = [G] + [DP] * 7
        L.orc_maxdiff_faces.argtypes = [G, C.c_int] + [DP] * 7
        L.orc_init_velocity.argtypes = [G, C.c_double, C.c_double, C.c_ulonglong, DP, DP, DP]
        L.orc_sim_new.restype = C.c_void_p
        L.orc_sim_new.argtypes = [G, C.POINTER(OrcBC), C.POINTER(OrcParams)]
        L.orc_sim_free.argtypes = [C.c_void_p]
        L.orc_sim_field.restype = DP
        L.orc_sim_field.argtypes = [C.c_void_p, C.c_int]
        L.orc_sim_start.argtypes = [C.c_void_p]
        L.orc_sim_step.argtypes = [C.c_void_p, DP]
        L.orc_adjoint_drag.restype = C.c_double
        L.orc_adjoint_drag.argtypes = [G, C.POINTER(OrcBC), C.POINTER(OrcParams), DP, DP, DP,
                                       C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                       C.c_double, DP, DP]
        L.orc_drag_objective.restype = C.c_double
        L.orc_drag_objective.argtypes = [G, C.POINTER(OrcBC), C.POINTER(OrcParams), DP, DP, DP,
                                         C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_adj_convdiff.argtypes = [G, C.c_double, C.c_double, C.c_int] + [DP] * 9
        L.orc_adj_bc_main.argtypes = [G, C.POINTER(OrcBC), C.c_int, C.c_double] + [DP] * 4

    # ---- adjoint of the time-averaged drag (cfd_adjoint.c) ----
    def adjoint_drag(self, g, bc, prm, u, v, w, mask, nsteps, adj_tol=1e-12):
        """(J, grad_inflow[ny], grad_mask[3]) with mask = (cx, cy, R, eps)."""
        gi = np.zeros(g.ny)
        gm = np.zeros(3)
        J = self.lib.orc_adjoint_drag(C.byref(g.c), C.byref(bc), C.byref(prm), ptr(u), ptr(v),
                                      ptr(w), *mask, nsteps, adj_tol, ptr(gi), ptr(gm))
        return J, gi, gm

    def drag_objective(self, g, bc, prm, u, v, w, mask, nsteps):
        return self.lib.orc_drag_objective(C.byref(g.c), C.byref(bc), C.byref(prm), ptr(u),
                                           ptr(v), ptr(w), *mask, nsteps)

    def adj_convdiff(self, g, dt, nu, u, v, w, lus, lvs, lws, z_periodic=0):
        lu, lv, lw = np.zeros(g.G), np.zeros(g.G), np.zeros(g.G)
        self.lib.orc_adj_convdiff(C.byref(g.c), dt, nu, z_periodic,
                                  *map(ptr, (u, v, w, lus, lvs, lws, lu, lv, lw)))
        return lu, lv, lw

    def adj_bc_main(self, g, bc, lu, lv, lw, advance=0, dt=0.0):
        """In place on (lu, lv, lw); returns the inflow adjoint (ghosted j)."""
        li = np.zeros(g.ny)
        self.lib.orc_adj_bc_main(C.byref(g.c), C.byref(bc), advance, dt,
                                 *map(ptr, (lu, lv, lw, li)))
        return li

    def grid(self, nxc, nyc, nzc, lx=20.0, ly=10.0, lz=0.5, kind=GRID_UNIFORM):
        return Grid(self.lib, nxc, nyc, nzc, lx, ly, lz, kind)

    # ---- stage wrappers (numpy in / numpy out) ----
    def convdiff(self, g, dt, nu, u, v, w, us, vs, ws, z_periodic=0):
        self.lib.orc_convdiff(C.byref(g.c), dt, nu, z_periodic, *map(ptr, (u, v, w, us, vs, ws)))

    def convdiff_les(self, g, dt, nu, viseff, u, v, w, us, vs, ws, z_periodic=0):
        self.lib.orc_convdiff_les(C.byref(g.c), dt, nu, ptr(viseff), z_periodic,
                                  *map(ptr, (u, v, w, us, vs, ws)))

    def smagorinsky(self, g, cs, u, v, w):
        """Eddy viscosity per interior cell (icelCal order) from T0."""
        coef, vis = np.zeros(g.N), np.zeros(g.N)
        self.lib.orc_smagorinsky_coef(C.byref(g.c), cs, ptr(coef))
        self.lib.orc_smagorinsky(C.byref(g.c), ptr(coef), *map(ptr, (u, v, w)), ptr(vis))
        return vis

    def bc_main(self, g, bc, u, v, w, p, advance=0, dt=0.0):
        self.lib.orc_bc_main(C.byref(g.c), C.byref(bc), advance, dt, *map(ptr, (u, v, w, p)))

    def bc_copy(self, g, u0, v0, w0, u1, v1, w1):
        self.lib.orc_bc_copy(C.byref(g.c), *map(ptr, (u0, v0, w0, u1, v1, w1)))

    def rhs(self, g, dt, us, vs, ws, precond=PRECOND_NONE, z_periodic=0):
        b = np.zeros(g.N)
        self.lib.orc_rhs(C.byref(g.c), dt, precond, z_periodic, ptr(us), ptr(vs), ptr(ws), ptr(b))
        return b

    def apply_A(self, g, x, precond=PRECOND_NONE, z_periodic=0):
        r = np.zeros(g.N)
        self.lib.orc_apply_A(C.byref(g.c), precond, z_periodic, ptr(np.ascontiguousarray(x)),
                             ptr(r))
        return r

    def apply_AT(self, g, y, precond=PRECOND_NONE, z_periodic=0):
        r = np.zeros(g.N)
        self.lib.orc_apply_AT(C.byref(g.c), precond, z_periodic, ptr(np.ascontiguousarray(y)),
                              ptr(r))
        return r

    def spe_vals(self, g):
        val = np.zeros(7 * g.N)
        self.lib.orc_spe_build(C.byref(g.c), ptr(val))
        return val

    def spe_apply_vals(self, g, val, x):
        r = np.zeros(g.N)
        self.lib.orc_spe_apply_vals(C.byref(g.c), ptr(val), ptr(x), ptr(r))
        return r

    def blocked_sum(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        return self.lib.orc_blocked_sum(ptr(v), v.size)

    def march_sum(self, g, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == g.N
        return self.lib.orc_march_sum(C.byref(g.c), ptr(v))

    def serial_sum(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        return self.lib.orc_serial_sum(ptr(v), v.size)

    def update_uf(self, g, dt, p, eta, us, vs, ws, u, v, w, f, z_periodic=0, convective=0):
        self.lib.orc_update_uf(C.byref(g.c), dt, z_periodic, convective,
                               *map(ptr, (p, eta, us, vs, ws, u, v, w, f)))

    def eta_cylinder_z(self, g, cx, cy, r, nsub=100):
        eta = np.zeros(g.G)
        self.lib.orc_eta_cylinder_z(C.byref(g.c), cx, cy, r, nsub, ptr(eta))
        return eta

    def eta_smoothed(self, g, cx, cy, r, eps):
        eta = np.zeros(g.G)
        self.lib.orc_eta_smoothed(C.byref(g.c), cx, cy, r, eps, ptr(eta))
        return eta

    def cd_cl(self, g, f, reduction=RED_SERIAL):
        cd, cl = C.c_double(), C.c_double()
        self.lib.orc_cd_cl(C.byref(g.c), reduction, ptr(f), C.byref(cd), C.byref(cl))
        return cd.value, cl.value

    def cfl(self, g, dt, u, v, w, z_periodic=0):
        return self.lib.orc_cfl(C.byref(g.c), dt, z_periodic, ptr(u), ptr(v), ptr(w))

    def maxdiff(self, g, u0, v0, w0, u3, v3, w3):
        """max(0, max over interior cells of T0 - T3) per component"""
        out = np.zeros(3)
        self.lib.orc_maxdiff(C.byref(g.c), *map(ptr, (u0, v0, w0, u3, v3, w3)), ptr(out))
        return tuple(out)

    def maxdiff_faces(self, g, u0, v0, w0, u3, v3, w3, z_periodic=0):
        """the same over the faces the projection computes"""
        out = np.zeros(3)
        self.lib.orc_maxdiff_faces(C.byref(g.c), z_periodic,
                                   *map(ptr, (u0, v0, w0, u3, v3, w3)), ptr(out))
        return tuple(out)

    def init_velocity(self, g, u0=1.0, amp=1e-3, seed=20261202):
        u, v, w = np.zeros(g.G), np.zeros(g.G), np.zeros(g.G)
        self.lib.orc_init_velocity(C.byref(g.c), u0, amp, seed, ptr(u), ptr(v), ptr(w))
        return u, v, w

    def sor(self, g, dt, us, vs, ws, p, omega=1.8, tol=1e-10, itmax=5000, redblack=1,
            z_periodic=0):
        err = C.c_double()
        it = self.lib.orc_sor(C.byref(g.c), redblack, z_periodic, dt, omega, tol, itmax,
                              *map(ptr, (us, vs, ws, p)), C.byref(err))
        return it, err.value


class BiCG:
    def __init__(self, oracle, g):
        self.o, self.g = oracle, g
        self.h = oracle.lib.orc_bicg_new(g.N)

    def solve(self, b, x, tol, precond=PRECOND_NONE, z_periodic=0, reduction=RED_SERIAL,
              maxit=5000):
        res = C.c_double()
        it = self.o.lib.orc_bicg_solve(self.h, C.byref(self.g.c), precond, z_periodic, reduction,
                                       tol, maxit, ptr(b), ptr(x), C.byref(res))
        return it, res.value

    def __del__(self):
        try:
            self.o.lib.orc_bicg_free(self.h)
        except Exception:
            pass


class Sim:
    """The oracle's whole time loop (orc_sim_*)."""

    FIELDS = {"u0": 0, "v0": 1, "w0": 2, "u1": 3, "v1": 4, "w1": 5, "u3": 6, "v3": 7,
              "w3": 8, "p": 9, "eta": 10, "f": 11, "x": 12, "b": 13}

    def __init__(self, oracle, g, bc, dt, re, solver=SOLVER_BICGSTAB, precond=PRECOND_NONE,
                 tol=1e-3, maxit=5000, omega=1.8, reduction=RED_SERIAL, skip_first=1, swap=0,
                 smag_cs=0.0):
        self.o, self.g, self.bc = oracle, g, bc
        self.prm = OrcParams(dt, 1.0 / re, solver, precond, tol, maxit, omega, reduction,
                             skip_first, swap, smag_cs)
        self.h = oracle.lib.orc_sim_new(C.byref(g.c), C.byref(bc), C.byref(self.prm))

    def field(self, name):
        which = self.FIELDS[name]
        n = 3 * self.g.G if name == "f" else (self.g.N if name in ("x", "b") else self.g.G)
        p = self.o.lib.orc_sim_field(self.h, which)
        return np.ctypeslib.as_array(p, shape=(n,))

    def start(self):
        self.o.lib.orc_sim_start(self.h)

    def step(self):
        out = np.zeros(11)
        self.o.lib.orc_sim_step(self.h, ptr(out))
        return dict(cd=out[0], cl=out[1], time=out[2], iters=int(out[3]), err=out[4],
                    cfl=out[5], restarts=int(out[6]), solved=int(out[7]),
                    maxdiff=tuple(out[8:11]))

    def __del__(self):
        try:
            self.o.lib.orc_sim_free(self.h)
        except Exception:
            pass


class RefHarness:
    """The reference's own functions (oracle/_ref/libcfdref.so)."""

    def __init__(self, nxc, nyc, nzc, lx=20.0, ly=10.0, lz=0.5, sml=0, re=100.0, dt=0.002,
                 precond=PRECOND_NONE, tol=1e-3, threads=1, path=REF_SO):
        L = self.lib = C.CDLL(path)
        L.ref_new.restype = C.c_void_p
        L.ref_new.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [C.c_int, C.c_double,
                                                                  C.c_double, C.c_int,
                                                                  C.c_double, C.c_int]
        for fn in ("ref_free", "ref_start", "ref_convdiff", "ref_bc_main", "ref_bc_copy",
                   "ref_rhs", "ref_transform", "ref_update", "ref_copyvel", "ref_eta"):
            getattr(L, fn).argtypes = [C.c_void_p]
        L.ref_field.restype = DP
        L.ref_field.argtypes = [C.c_void_p, C.c_int]
        L.ref_metrics.argtypes = [C.c_void_p, DP]
        L.ref_solve.argtypes = [C.c_void_p, DP]
        L.ref_cdcl.argtypes = [C.c_void_p, DP, DP]
        L.ref_apply_A.argtypes = [C.c_void_p, DP, DP]
        L.ref_sor_seq.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, DP]
        L.ref_step.argtypes = [C.c_void_p, C.c_int, DP]
        L.ref_maxdiff.argtypes = [C.c_void_p, DP]
        L.ref_set_maxit.argtypes = [C.c_void_p, C.c_int]
        L.ref_set_tol.argtypes = [C.c_void_p, C.c_double]
        L.ref_set_threads.argtypes = [C.c_int]
        L.ref_plot3d_q.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.ref_plot3d_x.argtypes = [C.c_void_p]
        self.nx, self.ny, self.nz = nxc + 4, nyc + 4, nzc + 4
        self.G = self.nx * self.ny * self.nz
        self.N = nxc * nyc * nzc
        self.h = L.ref_new(nxc, nyc, nzc, lx, ly, lz, sml, re, dt, precond, tol, threads)

    def field(self, name):
        which = Sim.FIELDS[name]
        n = 3 * self.G if name == "f" else (self.N if name in ("x", "b") else self.G)
        return np.ctypeslib.as_array(self.lib.ref_field(self.h, which), shape=(n,))

    def metrics(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        out = np.zeros((nx + 1) + (ny + 1) + (nz + 1) + 2 * (nx + ny + nz))
        self.lib.ref_metrics(self.h, ptr(out))
        sizes = [nx + 1, ny + 1, nz + 1, nx, ny, nz, nx, ny, nz]
        names = ["X", "Y", "Z", "Dx", "Dy", "Dz", "Dxs", "Dys", "Dzs"]
        res, o = {}, 0
        for n, s in zip(names, sizes):
            res[n] = out[o:o + s].copy()
            o += s
        return res

    def call(self, name):
        getattr(self.lib, "ref_" + name)(self.h)

    def solve(self):
        err = C.c_double()
        it = self.lib.ref_solve(self.h, C.byref(err))
        return it, err.value

    def cdcl(self):
        cd, cl = C.c_double(), C.c_double()
        self.lib.ref_cdcl(self.h, C.byref(cd), C.byref(cl))
        return cd.value, cl.value

    def apply_A(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.zeros(self.N)
        self.lib.ref_apply_A(self.h, ptr(x), ptr(r))
        return r

    def sor_seq(self, omega, tol, itmax):
        err = C.c_double()
        it = self.lib.ref_sor_seq(self.h, omega, tol, itmax, C.byref(err))
        return it, err.value

    def set_maxit(self, maxit):
        self.lib.ref_set_maxit(self.h, maxit)

    def set_tol(self, tol):
        self.lib.ref_set_tol(self.h, tol)

    def set_threads(self, threads):
        self.lib.ref_set_threads(threads)

    def maxdiff(self):
        """CheckSteadyStatebyMaxVal_omp on the harness's T0 and T3"""
        out = np.zeros(3)
        assert self.lib.ref_maxdiff(self.h, ptr(out)) == 0
        return tuple(out)

    def step(self, loop):
        out = np.zeros(6)
        self.lib.ref_step(self.h, loop, ptr(out))
        return dict(cd=out[0], cl=out[1], time=out[2], iters=int(out[3]), err=out[4],
                    solved=int(out[5]))

    def __del__(self):
        try:
            self.lib.ref_free(self.h)
        except Exception:
            pass


def ref_available():
    return os.path.exists(REF_SO)
