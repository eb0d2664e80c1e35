"""ctypes binding of include/cfdmx.h (the C-ABI of libcfdmx.so).

This is the Python view used by the tests and bench.py; C++ callers use the
adapter include/cfdmx_reference_api.hpp, which keeps the reference's own entry
point names.  Loading fails loudly when the HIP library has not been built.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcfdmx.so")

OK, ERR_ARG, ERR_HIP, ERR_STATE = 0, -1, -2, -3
GRID_UNIFORM, GRID_SML = 0, 1
SOLVER_BICGSTAB, SOLVER_RBSOR = 0, 1
PRECOND_NONE, PRECOND_JACOBI = 0, 1
PRECISION_FP64, PRECISION_MIXED, PRECISION_FP64_MG = 0, 1, 2

DP = C.POINTER(C.c_double)

# every function declared in include/cfdmx.h
EXPORTS = [
    "cfdmx_config_default", "cfdmx_create", "cfdmx_destroy", "cfdmx_last_error", "cfdmx_sizes",
    "cfdmx_grid_metrics", "cfdmx_upload_velocity", "cfdmx_download_velocity",
    "cfdmx_upload_pressure", "cfdmx_download_pressure", "cfdmx_upload_xresult",
    "cfdmx_download_xresult", "cfdmx_download_rhs", "cfdmx_upload_rhs", "cfdmx_upload_eta",
    "cfdmx_download_eta",
    "cfdmx_download_force", "cfdmx_set_inflow", "cfdmx_init_flow", "cfdmx_start",
    "cfdmx_set_eta_cylinder_z", "cfdmx_set_eta_smoothed", "cfdmx_convdiff", "cfdmx_bc_main",
    "cfdmx_bc_copy", "cfdmx_pressure_rhs", "cfdmx_solve_bicgstab", "cfdmx_solve_rbsor",
    "cfdmx_transform", "cfdmx_project_ib", "cfdmx_swap_t0_t3", "cfdmx_cd_cl", "cfdmx_cfl",
    "cfdmx_apply_operator", "cfdmx_apply_operator_transpose", "cfdmx_step",
    "cfdmx_synchronize", "cfdmx_profile", "cfdmx_profile_read", "cfdmx_set_tolerance", "cfdmx_set_physics", "cfdmx_upload_force",
    "cfdmx_drag_objective", "cfdmx_adjoint_drag", "cfdmx_plot3d_q", "cfdmx_plot3d_flush",
    "cfdmx_plot3d_x", "cfdmx_checkpoint_save", "cfdmx_checkpoint_load", "cfdmx_smagorinsky",
    "cfdmx_download_viseff", "cfdmx_cd_cl_noalloc",
]

# kernel classes of cfdmx_profile_read
KERNELS = {"convdiff": 0, "bc": 1, "rhs": 2, "bicg_init": 3, "bicg_A": 4, "bicg_B": 5,
           "bicg_C": 6, "bicg_check": 7, "sor_half": 8, "project": 9, "mg": 10, "pb": 11}


class AdjointInfo(C.Structure):
    _fields_ = [("forward_ms", C.c_double), ("reverse_ms", C.c_double),
                ("forward_iters", C.c_int), ("adjoint_iters", C.c_int), ("nsteps", C.c_int),
                ("checkpoints", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("grid_kind", C.c_int),
        ("re", C.c_double), ("dt", C.c_double),
        ("bc_num", (C.c_double * 3) * 6), ("bc_dir", (C.c_double * 3) * 6),
        ("z_periodic", C.c_int), ("outflow_convective", C.c_int),
        ("convective_speed", C.c_double), ("flux_correction", C.c_int),
        ("solver", C.c_int), ("precond", C.c_int), ("tol", C.c_double), ("maxit", C.c_int),
        ("sor_omega", C.c_double), ("skip_first_pressure", C.c_int), ("precision", C.c_int),
        ("inner_tol", C.c_double), ("xcd_remap", C.c_int), ("mg_graphs", C.c_int),
        ("smagorinsky_cs", C.c_double), ("steady_check", C.c_int),
        ("mg_extrapolate", C.c_int),
    ]


class StepInfo(C.Structure):
    _fields_ = [("cd", C.c_double), ("cl", C.c_double), ("time", C.c_double),
                ("iters", C.c_int), ("err", C.c_double), ("cfl", C.c_double),
                ("solved", C.c_int), ("restarts", C.c_int), ("resets", C.c_int),
                ("maxdiff", C.c_double * 3), ("unconverged", C.c_int)]

    def as_dict(self):
        return {n: (tuple(getattr(self, n)) if n == "maxdiff" else getattr(self, n))
                for n, _ in self._fields_}


_lib = None


def load_library(path=LIB_PATH):
    """Load libcfdmx.so; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m "
            "cfd_nssolver_forhpc_flowpassingcylinder_amd.build` (there is no CPU fallback)")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.cfdmx_config_default.argtypes = [C.POINTER(Config)]
    L.cfdmx_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.cfdmx_destroy.argtypes = [vp]
    L.cfdmx_last_error.restype = C.c_char_p
    L.cfdmx_last_error.argtypes = [vp]
    L.cfdmx_sizes.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.cfdmx_grid_metrics.argtypes = [vp, DP]
    L.cfdmx_upload_velocity.argtypes = [vp, C.c_int, DP, DP, DP]
    L.cfdmx_download_velocity.argtypes = [vp, C.c_int, DP, DP, DP]
    for n in ("cfdmx_upload_pressure", "cfdmx_download_pressure", "cfdmx_upload_xresult",
              "cfdmx_download_xresult", "cfdmx_download_rhs", "cfdmx_upload_rhs",
              "cfdmx_upload_eta", "cfdmx_upload_force",
              "cfdmx_download_eta", "cfdmx_download_force", "cfdmx_set_inflow"):
        getattr(L, n).argtypes = [vp, DP]
    L.cfdmx_init_flow.argtypes = [vp, C.c_double, C.c_double, C.c_ulonglong]
    L.cfdmx_plot3d_q.argtypes = [vp, C.c_char_p, C.c_double]
    L.cfdmx_plot3d_flush.argtypes = [vp]
    L.cfdmx_plot3d_x.argtypes = [vp, C.c_char_p]
    L.cfdmx_checkpoint_save.argtypes = [vp, C.c_char_p]
    L.cfdmx_smagorinsky.argtypes = [vp]
    L.cfdmx_download_viseff.argtypes = [vp, DP]
    L.cfdmx_checkpoint_load.argtypes = [vp, C.c_char_p]
    L.cfdmx_set_eta_cylinder_z.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int]
    L.cfdmx_set_eta_smoothed.argtypes = [vp] + [C.c_double] * 4
    for n in ("cfdmx_start", "cfdmx_convdiff", "cfdmx_bc_main", "cfdmx_bc_copy",
              "cfdmx_pressure_rhs", "cfdmx_transform", "cfdmx_project_ib",
              "cfdmx_swap_t0_t3", "cfdmx_synchronize"):
        getattr(L, n).argtypes = [vp]
    L.cfdmx_solve_bicgstab.argtypes = [vp, C.POINTER(C.c_int), DP]
    L.cfdmx_solve_rbsor.argtypes = [vp, C.POINTER(C.c_int), DP]
    L.cfdmx_cd_cl.argtypes = [vp, DP, DP]
    L.cfdmx_cd_cl_noalloc.argtypes = [vp, vp, vp]
    L.cfdmx_cfl.argtypes = [vp, DP]
    L.cfdmx_apply_operator.argtypes = [vp, DP, DP]
    L.cfdmx_apply_operator_transpose.argtypes = [vp, DP, DP]
    L.cfdmx_step.argtypes = [vp, C.c_int, C.POINTER(StepInfo)]
    L.cfdmx_profile.argtypes = [vp, C.c_int]
    L.cfdmx_set_tolerance.argtypes = [vp, C.c_double, C.c_int]
    L.cfdmx_set_physics.argtypes = [vp, C.c_double, C.c_double]
    L.cfdmx_profile_read.argtypes = [vp, C.c_int, DP, C.POINTER(C.c_longlong)]
    L.cfdmx_drag_objective.argtypes = [vp, C.c_int, DP]
    L.cfdmx_adjoint_drag.argtypes = [vp, C.c_int, C.c_int, DP, DP, DP, C.POINTER(AdjointInfo)]
    for n in EXPORTS:
        if n != "cfdmx_last_error":
            getattr(L, n).restype = C.c_int
    _lib = L
    return L


def default_config(**kw):
    L = load_library()
    cfg = Config()
    L.cfdmx_config_default(C.byref(cfg))
    for k, v in kw.items():
        if k in ("bc_num", "bc_dir"):
            for f in range(6):
                for c in range(3):
                    getattr(cfg, k)[f][c] = v[f][c]
        else:
            setattr(cfg, k, v)
    return cfg


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


class CfdmxError(RuntimeError):
    pass


class FlowSolver:
    """One device context (cfdmx_ctx) with numpy-based field movement."""

    def __init__(self, cfg=None, **kw):
        self.L = load_library()
        self.cfg = cfg if cfg is not None else default_config(**kw)
        h = C.c_void_p()
        rc = self.L.cfdmx_create(C.byref(self.cfg), C.byref(h))
        if rc != OK:
            raise CfdmxError(f"cfdmx_create: {rc}: {self.L.cfdmx_last_error(None).decode()}")
        self.h = h
        G, N = C.c_longlong(), C.c_longlong()
        self.L.cfdmx_sizes(self.h, C.byref(G), C.byref(N))
        self.G, self.N = G.value, N.value
        self.nx, self.ny, self.nz = self.cfg.nx + 4, self.cfg.ny + 4, self.cfg.nz + 4

    def _chk(self, rc, what):
        if rc != OK:
            raise CfdmxError(f"{what}: {rc}: {self.L.cfdmx_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.cfdmx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- data movement ----
    def metrics(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        out = np.zeros((nx + 1) + (ny + 1) + (nz + 1) + 2 * (nx + ny + nz))
        self._chk(self.L.cfdmx_grid_metrics(self.h, _p(out)), "grid_metrics")
        sizes = [nx + 1, ny + 1, nz + 1, nx, ny, nz, nx, ny, nz]
        names = ["X", "Y", "Z", "Dx", "Dy", "Dz", "Dxs", "Dys", "Dzs"]
        res, o = {}, 0
        for n, s in zip(names, sizes):
            res[n] = out[o:o + s].copy()
            o += s
        return res

    def upload_velocity(self, which, u, v, w):
        u, v, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w))
        self._chk(self.L.cfdmx_upload_velocity(self.h, which, _p(u), _p(v), _p(w)), "upload")

    def download_velocity(self, which):
        u, v, w = np.empty(self.G), np.empty(self.G), np.empty(self.G)
        self._chk(self.L.cfdmx_download_velocity(self.h, which, _p(u), _p(v), _p(w)),
                  "download")
        return u, v, w

    def _get(self, fn, n):
        a = np.empty(n)
        self._chk(getattr(self.L, fn)(self.h, _p(a)), fn)
        return a

    def _put(self, fn, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._chk(getattr(self.L, fn)(self.h, _p(a)), fn)

    def pressure(self):
        return self._get("cfdmx_download_pressure", self.G)

    def set_pressure(self, p):
        self._put("cfdmx_upload_pressure", p)

    def xresult(self):
        return self._get("cfdmx_download_xresult", self.N)

    def set_xresult(self, x):
        self._put("cfdmx_upload_xresult", x)

    def rhs(self):
        return self._get("cfdmx_download_rhs", self.N)

    def set_rhs(self, b):
        self._put("cfdmx_upload_rhs", b)

    def eta(self):
        return self._get("cfdmx_download_eta", self.G)

    def set_eta(self, eta):
        self._put("cfdmx_upload_eta", eta)

    def force(self):
        return self._get("cfdmx_download_force", 3 * self.G)

    def plot3d_q(self, path, time):
        """Stage the Plot3D q file (p, cell-centred u, v, w, eta); written by the next
        plot3d_q, plot3d_flush or close."""
        self._chk(self.L.cfdmx_plot3d_q(self.h, os.fsencode(path), time), "plot3d_q")

    def plot3d_flush(self):
        self._chk(self.L.cfdmx_plot3d_flush(self.h), "plot3d_flush")

    def plot3d_x(self, path):
        self._chk(self.L.cfdmx_plot3d_x(self.h, os.fsencode(path)), "plot3d_x")

    def smagorinsky(self):
        """Eddy viscosity from T0 (SmagorinskyModel); needs smagorinsky_cs > 0."""
        self._chk(self.L.cfdmx_smagorinsky(self.h), "smagorinsky")

    def viseff(self):
        return self._get("cfdmx_download_viseff", self.N)

    def checkpoint_save(self, path):
        """FP64 state of the run (fields, X_result history, step, time, solver state)."""
        self._chk(self.L.cfdmx_checkpoint_save(self.h, os.fsencode(path)), "checkpoint_save")

    def checkpoint_load(self, path):
        """Restore a checkpoint of a run with this configuration; later steps continue it
        bit for bit."""
        self._chk(self.L.cfdmx_checkpoint_load(self.h, os.fsencode(path)), "checkpoint_load")

    def set_inflow(self, profile):
        if profile is None:
            self._chk(self.L.cfdmx_set_inflow(self.h, None), "set_inflow")
        else:
            self._put("cfdmx_set_inflow", profile)

    # ---- setup ----
    def init_flow(self, u0=1.0, amp=1e-3, seed=20261202):
        self._chk(self.L.cfdmx_init_flow(self.h, u0, amp, seed), "init_flow")

    def start(self):
        self._chk(self.L.cfdmx_start(self.h), "start")

    def set_eta_cylinder_z(self, cx, cy, r, nsub=100):
        self._chk(self.L.cfdmx_set_eta_cylinder_z(self.h, cx, cy, r, nsub), "eta")

    def set_eta_smoothed(self, cx, cy, r, eps):
        self._chk(self.L.cfdmx_set_eta_smoothed(self.h, cx, cy, r, eps), "eta_smoothed")

    # ---- stages ----
    def convdiff(self):
        self._chk(self.L.cfdmx_convdiff(self.h), "convdiff")

    def bc_main(self):
        self._chk(self.L.cfdmx_bc_main(self.h), "bc_main")

    def bc_copy(self):
        self._chk(self.L.cfdmx_bc_copy(self.h), "bc_copy")

    def pressure_rhs(self):
        self._chk(self.L.cfdmx_pressure_rhs(self.h), "pressure_rhs")

    def solve_bicgstab(self):
        it, res = C.c_int(), C.c_double()
        self._chk(self.L.cfdmx_solve_bicgstab(self.h, C.byref(it), C.byref(res)), "bicgstab")
        return it.value, res.value

    def solve_rbsor(self):
        it, res = C.c_int(), C.c_double()
        self._chk(self.L.cfdmx_solve_rbsor(self.h, C.byref(it), C.byref(res)), "rbsor")
        return it.value, res.value

    def transform(self):
        self._chk(self.L.cfdmx_transform(self.h), "transform")

    def project_ib(self):
        self._chk(self.L.cfdmx_project_ib(self.h), "project_ib")

    def swap_t0_t3(self):
        self._chk(self.L.cfdmx_swap_t0_t3(self.h), "swap")

    def cd_cl(self):
        cd, cl = C.c_double(), C.c_double()
        self._chk(self.L.cfdmx_cd_cl(self.h, C.byref(cd), C.byref(cl)), "cd_cl")
        return cd.value, cl.value

    def cd_cl_noalloc(self, f_dev_ptr, cdcl_dev_ptr):
        """Cd, Cl of a force field already in device memory (address f_dev_ptr) into two
        doubles of device memory at cdcl_dev_ptr, on the context's stream, without any
        allocation; call synchronize() before reading the result on another stream."""
        self._chk(self.L.cfdmx_cd_cl_noalloc(self.h, C.c_void_p(f_dev_ptr),
                                             C.c_void_p(cdcl_dev_ptr)), "cd_cl_noalloc")

    def cfl(self):
        c = C.c_double()
        self._chk(self.L.cfdmx_cfl(self.h, C.byref(c)), "cfl")
        return c.value

    def apply_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.empty(self.N)
        self._chk(self.L.cfdmx_apply_operator(self.h, _p(x), _p(r)), "apply_operator")
        return r

    def apply_operator_transpose(self, y):
        """A^T y: the transpose of apply_operator's matrix (no reference counterpart)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        r = np.empty(self.N)
        self._chk(self.L.cfdmx_apply_operator_transpose(self.h, _p(y), _p(r)),
                  "apply_operator_transpose")
        return r

    def step(self, n=1):
        info = StepInfo()
        self._chk(self.L.cfdmx_step(self.h, n, C.byref(info)), "step")
        return info.as_dict()

    def synchronize(self):
        self._chk(self.L.cfdmx_synchronize(self.h), "synchronize")

    def set_force(self, f):
        self._put("cfdmx_upload_force", f)

    def set_physics(self, re, dt):
        self._chk(self.L.cfdmx_set_physics(self.h, re, dt), "set_physics")

    def set_tolerance(self, tol, maxit):
        self._chk(self.L.cfdmx_set_tolerance(self.h, tol, maxit), "set_tolerance")

    # ---- adjoint of the time-averaged drag ----
    def drag_objective(self, nsteps):
        """J = mean Cd over nsteps steps from the current state (state restored)."""
        J = C.c_double()
        self._chk(self.L.cfdmx_drag_objective(self.h, nsteps, C.byref(J)), "drag_objective")
        return J.value

    def adjoint_drag(self, nsteps, checkpoint_every=10, mask=True, inflow=True):
        """(J, dJ/d inflow profile (ghosted j), dJ/d (cx, cy, R), info dict)."""
        J = C.c_double()
        gi = np.zeros(self.ny) if inflow else None
        gm = np.zeros(3) if mask else None
        info = AdjointInfo()
        self._chk(self.L.cfdmx_adjoint_drag(self.h, nsteps, checkpoint_every, C.byref(J),
                                            _p(gi) if inflow else None,
                                            _p(gm) if mask else None, C.byref(info)),
                  "adjoint_drag")
        return J.value, gi, gm, info.as_dict()

    def profile(self, enable):
        self._chk(self.L.cfdmx_profile(self.h, 1 if enable else 0), "profile")

    def profile_read(self):
        """{kernel: (total_ms, launches)} for every kernel class"""
        out = {}
        for name, kid in KERNELS.items():
            ms, n = C.c_double(), C.c_longlong()
            self._chk(self.L.cfdmx_profile_read(self.h, kid, C.byref(ms), C.byref(n)),
                      "profile_read")
            out[name] = (ms.value, n.value)
        return out
