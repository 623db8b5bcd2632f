"""ctypes binding of libkofm.so (the C ABI declared in include/kofm.h).

The HIP library is required: importing works without it, but every call that
needs it raises ``KofmLibraryMissing``.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# KOFM_LIBRARY names another build of the library (for example one compiled
# with different code-generation flags); the default is the in-tree build.
LIB_PATH = os.environ.get("KOFM_LIBRARY") or os.path.join(HERE, "libkofm.so")

STORE_F32, STORE_F16, STORE_BF16, STORE_F64, STORE_F32X = 0, 1, 2, 3, 4
STORAGE_NAMES = {"f32": STORE_F32, "f16": STORE_F16, "bf16": STORE_BF16, "f64": STORE_F64,
                 "f32x": STORE_F32X}
OPTION_SAVE_MACRO_FIELDS = 0x2

FIELD_F, FIELD_RHO, FIELD_UX, FIELD_UY = 0, 1, 2, 3
FIELD_KAPPA, FIELD_OMEGA, FIELD_NUT, FIELD_SXY, FIELD_WALLD = 4, 5, 6, 7, 8
FIELD_TAU_NODE, FIELD_TAU_BL, FIELD_DIFF_K, FIELD_DIFF_OMEGA, FIELD_TEMP = 9, 10, 11, 12, 13
FIELD_NAMES = {
    "f": FIELD_F, "rho": FIELD_RHO, "ux": FIELD_UX, "uy": FIELD_UY, "k": FIELD_KAPPA,
    "omega": FIELD_OMEGA, "nut": FIELD_NUT, "S": FIELD_SXY, "wallD": FIELD_WALLD,
    "tau_node": FIELD_TAU_NODE, "tau_bl": FIELD_TAU_BL, "diff_k": FIELD_DIFF_K,
    "diff_omega": FIELD_DIFF_OMEGA, "T": FIELD_TEMP,
}
RED_SUM, RED_MAX, RED_MIN = 1, 2, 4
DT_QUERY, DT_HALVE, DT_DOUBLE, DT_INJECT_NAN = 0, 1, 2, 3

GEOM = {
    "nType": (0, np.int16), "dX": (1, np.float64), "ibb": (2, np.int32), "ibbQ": (3, np.float64),
    "ss": (4, np.int32), "ssMap": (5, np.int32), "ssInds": (6, np.int32),
    "lsMap": (7, np.uint16), "wallD": (8, np.float64), "cfss": (9, np.float64),
    "sInd": (10, np.int32), "weights": (11, np.float64), "bounds": (12, np.int32),
    "masses": (13, np.float64), "ibbSeg": (14, np.int32), "ibbDqDy": (15, np.float64),
}

# Every symbol the header declares; tests check the library exports them all.
EXPORTS = [
    "kofm_create", "kofm_create_host", "kofm_destroy", "kofm_last_error", "kofm_set_params", "kofm_build_geometry",
    "kofm_geometry_counts", "kofm_geometry_array", "kofm_init_state", "kofm_init_komega",
    "kofm_upload_state", "kofm_step", "kofm_time_steps", "kofm_wall_shear",
    "kofm_get_solver_stats", "kofm_reduce", "kofm_download", "kofm_save2file",
    "kofm_synchronize", "kofm_set_launch", "kofm_autotune", "kofm_set_graphs",
    "kofm_objective_run", "kofm_adjoint", "kofm_geometry_field_gradient",
    "kofm_ensemble_create", "kofm_ensemble_step", "kofm_ensemble_time_steps",
    "kofm_ensemble_set_block", "kofm_ensemble_set_launch", "kofm_ensemble_set_cache_budget",
    "kofm_ensemble_autotune",
    "kofm_ensemble_last_error",
    "kofm_ensemble_destroy",
    "kofm_reduce_device", "kofm_komega_time_step", "kofm_ini_poiseuille",
    "kofm_save_restart", "kofm_save_debug", "kofm_update_walls",
    "kofm_thermal_init", "kofm_thermal_step", "kofm_thermal_time_step", "kofm_set_temperature",
    "kofm_nusselt_init", "kofm_nusselt_sizes", "kofm_nusselt",
    "kofm_shear_removal", "kofm_removal_links",
]


class KofmLibraryMissing(RuntimeError):
    pass


class KofmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"kofm error {code}: {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("nX", C.c_int32), ("nY", C.c_int32), ("storage", C.c_int32),
                ("use_komega", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("tau0", C.c_double), ("Fval", C.c_double), ("nu", C.c_double),
                ("rho0", C.c_double), ("max_turb_visc", C.c_double), ("ko_dt", C.c_double),
                ("kappa_rel_tol", C.c_double), ("kappa_abs_tol", C.c_double),
                ("kappa_max_iters", C.c_int32), ("omega_rel_tol", C.c_double),
                ("omega_abs_tol", C.c_double), ("omega_max_iters", C.c_int32)]


class GeomParams(C.Structure):
    _fields_ = [("pipe_radius", C.c_double), ("eps", C.c_double), ("min_dx", C.c_double),
                ("wall_d_search_rad", C.c_int32), ("x_release_pos", C.c_double),
                ("x_min_val", C.c_double), ("x_stop_pos", C.c_double),
                ("cutoff_radius", C.c_double), ("index_radius_search", C.c_int32)]


class GeomCounts(C.Structure):
    _fields_ = [("nX", C.c_int32), ("nY", C.c_int32), ("XsizeFull", C.c_int32), ("D", C.c_int64),
                ("nIbb", C.c_int32), ("nSS", C.c_int32), ("nBLshear", C.c_int32),
                ("nIbbNodes", C.c_int32)]


class Objective(C.Structure):
    _fields_ = [("window", C.c_int32), ("x_in", C.c_int32), ("x_out", C.c_int32),
                ("flags", C.c_int32), ("shear_weight", C.c_double)]


ADJOINT_FROZEN_TURBULENCE = 1


class IniFlow(C.Structure):
    _fields_ = [("Re", C.c_double), ("nu", C.c_double), ("pipe_radius", C.c_double),
                ("rho", C.c_double), ("turbulent_profile", C.c_int32), ("perturb", C.c_int32),
                ("duplus_mult", C.c_double), ("epsilon_mult", C.c_double),
                ("beta_mult", C.c_double), ("alpha_mult", C.c_double),
                ("sigma_mult", C.c_double), ("seed", C.c_uint32)]


class SolverStats(C.Structure):
    _fields_ = [("kappa_iters", C.c_int32), ("omega_iters", C.c_int32),
                ("kappa_res0", C.c_double), ("kappa_res", C.c_double),
                ("omega_res0", C.c_double), ("omega_res", C.c_double), ("nan_flag", C.c_int32),
                ("steps_per_lb", C.c_int32), ("time_step", C.c_double),
                ("batches", C.c_int32), ("batch_rollbacks", C.c_int32),
                ("temp_iters", C.c_int32), ("temp_res0", C.c_double), ("temp_res", C.c_double),
                ("temp_steps_per_lb", C.c_int32), ("temp_time_step", C.c_double)]


class ThermalParams(C.Structure):
    _fields_ = [("alpha_fluid", C.c_double), ("alpha_foul", C.c_double), ("pr_turb", C.c_double),
                ("t_in", C.c_double), ("t_initial", C.c_double), ("dt", C.c_double),
                ("rel_tol", C.c_double), ("abs_tol", C.c_double), ("max_iters", C.c_int32)]


class NusseltParams(C.Structure):
    _fields_ = [("height", C.c_int32), ("wavy_start", C.c_int32), ("wavy_period", C.c_int32),
                ("wavy_periods", C.c_int32), ("k_soot", C.c_double), ("k_air", C.c_double)]


class RemovalParams(C.Structure):
    _fields_ = [("n_sizes", C.c_int32), ("tau_crit", C.c_void_p), ("drag_coeff", C.c_void_p),
                ("lift_coeff", C.c_void_p), ("mass", C.c_void_p), ("dt_wall", C.c_double),
                ("foul_size_switch", C.c_double), ("loc_stride", C.c_int32),
                ("timer_start", C.c_int32)]


class Particles(C.Structure):
    """Device addresses of the tracer's particle arrays."""
    _fields_ = [("pos", C.c_void_p), ("num_rep", C.c_void_p), ("type", C.c_void_p),
                ("dep_flag", C.c_void_p), ("dep_timer", C.c_void_p), ("loc", C.c_void_p),
                ("timer", C.c_void_p)]


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads libkofm.so (once) and declares the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise KofmLibraryMissing(
            f"{path} is not built; run komegafoulingmodel_amd/build.py (or __graft_entry__.build())")
    lib = C.CDLL(path)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    sig = {
        "kofm_create": (C.c_int, [C.POINTER(Grid), C.POINTER(P)]),
        "kofm_create_host": (C.c_int, [C.POINTER(Grid), C.POINTER(P)]),
        "kofm_destroy": (None, [P]),
        "kofm_last_error": (C.c_char_p, [P]),
        "kofm_set_params": (C.c_int, [P, C.POINTER(Params)]),
        "kofm_build_geometry": (C.c_int, [P, dp, C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                          C.POINTER(GeomParams)]),
        "kofm_geometry_counts": (C.c_int, [P, C.POINTER(GeomCounts)]),
        "kofm_update_walls": (C.c_int, [P, dp, C.c_int32]),
        "kofm_geometry_array": (C.c_int, [P, C.c_int32, P]),
        "kofm_init_state": (C.c_int, [P, dp, dp, dp]),
        "kofm_init_komega": (C.c_int, [P, C.c_double, C.c_double, C.c_double, C.c_double]),
        "kofm_upload_state": (C.c_int, [P, dp, dp, dp, dp, dp]),
        "kofm_step": (C.c_int, [P, C.c_int32, C.c_uint32]),
        "kofm_time_steps": (C.c_int, [P, C.c_int32, C.c_uint32, dp, dp]),
        "kofm_wall_shear": (C.c_int, [P]),
        "kofm_get_solver_stats": (C.c_int, [P, C.POINTER(SolverStats)]),
        "kofm_komega_time_step": (C.c_int, [P, C.c_int32, C.POINTER(C.c_int32), dp]),
        "kofm_thermal_init": (C.c_int, [P, C.POINTER(ThermalParams)]),
        "kofm_thermal_step": (C.c_int, [P, C.c_int32]),
        "kofm_thermal_time_step": (C.c_int, [P, C.c_int32, C.POINTER(C.c_int32), dp]),
        "kofm_set_temperature": (C.c_int, [P, dp]),
        "kofm_nusselt_init": (C.c_int, [P, C.POINTER(NusseltParams)]),
        "kofm_nusselt_sizes": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "kofm_nusselt": (C.c_int, [P, dp, dp, C.c_void_p, dp]),
        "kofm_shear_removal": (C.c_int, [P, C.POINTER(RemovalParams), C.POINTER(Particles),
                                         C.c_int64, C.c_int64, C.c_void_p]),
        "kofm_removal_links": (C.c_int, [P, C.c_double, dp, dp, dp, C.c_void_p]),
        "kofm_ini_poiseuille": (C.c_int, [P, C.POINTER(IniFlow), dp, dp, dp]),
        "kofm_reduce": (C.c_int, [P, C.c_int32, C.c_int32, dp]),
        "kofm_reduce_device": (C.c_int, [P, C.c_int32, C.c_int64, C.c_int32, P, P]),
        "kofm_download": (C.c_int, [P, C.c_int32, dp]),
        "kofm_save2file": (C.c_int, [P, C.c_char_p]),
        "kofm_save_restart": (C.c_int, [P, C.c_char_p]),
        "kofm_save_debug": (C.c_int, [P, C.c_char_p]),
        "kofm_synchronize": (C.c_int, [P]),
        "kofm_set_launch": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int64]),
        "kofm_set_graphs": (C.c_int, [P, C.c_int32]),
        "kofm_objective_run": (C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(Objective), dp]),
        "kofm_adjoint": (C.c_int, [P, dp, dp]),
        "kofm_geometry_field_gradient": (C.c_int, [P, dp, dp, dp]),
        "kofm_ensemble_create": (C.c_int, [C.POINTER(P), C.c_int32, C.POINTER(P)]),
        "kofm_ensemble_step": (C.c_int, [P, C.c_int32, C.c_uint32]),
        "kofm_ensemble_time_steps": (C.c_int, [P, C.c_int32, C.c_uint32, dp,
                                               C.POINTER(C.c_int64)]),
        "kofm_ensemble_set_block": (C.c_int, [P, C.c_int32]),
        "kofm_ensemble_set_cache_budget": (C.c_int, [P, C.c_int64, C.POINTER(C.c_int32)]),
        "kofm_ensemble_set_launch": (C.c_int, [P, C.c_int32, C.c_int32]),
        "kofm_ensemble_autotune": (C.c_int, [P, C.c_int32, C.POINTER(C.c_int32),
                                             C.POINTER(C.c_int32), C.POINTER(C.c_int64), dp]),
        "kofm_ensemble_last_error": (C.c_char_p, [P]),
        "kofm_ensemble_destroy": (None, [P]),
        "kofm_autotune": (C.c_int, [P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_int64), dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def reduce_device(field_ptr: int, elem_bytes: int, n: int, op: int, out_ptr: int,
                  stream: int = 0) -> None:
    """kofm_reduce_device: reduction of n values at a device address into the
    device double at out_ptr, enqueued on `stream` with nothing allocated and
    nothing synchronised (the caller synchronises before reading out_ptr)."""
    lib = load()
    rc = lib.kofm_reduce_device(C.c_void_p(field_ptr), elem_bytes, n, op, C.c_void_p(out_ptr),
                                C.c_void_p(stream) if stream else None)
    if rc != 0:
        raise KofmError(rc, lib.kofm_last_error(None).decode())


class Lattice:
    """One lattice on the GPU behind the C ABI.

    Host arrays are float64 in the reference layout (x fastest, XsizeFull
    columns, planes of XsizeFull*nY).
    """

    def __init__(self, nX: int, nY: int, storage: str | int = "f32", komega: bool = False,
                 host_only: bool = False):
        self.lib = load()
        st = STORAGE_NAMES[storage] if isinstance(storage, str) else int(storage)
        g = Grid(nX, nY, st, 1 if komega else 0)
        h = C.c_void_p()
        create = self.lib.kofm_create_host if host_only else self.lib.kofm_create
        rc = create(C.byref(g), C.byref(h))
        if rc != 0:
            raise KofmError(rc, self.lib.kofm_last_error(None).decode())
        self.h = h
        self.nX, self.nY, self.komega, self.storage = nX, nY, komega, st
        self.XF = ((nX + 127) // 128) * 128
        self.D = self.XF * nY

    def _check(self, rc: int):
        if rc != 0:
            raise KofmError(rc, self.lib.kofm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.kofm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, tau0, Fval, nu, rho0=1.0, max_turb_visc=10.0, ko_dt=1.0,
                   kappa_rel_tol=1e-4, kappa_abs_tol=1e-7, kappa_max_iters=1000,
                   omega_rel_tol=None, omega_abs_tol=None, omega_max_iters=None):
        p = Params(tau0, Fval, nu, rho0, max_turb_visc, ko_dt, kappa_rel_tol, kappa_abs_tol,
                   kappa_max_iters,
                   kappa_rel_tol if omega_rel_tol is None else omega_rel_tol,
                   kappa_abs_tol if omega_abs_tol is None else omega_abs_tol,
                   kappa_max_iters if omega_max_iters is None else omega_max_iters)
        self._check(self.lib.kofm_set_params(self.h, C.byref(p)))

    def build_geometry(self, C0: np.ndarray, BL: np.ndarray, **gp):
        C0 = np.ascontiguousarray(C0, dtype=np.float64)
        BL = np.ascontiguousarray(BL, dtype=np.int32)
        p = GeomParams(gp.get("pipe_radius", 0.5 * (self.nY - 4)), gp.get("eps", 1e-5),
                       gp.get("min_dx", 0.001), gp.get("wall_d_search_rad", 50),
                       gp["x_release_pos"], gp.get("x_min_val", gp["x_release_pos"] - 1.25),
                       gp["x_stop_pos"], gp.get("cutoff_radius", 5.0),
                       gp.get("index_radius_search", 30))
        self._check(self.lib.kofm_build_geometry(
            self.h, _dptr(C0), C0.shape[0], BL.ctypes.data_as(C.POINTER(C.c_int32)), BL.shape[0],
            C.byref(p)))
        self._nN = C0.shape[0]
        self._nBL = BL.shape[0]

    def update_walls(self, C: np.ndarray):
        """Moves the walls to the points C [nN, 2] (a fouling layer grows
        between the built walls and C): node types, links, spacings, wall
        distances and shear tables are updated."""
        C = np.ascontiguousarray(C, dtype=np.float64)
        self._check(self.lib.kofm_update_walls(self.h, _dptr(C), C.shape[0]))

    def counts(self) -> GeomCounts:
        c = GeomCounts()
        self._check(self.lib.kofm_geometry_counts(self.h, C.byref(c)))
        return c

    def geometry(self, name: str) -> np.ndarray:
        which, dt = GEOM[name]
        c = self.counts()
        n = {"nType": c.D, "dX": 4 * c.D, "ibb": c.nIbb, "ibbQ": c.nIbb, "ss": c.nSS,
             "ssMap": c.D, "ssInds": c.nX, "lsMap": 2 * c.XsizeFull, "wallD": c.D,
             "cfss": 16 * c.nSS, "sInd": 4 * c.nBLshear, "weights": 4 * c.nBLshear,
             "bounds": 4, "masses": 2, "ibbSeg": c.nIbb, "ibbDqDy": 2 * c.nIbb}[name]
        out = np.zeros(max(n, 1), dtype=dt)
        self._check(self.lib.kofm_geometry_array(self.h, which, out.ctypes.data_as(C.c_void_p)))
        return out[:n]

    def ini_poiseuille(self, Re, nu, pipe_radius, rho=1.0, turbulent=False, perturb=False,
                       duplus=0.025, epsilon=250.0, beta=100.0, alpha=55.75, sigma=0.00055,
                       seed=1):
        """(rho, ux, uy) of clVariablesLB::iniDists(Umax) on this geometry
        (host only); nodes it does not visit are zero."""
        out = [np.zeros(self.D) for _ in range(3)]
        p = IniFlow(Re, nu, pipe_radius, rho, 1 if turbulent else 0, 1 if perturb else 0, duplus,
                    epsilon, beta, alpha, sigma, seed)
        self._check(self.lib.kofm_ini_poiseuille(self.h, C.byref(p), _dptr(out[0]),
                                                 _dptr(out[1]), _dptr(out[2])))
        return tuple(out)

    def init_state(self, rho=None, ux=None, uy=None):
        a = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (rho, ux, uy)]
        self._check(self.lib.kofm_init_state(self.h, *[_dptr(v) for v in a]))

    def init_komega(self, k_ini=-1.0, omega_ini=-1.0, Re=0.0, geom_penalty_factor=1.0):
        self._check(self.lib.kofm_init_komega(self.h, k_ini, omega_ini, Re, geom_penalty_factor))

    def upload_state(self, f, k=None, omega=None, nut=None, S=None):
        arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64)
                for v in (f, k, omega, nut, S)]
        self._check(self.lib.kofm_upload_state(self.h, *[_dptr(v) for v in arrs]))

    def step(self, n: int = 1, options: int = 0):
        self._check(self.lib.kofm_step(self.h, n, options))

    def time_steps(self, n: int, options: int = 0, kernel: bool = True):
        """(ms per step, ms per collide-stream kernel per step).  With
        kernel=False no events sit inside the timed window and the kernel
        figure is None."""
        a, b = C.c_double(), C.c_double()
        self._check(self.lib.kofm_time_steps(self.h, n, options, C.byref(a),
                                             C.byref(b) if kernel else None))
        return a.value, (b.value if kernel else None)

    def wall_shear(self):
        self._check(self.lib.kofm_wall_shear(self.h))

    def solver_stats(self) -> SolverStats:
        s = SolverStats()
        self._check(self.lib.kofm_get_solver_stats(self.h, C.byref(s)))
        return s

    def komega_time_step(self, action: int = 0):
        """KOFM_DT_* action (0 query, 1 halve, 2 double, 3 inject a
        non-finite residual into the next solve); returns (sub-steps per LB
        step, time step) after it."""
        n, dt = C.c_int32(), C.c_double()
        self._check(self.lib.kofm_komega_time_step(self.h, action, C.byref(n), C.byref(dt)))
        return n.value, dt.value

    def thermal_init(self, *, alpha_fluid: float, alpha_foul: float, pr_turb: float = 0.9,
                     t_in: float = 2.0, t_initial: float = 0.5, dt: float = 1.0,
                     rel_tol: float = 1e-4, abs_tol: float = 1e-7, max_iters: int = 1000):
        """kofm_thermal_init (clVariablesFD::ini): T = t_initial on the nodes
        fluid in the clean geometry; needs k-omega and a built geometry."""
        q = ThermalParams(alpha_fluid, alpha_foul, pr_turb, t_in, t_initial, dt, rel_tol, abs_tol,
                          max_iters)
        self._check(self.lib.kofm_thermal_init(self.h, C.byref(q)))

    def thermal_step(self, n: int = 1):
        """n calls of clVariablesFD::Solve with the current u and nu_t."""
        self._check(self.lib.kofm_thermal_step(self.h, n))

    def thermal_time_step(self, action: int = 0):
        """KOFM_DT_* action on the temperature time step; returns (sub-steps
        per call, time step) after it."""
        n, dt = C.c_int32(), C.c_double()
        self._check(self.lib.kofm_thermal_time_step(self.h, action, C.byref(n), C.byref(dt)))
        return n.value, dt.value

    def set_temperature(self, T: np.ndarray):
        T = np.ascontiguousarray(T, dtype=np.float64)
        assert T.size == self.D
        self._check(self.lib.kofm_set_temperature(self.h, _dptr(T)))

    def nusselt_init(self, *, height: int, wavy_start: int, wavy_period: int, wavy_periods: int,
                     k_soot: float, k_air: float):
        """kofm_nusselt_init (clVariablesFD::iniNuCoeffs) over the particle
        domain of the built geometry; needs the temperature solve."""
        q = NusseltParams(height, wavy_start, wavy_period, wavy_periods, k_soot, k_air)
        self._check(self.lib.kofm_nusselt_init(self.h, C.byref(q)))

    def nusselt_sizes(self):
        """(columns, period boundaries) of the Nusselt record."""
        c, l = C.c_int32(), C.c_int32()
        self._check(self.lib.kofm_nusselt_sizes(self.h, C.byref(c), C.byref(l)))
        return c.value, l.value

    def nusselt(self, nodes: bool = False) -> dict:
        """One Nusselt record: nu [5 columns], mean [2 locations] and, with
        nodes, the nearest-node slots (nodes [2 columns], vec [4 columns])."""
        n, locs = self.nusselt_sizes()
        out = dict(nu=np.zeros(5 * n), mean=np.zeros(2 * locs))
        if nodes:
            out["nodes"] = np.zeros(2 * n, np.int32)
            out["vec"] = np.zeros(4 * n)
        self._check(self.lib.kofm_nusselt(
            self.h, _dptr(out["nu"]), _dptr(out["mean"]),
            out["nodes"].ctypes.data_as(C.c_void_p) if nodes else None,
            _dptr(out["vec"]) if nodes else None))
        return out

    def shear_removal(self, particles: dict, bl_dep: int, offset: int, count: int, *,
                      tau_crit, drag_coeff, lift_coeff, mass, dt_wall: float = 3.0,
                      foul_size_switch: float = 0.5, loc_stride: int, timer_start: int):
        """kofm_shear_removal (TR_shear_removal) on the particle arrays at the
        device addresses in `particles` (pos, num_rep, type, dep_flag,
        dep_timer, loc, timer) and the deposit counters at bl_dep; enqueued on
        the context's stream."""
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        tc, dc, lc, mp = f64(tau_crit), f64(drag_coeff), f64(lift_coeff), f64(mass)
        q = RemovalParams(len(mp), tc.ctypes.data, dc.ctypes.data, lc.ctypes.data, mp.ctypes.data,
                          dt_wall, foul_size_switch, loc_stride, timer_start)
        par = Particles(*(particles[k] for k in ("pos", "num_rep", "type", "dep_flag",
                                                 "dep_timer", "loc", "timer")))
        self._check(self.lib.kofm_shear_removal(self.h, C.byref(q), C.byref(par), offset, count,
                                                bl_dep))

    def removal_links(self, foul_size_switch: float = 0.5) -> dict:
        """kofm_removal_links: the tracer's link arrays for the walls now."""
        nBL = self._nBL
        out = dict(p0=np.zeros(2 * nBL), vN=np.zeros(2 * nBL), len=np.zeros(nBL),
                   surf=np.zeros(nBL, np.int16))
        self._check(self.lib.kofm_removal_links(self.h, foul_size_switch, _dptr(out["p0"]),
                                                _dptr(out["vN"]), _dptr(out["len"]),
                                                out["surf"].ctypes.data_as(C.c_void_p)))
        return out

    def reduce(self, field: str, op: int = RED_SUM) -> float:
        out = C.c_double()
        self._check(self.lib.kofm_reduce(self.h, FIELD_NAMES[field], op, C.byref(out)))
        return out.value

    def download(self, field: str) -> np.ndarray:
        c = self.counts()
        n = {"f": 9 * self.D, "S": 3 * self.D, "tau_node": c.nSS, "tau_bl": c.nBLshear}.get(field, self.D)
        out = np.zeros(max(n, 1), dtype=np.float64)
        self._check(self.lib.kofm_download(self.h, FIELD_NAMES[field], _dptr(out)))
        return out[:n]

    def save2file(self, directory: str):
        self._check(self.lib.kofm_save2file(self.h, directory.encode()))

    def save_restart(self, directory: str):
        """lbf.bin, lbkappa.bin, lbomega.bin, lsc.bin (saveRestartFiles)."""
        self._check(self.lib.kofm_save_restart(self.h, directory.encode()))

    def save_debug(self, directory: str):
        """nType.txt, dXarr_0..3.txt, lbnut.txt, wallD.txt (saveDebug)."""
        self._check(self.lib.kofm_save_debug(self.h, directory.encode()))

    def set_launch(self, block_x: int, nodes_per_lane: int = 1, plane_pad: int = 1088,
                   steps_per_pass: int = 1):
        self._check(self.lib.kofm_set_launch(self.h, block_x, nodes_per_lane, steps_per_pass,
                                             plane_pad))

    def autotune(self, steps: int = 50):
        """Picks the collide-stream work-group width, nodes per lane, steps per
        pass and plane padding; returns (block_x, nodes_per_lane, plane_pad,
        ms_per_step, steps_per_pass)."""
        b, n, sp, pad, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int64(), C.c_double()
        self._check(self.lib.kofm_autotune(self.h, steps, C.byref(b), C.byref(n), C.byref(sp),
                                           C.byref(pad), C.byref(ms)))
        return b.value, n.value, pad.value, ms.value, sp.value

    def set_graphs(self, on: bool):
        self._check(self.lib.kofm_set_graphs(self.h, 1 if on else 0))

    def objective_run(self, nsteps: int, interval: int, window: int, x_in: int = 0,
                      x_out: int | None = None, shear_weight: float = 0.1,
                      frozen_turbulence: bool = False) -> float:
        """Runs nsteps accumulating the wall objective J over the last `window`
        steps and keeping checkpoints every `interval` steps; returns J.  With
        frozen_turbulence the following adjoint holds the eddy viscosity at
        its recorded values."""
        o = Objective(window, x_in, self.nX - 1 if x_out is None else x_out,
                      ADJOINT_FROZEN_TURBULENCE if frozen_turbulence else 0, shear_weight)
        J = C.c_double()
        self._check(self.lib.kofm_objective_run(self.h, nsteps, interval, C.byref(o), C.byref(J)))
        return J.value

    def adjoint(self):
        """dJ/dy of every wall point and dJ/dq of every boundary link for the
        last objective_run."""
        dy = np.zeros(self._nN, dtype=np.float64)
        dq = np.zeros(max(self.counts().nIbb, 1), dtype=np.float64)
        self._check(self.lib.kofm_adjoint(self.h, _dptr(dy), _dptr(dq)))
        return dy, dq[:self.counts().nIbb]

    def geometry_field_gradient(self, wall_d_bar=None, dx_bar=None):
        """dJ/dy of the wall points for given sensitivities of the wall distance
        field (D values) and the axis wall spacings (4*D values)."""
        dy = np.zeros(self._nN, dtype=np.float64)
        wd = None if wall_d_bar is None else np.ascontiguousarray(wall_d_bar, dtype=np.float64)
        dx = None if dx_bar is None else np.ascontiguousarray(dx_bar, dtype=np.float64)
        self._check(self.lib.kofm_geometry_field_gradient(
            self.h, None if wd is None else _dptr(wd), None if dx is None else _dptr(dx), _dptr(dy)))
        return dy

    def synchronize(self):
        self._check(self.lib.kofm_synchronize(self.h))


class Ensemble:
    """Several Lattice objects (same storage, k-omega off) stepped together,
    one launch per step."""

    def __init__(self, members):
        self.lib = load()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        h = C.c_void_p()
        rc = self.lib.kofm_ensemble_create(arr, len(self.members), C.byref(h))
        if rc != 0:
            raise KofmError(rc, self.lib.kofm_ensemble_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise KofmError(rc, self.lib.kofm_ensemble_last_error(self.h).decode())

    def step(self, n: int = 1, options: int = 0):
        self._check(self.lib.kofm_ensemble_step(self.h, n, options))

    def time_steps(self, n: int, options: int = 0):
        ms, fl = C.c_double(), C.c_int64()
        self._check(self.lib.kofm_ensemble_time_steps(self.h, n, options, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value

    def set_block(self, block_x: int):
        self._check(self.lib.kofm_ensemble_set_block(self.h, block_x))

    def set_launch(self, block_x: int, steps_per_launch: int = 1):
        self._check(self.lib.kofm_ensemble_set_launch(self.h, block_x, steps_per_launch))

    def autotune(self, steps: int = 20):
        """Times each launch shape and cache budget on the members (their
        state is restored), keeps the fastest; returns (block_x,
        steps_per_launch, cache_budget, ms_per_step)."""
        b, s, g, ms = C.c_int32(), C.c_int32(), C.c_int64(), C.c_double()
        self._check(self.lib.kofm_ensemble_autotune(self.h, steps, C.byref(b), C.byref(s),
                                                     C.byref(g), C.byref(ms)))
        return b.value, s.value, g.value, ms.value

    def set_cache_budget(self, nbytes: int) -> int:
        """Groups the members under a population-byte budget (0: one group);
        returns the number of groups."""
        g = C.c_int32()
        self._check(self.lib.kofm_ensemble_set_cache_budget(self.h, nbytes, C.byref(g)))
        return g.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.kofm_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
