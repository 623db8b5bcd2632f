"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The FP64 CPU restatement of the reference step (oracle/oracle.c).  Only the
tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
use it, always as the checker; the product library never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


class OrcPar(C.Structure):
    _fields_ = [("nX", C.c_int), ("nY", C.c_int), ("XF", C.c_int), ("tau0", C.c_double),
                ("Fval", C.c_double), ("nu", C.c_double), ("maxNut", C.c_double),
                ("dt", C.c_double), ("options", C.c_int), ("threads", C.c_int)]


class OrcSolver(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("maxit", C.c_int),
                ("last_iters_k", C.c_int), ("last_iters_o", C.c_int),
                ("steps_per_lb", C.c_int), ("dt_cur", C.c_double), ("since_change", C.c_int),
                ("inject_nan", C.c_int), ("halvings", C.c_int)]


class OrcTSolver(C.Structure):
    _fields_ = [("alpha_fluid", C.c_double), ("alpha_foul", C.c_double),
                ("pr_turb_inv", C.c_double), ("t_in", C.c_double), ("rtol", C.c_double),
                ("atol", C.c_double), ("maxit", C.c_int), ("steps_per_lb", C.c_int),
                ("dt", C.c_double), ("since_change", C.c_int), ("last_iters", C.c_int),
                ("halvings", C.c_int), ("inject_nan", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.orc_reduce.restype = C.c_double
        _lib.orc_reduce.argtypes = [C.c_void_p, C.c_long, C.c_int]
        _lib.orc_max_threads.restype = C.c_int
    return _lib


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def geometry(nX, nY, C0, BL, *, eps=1e-5, min_dx=0.001, x_release_pos=10.0, x_stop_pos=None,
             x_min_val=None, wall_d_search_rad=50, cutoff_radius=5.0, index_radius_search=30,
             pipe_radius=None):
    """orc_geometry + shear coefficient setup. Returns a dict of numpy arrays."""
    lib = load()
    XF = ((nX + 127) // 128) * 128
    D = XF * nY
    C0 = np.ascontiguousarray(C0, dtype=np.float64)
    BL = np.ascontiguousarray(BL, dtype=np.int32)
    if x_stop_pos is None:
        x_stop_pos = float(nX - 12)
    if x_min_val is None:
        x_min_val = x_release_pos - 1.25
    ssMin = max(int(np.floor(x_release_pos)) - 10, 0)
    ssMax = int(np.ceil(x_stop_pos)) + 10
    g = dict(nType=np.zeros(D, np.int16), dX=np.zeros(4 * D), ss=np.zeros(D, np.int32),
             ssMap=np.zeros(D, np.int32), ssInds=np.zeros(nX, np.int32),
             lsMap=np.zeros(2 * XF, np.uint16), wallD=np.zeros(D))
    cap = 64 * nX + 16 * nY + 1024
    ibb = np.zeros(cap, np.int32)
    q = np.zeros(cap)
    nSS = C.c_int()
    nIbb = C.c_int()
    rc = lib.orc_geometry(nX, XF, nY, p(C0), C0.shape[0], p(BL), BL.shape[0], C.c_double(eps),
                          C.c_double(min_dx), ssMin, ssMax, wall_d_search_rad, p(g["nType"]),
                          p(g["dX"]), p(g["ss"]), p(g["ssMap"]), p(g["ssInds"]), C.byref(nSS),
                          p(g["lsMap"]), p(g["wallD"]), p(ibb), p(q), cap, C.byref(nIbb))
    if rc != 0:
        raise RuntimeError(f"oracle geometry failed with {rc}")
    assert nIbb.value <= cap
    g["ss"] = g["ss"][:nSS.value].copy()
    g["ibb"] = ibb[:nIbb.value].copy()
    g["ibbQ"] = q[:nIbb.value].copy()
    # wall-shear coefficients
    n = nSS.value
    g["cfss"] = np.zeros(max(16 * n, 1))
    lib.orc_shear_coeffs1(XF, p(g["ss"]), n, p(g["nType"]), C.c_double(cutoff_radius), p(g["cfss"]))
    g["cfss"] = g["cfss"][:16 * n]
    nBL = BL.shape[0]
    bounds = np.zeros(4, np.int32)
    P0 = np.zeros(2 * nBL)
    vN = np.zeros(2 * nBL)
    blLen = np.zeros(nBL)
    rc = lib.orc_shear_bl_setup(p(C0), p(BL), nBL, C.c_double(x_release_pos),
                                C.c_double(x_min_val), C.c_double(x_stop_pos), p(bounds), p(P0),
                                p(vN), p(blLen))
    if rc != 0:
        raise RuntimeError("oracle shear link setup failed")
    nb = int(bounds[1] + bounds[3])
    g["bounds"] = bounds
    g["weights"] = np.zeros(4 * nb)
    g["sInd"] = np.zeros(4 * nb, np.int32)
    lib.orc_shear_coeffs2(XF, p(bounds), p(P0), p(vN), p(blLen), p(g["ss"]), p(g["ssInds"]),
                          ssMin, ssMax, index_radius_search, C.c_double(cutoff_radius),
                          p(g["weights"]), p(g["sInd"]))
    g.update(nX=nX, nY=nY, XF=XF, D=D)
    return g


def nusselt(nX, nY, nType, C0, C_now, BL, *, x_min_val, x_stop_pos, height, wavy_start, wavy_period,
            wavy_periods, k_soot, k_air, T, ux, uy, dX, dX0):
    """orc_nusselt: one Nusselt record from the given fields (FP64); C0 are
    the walls the geometry was built from, C_now the walls now.

    Returns dict(nu [5 n], mean [2 (wavy_periods + 1)], nodes [2 n], vec [4 n]).
    """
    lib = load()
    XF = ((nX + 127) // 128) * 128
    n = int(np.floor(x_stop_pos)) - int(np.floor(x_min_val))
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    T, ux, uy, dX, dX0 = f64(T), f64(ux), f64(uy), f64(dX), f64(dX0)
    C0, C_now = f64(C0).ravel(), f64(C_now).ravel()
    BL = np.ascontiguousarray(BL, dtype=np.int32).ravel()
    nType = np.ascontiguousarray(nType, dtype=np.int16)
    out = dict(nu=np.zeros(5 * n), mean=np.zeros(2 * (wavy_periods + 1)),
               nodes=np.zeros(2 * n, np.int32), vec=np.zeros(4 * n))
    lib.orc_nusselt.argtypes = ([C.c_int] * 3 + [C.c_void_p] * 4 + [C.c_int, C.c_double,
                                C.c_double] + [C.c_int] * 4 + [C.c_double] * 2
                                + [C.c_void_p] * 9)
    rc = lib.orc_nusselt(nX, XF, nY, p(nType), p(C0), p(C_now), p(BL), len(BL) // 2,
                         float(x_min_val), float(x_stop_pos), height, wavy_start, wavy_period,
                         wavy_periods, float(k_soot), float(k_air), p(T), p(ux), p(uy), p(dX),
                         p(dX0), p(out["nu"]), p(out["mean"]), p(out["nodes"]), p(out["vec"]))
    if rc != 0:
        raise ValueError("orc_nusselt: no boundary links bracket the column range")
    return out


def removal_links(C_now, C0, BL, foul_size_switch):
    """orc_removal_links: the tracer's link arrays (p0 [nBL, 2], vN [nBL, 2],
    len [nBL], surf [nBL]) for the walls C_now built from C0."""
    lib = load()
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
    C_now, C0 = f64(C_now), f64(C0)
    BL = np.ascontiguousarray(BL, dtype=np.int32).ravel()
    nBL = len(BL) // 2
    out = dict(p0=np.zeros(2 * nBL), vN=np.zeros(2 * nBL), len=np.zeros(nBL),
               surf=np.zeros(nBL, np.int16))
    lib.orc_removal_links.argtypes = [C.c_int] + [C.c_void_p] * 3 + [C.c_double] + [C.c_void_p] * 4
    lib.orc_removal_links(nBL, p(C_now), p(C0), p(BL), float(foul_size_switch), p(out["p0"]),
                          p(out["vN"]), p(out["len"]), p(out["surf"]))
    return out


def shear_removal(links, tau_link, par, *, tau_crit, drag_coeff, lift_coeff, mass, dt_wall,
                  x_min_val, x_max_val, x_ind_start, loc_stride, timer_start, bl_dep, offset,
                  count):
    """orc_shear_removal on host copies: `par` holds pos [2 n], num_rep, type,
    dep_flag, dep_timer, loc, timer (updated in place), bl_dep likewise."""
    lib = load()
    n_sizes = len(mass)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    tc, dc, lc, mp, tau = f64(tau_crit), f64(drag_coeff), f64(lift_coeff), f64(mass), f64(tau_link)
    lib.orc_shear_removal.argtypes = ([C.c_int] + [C.c_void_p] * 4 + [C.c_double] * 3
                                      + [C.c_int] * 4 + [C.c_void_p] * 13 + [C.c_long] * 2)
    lib.orc_shear_removal(n_sizes, p(tc), p(dc), p(lc), p(mp), float(dt_wall), float(x_min_val),
                          float(x_max_val), x_ind_start, loc_stride, timer_start,
                          len(links["len"]), p(links["p0"]), p(links["vN"]), p(links["len"]),
                          p(links["surf"]), p(tau), p(par["pos"]), p(par["num_rep"]),
                          p(par["type"]), p(par["dep_flag"]), p(par["dep_timer"]), p(par["loc"]),
                          p(par["timer"]), p(bl_dep), int(offset), int(count))


def feq(rho, ux, uy):
    lib = load()
    out = np.zeros(9)
    lib.orc_feq(C.c_double(rho), C.c_double(ux), C.c_double(uy), p(out))
    return out


class OracleLattice:
    """FP64 reference-algorithm lattice on the CPU."""

    def __init__(self, geo: dict, *, tau0, Fval, nu, komega=False, max_turb_visc=10.0, dt=1.0,
                 rtol=1e-16, atol=0.0, maxit=3000, threads=0):
        self.lib = load()
        self.g = geo
        nX, nY, XF, D = geo["nX"], geo["nY"], geo["XF"], geo["D"]
        self.par = OrcPar(nX, nY, XF, tau0, Fval, nu, max_turb_visc, dt, 0x2, threads)
        self.sv = OrcSolver(rtol, atol, maxit, 0, 0, max(1, int(1.0 / dt)), dt, 0, 0, 0)
        self.komega = komega
        self.D = D
        self.FA = np.zeros(9 * D)
        self.FB = np.zeros(9 * D)
        self.rho = np.zeros(D)
        self.ux = np.zeros(D)
        self.uy = np.zeros(D)
        self.K = np.zeros(D)
        self.O = np.zeros(D)
        self.nut = np.zeros(D)
        self.S = np.zeros(3 * D)
        self.work = np.zeros(26 * D)
        self.dX = np.ascontiguousarray(geo["dX"])
        self.wallD = np.ascontiguousarray(geo["wallD"])
        self.nType = geo["nType"]
        self.ibb = geo["ibb"]
        self.ibbQ = geo["ibbQ"]

    def init_state(self, rho, ux, uy):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        ux = np.ascontiguousarray(ux, dtype=np.float64)
        uy = np.ascontiguousarray(uy, dtype=np.float64)
        self.lib.orc_init_dists(self.par.nX, self.par.XF, self.par.nY, p(self.nType), p(rho), p(ux),
                                p(uy), p(self.FA))
        fluid = (self.nType.astype(np.int32) & 1) == 0
        self.rho[:] = np.where(fluid, rho, 0.0)
        self.ux[:] = np.where(fluid, ux, 0.0)
        self.uy[:] = np.where(fluid, uy, 0.0)
        self.S[:] = 0.0

    def init_komega(self, k, w):
        fluid = (self.nType.astype(np.int32) & 0x2) != 0
        self.K[:] = np.where(fluid, k, 0.0)
        self.O[:] = np.where(fluid, w, 0.0)
        self.nut[:] = np.where(fluid, k / w, 0.0)
        self.S[:] = 0.0

    def step(self, n=1):
        rc = self.lib.orc_run(C.byref(self.par), p(self.nType), p(self.ibb), p(self.ibbQ),
                              len(self.ibb), p(self.FA), p(self.FB), p(self.rho), p(self.ux),
                              p(self.uy), p(self.K), p(self.O), p(self.wallD), p(self.nut),
                              p(self.S), p(self.dX), 1 if self.komega else 0, C.byref(self.sv),
                              p(self.work), n)
        if rc != 0:
            raise RuntimeError(f"oracle step failed ({rc})")

    def shear(self):
        g = self.g
        n = len(g["ss"])
        tau = np.zeros(max(n, 1))
        self.lib.orc_shear1(C.byref(self.par), p(g["ss"]), n, p(g["cfss"]), p(self.FA),
                            p(self.rho), p(self.ux), p(self.uy),
                            p(self.nut) if self.komega else None, p(tau))
        nb = int(g["bounds"][1] + g["bounds"][3])
        bl = np.zeros(max(nb, 1))
        self.lib.orc_shear2(p(g["bounds"]), p(g["weights"]), p(g["sInd"]), p(tau), p(bl))
        return tau[:n], bl[:nb]

    def reduce(self, a, op=1):
        return self.lib.orc_reduce(p(np.ascontiguousarray(a)), len(a), op)

    # --- temperature (clVariablesFD) ---------------------------------------
    def init_thermal(self, *, alpha_fluid, alpha_foul, pr_turb, t_in, t_initial, dt=1.0,
                     rtol=1e-4, atol=1e-7, maxit=1000, nType=None):
        """clVariablesFD::ini: T = Tinitial on the nodes that are fluid in the
        clean geometry, 0 elsewhere; stepsPerLB = (int)(1 / dTtfd).  nType
        (default the geometry's) carries the fouling layer: solid now, fluid
        in the clean geometry."""
        self.tnType = np.ascontiguousarray(self.nType if nType is None else nType, dtype=np.int16)
        m0fluid = (self.tnType.astype(np.int32) & 0x8) != 0
        self.T = np.where(m0fluid, float(t_initial), 0.0)
        self.ts = OrcTSolver(alpha_fluid, alpha_foul, 1.0 / pr_turb, t_in, rtol, atol, maxit,
                             int(1.0 / dt), dt, 0, 0, 0, 0)
        self.twork = np.zeros(16 * self.D)
        self.dX0 = self.dX.copy()

    def t_step(self, n=1):
        """n calls of clVariablesFD::Solve with the current u and nu_t."""
        nut = self.nut if self.komega else np.zeros(self.D)
        for _ in range(n):
            rc = self.lib.orc_t_step(C.byref(self.par), p(self.tnType), p(self.dX0), p(self.T),
                                     p(nut), p(self.ux), p(self.uy), C.byref(self.ts),
                                     p(self.twork))
            if rc != 0:
                raise RuntimeError(f"oracle temperature step failed ({rc})")
