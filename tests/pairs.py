"""Shared helpers: the same seeded workload on the GPU library and on the oracle."""
from __future__ import annotations

import numpy as np

import oracle
from komegafoulingmodel_amd import channel, lib


def rel_err(a, b):
    """max |a - b| / max |b| (normwise relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))


def vec_rel_err(ax, ay, bx, by):
    """max |a - b| / max |b| for 2-vectors given by components."""
    d = np.hypot(np.asarray(ax) - bx, np.asarray(ay) - by)
    s = np.max(np.hypot(bx, by))
    return float(np.max(d) / (s if s > 0 else 1.0))


def fluid_mask(nType):
    return (np.asarray(nType).astype(np.int32) & 1) == 0


def make_pair(w: channel.Workload, storage="f32", *, solver_rtol=1e-13, solver_maxit=400,
              oracle_only=False):
    g = oracle.geometry(w.nX, w.nY, w.C0, w.BL, **w.geom_params())
    orc = oracle.OracleLattice(g, tau0=w.tau0, Fval=w.Fval, nu=w.nu, komega=w.komega)
    rho, ux, uy = w.initial_velocity(g["nType"])
    orc.init_state(rho, ux, uy)
    if w.komega:
        orc.init_komega(w.k_ini, w.omega_ini)
    if oracle_only:
        return None, orc
    L = lib.Lattice(w.nX, w.nY, storage=storage, komega=w.komega)
    L.set_params(w.tau0, w.Fval, w.nu, kappa_rel_tol=solver_rtol, kappa_abs_tol=0.0,
                 kappa_max_iters=solver_maxit)
    L.build_geometry(w.C0, w.BL, **w.geom_params())
    L.init_state(rho, ux, uy)
    if w.komega:
        L.init_komega(w.k_ini, w.omega_ini, w.Re)
    return L, orc
