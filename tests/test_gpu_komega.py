"""GPU parity of the k-omega step beyond the first steps, and its time step
control, through the C ABI against the FP64 oracle.

* A clamp-free case: flat walls at half-integer rows (every wall spacing
  0.5), k-omega on.  k and omega stay positive on every fluid node over the
  whole run (the oracle's minima are printed), so neither clamp of the
  collision acts and the step is smooth: FP32 populations (with the FP64
  k-omega update) are held to 1e-5 and FP64 to 1e-10 after 500 steps.
* Time step control (kOmega.cpp:31-73, 387-414): a non-finite residual,
  injected from the host into the third step's solve on both sides,
  restores k and omega, halves the time step and doubles the sub-steps; 20
  quiet steps later the time step doubles back.  The GPU follows the oracle
  through all of it.
* "Geometric Penalty Factor" scales the mean velocity of the turbulence
  intensity estimate (kOmega.cpp:530-570) when the flow is at rest.
"""
import math

import numpy as np
import pytest

from komegafoulingmodel_amd import channel, lib
from pairs import fluid_mask, make_pair, rel_err, vec_rel_err

pytestmark = pytest.mark.gpu


def _clamp_free():
    nu, Re, H = 0.01, 500.0, 56.0
    return channel.wavy_channel(256, 72, H=H, amp=0.0, period=128.0, nu=nu, Re=Re,
                                Fval=channel.turbulent_forcing(H, nu, Re), seed=7)


def _errors(L, orc, m):
    errs = {k: rel_err(L.download(k)[m], getattr(orc, a)[m])
            for k, a in (("k", "K"), ("omega", "O"), ("nut", "nut"))}
    errs["u"] = vec_rel_err(L.download("ux")[m], L.download("uy")[m], orc.ux[m], orc.uy[m])
    errs["f"] = rel_err(L.download("f"), orc.FA)
    return errs


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-10)])
def test_komega_clamp_free_matches_oracle_over_500_steps(storage, tol):
    w = _clamp_free()
    L, orc = make_pair(w, storage, solver_rtol=1e-15, solver_maxit=2000)
    m = fluid_mask(orc.nType)
    for total, n in ((100, 100), (500, 400)):
        L.step(n, 2)
        orc.step(n)
        assert orc.K[m].min() > 0.0 and orc.O[m].min() > 0.0
        errs = _errors(L, orc, m)
        print(storage, "step", total, {k: f"{v:.2e}" for k, v in errs.items()},
              f"oracle min k {orc.K[m].min():.3e} min omega {orc.O[m].min():.3e}", flush=True)
        for k, v in errs.items():
            assert v < tol, (storage, total, k, v)


def test_komega_time_step_control_matches_oracle():
    w = _clamp_free()
    L, orc = make_pair(w, "f64", solver_rtol=1e-15, solver_maxit=2000)
    m = fluid_mask(orc.nType)
    L.step(2, 2)
    orc.step(2)
    assert L.komega_time_step(lib.DT_QUERY) == (1, 1.0)
    L.komega_time_step(lib.DT_INJECT_NAN)
    orc.sv.inject_nan = 1
    L.step(1, 2)
    orc.step(1)
    st = L.solver_stats()
    assert (st.nan_flag, st.steps_per_lb, st.time_step) == (1, 2, 0.5), (st.nan_flag, st.steps_per_lb)
    assert (orc.sv.halvings, orc.sv.steps_per_lb) == (1, 2)
    errs = _errors(L, orc, m)
    print("after the halved step", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < 1e-10, ("halved step", k, v)
    # 19 more steps at the halved time step, the 20th doubles it back
    L.step(19, 2)
    orc.step(19)
    assert L.komega_time_step(lib.DT_QUERY) == (2, 0.5)
    L.step(1, 2)
    orc.step(1)
    assert L.komega_time_step(lib.DT_QUERY) == (1, 1.0) and orc.sv.steps_per_lb == 1
    L.step(5, 2)
    orc.step(5)
    errs = _errors(L, orc, m)
    print("26 steps after the halving", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < 1e-10, ("after doubling back", k, v)
    assert L.solver_stats().nan_flag == 0


def test_komega_halve_restores_the_start_of_the_last_step():
    w = _clamp_free()
    L, _ = make_pair(w, "f64")
    m = fluid_mask(L.geometry("nType"))
    L.step(3, 2)
    k0, o0 = L.download("k"), L.download("omega")
    L.step(1, 2)
    assert not np.array_equal(L.download("k")[m], k0[m])
    assert L.komega_time_step(lib.DT_HALVE) == (2, 0.5)
    assert np.array_equal(L.download("k")[m], k0[m]) and np.array_equal(L.download("omega")[m], o0[m])
    assert L.komega_time_step(lib.DT_HALVE) == (4, 0.25)
    assert L.komega_time_step(lib.DT_DOUBLE) == (2, 0.5)
    assert L.komega_time_step(lib.DT_DOUBLE) == (1, 1.0)
    assert L.komega_time_step(lib.DT_DOUBLE) == (1, 1.0)  # doubleTimeStep stops at one sub-step


@pytest.mark.parametrize("gpf", [1.0, 2.5])
def test_init_komega_geometric_penalty_factor(gpf):
    """From rest, umean = Re nu / R / 4 * gpf, Re_ini = 4 R umean / nu,
    I = 0.16 Re_ini^(-1/8), k = 1.5 (I umean)^2, omega = sqrt(k) / 0.5477225575 / (0.036 R)."""
    w = channel.flat_channel(256, 40, Fval=1e-6, noise=0.0)
    L = lib.Lattice(w.nX, w.nY, storage="f32", komega=True)
    L.set_params(w.tau0, w.Fval, w.nu)
    L.build_geometry(w.C0, w.BL, **w.geom_params())
    L.init_state()
    Re, R, nu = 2000.0, w.pipe_radius, w.nu
    L.init_komega(-1.0, -1.0, Re, gpf)
    um = Re * nu / R / 4.0 * gpf
    ti = 0.16 * (R * 4.0 * um / nu) ** (-1.0 / 8.0)
    k = 1.5 * (ti * um) ** 2
    om = math.sqrt(k) / 0.5477225575 / (0.036 * R)
    fl = (L.geometry("nType").astype(np.int32) & 0x2) != 0
    assert np.allclose(L.download("k")[fl], k, rtol=1e-14, atol=0)
    assert np.allclose(L.download("omega")[fl], om, rtol=1e-14, atol=0)
    assert np.allclose(L.download("nut")[fl], k / om, rtol=1e-14, atol=0)
