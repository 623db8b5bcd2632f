"""GPU parity of the temperature solve (clVariablesFD) through the C ABI
against the FP64 oracle.

Each step is the reference's order (LBLS.cpp:190-200): the LB collide-stream
with the k-omega solve, then clVariablesFD::Solve with the new u and nu_t.

* A clamp-free k-omega channel (flat walls at half-integer rows): the
  temperature follows the oracle over 100 steps, FP64 populations to 1e-9 and
  FP32 populations to 1e-5.
* The wavy channel whose near-wall k rows are not diagonally dominant: T
  swings far outside [0, T_in] on the first step (as in the oracle, which
  restates the reference's rows) and both solves still agree.
* A fouling layer grown by kofm_update_walls: its nodes are rows of the
  temperature system with the soot diffusivity, the spacings are those of the
  clean walls (dXArr0), and the oracle is started from the updated lattice's
  state.
* Time step control: a failed solve injected from the host halves the time
  step on both sides, and 20 quiet calls later it doubles back.
"""
import numpy as np
import pytest

import oracle
from komegafoulingmodel_amd import channel, lib
from pairs import fluid_mask, make_pair, rel_err

pytestmark = pytest.mark.gpu

TP = dict(alpha_fluid=0.01 / 0.708, alpha_foul=0.004, pr_turb=0.708, t_in=2.0, t_initial=0.5)


def _clamp_free():
    nu, Re, H = 0.01, 500.0, 56.0
    return channel.wavy_channel(256, 72, H=H, amp=0.0, period=128.0, nu=nu, Re=Re,
                                Fval=channel.turbulent_forcing(H, nu, Re), seed=7)


def _wavy():
    nu, Re, H = 0.01, 500.0, 56.0
    return channel.wavy_channel(256, 72, H=H, amp=4.0, period=128.0, nu=nu, Re=Re,
                                Fval=channel.turbulent_forcing(H, nu, Re), seed=7)


def _pair(w, storage, rtol=1e-13):
    L, orc = make_pair(w, storage, solver_rtol=1e-14, solver_maxit=2000)
    L.thermal_init(**TP, rel_tol=rtol, abs_tol=0.0, max_iters=2000)
    orc.init_thermal(**TP, rtol=rtol, atol=0.0, maxit=2000)
    return L, orc


def _rows(orc):
    return (orc.tnType.astype(np.int32) & 0x4) == 0


def _run(L, orc, n):
    for _ in range(n):
        L.step(1, lib.OPTION_SAVE_MACRO_FIELDS)
        L.thermal_step(1)
        orc.step(1)
        orc.t_step(1)


@pytest.mark.parametrize("storage,tol", [("f64", 1e-9), ("f32", 1e-5)])
def test_temperature_matches_oracle_on_a_clamp_free_channel(storage, tol):
    w = _clamp_free()
    L, orc = _pair(w, storage)
    rows = _rows(orc)
    for total, n in ((10, 10), (100, 90)):
        _run(L, orc, n)
        T = L.download("T")
        st = L.solver_stats()
        e = rel_err(T[rows], orc.T[rows])
        print(f"{storage} step {total}: T error {e:.2e}, T in [{orc.T[rows].min():.4f}, "
              f"{orc.T[rows].max():.4f}], iterations gpu {st.temp_iters} oracle "
              f"{orc.ts.last_iters}", flush=True)
        assert e < tol, (storage, total, e)
        assert st.temp_steps_per_lb == 1 and orc.ts.halvings == 0


def test_temperature_on_wavy_walls_matches_oracle():
    w = _wavy()
    L, orc = _pair(w, "f64")
    rows = _rows(orc)
    for s in range(1, 6):
        _run(L, orc, 1)
        T = L.download("T")
        e = rel_err(T[rows], orc.T[rows])
        print(f"wavy step {s}: T error {e:.2e}, T in [{orc.T[rows].min():.3f}, "
              f"{orc.T[rows].max():.3f}]", flush=True)
        assert e < 1e-8, (s, e)


def test_temperature_through_a_fouling_layer_matches_oracle():
    from test_gpu_walls import _grow, _workload
    w = _workload(True)
    L, _ = make_pair(w, "f64", solver_rtol=1e-14, solver_maxit=2000)
    L.step(10, 2)
    C = _grow(w.C0, w.nX, 3.3, 97.0, 14.0)
    L.update_walls(C)
    nType = L.geometry("nType")
    foul = ((nType.astype(np.int32) & 0x1) != 0) & ((nType.astype(np.int32) & 0x8) != 0)
    assert foul.sum() > 20
    L.thermal_init(**TP, rel_tol=1e-13, abs_tol=0.0, max_iters=2000)
    # the oracle: the LB and k-omega on the geometry of the new walls, the
    # temperature rows on the updated node types with the clean walls'
    # spacings
    g = oracle.geometry(w.nX, w.nY, C, w.BL, **w.geom_params())
    g0 = oracle.geometry(w.nX, w.nY, w.C0, w.BL, **w.geom_params())
    orc = oracle.OracleLattice(g, tau0=w.tau0, Fval=w.Fval, nu=w.nu, komega=True, rtol=1e-14,
                               maxit=2000)
    orc.FA[:] = L.download("f")
    for attr, name in (("rho", "rho"), ("ux", "ux"), ("uy", "uy"), ("K", "k"), ("O", "omega"),
                       ("nut", "nut"), ("S", "S")):
        getattr(orc, attr)[:] = L.download(name)
    D = g["D"]
    orc.work[:D] = L.download("diff_k")
    orc.work[D:2 * D] = L.download("diff_omega")
    orc.init_thermal(**TP, rtol=1e-13, atol=0.0, maxit=2000, nType=nType)
    orc.dX0[:] = g0["dX"]
    rows = _rows(orc)
    assert np.all(rows[foul])
    _run(L, orc, 8)
    T = L.download("T")
    e = rel_err(T[rows], orc.T[rows])
    ef = rel_err(T[foul], orc.T[foul])
    print(f"fouled: T error {e:.2e} (fouling layer {ef:.2e}), layer T in "
          f"[{orc.T[foul].min():.4f}, {orc.T[foul].max():.4f}]", flush=True)
    assert e < 1e-9
    L.close()


def test_temperature_time_step_control_matches_oracle():
    w = _clamp_free()
    L, orc = _pair(w, "f64")
    rows = _rows(orc)
    _run(L, orc, 2)
    L.thermal_time_step(lib.DT_INJECT_NAN)
    orc.ts.inject_nan = 1
    _run(L, orc, 1)
    n, dt = L.thermal_time_step(lib.DT_QUERY)
    assert (n, dt) == (2, 0.5) and orc.ts.steps_per_lb == 2 and orc.ts.halvings == 1
    assert rel_err(L.download("T")[rows], orc.T[rows]) < 1e-9
    _run(L, orc, 20)
    n, dt = L.thermal_time_step(lib.DT_QUERY)
    assert (n, dt) == (1, 1.0) and orc.ts.steps_per_lb == 1
    e = rel_err(L.download("T")[rows], orc.T[rows])
    print(f"after halving and doubling back: T error {e:.2e}", flush=True)
    assert e < 1e-9


def test_temperature_needs_komega_and_a_geometry():
    w = channel.fouled_channel(128, 40, Fval=1e-5, seed=5, amp=(0.0, 2.0), width=(10.0, 40.0))
    L, _ = make_pair(w, "f32")
    with pytest.raises(lib.KofmError):
        L.thermal_init(**TP)
    with pytest.raises(lib.KofmError):
        L.thermal_step(1)
    L.close()
