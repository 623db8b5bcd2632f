"""GPU parity: libkofm.so (HIP, gfx950) against the FP64 oracle on the same
seeded inputs, through the C ABI.

Tolerances (normwise relative, max|gpu - oracle| / max|oracle|; the velocity
is compared as a vector, max|u_gpu - u_oracle| / max|u_oracle|, because a
single component can decay to the FP32 noise floor):
  FP32 storage + FP32 arithmetic: 1e-5 for populations, rho, u, k, omega,
  nu_t and wall shear after the stated steps;
  FP64 storage + FP64 arithmetic: 1e-11;
  FP16 / BF16 storage: reported, bounded loosely (2e-3 / 2e-2 on u).
"""
import numpy as np
import pytest

from komegafoulingmodel_amd import channel
from pairs import fluid_mask, make_pair, rel_err, vec_rel_err

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-11


def _mini_flat():
    return channel.flat_channel(256, 24, Fval=1e-5, noise=1e-4, seed=11)


def _mini_fouled():
    return channel.fouled_channel(256, 64, Fval=1e-5, seed=5, amp=(0.0, 4.0), width=(10.0, 60.0))


def _mini_wavy_ko():
    nu, Re, H = 0.01, 2000.0, 40.0
    return channel.wavy_channel(256, 56, H=H, amp=4.0, period=128.0, nu=nu, Re=Re,
                                Fval=channel.turbulent_forcing(H, nu, Re), seed=7)


def _compare_lb(L, orc, tol, label):
    m = fluid_mask(orc.nType)
    f = L.download("f")
    errs = {
        "f": rel_err(f, orc.FA),
        "rho": rel_err(L.download("rho")[m], orc.rho[m]),
        "u": vec_rel_err(L.download("ux")[m], L.download("uy")[m], orc.ux[m], orc.uy[m]),
    }
    print(label, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < tol, (label, k, v)
    return errs


@pytest.mark.parametrize("make,steps", [(_mini_flat, 300), (_mini_fouled, 300)])
def test_lb_fp32_matches_oracle(make, steps):
    w = make()
    L, orc = make_pair(w, "f32")
    # first step has no pending wall correction, later ones do
    L.step(1, 2)
    orc.step(1)
    _compare_lb(L, orc, TOL32, f"{w.name} step1")
    L.step(steps - 1, 2)
    orc.step(steps - 1)
    _compare_lb(L, orc, TOL32, f"{w.name} step{steps}")


def test_lb_fp64_matches_oracle():
    w = _mini_fouled()
    L, orc = make_pair(w, "f64")
    L.step(200, 2)
    orc.step(200)
    _compare_lb(L, orc, TOL64, "fouled f64")


@pytest.mark.parametrize("storage,tol_u", [("f16", 2e-3), ("bf16", 2e-2)])
def test_lb_16bit_storage_error_report(storage, tol_u):
    w = _mini_fouled()
    L, orc = make_pair(w, storage)
    L.step(200, 2)
    orc.step(200)
    m = fluid_mask(orc.nType)
    eu = vec_rel_err(L.download("ux")[m], L.download("uy")[m], orc.ux[m], orc.uy[m])
    ef = rel_err(L.download("f"), orc.FA)
    print(storage, f"ux {eu:.2e} f {ef:.2e}", flush=True)
    assert eu < tol_u and ef < tol_u


def test_wall_shear_matches_oracle():
    w = _mini_fouled()
    L, orc = make_pair(w, "f32")
    L.step(150, 2)
    orc.step(150)
    L.wall_shear()
    tn, tb = orc.shear()
    gn, gb = L.download("tau_node"), L.download("tau_bl")
    en, eb = rel_err(gn, tn), rel_err(gb, tb)
    print(f"shear node {en:.2e} link {eb:.2e}", flush=True)
    assert en < TOL32 and eb < TOL32


def test_reductions_match_oracle():
    w = _mini_fouled()
    L, orc = make_pair(w, "f64")
    L.step(50, 2)
    orc.step(50)
    for field, arr in (("rho", orc.rho), ("ux", orc.ux), ("uy", orc.uy)):
        for op in (1, 2, 4):
            a = L.reduce(field, op)
            b = orc.reduce(arr, op)
            assert abs(a - b) <= 1e-11 * max(1.0, abs(b)), (field, op, a, b)


def test_komega_fp32_matches_oracle():
    w = _mini_wavy_ko()
    L, orc = make_pair(w, "f32")
    m = fluid_mask(orc.nType)
    for n in (1, 4):
        L.step(n, 2)
        orc.step(n)
        st = L.solver_stats()
        errs = {k: rel_err(L.download(k)[m], getattr(orc, a)[m])
                for k, a in (("k", "K"), ("omega", "O"), ("nut", "nut"))}
        errs["u"] = vec_rel_err(L.download("ux")[m], L.download("uy")[m], orc.ux[m], orc.uy[m])
        errs["f"] = rel_err(L.download("f"), orc.FA)
        print("ko", n, {k: f"{v:.2e}" for k, v in errs.items()},
              "iters", st.kappa_iters, st.omega_iters, flush=True)
        for k, v in errs.items():
            assert v < TOL32, (k, v)


def test_komega_fp64_matches_oracle():
    w = _mini_wavy_ko()
    L, orc = make_pair(w, "f64", solver_rtol=1e-15, solver_maxit=2000)
    m = fluid_mask(orc.nType)
    L.step(3, 2)
    orc.step(3)
    for k, a in (("k", "K"), ("omega", "O"), ("nut", "nut"), ("ux", "ux")):
        e = rel_err(L.download(k)[m], getattr(orc, a)[m])
        print("ko f64", k, f"{e:.2e}", flush=True)
        assert e < 1e-9, (k, e)
