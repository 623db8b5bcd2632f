"""GPU tests of the restart path and of the reference-style driver.

* Restart (Calc_Ro_J, lbKernels.cl:137-164): uploading the populations a
  context downloaded gives the moments of those populations (rho = sum f,
  u = j / rho, no force shift) and continues bit-identically.
* save2file writes lbf.bin as the FP64 FA array in the reference layout.
* kofm_run (the LBLS.cpp sequence over the C++ facades) reads
  RunParam.yaml, lsc0.txt and lsbl.txt, runs to Stop Time and writes the
  populations that the FP64 oracle reaches from the same initial state.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from komegafoulingmodel_amd import channel, lib
from pairs import fluid_mask, make_pair, rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mini():
    return channel.fouled_channel(256, 64, Fval=1e-5, seed=5, amp=(0.0, 4.0), width=(10.0, 60.0))


def test_restart_from_downloaded_populations_is_exact():
    w = _mini()
    L, orc = make_pair(w, "f32")
    L.step(37, 2)
    f = L.download("f")
    L2 = lib.Lattice(w.nX, w.nY, storage="f32")
    L2.set_params(w.tau0, w.Fval, w.nu)
    L2.build_geometry(w.C0, w.BL, **w.geom_params())
    L2.upload_state(f)
    F = f.reshape(9, -1)
    m = fluid_mask(orc.nType)
    rho = F.sum(axis=0)
    jx = (F[1] - F[2]) + (F[5] - F[6]) + (F[7] - F[8])
    jy = (F[3] - F[4]) + (F[5] - F[6]) - (F[7] - F[8])
    assert rel_err(L2.download("rho")[m], rho[m]) < 1e-6
    u = np.hypot(jx / np.where(m, rho, 1.0), jy / np.where(m, rho, 1.0))[m]
    du = np.hypot(L2.download("ux")[m] - (jx / np.where(m, rho, 1.0))[m],
                  L2.download("uy")[m] - (jy / np.where(m, rho, 1.0))[m])
    assert np.max(du) < 1e-5 * np.max(u)
    L.step(20, 2)
    L2.step(20, 2)
    assert np.array_equal(L.download("f"), L2.download("f"))
    assert np.array_equal(L.download("ux"), L2.download("ux"))


def test_save2file_writes_the_reference_layout(tmp_path):
    w = _mini()
    L, _ = make_pair(w, "f32")
    L.step(12, 2)
    L.save2file(str(tmp_path))
    lbf = np.fromfile(tmp_path / "lbf.bin", dtype=np.float64)
    assert lbf.shape == (9 * L.D,)
    assert np.array_equal(lbf, L.download("f"))


def test_driver_runs_the_reference_sequence(tmp_path):
    """The laminar sequence.  A thermal document holding only "Dummy", as the
    reference's RunParam.yaml has, leaves "Thermal Solver" at its default, on
    (ConstDef.h:222), and the temperature rows read nu_t (tfdKernels.cl:
    230-233), so with k-omega off the solve is switched off explicitly."""
    w = _mini()
    load, out = tmp_path / "load", tmp_path / "out"
    load.mkdir()
    out.mkdir()
    channel.write_load_files(str(load), w.C0, w.BL)
    (load / "RunParam.yaml").write_text(f"""# driver test
--- # System Parameters
ParamsName: 1
nX: !!int {w.nX}
nY: !!int {w.nY}
Pipe Radius: !!float {w.pipe_radius!r}
Stop Time: !!int 40
Restart Run: False

--- # LS Parameters
ParamsName: 5
LS Spacing: !!float 1.0
nN: !!int {w.C0.shape[0]}

--- # Fluid Parameters
ParamsName: 2
Tau: !!float {w.tau0!r}
Fval: !!float {w.Fval!r}
Use Turb Model: False
Ini LB Steps: !!int 0
Ini Poiseuille: False
load Vel Text: False

--- # Thermal Parameters
ParamsName: 3
Dummy: !!int 0
Thermal Solver: False

--- # Particle Parameters
ParamsName: 6
X Release Pos: !!float 10.0
X Stop Pos: !!float {float(w.nX - 12)!r}
""")
    exe = os.path.join(ROOT, "komegafoulingmodel_amd", "kofm_run")
    res = subprocess.run([exe, str(load), str(out)], capture_output=True, text=True, timeout=120)
    print(res.stdout[-400:], res.stderr[-400:], flush=True)
    assert res.returncode == 0
    g = oracle.geometry(w.nX, w.nY, w.C0, w.BL, **w.geom_params())
    orc = oracle.OracleLattice(g, tau0=w.tau0, Fval=w.Fval, nu=w.nu)
    fluid = fluid_mask(g["nType"])
    orc.init_state(np.where(fluid, 1.0, 0.0), np.zeros(g["D"]), np.zeros(g["D"]))
    orc.step(40)
    lbf = np.fromfile(out / "lbf.bin", dtype=np.float64)
    err = rel_err(lbf, orc.FA)
    print(f"driver lbf vs oracle after 40 steps {err:.2e}", flush=True)
    assert err < 1e-5


def test_driver_ini_poiseuille_and_save_macros_on_start(tmp_path):
    """"Ini Poiseuille" (clVariablesLB.cpp:325-329, iniDists(Umax) :427-496)
    through kofm_run: with the laminar profile and no perturbation the
    initial populations are feq(Rho, u(y), 0) with the parabola
    u(y) = 1.5 Re nu / R (1 - ((y - 0.5) - R)^2 / R^2) of a channel whose walls
    sit at y = 0.5 and nY - 1.5; "Save Macros On Start" writes lbux.txt
    (one row per x, nY tab-separated values, 6 significant digits)."""
    w = channel.flat_channel(256, 40, Fval=1e-6, noise=0.0)
    load, out = tmp_path / "load", tmp_path / "out"
    load.mkdir()
    out.mkdir()
    channel.write_load_files(str(load), w.C0, w.BL)
    Re = 5.0  # u_max = 1.5 Re nu / R = 0.039
    (load / "RunParam.yaml").write_text(f"""# driver test
--- # System Parameters
ParamsName: 1
nX: !!int {w.nX}
nY: !!int {w.nY}
Pipe Radius: !!float {w.pipe_radius!r}
Stop Time: !!int 0

--- # Fluid Parameters
ParamsName: 2
Tau: !!float {w.tau0!r}
Re: !!float {Re!r}
Use Turb Model: False
Ini LB Steps: !!int 0
Ini Poiseuille: True
Ini Turb Velocity: False
Perturb Vel Field: False
load Vel Text: False
Save Macros On Start: True

--- # Particle Parameters
ParamsName: 6
X Release Pos: !!float 10.0
X Stop Pos: !!float {float(w.nX - 12)!r}
""")
    exe = os.path.join(ROOT, "komegafoulingmodel_amd", "kofm_run")
    res = subprocess.run([exe, str(load), str(out)], capture_output=True, text=True, timeout=120)
    print(res.stdout[-400:], res.stderr[-400:], flush=True)
    assert res.returncode == 0
    R, nu = w.pipe_radius, w.nu
    umax = 1.5 * Re * nu / R
    XF, D = w.XF, w.D
    lbf = np.fromfile(out / "lbf.bin", dtype=np.float64).reshape(9, D)
    ux_txt = np.loadtxt(out / "lbux.txt")
    assert ux_txt.shape == (w.nX, w.nY)
    for y in range(1, w.nY - 1):
        s = (y - 0.5) - R
        u = umax * (1.0 - s * s / (R * R))
        fe = oracle.feq(1.0, u, 0.0)
        got = lbf[:, y * XF:y * XF + w.nX]
        assert np.max(np.abs(got - fe[:, None])) < 1e-8, y
        assert np.allclose(ux_txt[:, y], u, rtol=1e-5, atol=1e-12), y
    assert np.all(ux_txt[:, 0] == 0.0) and np.all(ux_txt[:, -1] == 0.0)


def _driver_yaml(w, stop, extra_sys="", extra_fluid="", extra_docs=""):
    return f"""# driver test
--- # System Parameters
ParamsName: 1
nX: !!int {w.nX}
nY: !!int {w.nY}
Pipe Radius: !!float {w.pipe_radius!r}
Stop Time: !!int {stop}
{extra_sys}
--- # LS Parameters
ParamsName: 5
LS Spacing: !!float 1.0
nN: !!int {w.C0.shape[0]}

--- # Fluid Parameters
ParamsName: 2
Tau: !!float {w.tau0!r}
Fval: !!float {w.Fval!r}
Use Turb Model: False
Ini LB Steps: !!int 0
Ini Poiseuille: False
load Vel Text: False
{extra_fluid}
--- # Particle Parameters
ParamsName: 6
X Release Pos: !!float 10.0
X Stop Pos: !!float {float(w.nX - 12)!r}
{extra_docs}"""


def _run_driver(load, out):
    exe = os.path.join(ROOT, "komegafoulingmodel_amd", "kofm_run")
    res = subprocess.run([exe, str(load), str(out)], capture_output=True, text=True, timeout=120)
    print(res.stdout[-300:], res.stderr[-300:], flush=True)
    assert res.returncode == 0
    return res


def test_driver_dumps_restart_files_and_restarts_from_them(tmp_path):
    """The dump sequence of clProblem::step (clProblem.cpp:130-168): every
    dump writes the save2file text, the held time data and RunParam.yaml
    (Current Time, Restart Run) into the output directory; every "Save Bin
    Step" dumps the restart files (lbf.bin, lsc.bin, saveRestartFiles) and
    RunParam.yaml go to the load directory.  Time data rows are held and
    written every "lbOut Max Lines" rows.  A run of 40 steps restarted from
    its load directory to 60 steps ends bit-identical to 60 steps in one
    run."""
    w = _mini()
    sys_extra = "Dump Timestep: !!int 10\nSave Avgs Timestep: !!int 5\nSave Bin Step: !!int 2\n"
    fl_extra = "lbOut Max Lines: !!int 3\n"
    # one run of 60 steps
    la, oa = tmp_path / "la", tmp_path / "oa"
    la.mkdir(), oa.mkdir()
    channel.write_load_files(str(la), w.C0, w.BL)
    (la / "RunParam.yaml").write_text(_driver_yaml(w, 60, sys_extra, fl_extra))
    _run_driver(la, oa)
    # 40 steps, then a restart from the load directory to 60
    lb, ob = tmp_path / "lb", tmp_path / "ob"
    lb.mkdir(), ob.mkdir()
    channel.write_load_files(str(lb), w.C0, w.BL)
    (lb / "RunParam.yaml").write_text(_driver_yaml(w, 40, sys_extra, fl_extra))
    _run_driver(lb, ob)
    y = (ob / "RunParam.yaml").read_text()
    assert "Current Time: !!int 40" in y and "Restart Run: True" in y
    assert (lb / "RunParam.yaml").read_text() == y
    assert np.array_equal(np.fromfile(lb / "lbf.bin"), np.fromfile(ob / "lbf.bin"))
    # lsc.bin holds the wall points as read from lsc0.txt
    assert np.array_equal(np.fromfile(lb / "lsc.bin").reshape(-1, 2), np.loadtxt(lb / "lsc0.txt"))
    rows = np.loadtxt(ob / "redUx.txt")
    assert rows.shape == (8, 2) and list(rows[:, 0]) == [5, 10, 15, 20, 25, 30, 35, 40]
    (lb / "RunParam.yaml").write_text(y.replace("Stop Time: !!int 40", "Stop Time: !!int 60"))
    ob2 = tmp_path / "ob2"
    ob2.mkdir()
    _run_driver(lb, ob2)
    assert np.array_equal(np.fromfile(ob2 / "lbf.bin"), np.fromfile(oa / "lbf.bin"))
    assert "Current Time: !!int 60" in (ob2 / "RunParam.yaml").read_text()


def test_save_debug_and_restart_file_layouts(tmp_path):
    """kofm_save_debug writes nType.txt (integers), dXarr_0..3.txt, lbnut.txt
    and wallD.txt in the save2file layout (one row per x, nY values);
    kofm_save_restart writes lbf.bin, lbkappa.bin, lbomega.bin and lsc.bin."""
    nu, Re, H = 0.01, 500.0, 56.0
    w = channel.wavy_channel(256, 72, H=H, amp=4.0, period=128.0, nu=nu, Re=Re,
                             Fval=channel.turbulent_forcing(H, nu, Re), seed=7)
    L, _ = make_pair(w, "f32")
    L.step(3, 2)
    L.save_debug(str(tmp_path))
    L.save_restart(str(tmp_path))
    nx, ny, XF = w.nX, w.nY, w.XF

    def grid(a):
        return np.asarray(a).reshape(-1, XF)[:ny, :nx].T

    nt = np.loadtxt(tmp_path / "nType.txt", dtype=np.int64)
    assert np.array_equal(nt, grid(L.geometry("nType").astype(np.int64)))
    dX = L.geometry("dX").reshape(4, -1)
    for k in range(4):
        assert np.allclose(np.loadtxt(tmp_path / f"dXarr_{k}.txt"), grid(dX[k]), rtol=1e-5, atol=0)
    assert np.allclose(np.loadtxt(tmp_path / "lbnut.txt"), grid(L.download("nut")), rtol=1e-5, atol=1e-300)
    assert np.allclose(np.loadtxt(tmp_path / "wallD.txt"), grid(L.download("wallD")), rtol=1e-5, atol=0)
    assert np.array_equal(np.fromfile(tmp_path / "lbf.bin"), L.download("f"))
    assert np.array_equal(np.fromfile(tmp_path / "lbkappa.bin"), L.download("k"))
    assert np.array_equal(np.fromfile(tmp_path / "lbomega.bin"), L.download("omega"))
    assert np.array_equal(np.fromfile(tmp_path / "lsc.bin").reshape(-1, 2), w.C0)


def test_driver_runs_the_temperature_solve_and_restarts_it(tmp_path):
    """kofm_run with a thermal document and k-omega on: clVariablesFD::Solve
    after every LB step ("LB Steps Per T" 1, LBLS.cpp:197-200), temp.txt in
    the save2file layout (one row per x, nY values), and temp.bin with the
    restart files; a run of 12 steps restarted from its load directory
    continues from temp.bin.  With k-omega on, a restart is not bit-identical
    in general: the strain that the SST limiter takes from the previous step
    is not among the reference's restart files (lbf, lbkappa, lbomega), so
    the first restarted step sees none (on this slow channel the limiter
    does not act and the runs agree bit for bit).  The check is that the
    restarted temperature stays within 1 % of the change the first 12 steps
    made; restarted from the initial field it would be far off.

    "Calculate Nu" is on by default, with the wavy section given: Nu.txt and
    NuMean.txt get one row every "Save Nu Timestep" steps (time, then 5
    values per column of the particle domain, 2 per period boundary), held
    "Max Nu Output Lines" at a time; the restarted run resumes the cadence
    from RunParam.yaml's nextSaveNuTime."""
    w = _mini()
    fl = ("Use Turb Model: True\nKappa Max Rel Tol: !!float 1e-12\n"
          "Kappa Max Abs Tol: !!float 0.0\nKappa Max Iterations: !!int 400\n")
    th = ("--- # Thermal Parameters\nParamsName: 3\nThermal Solver: True\n"
          "Tinitial: !!float 0.5\nTinlet Sim: !!float 2.0\nTemp Max Rel Tol: !!float 1e-12\n"
          "Temp Max Abs Tol: !!float 0.0\nMax Nu Output Lines: !!int 2\n")
    sys_extra = ("Dump Timestep: !!int 6\nSave Avgs Timestep: !!int 6\nSave Bin Step: !!int 2\n"
                 "Wavy Start Loc: !!int 0\nPeriod Length: !!int 100\nNum Periods: !!int 2\n"
                 "Save Nu Timestep: !!int 4\n")
    la, oa = tmp_path / "la", tmp_path / "oa"
    la.mkdir(), oa.mkdir()
    channel.write_load_files(str(la), w.C0, w.BL)
    (la / "RunParam.yaml").write_text(_driver_yaml(w, 24, sys_extra, fl, th))
    res = _run_driver(la, oa)
    assert "temperature on Nu on" in res.stdout
    n = w.nX - 12 - 8  # columns floor(X Min Val) = 8 to floor(X Stop Pos) - 1
    nu = np.loadtxt(oa / "Nu.txt", ndmin=2)
    mean = np.loadtxt(oa / "NuMean.txt", ndmin=2)
    assert nu.shape == (6, 1 + 5 * n) and mean.shape == (6, 1 + 2 * 3)
    assert np.array_equal(nu[:, 0], [4, 8, 12, 16, 20, 24])
    assert np.array_equal(mean[:, 0], nu[:, 0])
    assert np.all(np.isfinite(nu[:, 1::5])) and np.all(nu[:, 1::5] > 0.0)
    assert np.all(mean[:, 2::2] > 0.0)
    T = np.loadtxt(oa / "temp.txt")
    assert T.shape == (w.nX, w.nY)
    g = oracle.geometry(w.nX, w.nY, w.C0, w.BL, **w.geom_params())
    rows = ((g["nType"].astype(np.int32) & 0x4) == 0).reshape(w.nY, w.XF)[:, :w.nX].T
    walls = ~rows
    assert np.all(T[walls] == 0.0)
    assert np.any(np.abs(T[rows] - 0.5) > 1e-3)
    assert np.all(T[rows] > -1e-9) and np.all(T[rows] <= 2.0 + 1e-9)
    lb, ob = tmp_path / "lb", tmp_path / "ob"
    lb.mkdir(), ob.mkdir()
    channel.write_load_files(str(lb), w.C0, w.BL)
    (lb / "RunParam.yaml").write_text(_driver_yaml(w, 12, sys_extra, fl, th))
    _run_driver(lb, ob)
    assert (lb / "temp.bin").exists()
    y = (lb / "RunParam.yaml").read_text()
    (lb / "RunParam.yaml").write_text(y.replace("Stop Time: !!int 12", "Stop Time: !!int 24"))
    ob2 = tmp_path / "ob2"
    ob2.mkdir()
    _run_driver(lb, ob2)
    assert np.array_equal(np.loadtxt(ob2 / "Nu.txt", ndmin=2)[:, 0], [16, 20, 24])
    T2 = np.loadtxt(ob2 / "temp.txt")
    gap = np.max(np.abs(T2[rows] - T[rows])) / np.max(np.abs(T[rows]))
    T0 = np.loadtxt(ob / "temp.txt")
    moved = np.max(np.abs(T0[rows] - 0.5))
    print(f"restarted temperature against one run: {gap:.2e}; change over the first 12 "
          f"steps {moved:.2e}", flush=True)
    assert gap < 1e-2 * moved
