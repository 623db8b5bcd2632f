"""CPU tests of the host side: the C ABI library loads and exports every
symbol include/kofm.h declares, the host geometry builder (run through a
host-only context) is bit-identical to the oracle's restatement, and the
channel recipe reproduces the reference's MakeChannel layout."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from komegafoulingmodel_amd import channel, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "kofm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(kofm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    names = _header_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert sorted(lib.EXPORTS) == names


def test_create_reports_missing_device_without_crashing():
    L = lib.load()
    g = lib.Grid(64, 16, 0, 0)
    h = C.c_void_p()
    rc = L.kofm_create(C.byref(g), C.byref(h))
    if rc == 0:  # a GPU is present
        L.kofm_destroy(h)
    else:
        assert rc == -2 and b"HIP" in L.kofm_last_error(None)


def test_argument_errors_are_codes():
    L = lib.load()
    g = lib.Grid(2, 2, 0, 0)
    h = C.c_void_p()
    assert L.kofm_create_host(C.byref(g), C.byref(h)) == -1
    hl = lib.Lattice(64, 16, host_only=True)
    with pytest.raises(lib.KofmError) as e:
        hl.step(1)
    assert e.value.code == -3
    with pytest.raises(lib.KofmError):
        hl.set_params(tau0=2.5, Fval=0.0, nu=0.1)


WORKLOADS = {
    "flat": lambda: channel.flat_channel(256, 30),
    "fouled": lambda: channel.fouled_channel(384, 72, seed=13, amp=(0, 6), width=(10, 80)),
    "wavy": lambda: channel.wavy_channel(256, 96, H=80, amp=6, period=128, nu=0.01, Re=2000,
                                         Fval=1e-6),
    "fouled_padded": lambda: channel.fouled_channel(200, 40, seed=2, amp=(0, 3), width=(10, 40)),
}


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_host_geometry_matches_oracle_bit_exact(name):
    w = WORKLOADS[name]()
    g = oracle.geometry(w.nX, w.nY, w.C0, w.BL, **w.geom_params())
    L = lib.Lattice(w.nX, w.nY, host_only=True)
    L.build_geometry(w.C0, w.BL, **w.geom_params())
    c = L.counts()
    assert c.XsizeFull == g["XF"] and c.D == g["D"]
    for k in ["nType", "dX", "ibb", "ibbQ", "ss", "ssMap", "ssInds", "lsMap", "wallD", "cfss",
              "sInd", "weights", "bounds"]:
        assert np.array_equal(L.geometry(k), g[k]), k
    m = L.geometry("masses")
    nSf = 0  # fresh run: no fouled-away nodes
    assert m[0] == 2 * int(w.pipe_radius) * w.nX - nSf and m[1] == 0


def test_geometry_rejects_fluid_on_edge_rows():
    """Rows 0 and nY-1 must be solid (clVariablesLS.cpp:935-937)."""
    C0, BL = channel.wall_points(64, lambda x: np.full_like(x, -0.5), lambda x: np.full_like(x, 20.5))
    L = lib.Lattice(64, 24, host_only=True)
    with pytest.raises(lib.KofmError) as e:
        L.build_geometry(C0, BL, x_release_pos=10.0, x_stop_pos=40.0)
    assert e.value.code == -20


def test_channel_recipe_matches_makechannel():
    """MakeChannel_Test.m:68-113: x from -0.75 in unit steps past nX - 0.5,
    top wall back to -0.75; nN = 2 (nX + 2), nBL = nN - 2 (the shipped
    RunParam.yaml pairs nX = 2000 with nN = 4004)."""
    C0, BL = channel.wall_points(2000, lambda x: 1.5 + 0 * x, lambda x: 201.5 + 0 * x)
    assert C0.shape == (4004, 2) and BL.shape == (4002, 2)
    assert C0[0, 0] == -0.75 and C0[2001, 0] == 2000.25 and C0[2002, 0] == 2000.25
    assert C0[-1, 0] == -0.75
    assert np.all(BL[:2001, 1] - BL[:2001, 0] == 1) and BL[2001, 0] == 2002


def test_write_load_files(tmp_path):
    w = channel.flat_channel(128, 20)
    channel.write_load_files(str(tmp_path), w.C0, w.BL)
    c = np.loadtxt(tmp_path / "lsc0.txt")
    b = np.loadtxt(tmp_path / "lsbl.txt", dtype=np.int64)
    assert np.array_equal(c, w.C0) and np.array_equal(b, w.BL)


def test_runparam_yaml_through_the_facades():
    """The shipped load/RunParam.yaml (kept as tests/golden/RunParam.yaml)
    read by the C++ facades: Tau is the relaxation frequency and
    Mu = (1/Tau - 0.5)/3, Fval = 3 Re Mu^2 / R^3 (clVariablesLB.cpp:566-600),
    XsizeFull = ceil(nX/128)*128 (clProblem.cpp:469)."""
    import subprocess
    exe = os.path.join(ROOT, "komegafoulingmodel_amd", "kofm_run")
    out = subprocess.run([exe, "--params", os.path.join(ROOT, "tests", "golden")],
                         capture_output=True, text=True, check=True).stdout
    v = {k: float(x) for k, x in (line.split() for line in out.strip().splitlines())}
    tau, R, Re = 1.99445983, 100.0, 10800.0
    mu = (1 / tau - 0.5) / 3
    assert v["nX"] == 2000 and v["nY"] == 204 and v["XsizeFull"] == 2048
    assert v["Channel_Height"] == 200 and v["Pipe_radius"] == R
    assert v["tau0"] == tau and abs(v["MuVal"] - mu) < 1e-18
    assert abs(v["Fval"] - Re * 3 * mu * mu / R ** 3) < 1e-22
    assert v["UseTurbModel"] == 1 and v["trStepsWall"] == 3
