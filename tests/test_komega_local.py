"""CPU check of the k-omega adjoint's local transposes.

The adjoint kernels apply the transposed Jacobians of the per-node k-omega
maps (diffusivities, row assembly, eddy viscosity, strain) by forward-mode
evaluation of the same templated functions the forward kernels run
(csrc/komega_local.hpp).  komega_local_check.cpp compares J^T yb with
central differences of <yb, f(x)> for random realistic inputs, including
near-wall spacings (the difference-weight fallback and clamps) and wall
omega rows.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "komegafoulingmodel_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_local_transposes_match_finite_differences(tmp_path):
    exe = tmp_path / "komega_local_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "komega_local_check.cpp"),
                    "-o", str(exe)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(res.stdout[-2000:])
    assert res.returncode == 0
