"""The C++ host layer's GPU cases: binding_system_interactions, lbfgs_step,
free_ligand::apply_parameters, monte_carlo_searcher and pose_selection over
libpxdock.so, against the reference gtest vectors (protenix-dock_amd/cpp/test_host.cpp)."""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "protenix-dock_amd", "cpp", "_build", "test_host")


def test_gpu_cpp_host_all_cases():
    assert os.path.exists(BIN), "build the C++ host tests first (__graft_entry__.build())"
    print("[cpp] running", BIN, flush=True)
    out = subprocess.run([BIN, GOLDEN], capture_output=True, text=True, timeout=240)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 failures" in out.stdout
