"""Runs the C++ drop-in program built against include/mfgpu.hpp: solver code
written like Solid::solve_linear_system on the device operator."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "matrix-free-dealii-precice_amd",
                   "_build", "drop_in_test")


def test_cpp_drop_in(cuda):
    assert os.path.exists(EXE), "drop-in program not built"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "DROPIN OK" in r.stdout
