"""GPU test of the C++ host facade: runs tests/cpp/test_host.cpp (built by
__graft_entry__.build() into fem_amd/_build/test_host), which drives
NumericalCoarseningUsingMultigrid, FEM3DFun (adjoint vs central differences),
gradientDescent and MaterialOptimizer through the C ABI."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fem_amd", "_build", "test_host")


def test_cpp_host_facade():
    assert os.path.exists(EXE), "build the facade test first (__graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
