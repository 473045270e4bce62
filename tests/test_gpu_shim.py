"""The C++ drop-in (include/flexmesh/fm_engine.hpp) called the way engine code
calls the reference: pattern and colouring bit-exact, assembly and PCG against
the FP64 oracle, NonlinearNewtonSolver against the closed-form uniaxial state."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.gpu
def test_cpp_engine_shim():
    exe = os.path.join(HERE, "cpp", "test_engine_shim")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/test_engine_shim is not built (run __graft_entry__.build())")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL OK" in p.stdout
