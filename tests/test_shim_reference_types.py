"""The C++ drop-in compiled against the reference's real headers.

tests/cpp/shim_reference_types.cpp instantiates every template of
include/flexmesh/fm_engine.hpp with the reference's own MeshTopology,
ProblemDescriptor, SparseCSR, NeoHookeanMaterial, BoundaryConditions, LoadStep
and ZZErrorEstimator.  The reference includes its headers as fem/... and
core/... while the directories are FEM/ and CORE/; a temporary include
directory holds two symlinks to them.  Compile only (no link, no run); skipped
where the reference sources are not present.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("REF", "/root/reference")


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "FEM")) or shutil.which("g++") is None,
                    reason="reference sources or g++ not present")
def test_shim_compiles_against_reference_types(tmp_path):
    inc = tmp_path / "inc"
    inc.mkdir()
    os.symlink(os.path.join(REF, "FEM"), inc / "fem")
    os.symlink(os.path.join(REF, "CORE"), inc / "core")
    src = os.path.join(ROOT, "tests", "cpp", "shim_reference_types.cpp")
    r = subprocess.run(["g++", "-std=c++20", "-O0", "-w", "-c", f"-I{inc}", f"-I{ROOT}/include", src, "-o",
                        str(tmp_path / "shim.o")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
