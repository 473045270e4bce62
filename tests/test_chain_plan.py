"""CPU check of the chain-bundle sweep plan: tests/cpp/test_chain_plan plans
the ILU(0) sweeps of assembled box-mesh tangents (cube, slab, slender beam,
single line, lanes per node 4 and 8), replays each plan with the data flow of
the sweep kernels and compares with the oracle's sequential ILU(0) apply; a
randomly renumbered mesh must be refused (the level-scheduled sweeps take it)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_chain_plan_replay_matches_sequential_apply():
    exe = os.path.join(HERE, "cpp", "test_chain_plan")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/test_chain_plan is not built (run __graft_entry__.build())")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL OK" in p.stdout
