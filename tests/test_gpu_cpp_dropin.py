"""The reference's whole-step tests (tests/SimulationTests.cpp) compiled
against the C++ drop-in vapor::hip::MacGridFluidSolver3D and run on the GPU."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "cpp" / "_build" / "simulation_tests"


@pytest.mark.parametrize("case", ["", "seeded_replay", "plume"])
def test_reference_simulation_tests_on_dropin(case):
    assert EXE.exists(), "build first: __graft_entry__.build()"
    print(f"[gpu] C++ drop-in simulation_tests {case or 'baselines'}", flush=True)
    args = [str(EXE)] + ([case] if case else [])
    proc = subprocess.run(args, capture_output=True, text=True, timeout=240)
    print(proc.stdout[-2000:], proc.stderr[-2000:], flush=True)
    assert proc.returncode == 0, proc.stderr
    if not case:
        assert proc.stdout.count("PASS") == 16
