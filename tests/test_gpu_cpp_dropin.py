"""The C++ drop-in (include/vapor_mac3d.hpp) on the GPU: tests/cpp/dropin_tests
in its four groups. The published whole-step baselines come from the data
file tests/golden/baselines.json (SimulationTests.cpp:17-34 of the reference);
the plume bounds are SimulationTests.cpp:79-99's; the stage-wiring checks
follow BackendTests.cpp:62-143 (dt/2 and K/2 per projection under
reflection, call counts, ownership, reset, null rejection, box changes)."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

HERE = Path(__file__).resolve().parent
EXE = HERE / "cpp" / "_build" / "dropin_tests"
BASELINES = HERE / "golden" / "baselines.json"


@pytest.mark.parametrize("group", ["baselines", "replay", "plume", "stages"])
def test_cpp_dropin(group):
    assert EXE.exists(), "build first: __graft_entry__.build()"
    print(f"[gpu] C++ drop-in checks: {group}", flush=True)
    args = [str(EXE), group] + ([str(BASELINES)] if group == "baselines" else [])
    proc = subprocess.run(args, capture_output=True, text=True, timeout=280)
    print(proc.stdout[-4000:], proc.stderr[-2000:], flush=True)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr
    assert "FAIL" not in proc.stdout
    if group == "baselines":
        assert proc.stdout.count("ok ") == 1 + 16 * 6


EXE2D = HERE / "cpp" / "_build" / "dropin2d_tests"


@pytest.mark.parametrize("group", ["grid", "advection", "projection", "plume", "reflection", "historical"])
def test_cpp_dropin_mac2d(group):
    """include/vapor_mac2d.hpp from C++ (tests/cpp/dropin2d_tests.cpp): the
    bounds of vapor's Mac2DTests.cpp and Historical2DTests.cpp."""
    assert EXE2D.exists(), "build first: __graft_entry__.build()"
    print(f"[gpu] C++ 2D drop-in checks: {group}", flush=True)
    proc = subprocess.run([str(EXE2D), group], capture_output=True, text=True, timeout=280)
    print(proc.stdout[-4000:], proc.stderr[-2000:], flush=True)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr
    assert "FAIL" not in proc.stdout and proc.stdout.count("ok ") >= 3


WIRING = HERE.parent / "oracle" / "_ref" / "stage_wiring_tests"


@pytest.mark.skipif(not WIRING.exists(), reason="built only where the reference tree was present at build time")
def test_reference_side_binding_in_vapors_solver():
    """include/vapor_hip_stages.hpp installed into vapor's own CPU solver
    (compiled from the reference sources): wiring and agreement with vapor's
    CPU stages within 2e-5 (tests/cpp/reference_stage_wiring.cpp)."""
    print("[gpu] reference-side stage binding in vapor's solver", flush=True)
    proc = subprocess.run([str(WIRING)], capture_output=True, text=True, timeout=280)
    print(proc.stdout[-4000:], proc.stderr[-2000:], flush=True)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr
    assert "FAIL" not in proc.stdout and proc.stdout.count("ok ") >= 30
