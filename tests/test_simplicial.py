"""The simplicial 3D compute backend (include/vapor_simplicial3d.h) inside
vapor's own simplicial solver.

vapor's SimplicialFluidSolver3D hands circulation advection, smoke advection
and flux recovery to a SimplicialCompute3D backend
(src/solvers/simplicial3d/SimplicialCompute3D.hpp:7-13). The programs run
here are vapor's solver compiled from its sources with this library's backend
installed through include/vapor_gpu_simplicial3d.hpp, stepped beside the same
solver on its CPU path (tests/cpp/reference_simplicial_tests.cpp):

* simplicial_hostcheck: the backend's C ABI implemented on the CPU from the
  device code (tests/hostcheck/simplicial_host_abi.cpp), under
  AddressSanitizer - the binding, the mesh checks and the device arithmetic
  against vapor's CPU path, bit for bit, without a GPU;
* simplicial_tests: the HIP library on the GPU - bit-identical trajectories
  in the reference recovery order, vapor's own GPU bars
  (tests/SimplicialGpu3DTests.cpp:57-89) in the fast order, vapor's motion
  check, the error paths.

The meshes are vapor's own example meshes (examples/simplicial3d/*.tet,
examples/simplicial3d-bunny/bunny-fluid.tet), copied under
tests/golden/simplicial. Both programs link vapor's solver sources, so they
exist only where the reference tree was present at build time.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
REF = HERE.parent / "oracle" / "_ref"
HOSTCHECK = REF / "simplicial_hostcheck"
GPU = REF / "simplicial_tests"
MESH = HERE / "golden" / "simplicial"
BOX, DENSE, DENSE16, BUNNY = (MESH / n for n in ("box.tet", "box-dense.tet", "box-dense16.tet", "bunny-fluid.tet"))
NOT_BUILT = "built only where the reference tree was present at build time"


def run(exe: Path, *args: str, timeout: int = 280) -> str:
    print(f"[simplicial] {exe.name} {' '.join(args)}", flush=True)
    proc = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=timeout)
    print(proc.stdout[-4000:], proc.stderr[-2000:], flush=True)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr[-2000:]
    return proc.stdout


@pytest.mark.skipif(not HOSTCHECK.exists(), reason=NOT_BUILT)
@pytest.mark.parametrize("mesh,kind,steps", [(BOX, "box", 10), (DENSE, "box", 2), (BUNNY, "domain", 2),
                                             (BUNNY, "nobins", 1)],
                         ids=["box", "box-dense", "bunny", "bunny-without-bins"])
def test_device_code_on_the_host_is_bit_identical(mesh, kind, steps):
    out = run(HOSTCHECK, "exact", str(mesh), kind, str(steps))
    assert f"PASS exact: {steps} steps bit-identical" in out


@pytest.mark.skipif(not HOSTCHECK.exists(), reason=NOT_BUILT)
@pytest.mark.parametrize("mesh,kind", [(BOX, "box"), (BUNNY, "domain")], ids=["box", "bunny"])
def test_host_error_paths(mesh, kind):
    out = run(HOSTCHECK, "errors", str(mesh), kind)
    assert "PASS errors" in out and out.count("Non-finite CG RHS") == 2


@pytest.mark.gpu
@pytest.mark.skipif(not GPU.exists(), reason=NOT_BUILT)
@pytest.mark.parametrize("mesh,kind,steps", [(BOX, "box", 20), (DENSE, "box", 10), (BUNNY, "domain", 5),
                                             (BUNNY, "nobins", 3)],
                         ids=["box", "box-dense", "bunny", "bunny-without-bins"])
def test_gpu_backend_is_bit_identical_in_reference_order(mesh, kind, steps):
    """Seeded circulation and every step's flux, potential, vorticity, vertex
    and cell smoke equal vapor's CPU path bit for bit; same CG iterations and
    residuals (IC(0) on the bunny, Jacobi on boxes). "nobins" builds the bunny
    through vapor's Mesh(vertices, tets), which has no spatial bins, so
    locate and clipSegment take their scans over every tet and boundary face."""
    out = run(GPU, "exact", str(mesh), kind, str(steps))
    assert f"PASS exact: {steps} steps bit-identical" in out


@pytest.mark.gpu
@pytest.mark.skipif(not GPU.exists(), reason=NOT_BUILT)
@pytest.mark.parametrize("mesh,kind,steps", [(DENSE, "box", 60), (BUNNY, "domain", 30)], ids=["box-dense", "bunny"])
def test_gpu_backend_meets_vapors_gpu_bars(mesh, kind, steps):
    """vapor's own GPU backend test, fast recovery order: circulation <= 1e-8,
    flux <= 1e-5, density <= 1e-3, boundary flux <= 1e-10, divergence <= 1e-8
    each step; resize <= 1e-5; reset replay <= 1e-10."""
    out = run(GPU, "parity", str(mesh), kind, str(steps))
    assert "PASS resize and reset" in out


@pytest.mark.gpu
@pytest.mark.skipif(not GPU.exists(), reason=NOT_BUILT)
def test_gpu_smoke_rises():
    """vapor's motion check (tests/SimplicialGpu3DTests.cpp:26-41): 90 frames
    on box-dense, emitter off after 30, the smoke's centre rises by > 0.03."""
    out = run(GPU, "motion", str(DENSE), "90")
    assert "GPU motion" in out


@pytest.mark.gpu
@pytest.mark.skipif(not GPU.exists(), reason=NOT_BUILT)
@pytest.mark.parametrize("mesh,kind", [(BOX, "box"), (BUNNY, "domain")], ids=["box", "bunny"])
def test_gpu_error_paths(mesh, kind):
    """Non-finite right-hand sides, bad timesteps and bad CG parameters raise
    the reference's exceptions; the backend keeps working afterwards."""
    out = run(GPU, "errors", str(mesh), kind)
    assert "PASS errors" in out and out.count("Non-finite CG RHS") == 2
