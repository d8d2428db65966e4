"""Snapshot I/O and the device-resident hand-off of the fields.

- Snapshots (vgx_sim_save_snapshot / vgx_sim_load_snapshot): raw FP32 files
  in the reference layouts plus a JSON header. The density file holds the
  bytes vapor's raw density dump writes (ReflectionComparison.hpp:74-75: the
  n^3 floats of MacGridState3D::density(), native little-endian). A snapshot
  loaded into a fresh simulation of the same grid continues the trajectory
  bit for bit, warm-started tolerance solves included; a mismatched grid or a
  damaged file is rejected (GpuMacSimulation3D::upload's checks,
  GpuMacSimulation3D.cpp:110-113) without changing the state.
- Device fields (vgx_sim_device_fields; RenderData.hpp:11-13 with a device
  address in place of the GL buffer, GpuMacSimulation3D.cpp:107-109): the
  addresses hold the state a download returns, read here by hipMemcpy."""
import ctypes

import numpy as np
import pytest

from helpers import bits_equal
from vapor_amd import mac3d
from vapor_amd.train_warm_start import scene

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def say(*args):
    print("[gpu io]", *args, flush=True)


def stepped(n, steps, seed=3):
    sim = mac3d.GpuMacSimulation3D(n, 1.0, scene(seed, mac3d))
    for _ in range(steps):
        sim.advance(DT)
    return sim


def test_snapshot_continues_the_trajectory_bitwise(tmp_path):
    n = 32
    say("snapshot round trip, 32^3 reflection MacCormack, tolerance solves with warm start")
    a = stepped(n, 3)
    base = tmp_path / "state"
    a.saveSnapshot(base)
    snap = mac3d.read_snapshot(base)
    header = snap["header"]
    assert header["resolution"] == n and header["box_size"] == 1.0
    rho, u, v, w = a.fields()
    assert bits_equal(snap["density"], rho) and bits_equal(snap["velocity_x"], u)
    assert bits_equal(snap["velocity_y"], v) and bits_equal(snap["velocity_z"], w)
    assert bits_equal(snap["pressure"], a.pressure())
    # vapor's raw density dump: n^3 native floats, nothing else
    raw = (tmp_path / "state.density").read_bytes()
    assert len(raw) == 4 * n ** 3 and raw == rho.astype("<f4").tobytes()

    b = mac3d.GpuMacSimulation3D(n, 1.0, scene(3, mac3d))
    b.loadSnapshot(base)
    for step in range(3):
        a.advance(DT)
        b.advance(DT)
        for k in (0, 1):
            assert a.lastProjection(k).iterations == b.lastProjection(k).iterations, (step, k)
    for x, y in zip(a.fields(), b.fields()):
        assert bits_equal(x, y)
    assert bits_equal(a.pressure(), b.pressure())
    say("continued 3 steps bit for bit; sweeps", a.lastProjection(0).iterations)
    a.close()
    b.close()


def test_snapshot_rejections_leave_the_state(tmp_path):
    a = stepped(16, 2)
    base = tmp_path / "s16"
    a.saveSnapshot(base)
    other = stepped(24, 1)
    before = other.fields()
    with pytest.raises(ValueError, match="resolution"):
        other.loadSnapshot(base)
    boxed = mac3d.GpuMacSimulation3D(16, 2.0, scene(3, mac3d))
    with pytest.raises(ValueError, match="box"):
        boxed.loadSnapshot(base)
    target = stepped(16, 1, seed=5)
    kept = target.fields()
    with open(str(base) + ".vely", "r+b") as f:  # a truncated field file
        f.truncate(100)
    with pytest.raises(ValueError, match="size"):
        target.loadSnapshot(base)
    with pytest.raises(ValueError):
        target.loadSnapshot(tmp_path / "missing")
    for x, y in zip(kept, target.fields()):
        assert bits_equal(x, y)
    for x, y in zip(before, other.fields()):
        assert bits_equal(x, y)
    for s in (a, other, boxed, target):
        s.close()


def test_device_fields_hold_the_state():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    device_to_host = 2
    n = 24
    sim = stepped(n, 2)
    for step in range(2):
        ptrs = sim.deviceFields()
        assert all(ptrs[k] for k in ptrs)
        want = dict(zip(("rho", "u", "v", "w"), sim.fields()))
        want["p"] = sim.pressure()
        for name, expected in want.items():
            got = np.empty_like(expected)
            assert hip.hipMemcpy(got.ctypes.data, ptrs[name], got.nbytes, device_to_host) == 0
            assert bits_equal(got, expected), (step, name)
        sim.advance(DT)
    say("device addresses hold the downloaded state, before and after a step")
    sim.close()
