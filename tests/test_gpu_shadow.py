"""Shadow runs (vgx_shadow_*): the same scene stepped by plain per-operation
kernels in a chosen precision.

The FP32 shadow is vapor's FP32 step by a second route through the same
templates the FP64 shadow instantiates in double; agreeing with the FP32 path
bit for bit (which itself agrees with the reference's fixtures and the oracle)
pins those templates. The FP64 shadow is then checked by what FP64 arithmetic
must change and what it must not: it meets tolerances the FP32 sweeps cannot
reach, its fields stay within FP32 rounding of the FP32 path over short runs,
and the FP32 path's pressure moves toward it as FP64 refinement tightens the
FP32 solve."""
import numpy as np
import pytest

import oracle
from helpers import bits_equal, max_abs
from vapor_amd import mac3d

pytestmark = pytest.mark.gpu


def say(*args):
    print("[gpu shadow]", *args, flush=True)


def scene(n_iters=24, **extra):
    p = mac3d.MacGridParameters3D(mac3d.ReflectionMacCormack)
    p.update(emitter_radius=0.15, emitter2_enabled=1, emitter2_radius=0.1, emitter2_center_x=0.7,
             emitter2_center_y=0.3, emitter2_center_z=0.2, project_iterations=n_iters, emitter_profile=1)
    p.update(**extra)
    return p


@pytest.mark.parametrize("n,algorithm,tolerance", [(20, mac3d.ReflectionMacCormack, False),
                                                   (13, mac3d.SimpleSemiLagrangian, False),
                                                   (16, mac3d.ReflectionMacCormack, True)])
def test_fp32_shadow_is_the_fp32_step(n, algorithm, tolerance):
    say("fp32 shadow", n, algorithm, tolerance)
    p = scene()
    base = mac3d.MacGridParameters3D(algorithm)
    p.update(reflection=base.reflection, maccormack_vel=base.maccormack_vel, maccormack_smoke=base.maccormack_smoke,
             smoke_decay=0.05)
    if tolerance:
        p.update(sor_omega=1.8, solve_mode=1, tolerance=1e-5, max_sweeps=4000, check_interval=8, warm_start=1)
    sim = mac3d.GpuMacSimulation3D(n, 1.0, p)
    sim.addVelocityNoise(0.05, 7)
    shadow = mac3d.ShadowRun3D(sim, 32)
    for step in range(4):
        sim.advance(1 / 60)
        shadow.advance(1 / 60)
        r, q = sim.lastProjection(), shadow.lastProjection()
        assert r.iterations == q.iterations, (step, r.iterations, q.iterations)
    errors = shadow.compare(sim)
    for name, e in errors.items():
        say(name, e)
        assert e["ref_max"] > 0.0, name
    got = sim.fields() + (sim.pressure(),)
    want = shadow.fields()
    for name, a, b in zip(mac3d.ShadowRun3D.FIELDS, got, want):
        if tolerance:  # FP64 residual partials summed in another order
            assert max_abs(a, b) <= 1e-5 * max(1.0, float(np.max(np.abs(b)))), name
        else:
            assert bits_equal(a, b.astype(np.float32)), (name, max_abs(a, b))


def test_fp32_shadow_matches_the_oracle():
    """The shadow's FP32 instantiation against the oracle directly."""
    say("fp32 shadow vs oracle")
    n = 12
    p = scene(16, emitter_profile=0)
    sim = mac3d.GpuMacSimulation3D(n, 1.0, p)
    ref = oracle.OracleSim(n, 1.0, p)
    shadow = mac3d.ShadowRun3D(sim, 32)
    for _ in range(3):
        shadow.advance(1 / 60)
        ref.advance(1 / 60)
    for got, want in zip(shadow.fields()[:4], ref.fields()):
        assert bits_equal(got.astype(np.float32), want), max_abs(got, want)


def test_fp64_shadow_tracks_the_fp32_path():
    say("fp64 shadow")
    n = 24
    p = scene(sor_omega=1.8, solve_mode=1, tolerance=1e-7, max_sweeps=20000, check_interval=16, warm_start=1,
              refinement_steps=5)
    sim = mac3d.GpuMacSimulation3D(n, 1.0, p)
    sim.addVelocityNoise(0.05, 11)
    shadow = mac3d.ShadowRun3D(sim, 64)
    for _ in range(4):
        sim.advance(1 / 60)
        shadow.advance(1 / 60)
        for k in (0, 1):
            q = shadow.lastProjection(k)
            assert q.converged and q.final_residual_l2 <= 1e-7, q.final_residual_l2
            r = sim.lastProjection(k)
            assert r.final_residual_l2 <= 1e-7, r.final_residual_l2
    errors = shadow.compare(sim)
    for name, e in errors.items():
        say(name, e)
        rel = e["max_abs"] / e["ref_max"]
        # FP32 rounding of a few steps, and the two solves stopping at
        # different iterates within the 1e-7 tolerance.
        assert rel < (1e-3 if name == "p" else 1e-4), (name, rel)
    # FP64 arithmetic differs from FP32: the fields are not bit-identical.
    assert errors["u"]["max_abs"] > 0.0


def test_fp64_shadow_reaches_tolerances_fp32_cannot():
    say("fp64 tight tolerance")
    n = 16
    p = scene(sor_omega=1.8, solve_mode=1, tolerance=1e-11, max_sweeps=6000, check_interval=16)
    sim = mac3d.GpuMacSimulation3D(n, 1.0, p)
    shadow = mac3d.ShadowRun3D(sim, 64)
    shadow.advance(1 / 60)
    sim.advance(1 / 60)
    q, r = shadow.lastProjection(), sim.lastProjection()
    say("fp64", q.final_residual_l2, q.iterations, q.divergence_after, "fp32", r.final_residual_l2, r.iterations)
    assert q.converged and q.final_residual_l2 <= 1e-11
    assert not r.converged and r.final_residual_l2 > 1e-9
    assert q.divergence_after < 1e-2 * r.divergence_after


def test_refinement_moves_the_fp32_pressure_toward_fp64():
    say("refinement toward fp64")
    n = 16
    errs = {}
    for tol, refine in ((1e-5, 0), (1e-8, 5)):
        p = scene(sor_omega=1.8, solve_mode=1, tolerance=tol, max_sweeps=20000, check_interval=16,
                  refinement_steps=refine)
        sim = mac3d.GpuMacSimulation3D(n, 1.0, p)
        shadow_params = p.copy()
        shadow_params.update(tolerance=1e-12)
        exact = mac3d.GpuMacSimulation3D(n, 1.0, shadow_params)
        shadow = mac3d.ShadowRun3D(exact, 64)
        sim.advance(1 / 60)
        shadow.advance(1 / 60)
        errs[tol] = shadow.compare(sim)["p"]
        say(tol, errs[tol])
    assert errs[1e-8]["rms"] < 0.1 * errs[1e-5]["rms"]
